"""Pins of the numpy PISCES oracle (oracle/pisces.py) against the reference's
PISCES and MultiBand tests (element conservation of the transitions at the
reference's initial values, zero state, euphotic depth and mixed-layer mean,
MultiBand closed forms) and against the host package's parameter vector.
CPU only."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
import pisces as OP
from oceanbiome import pisces as HP
from oceanbiome.grid import RectilinearGrid

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "reference_values.json")))


def _box_state(values):
    return {k: np.array([float(values.get(k, 0.0))]) for k in OP.TRACERS}


def _box_aux(q, Omega):
    one = np.array([1.0])
    return dict(z=-5.0 * one, zmxl=-10.0 * one, zeu=-10.0 * one, kappa=1.0 * one, mlPAR=300.0 * one,
                Omega=Omega * one, PAR=300.0 * one, PAR1=100.0 * one, PAR2=100.0 * one, PAR3=100.0 * one,
                wPOCc=0.0 * one, wGOCc=0.0 * one)


def _conservation_params():
    q = OP.default_parameters()
    q["fe_excess"] = 0.0   # SimpleIron(excess_scavenging_enhancement = 0)
    q["fix_max"] = 0.0     # NitrateAmmonia(maximum_fixation_rate = 0)
    return q


def test_host_and_oracle_parameters_agree():
    """The device receives the host package's OBME_PP_* vector; the oracle's
    own defaults must be the same numbers (two independent restatements of
    the reference constructors)."""
    grid = RectilinearGrid(size=(2, 2, 4), extent=(1, 1, 40))
    m = HP.PISCES(grid)
    q_host = OP.params_from_vector(m.param_vector())
    q_orc = OP.default_parameters()
    for cls in ("nano", "diatoms", "micro", "meso"):
        for k, v in q_orc[cls].items():
            if k == "pref":
                assert q_host[cls]["pref"] == v
            else:
                assert q_host[cls][k] == pytest.approx(v, rel=1e-15), (cls, k)
    for k, v in q_orc.items():
        if isinstance(v, (int, float)):
            assert q_host[k] == pytest.approx(v, rel=1e-15), k
    assert q_host["doc_agg"] == pytest.approx(q_orc["doc_agg"], rel=1e-15)
    assert q_host["pom_agg"] == pytest.approx(q_orc["pom_agg"], rel=1e-15)
    assert len(m.param_vector()) == HP.PARAM_COUNT == 159


def test_zero_state_has_zero_transitions():
    """test_PISCES.jl:62-66: a zero state stays zero."""
    q = _conservation_params()
    F = _box_state({})
    G = OP.tendencies(q, F, _box_aux(q, 1.0), 0.0)
    assert set(G) == set(OP.WRITTEN)
    for k, v in G.items():
        assert np.all(v == 0.0), k


def test_element_conservation_at_reference_values():
    """test_PISCES.jl:68-93: carbon, iron, silicon, phosphorus and nitrogen
    sums of the transitions vanish at PISCES_INITIAL_VALUES."""
    c = REF["pisces_conservation"]
    q = _conservation_params()
    F = _box_state(c["initial_values"])
    cc = HP.PISCES(RectilinearGrid(size=(1, 1, 2), extent=(1, 1, 10))).carbon_chemistry
    solver = cc.solver.struct(cc.initial_pH_guess)
    Om = O.calcite_saturation(solver, 2205.0, 2566.0, 10.0, 35.0, P=5 * 9.80665 * 1026 / 100000, silicate=8.5)
    G = OP.tendencies(q, F, _box_aux(q, Om), 0.0)
    g = {k: float(v[0]) for k, v in G.items()}
    carbon = sum(g[k] for k in ("P", "D", "Z", "M", "DOC", "POC", "GOC", "DIC", "CaCO3"))
    iron = (g["PFe"] + g["DFe"] + g["Z"] * q["micro"]["iron_ratio"] + g["M"] * q["meso"]["iron_ratio"] + g["SFe"]
            + g["BFe"] + g["Fe"])
    silicon = g["DSi"] + g["Si"] + g["PSi"]
    thP, thN = q["thetaP"], q["thetaN"]
    phosphate = sum(g[k] * thP for k in ("P", "D", "Z", "M", "DOC", "POC", "GOC")) + g["PO4"]
    nitrogen = g["NH4"] + g["NO3"] + sum(g[k] * thN for k in ("P", "D", "Z", "M", "DOC", "POC", "GOC"))
    tol = c["atol"]
    assert abs(carbon) <= tol["carbon"], carbon
    assert abs(iron) <= tol["iron"], iron
    assert abs(silicon) <= tol["silicon"], silicon
    assert abs(phosphate) <= tol["phosphate"], phosphate
    assert abs(nitrogen) <= tol["nitrogen"], nitrogen
    # and the transitions are not trivially zero
    assert abs(g["P"]) > 1e-9 and abs(g["POC"]) > 1e-9 and abs(g["Fe"]) > 1e-12


def test_conservation_random_states():
    """The same element balances at random states (the invariants pin every
    exchange term between compartments)."""
    rng = np.random.default_rng(9)
    q = _conservation_params()
    base = REF["pisces_conservation"]["initial_values"]
    n = 64
    F = {k: float(base[k]) * np.exp(0.5 * rng.standard_normal(n)) for k in OP.TRACERS}
    F["T"] = rng.uniform(2, 28, n)
    F["S"] = np.full(n, 35.0)
    one = np.ones(n)
    A = dict(z=rng.uniform(-200, -1, n), zmxl=-30.0 * one, zeu=-40.0 * one, kappa=one, mlPAR=rng.uniform(1, 300, n),
             Omega=rng.uniform(0.5, 4.0, n), PAR=rng.uniform(0, 300, n), PAR1=rng.uniform(0, 100, n),
             PAR2=rng.uniform(0, 100, n), PAR3=rng.uniform(0, 100, n), wPOCc=-2 / OP.DAY * one,
             wGOCc=-rng.uniform(30, 60, n) / OP.DAY)
    G = OP.tendencies(q, F, A, 3.0e6)
    carbon = sum(G[k] for k in ("P", "D", "Z", "M", "DOC", "POC", "GOC", "DIC", "CaCO3"))
    scale = sum(abs(G[k]) for k in ("P", "D", "Z", "M", "DOC", "POC", "GOC", "DIC", "CaCO3"))
    assert np.all(np.abs(carbon) <= 1e-13 * scale)
    silicon = G["DSi"] + G["Si"] + G["PSi"]
    assert np.all(np.abs(silicon) <= 1e-13 * (abs(G["DSi"]) + abs(G["Si"]) + abs(G["PSi"])))


def test_euphotic_depth_and_mixed_layer_mean():
    """test_PISCES.jl:97-128: PAR = 3 exp(z/10) on 10 cells over 100 m gives
    zeu = -10 log(1000) and a mixed-layer (25 m) mean near 3 (10/25)(1 - e^-2.5)."""
    grid = RectilinearGrid(size=(1, 1, 10), extent=(1, 1, 100))
    zc, zf = np.asarray(grid.zc), np.asarray(grid.zf)
    PAR = 3 * np.exp(zc / 10)[:, None]
    zeu = OP.euphotic_depth(PAR, zc, 3.0)
    assert zeu[0] == pytest.approx(-10 * math.log(1000), rel=1e-12)
    ml = OP.mixed_layer_mean(PAR, zf, -25.0)
    assert ml[0] == pytest.approx(3 * 10 / 25 * (1 - math.exp(-25 / 10)), rel=0.1)
    # no cell below the cutoff: the depth below the bottom cell
    zeu2 = OP.euphotic_depth(np.full((10, 1), 2.0), zc, 3.0)
    assert zeu2[0] == zc[0] - (zc[1] - zc[0])


def test_multi_band_closed_forms():
    """test_light.jl:52-106: Chl = 2 with one band and with two bands."""
    grid = RectilinearGrid(size=(1, 1, 20), extent=(1, 1, 40))
    zc = np.asarray(grid.zc)
    Chl = np.full((20, 1), 2.0)
    kw, e, chi = HP.morel_band_coefficients(((1, 2),), lam=np.array([1.0, 2.0]), kw=np.array([0.01, 0.01]),
                                            e=np.array([2.0, 2.0]), chi=np.array([0.1, 0.1]))
    out = OP.multi_band_PAR(100.0, zc, Chl, kw, e, chi, [1.0])
    assert out[0, :, 0] == pytest.approx(100 * np.exp(zc * (0.01 + 0.1 * 2 ** 2)), rel=1e-12)
    kw, e, chi = HP.morel_band_coefficients(((1, 2), (8, 9)), lam=np.array([1.0, 2.0, 8.0, 9.0]),
                                            kw=np.array([0.01, 0.01, 0.02, 0.02]), e=np.array([2.0, 2.0, 1.5, 1.5]),
                                            chi=np.array([0.1, 0.1, 0.2, 0.2]))
    out = OP.multi_band_PAR(100.0, zc, Chl, kw, e, chi, [0.5, 0.5])
    assert out[0, :, 0] == pytest.approx(100 * np.exp(zc * (0.01 + 0.1 * 2 ** 2)) / 2, rel=1e-12)
    assert out[1, :, 0] == pytest.approx(100 * np.exp(zc * (0.02 + 0.2 * 2 ** 1.5)) / 2, rel=1e-12)


def test_morel_default_bands():
    """Band means of the Morel table lie within the table's range over each band."""
    kw, e, chi = HP.morel_band_coefficients()
    lam = HP.MOREL_LAMBDA
    for n, (b0, b1) in enumerate(((400, 500), (500, 600), (600, 700))):
        sel = (lam >= b0) & (lam <= b1)
        assert HP.MOREL_KW[sel].min() <= kw[n] <= HP.MOREL_KW[sel].max()
        assert HP.MOREL_E[sel].min() <= e[n] <= HP.MOREL_E[sel].max()
        assert HP.MOREL_CHI[sel].min() <= chi[n] <= HP.MOREL_CHI[sel].max()


def test_day_length():
    """CBMDayLength (Utils.jl:17-33): about 12 h at the equinoxes, long
    summer days in the north at 45 N, and the argument order matters."""
    d = OP.DAY
    assert 14.5 * 3600 < OP.cbm_day_length(172 * d, 45.0) < 16.5 * 3600
    assert 7.5 * 3600 < OP.cbm_day_length(355 * d, 45.0) < 9.5 * 3600
    assert abs(OP.cbm_day_length(80 * d, 0.0) - 12.1 * 3600) < 0.3 * 3600


def test_pisces_conserved_groups_match_host():
    """The oracle's conserved groups (PISCES/coupling_utils.jl:11-35) and the
    host description's (two restatements) agree in order and factors."""
    grid = RectilinearGrid(size=(2, 2, 4), extent=(1, 1, 40))
    m = HP.PISCES(grid, scale_negatives=True)
    q = OP.params_from_vector(m.param_vector())
    ascii_ = {"CaCO₃": "CaCO3", "NO₃": "NO3", "NH₄": "NH4", "PO₄": "PO4", "O₂": "O2"}
    host = [([ascii_.get(n, n) for n in names], list(f)) for names, f in m.conserved_tracers()]
    orc = [(list(names), list(f)) for names, f, _ in OP.conserved_groups(q)]
    assert [h[0] for h in host] == [o[0] for o in orc]
    for (hn, hf), (on, of) in zip(host, orc):
        assert hf == pytest.approx(of, rel=1e-15)
    s = m.config().scaling
    assert s.n_groups == 5 and [s.n_members[g] for g in range(5)] == [9, 7, 8, 3, 9]


def test_pisces_scaling_zeroes_negatives_and_keeps_the_last_group_total():
    """Groups share members (P is in the carbon, phosphate and nitrogen
    groups), so each later group rescales members of the earlier ones: a
    group keeps the total it sees when it runs (checked for the last,
    nitrogen), and no member is left negative."""
    rng = np.random.default_rng(2)
    q = OP.default_parameters()
    base = REF["pisces_conservation"]["initial_values"]
    n = 256
    F = {k: float(base.get(k, 1.0)) * np.exp(0.3 * rng.standard_normal(n)) for k in OP.TRACERS}
    for k in ("P", "Z", "POC", "NH4", "Fe", "DSi"):
        flip = rng.random(n) < 0.2
        F[k][flip] = -0.1 * F[k][flip]
    groups = OP.conserved_groups(q)
    names, factors, _ = groups[-1]
    OP.scale_negatives(F, groups[:-1])
    before = sum(F[k] * s for k, s in zip(names, factors))   # the total the last group sees
    OP.scale_negatives(F, groups[-1:])
    after = sum(F[k] * s for k, s in zip(names, factors))
    ok = before > 0
    assert np.allclose(after[ok], before[ok], rtol=1e-13)
    for g_names, _, _ in groups:
        for k in g_names:
            assert np.all((F[k] >= 0) | np.isnan(F[k])), k


def test_oxygen_gas_exchange_known_answers():
    """test_gasexchange_carbon_chem.jl:157-185: the O2 Schmidt number at
    20 °C is 568 ± 1 and the air-side O2 at 25 °C, S = 35 is 200 ± 50; the
    flux is k (O2 - C_air α) with k from the CO2 path's Ho06 k660."""
    from oceanbiome import OxygenGasExchangeBoundaryCondition, CarbonDioxideGasExchangeBoundaryCondition
    o2 = OxygenGasExchangeBoundaryCondition().struct()
    assert abs(O.o2_schmidt(o2, 20.0) - 568) <= 1
    air = O.o2_air_concentration(o2, 25.0, 35.0)
    assert abs(air - 200) <= 50
    # zero flux at saturation, linear in the water concentration
    assert O.o2_flux(o2, air, 25.0, 35.0) == pytest.approx(0.0, abs=1e-18)
    f1, f2 = O.o2_flux(o2, air + 10, 25.0, 35.0), O.o2_flux(o2, air + 20, 25.0, 35.0)
    assert f2 == pytest.approx(2 * f1, rel=1e-12)
    # k equals the CO2 path's k660/√(Sc/660) with the O2 Schmidt number
    co2 = CarbonDioxideGasExchangeBoundaryCondition().struct()
    k660 = 0.266 / 3600 / 100 * 2.0 ** 2
    assert f1 / 10 == pytest.approx(k660 / math.sqrt(O.o2_schmidt(o2, 25.0) / 660), rel=1e-12)
    assert co2.transfer_velocity[2] == pytest.approx(0.266 / 3600 / 100, rel=1e-15)


def _surface_case(surface, t, Nx=5, Ny=3, Nz=12, seed=4):
    from oceanbiome import workloads
    wl = workloads.pisces_3d(Nx, Ny, Nz, nsteps=1, seed=seed, surface_PAR=surface)
    g = wl.grid
    m = wl.bgc
    q = OP.params_from_vector(m.param_vector())
    light = dict(surface_PAR=m.surface_PAR, kw=m.kw, e=m.e, chi=m.chi, division=m.division)
    F = {n: g.interior(a).copy() for n, a in zip(OP.TRACERS, wl.tracers)}
    cc = m.carbon_chemistry
    return OP.update_state(q, light, F, np.asarray(g.zc), np.asarray(g.zf), 0.0 if t is None else t,
                           cc.solver.struct(cc.initial_pH_guess), nthreads=2), g


def test_surface_PAR_modes_reduce_to_the_constant():
    """default_surface_PAR (Light.jl:47) at t = 0 is the constant 100 and at
    t = 3 h the constant 100 cos(π/4); a uniform 2-D PAR⁰ is the constant;
    a 2-D PAR⁰ gives each column the light of its own constant; at night
    (t = 12 h, PAR⁰ = 0) the light is zero, the euphotic depth falls back to
    the level below the bottom (compute_euphotic_depth.jl:33: log 0 - log 0
    is not finite) and the mixed-layer light is zero."""
    keys = ("PAR", "PAR1", "PAR2", "PAR3", "zeu", "mlPAR")
    A0, _ = _surface_case(100.0, 0.0)
    Ad, _ = _surface_case("default", 0.0)
    for k in keys:
        assert np.array_equal(A0[k], Ad[k]), k
    A3, _ = _surface_case(100.0 * O.default_surface_PAR(3 * 3600.0) / 100.0, 3 * 3600.0)
    Ad3, _ = _surface_case("default", 3 * 3600.0)
    for k in keys:
        assert np.array_equal(A3[k], Ad3[k]), k
    Au, _ = _surface_case(np.full((3, 5), 70.0), 0.0)
    Ac, _ = _surface_case(70.0, 0.0)
    for k in keys:
        assert np.array_equal(Au[k], Ac[k]), k
    rng = np.random.default_rng(1)
    arr = rng.uniform(20.0, 200.0, (3, 5))
    Aa, g = _surface_case(arr, 0.0)
    for j in range(3):
        for i in range(5):
            Ai, _ = _surface_case(float(arr[j, i]), 0.0)
            for k in keys:
                assert np.array_equal(Aa[k][..., j, i], Ai[k][..., j, i]), (k, j, i)
    An, g = _surface_case("default", 12 * 3600.0)
    assert not np.any(An["PAR"]) and not np.any(An["mlPAR"])
    z0 = g.zc[0] - (g.zc[1] - g.zc[0])
    assert np.all(An["zeu"] == z0)
