"""The Oxygen component of LOBSTER (NutrientsPlanktonDetritus/oxygen.jl) in
the oracle.  Pinned by: the reference's own test properties
(test/test_NutrientsPlanktonDetritus.jl:41-145 with oxygen = Oxygen(): the
O2 tracer is materialised last, a zero state stays zero, nitrogen is
conserved over 100 steps of 1 s); the one-way coupling the component
documents (oxygen.jl:1-12: every other tendency is unchanged, bit for bit);
and a stoichiometric identity that holds for any state and any ratios Rp,
Rn: G_O2 + Rp G_NO3 + (Rp - Rn) G_NH4 = Rp mu0 L_PAR P L_NH4 / 2, the
ammonium-limited share of growth, with L_PAR = PAR / (kPAR + PAR) and
L_NH4 = NH4 / (kNH4 + NH4) formed here from the parameters (a dropped term,
a wrong sign or swapped ratios fail it).  NPZD's method (oxygen.jl:35-46,
respiration = the plankton's inorganic nitrogen waste alone, as Julia
parses it): G_O2 + (Rp - Rn) G_N = Rp muP + (Rp - Rn) (r D - (1 - ax g) muP),
muP formed here from the parameters; a respiration that also held the
detritus waste r D fails it.  The O2 gas exchange on that
tracer: the C flux against the restatement the PISCES coupling uses, zero
at saturation, and the top cell's O2 tendency as the only change, by
-F / dz.  CPU only."""
import json
import os

import numpy as np
import pytest

import oracle as O
from oceanbiome import LOBSTER, NPZD, CarbonateSystem, Oxygen, RectilinearGrid, abi
from oceanbiome.models import TwoBandPhotosyntheticallyActiveRadiation, TwoParticleAndDissolved

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "reference_values.json")))


def _box():
    return RectilinearGrid(size=(1, 1, 1), extent=(1.0, 1.0, 2.0), halo=(1, 1, 1),
                           topology=("Periodic", "Periodic", "Bounded"))


def _lobster(grid, carbonate, oxygen=True, sinking=False):
    det = TwoParticleAndDissolved()
    if not sinking:
        det.small_particle_sinking_speed = 0.0
        det.large_particle_sinking_speed = 0.0
    return LOBSTER(grid, detritus=det, carbonate_system=CarbonateSystem() if carbonate else None,
                   oxygen=Oxygen() if oxygen else None,
                   light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=100.0),
                   temperature=12.0, salinity=35.0)


def _state(rng, carbonate):
    v = [10 * rng.random(), rng.random(), rng.random(), rng.random(), rng.random(), rng.random(), rng.random()]
    if carbonate:
        v += [2000 * rng.random(), 2000 * rng.random()]
    return np.array(v + [300 * rng.random()])


def test_host_description():
    """oxygen.jl:14-19: defaults Rp = 10.75, Rn = 2, tracer O2 after DIC, Alk."""
    g = _box()
    ox = Oxygen()
    assert (ox.respiration_oxygen_nitrogen_ratio, ox.nitrification_oxygen_nitrogen_ratio) == (10.75, 2.0)
    assert _lobster(g, True).required_biogeochemical_tracers()[-3:] == ("DIC", "Alk", "O₂")
    assert _lobster(g, False).required_biogeochemical_tracers()[-2:] == ("DOM", "O₂")
    cfg = _lobster(g, True).config()
    assert cfg.oxygen.enabled == 1 and cfg.oxygen.respiration_oxygen_nitrogen_ratio == 10.75
    assert _lobster(g, True, oxygen=False).config().oxygen.enabled == 0
    assert NPZD(g, oxygen=Oxygen()).required_biogeochemical_tracers() == ("N", "P", "Z", "T", "D", "O₂")


@pytest.mark.parametrize("carbonate", [False, True])
def test_zero_state_stays_zero(carbonate):
    """test_NutrientsPlanktonDetritus.jl:105-117 with oxygen = Oxygen()."""
    g = _box()
    bgc = _lobster(g, carbonate)
    n = len(bgc.required_biogeochemical_tracers())
    out = O.npd_tendencies_oxygen(int(carbonate), bgc.npd_struct(), Oxygen().struct(), np.zeros(n), 100.0)
    assert out.shape == (n,) and np.all(out == 0.0)
    tr = [g.new_field() for _ in range(n)]
    O.grid_run(bgc.config(), tr, 10, 60.0, nthreads=1)
    assert all(np.all(t == 0.0) for t in tr)


@pytest.mark.parametrize("carbonate", [False, True])
def test_one_way_coupling(carbonate):
    """oxygen.jl:9-12: O2 does not affect any other group."""
    rng = np.random.default_rng(7)
    g = _box()
    bgc = _lobster(g, carbonate)
    for _ in range(50):
        v = _state(rng, carbonate)
        with_o2 = O.npd_tendencies_oxygen(int(carbonate), bgc.npd_struct(), Oxygen().struct(), v, 80.0)
        without = O.npd_tendencies(abi.OBME_LOBSTER, int(carbonate), bgc.npd_struct(), v[:-1], 80.0)
        np.testing.assert_array_equal(with_o2[:-1], without)
        # the O2 value itself is not read
        v2 = v.copy()
        v2[-1] = -1e6
        np.testing.assert_array_equal(
            O.npd_tendencies_oxygen(int(carbonate), bgc.npd_struct(), Oxygen().struct(), v2, 80.0), with_o2)


@pytest.mark.parametrize("carbonate", [False, True])
def test_stoichiometric_identity(carbonate):
    rng = np.random.default_rng(11)
    g = _box()
    bgc = _lobster(g, carbonate)
    q = bgc.npd_struct()
    for it in range(200):
        Rp, Rn = 20 * rng.random(), 5 * rng.random()
        ox = abi.Oxygen(enabled=1, respiration_oxygen_nitrogen_ratio=Rp, nitrification_oxygen_nitrogen_ratio=Rn)
        v = _state(rng, carbonate)
        PAR = 200 * rng.random()
        G = O.npd_tendencies_oxygen(int(carbonate), q, ox, v, PAR)
        NO3, NH4, P = v[0], v[1], v[2]
        L_PAR = PAR / (q.light_half_saturation + PAR)          # MondoLightLimitation, LOBSTER's default
        L_NH4 = NH4 / (q.ammonia_half_saturation + NH4)
        expect = Rp * q.phytoplankton_maximum_growth_rate * L_PAR * P * L_NH4 / 2
        lhs = G[-1] + Rp * G[0] + (Rp - Rn) * G[1]
        scale = abs(G[-1]) + Rp * abs(G[0]) + abs(Rp - Rn) * abs(G[1]) + abs(expect)
        assert abs(lhs - expect) <= 1e-13 * scale, (it, lhs, expect)
    # Rp = 0, Rn = 1: the O2 tendency is the NH4 tendency itself
    ox = abi.Oxygen(enabled=1, respiration_oxygen_nitrogen_ratio=0.0, nitrification_oxygen_nitrogen_ratio=1.0)
    v = _state(rng, carbonate)
    G = O.npd_tendencies_oxygen(int(carbonate), q, ox, v, 50.0)
    assert G[-1] == G[1]


@pytest.mark.parametrize("carbonate", [False, True])
def test_box_nitrogen_conservation_with_oxygen(carbonate):
    """test_NutrientsPlanktonDetritus.jl:119-145 with oxygen: one cell, no
    sinking, 100 steps of 1 s from random values; the nitrogen sum holds and
    O2 moves."""
    rng = np.random.default_rng(42)
    g = _box()
    bgc = _lobster(g, carbonate)
    v = _state(rng, carbonate)
    tr = []
    for x in v:
        f = g.new_field()
        g.interior(f)[...] = x
        tr.append(f)
    names = [n.replace("₃", "3").replace("₄", "4") for n in bgc.required_biogeochemical_tracers()]
    Nn = REF["conservation"]["nitrogen_LOBSTER"]
    total = lambda: sum(float(g.interior(tr[names.index(n)])[0, 0, 0]) for n in Nn)  # noqa: E731
    N0 = total()
    O.grid_run(bgc.config(), tr, 100, 1.0, nthreads=1)
    assert total() == pytest.approx(N0, rel=1e-13)
    assert g.interior(tr[-1])[0, 0, 0] != v[-1]


# ---- the O2 gas exchange on LOBSTER's O2 (GasExchange.jl:127-146) ----------

def test_o2_flux_in_c_matches_the_pisces_restatement():
    """orc_o2_flux (the NPD coupling) against oracle.o2_flux (the PISCES
    coupling, pinned in the PISCES tests); zero at saturation; slope k."""
    from oceanbiome import OxygenGasExchangeBoundaryCondition
    g = OxygenGasExchangeBoundaryCondition().struct()
    rng = np.random.default_rng(5)
    for _ in range(100):
        O2, T, S = 400 * rng.random(), 30 * rng.random(), 30 + 8 * rng.random()
        f = O.lib().orc_o2_flux(g, O2, T, S)
        assert f == pytest.approx(O.o2_flux(g, O2, T, S), rel=1e-14, abs=1e-18)
        sat = O.o2_air_concentration(g, T, S)
        assert abs(O.lib().orc_o2_flux(g, sat, T, S)) <= 1e-14 * sat
        k = O.lib().orc_o2_flux(g, O2 + 1.0, T, S) - f
        assert k > 0


def test_o2_flux_is_the_only_difference_at_the_top():
    from oceanbiome import OxygenGasExchangeBoundaryCondition
    grid = RectilinearGrid(size=(3, 2, 6), extent=(3.0, 2.0, 30.0), halo=(1, 1, 1),
                           topology=("Periodic", "Periodic", "Bounded"))
    rng = np.random.default_rng(9)

    def model(o2):
        return LOBSTER(grid, carbonate_system=CarbonateSystem(), oxygen=Oxygen(),
                       gas_exchange=OxygenGasExchangeBoundaryCondition() if o2 else None,
                       light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=50.0),
                       temperature=14.0, salinity=34.0)
    tr = []
    for lo, hi in [(5, 12), (0.1, 1), (0.05, 0.5), (0.05, 0.3), (0.5, 3), (0.5, 3), (0.1, 1),
                   (2000, 2200), (2300, 2400), (150, 300)]:
        f = grid.new_field()
        grid.interior(f)[...] = rng.uniform(lo, hi, (6, 2, 3))
        tr.append(f)
    G0 = [np.zeros_like(a) for a in tr]
    G1 = [np.zeros_like(a) for a in tr]
    O.grid_tendencies(model(False).config(), tr, G0)
    O.grid_tendencies(model(True).config(), tr, G1)
    for n in range(9):
        np.testing.assert_array_equal(G1[n], G0[n])
    d = grid.interior(G1[9]) - grid.interior(G0[9])
    assert np.all(d[:-1] == 0)
    g = OxygenGasExchangeBoundaryCondition().struct()
    top = grid.interior(tr[9])[-1]
    expect = -np.vectorize(lambda c: O.o2_flux(g, c, 14.0, 34.0))(top) / 5.0
    np.testing.assert_allclose(d[-1], expect, rtol=1e-12)


def test_host_gas_exchange_pairs():
    from oceanbiome import CarbonDioxideGasExchangeBoundaryCondition, OxygenGasExchangeBoundaryCondition
    g = _box()
    bgc = LOBSTER(g, carbonate_system=CarbonateSystem(), oxygen=Oxygen(),
                  gas_exchange=(CarbonDioxideGasExchangeBoundaryCondition(), OxygenGasExchangeBoundaryCondition()))
    cfg = bgc.config()
    assert cfg.co2.enabled == 1 and cfg.o2.enabled == 1 and cfg.o2.air_concentration == 9352.7
    with pytest.raises(ValueError):
        LOBSTER(g, gas_exchange=OxygenGasExchangeBoundaryCondition())
    with pytest.raises(ValueError):   # NPZD's T is a tracer: its O2 flux is not built
        NPZD(g, oxygen=Oxygen(), gas_exchange=OxygenGasExchangeBoundaryCondition())
    with pytest.raises(NotImplementedError):
        LOBSTER(g, oxygen=Oxygen(),
                gas_exchange=(OxygenGasExchangeBoundaryCondition(), OxygenGasExchangeBoundaryCondition()))


def test_npzd_oxygen_identity_and_coupling():
    rng = np.random.default_rng(23)
    g = _box()
    bgc = NPZD(g, oxygen=Oxygen())
    q = bgc.npd_struct()
    for it in range(200):
        Rp, Rn = 20 * rng.random(), 5 * rng.random()
        ox = abi.Oxygen(enabled=1, respiration_oxygen_nitrogen_ratio=Rp, nitrification_oxygen_nitrogen_ratio=Rn)
        N, P, Z, T, D = 10 * rng.random(), rng.random(), rng.random(), 5 + 20 * rng.random(), rng.random()
        PAR = 200 * rng.random()
        v = np.array([N, P, Z, T, D, 300 * rng.random()])
        G = O.npd_tendencies_oxygen(0, q, ox, v, PAR, model=abi.OBME_NPZD)
        without = O.npd_tendencies(abi.OBME_NPZD, 0, q, v[:-1], PAR)
        np.testing.assert_array_equal(G[:-1], without)
        L_PAR = PAR / np.sqrt(PAR ** 2 + q.light_half_saturation ** 2)   # AnalyticalLightLimitation, NPZD's default
        L_N = N / (N + q.nitrate_half_saturation)
        Lt = q.temperature_coefficient ** (T / 10) if q.use_temperature_coefficient else 1.0
        muP = q.phytoplankton_maximum_growth_rate * L_PAR * L_N * Lt * P
        uptake = (1 - q.ammonia_fraction_of_exudate * q.phytoplankton_exudation_fraction) * muP
        expect = Rp * muP + (Rp - Rn) * (q.remineralisation_rate * D - uptake)
        lhs = G[-1] + (Rp - Rn) * G[0]
        scale = abs(G[-1]) + abs(Rp - Rn) * (abs(G[0]) + q.remineralisation_rate * D + uptake) + Rp * muP
        assert abs(lhs - expect) <= 1e-13 * scale, (it, lhs, expect)
    tr = [g.new_field() for _ in range(6)]
    O.grid_run(bgc.config(), tr, 10, 60.0, nthreads=1)
    assert all(np.all(t == 0.0) for t in tr)
