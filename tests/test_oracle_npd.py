"""Pins of the FP64 oracle's NPZD / LOBSTER step: tendencies, TwoBand PAR,
sinking and RK3, against the reference's own tests (conservation, the
two-cell light closed form, the sinking profile) and against closed forms
(the RK3 amplification polynomial of a linear column).  CPU only."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from oceanbiome import abi
from oceanbiome.grid import RectilinearGrid
from oceanbiome.models import (LOBSTER, NPZD, CarbonateSystem, Detritus, PhytoZoo, TwoParticleAndDissolved,
                               TwoBandPhotosyntheticallyActiveRadiation)
from oceanbiome.units import day

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "reference_values.json")))


def _box(Nz=1, extent_z=2.0):
    return RectilinearGrid(size=(1, 1, Nz), extent=(1.0, 1.0, extent_z), halo=(1, 1, 1),
                           topology=("Periodic", "Periodic", "Bounded"))


def _lobster(grid, carbonate=True, sinking=False, open_bottom=True, PAR0=100.0):
    det = TwoParticleAndDissolved(open_bottom=open_bottom)
    if not sinking:
        det.small_particle_sinking_speed = 0.0
        det.large_particle_sinking_speed = 0.0
    return LOBSTER(grid, detritus=det, carbonate_system=CarbonateSystem() if carbonate else None,
                   light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=PAR0),
                   temperature=12.0, salinity=35.0)


def _npzd(grid, sinking=True, open_bottom=True, PAR0=100.0):
    b = NPZD(grid, detritus=Detritus(open_bottom=open_bottom),
             light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=PAR0))
    if not sinking:
        b.plankton.phytoplankton_sinking_speed = 0.0
        b.detritus.sinking_speed = 0.0
    return b


def _ascii(name):
    return name.replace("\u2083", "3").replace("\u2084", "4")


def _names(bgc):
    return [_ascii(n) for n in bgc.required_biogeochemical_tracers()]


def _set(grid, bgc, values):
    """Parent arrays (tracer order) holding `values[name]` (scalar or column)."""
    out = []
    for name in _names(bgc):
        f = grid.new_field()
        grid.interior(f)[...] = np.asarray(values.get(name, 0.0), dtype=np.float64).reshape(-1, 1, 1)
        out.append(f)
    return out


def _column_sum(grid, bgc, tr, names, scale=None):
    idx = {n: q for q, n in enumerate(_names(bgc))}
    dz = np.asarray(grid.dz)
    s = 0.0
    for q, n in enumerate(names):
        col = grid.interior(tr[idx[n]])[:, 0, 0]
        s += (1.0 if scale is None else scale[q]) * float(np.sum(col * dz))
    return s


# ---- tendencies -------------------------------------------------------------

@pytest.mark.parametrize("kind", ["npzd", "lobster", "lobster_carb"])
def test_zero_state_has_zero_tendencies(kind):
    """test_NutrientsPlanktonDetritus.jl:105-117: a zero state stays zero."""
    g = _box()
    bgc = _npzd(g) if kind == "npzd" else _lobster(g, carbonate=(kind == "lobster_carb"))
    n = len(bgc.required_biogeochemical_tracers())
    out = O.npd_tendencies(bgc.model, 1 if bgc.carbonate_system else 0, bgc.npd_struct(), np.zeros(n), 100.0)
    assert np.all(out == 0.0)
    tr = _set(g, bgc, {})
    O.grid_run(bgc.config(), tr, 10, 60.0, nthreads=1)
    assert all(np.all(t == 0.0) for t in tr)


def _random_state(rng, bgc):
    names = _names(bgc)
    v = {"N": 10 * rng.random(), "NO3": 10 * rng.random(), "NH4": rng.random(), "P": rng.random(),
         "Z": rng.random(), "D": rng.random(), "sPOM": rng.random(), "bPOM": rng.random(), "DOM": rng.random(),
         "DIC": 2000 * rng.random(), "Alk": 2000 * rng.random(), "T": 5 + 20 * rng.random()}
    return np.array([v[n] for n in names]), names


@pytest.mark.parametrize("kind", ["npzd", "lobster", "lobster_carb"])
def test_tendencies_conserve_nitrogen_and_carbon(kind):
    """Per evaluation, the nitrogen tendencies sum to zero, and with the
    CarbonateSystem the Redfield-weighted carbon tendencies too
    (coupling_utils.jl:1-50 groups; the reference checks the same sums after
    100 steps)."""
    rng = np.random.default_rng(42)
    g = _box()
    bgc = _npzd(g) if kind == "npzd" else _lobster(g, carbonate=(kind == "lobster_carb"))
    p = bgc.npd_struct()
    R, rho = bgc.plankton.redfield_ratio, bgc.plankton.carbon_calcite_ratio
    for _ in range(50):
        vals, names = _random_state(rng, bgc)
        PAR = 200 * rng.random()
        G = dict(zip(names, O.npd_tendencies(bgc.model, 1 if bgc.carbonate_system else 0, p, vals, PAR)))
        Nn = REF["conservation"]["nitrogen_NPZD" if kind == "npzd" else "nitrogen_LOBSTER"]
        terms = [G[n] for n in Nn]
        assert abs(sum(terms)) <= 1e-15 * sum(abs(t) for t in terms) * len(terms) + 1e-30
        if kind == "lobster_carb":
            sc = {"P": (1 + rho) * R, "Z": R, "DIC": 1.0, "sPOM": R, "bPOM": R, "DOM": R}
            terms = [sc[n] * G[n] for n in sc]
            assert abs(sum(terms)) <= 1e-14 * sum(abs(t) for t in terms) + 1e-30
        if "T" in G:
            assert G["T"] == 0.0


@pytest.mark.parametrize("kind", ["npzd", "lobster", "lobster_carb"])
def test_box_conservation_100_steps(kind):
    """test_NutrientsPlanktonDetritus.jl:119-145: one cell, no sinking, 100
    steps of 1 s from random values; N (and C) sums unchanged."""
    rng = np.random.default_rng(42)
    g = _box()
    bgc = _npzd(g, sinking=False) if kind == "npzd" else _lobster(g, carbonate=(kind == "lobster_carb"))
    vals, names = _random_state(rng, bgc)
    tr = _set(g, bgc, dict(zip(names, vals)))
    Nn = REF["conservation"]["nitrogen_NPZD" if kind == "npzd" else "nitrogen_LOBSTER"]
    R, rho = bgc.plankton.redfield_ratio, bgc.plankton.carbon_calcite_ratio
    cn = ["P", "Z", "DIC", "sPOM", "bPOM", "DOM"]
    cs = [(1 + rho) * R, R, 1.0, R, R, R]
    N0 = _column_sum(g, bgc, tr, Nn)
    C0 = _column_sum(g, bgc, tr, cn, cs) if kind == "lobster_carb" else None
    O.grid_run(bgc.config(), tr, 100, 1.0, nthreads=1)
    assert _column_sum(g, bgc, tr, Nn) == pytest.approx(N0, rel=1e-13)
    if C0 is not None:
        assert _column_sum(g, bgc, tr, cn, cs) == pytest.approx(C0, rel=1e-13)
    # and the state did change
    assert not np.allclose(g.interior(tr[names.index("P")]), vals[names.index("P")], rtol=1e-12, atol=0)


def test_column_with_closed_bottom_conserves_nitrogen():
    """Sinking in flux form with a closed bottom moves mass but loses none:
    sum(c dz) of the nitrogen group is invariant (flux-form divergence, zero
    flux through the top face and, closed, through the bottom face)."""
    rng = np.random.default_rng(1)
    g = _box(Nz=24, extent_z=48.0)
    bgc = _npzd(g, sinking=True, open_bottom=False)
    vals = {"N": rng.uniform(1, 10, 24), "P": rng.uniform(0.01, 0.1, 24), "Z": rng.uniform(0.01, 0.1, 24),
            "D": rng.uniform(0.0, 0.5, 24), "T": np.full(24, 12.0)}
    tr = _set(g, bgc, vals)
    Nn = REF["conservation"]["nitrogen_NPZD"]
    N0 = _column_sum(g, bgc, tr, Nn)
    O.grid_run(bgc.config(), tr, 50, 600.0, nthreads=1)
    assert _column_sum(g, bgc, tr, Nn) == pytest.approx(N0, rel=1e-13)


# ---- light ----------------------------------------------------------------

def test_two_band_two_cell_closed_form():
    """test_light.jl:10-50: two cells of 1 m with P = 2.5 + z."""
    c = REF["two_band"]
    par = TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=c["PAR0"])
    s = par.struct()
    kr, kb, xr, xb = par.water_red_attenuation, par.water_blue_attenuation, \
        par.chlorophyll_red_attenuation, par.chlorophyll_blue_attenuation
    er, eb, r, R = par.chlorophyll_red_exponent, par.chlorophyll_blue_exponent, par.pigment_ratio, \
        par.phytoplankton_chlorophyll_ratio
    Ir = [(2.0 * R / r) ** er * 0.5]
    Ib = [(2.0 * R / r) ** eb * 0.5]
    Ir.append(Ir[0] + (2.0 * R / r) ** er * 0.5 + (1.0 * R / r) ** er * 0.5)
    Ib.append(Ib[0] + (2.0 * R / r) ** eb * 0.5 + (1.0 * R / r) ** eb * 0.5)
    expected = 100.0 * np.array([math.exp(-0.5 * kr - Ir[0] * xr) + math.exp(-0.5 * kb - Ib[0] * xb),
                                 math.exp(-1.5 * kr - Ir[1] * xr) + math.exp(-1.5 * kb - Ib[1] * xb)]) / 2
    got = O.two_band_column(s, np.array(c["zc"]), np.array(c["zf"]), np.array(c["P"]), c["PAR0"])
    # the oracle returns bottom to top; the test's expectation is top first
    assert got[::-1] == pytest.approx(expected, rel=1e-14)


def test_two_band_limits():
    """No chlorophyll: PAR = PAR0 (e^{kr z} + e^{kb z}) / 2 (pure water); PAR
    is monotone in depth with chlorophyll."""
    par = TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=80.0)
    s = par.struct()
    g = _box(Nz=16, extent_z=64.0)
    zc, zf = np.asarray(g.zc), np.asarray(g.zf)
    got = O.two_band_column(s, zc, zf, np.zeros(16), 80.0)
    want = 80.0 * (np.exp(par.water_red_attenuation * zc) + np.exp(par.water_blue_attenuation * zc)) / 2
    assert got == pytest.approx(want, rel=1e-14)
    got = O.two_band_column(s, zc, zf, np.full(16, 0.3), 80.0)
    assert np.all(np.diff(got) > 0)


def test_default_surface_PAR():
    """default_surface_PAR (Light/Light.jl:217-219): 100 max(0, cos(pi t / 12h))."""
    assert O.default_surface_PAR(0.0) == 100.0
    assert O.default_surface_PAR(6 * 3600.0) == pytest.approx(0.0, abs=1e-12)
    assert O.default_surface_PAR(12 * 3600.0) == 0.0
    assert O.default_surface_PAR(3 * 3600.0) == pytest.approx(100 * math.cos(math.pi / 4), rel=1e-15)


# ---- sinking --------------------------------------------------------------

def test_sinking_profile():
    """sinking_velocity_fields.jl:16: w = -speed on faces 1..Nz when the
    bottom is open; closed, w = -speed (1 - exp((1-k)/2)), so the bottom
    face has w = 0 (test_utils.jl:50-70); the top face stays 0."""
    c = REF["sinking_profile"]
    Nz, sp = c["Nz"], c["speed"]
    w = O.sinking_profile(sp, Nz, open_bottom=True)
    assert np.all(w[:Nz] == -sp) and w[Nz] == 0.0
    w = O.sinking_profile(sp, Nz, open_bottom=False)
    k = np.arange(1, Nz + 1)
    assert w[0] == 0.0 and w[Nz] == 0.0
    assert w[:Nz] == pytest.approx(-sp * (1 - np.exp((1 - k) / 2.0)), rel=1e-15)


def test_sinking_divergence_budget():
    """Upwind flux form: sum(div dz) = F_top - F_bottom = -w_bottom c_bottom."""
    rng = np.random.default_rng(2)
    Nz = 12
    dz = rng.uniform(0.5, 3.0, Nz)
    c = rng.uniform(0.0, 1.0, Nz)
    for ob in (True, False):
        w = O.sinking_profile(1e-4, Nz, open_bottom=ob)
        div = O.sinking_divergence(w, c, dz)
        assert float(np.sum(div * dz)) == pytest.approx(-w[0] * c[0], abs=1e-18)
        # the upwind cell of each face is the one above it (w < 0)
        want = np.array([((w[k + 1] * (c[k + 1] if k + 1 < Nz else c[k])) - w[k] * c[k]) / dz[k]
                         for k in range(Nz)])
        assert div == pytest.approx(want, rel=1e-15, abs=1e-30)


# ---- RK3 --------------------------------------------------------------------

def test_rk3_amplification_polynomial_on_a_linear_column():
    """With P = Z = 0 the NPZD detritus obeys the linear ODE dD/dt = A D,
    A = -r I - M (remineralisation and the upwind sinking operator M), and N
    receives r D.  Any 3-stage, third-order Runge-Kutta step advances a linear
    autonomous system by I + hA + (hA)^2/2 + (hA)^3/6: a wrong gamma or zeta
    of the low-storage scheme (timesteppers.jl:79-92) changes that
    polynomial."""
    Nz = 10
    g = _box(Nz=Nz, extent_z=20.0)
    bgc = _npzd(g, sinking=True, open_bottom=True)
    rng = np.random.default_rng(5)
    D0 = rng.uniform(0.1, 1.0, Nz)
    N0 = rng.uniform(1.0, 2.0, Nz)
    tr = _set(g, bgc, {"N": N0, "D": D0, "T": np.full(Nz, 10.0)})
    h = 20000.0
    O.grid_run(bgc.config(), tr, 1, h, nthreads=1)
    r = bgc.detritus.remineralisation_rate
    w = bgc.detritus.sinking_speed
    dz = np.asarray(g.dz)
    M = np.zeros((Nz, Nz))   # div = M c, upwind with w = -speed on faces 0..Nz-1, 0 at the top
    for k in range(Nz):
        if k + 1 < Nz:
            M[k, k + 1] += -w / dz[k]
        M[k, k] -= -w / dz[k]
    A = -r * np.eye(Nz) - M
    hA = h * A
    P3 = np.eye(Nz) + hA + hA @ hA / 2 + hA @ hA @ hA / 6
    names = _names(bgc)
    D1 = g.interior(tr[names.index("D")])[:, 0, 0]
    assert D1 == pytest.approx(P3 @ D0, rel=1e-13)
    # the P, Z columns stay zero, and N gains exactly what remineralisation gave
    assert np.all(g.interior(tr[names.index("P")]) == 0.0)
    # N1 - N0 = r h (D0 + hA D0/2 + (hA)^2 D0/6)
    dN = r * h * (D0 + hA @ D0 / 2 + hA @ hA @ D0 / 6)
    N1 = g.interior(tr[names.index("N")])[:, 0, 0]
    assert N1 - N0 == pytest.approx(dN, rel=1e-11)


def test_rk3_stage_clocks_with_diurnal_PAR():
    """The surface PAR of each stage is evaluated at the stage clock
    t, t + 8/15 h, t + 2/3 h: at t0 = 6h - 8/15 h the second stage sees
    cos(pi/2) = 0 light; a run from a time when all three stage clocks are in
    the dark equals a run with PAR0 = 0."""
    g = _box(Nz=4, extent_z=8.0)
    vals = {"N": np.full(4, 5.0), "P": np.full(4, 0.1), "Z": np.full(4, 0.05), "D": np.full(4, 0.1),
            "T": np.full(4, 10.0)}
    b_dark = _npzd(g, PAR0=0.0)
    b_diur = NPZD(g, light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR="default"))
    t_dark = 13 * 3600.0
    a = _set(g, b_dark, vals)
    b = _set(g, b_diur, vals)
    O.grid_run(b_dark.config(), a, 1, 600.0, t0=t_dark, nthreads=1)
    O.grid_run(b_diur.config(), b, 1, 600.0, t0=t_dark, nthreads=1)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
