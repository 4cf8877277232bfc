"""The oracle's SimpleMultiG sediment (Models/Sediments/simple_multi_G.jl,
nitrogen form) pinned by what the mathematics fixes:

- the coupled NH4 + NO3 tendencies sum to the remineralisation
  λs Ns + λf Nf exactly (the 0.8 pₙ′ Cr terms cancel: nitrogen moves between
  the sediment and the water, it is not made or lost);
- the field tendencies sum to the incoming flux minus that remineralisation
  (fs + ff = 1), and decay exponentially without a flux;
- at the state where every logarithm of the three fractions vanishes
  (Cr·day = k·day = NO3 = NH4 = O2 = 1) each fraction reduces to e^A times
  its prefactor, which fixes the couplings and the solid deposition
  fraction 0.223 r^0.336 in closed form;
- no oxygen: no nitrification, no denitrification, all the remineralised
  nitrogen to NH4, no oxygen demand; an empty sediment couples nothing;
- the reference's own integration test (test/test_sediments.jl:36-75):
  LOBSTER + O2 on 3×3×50 cells, 500 m, sPOM = bPOM = 1, NO3 = 10, NH4 = 1,
  O2 = 1000, 100 steps of 1 s: total nitrogen (tracers by volume plus the
  sediment's fields by area) kept to 2e-7 relative, and a sediment that is
  no longer empty.
A mistake planted in a copy of the oracle (a dropped 0.8, the decay of the
wrong field) fails the first two checks."""
import math
import shutil

import numpy as np
import pytest

import mutants
import oracle as O
from oceanbiome import LOBSTER, Oxygen, RectilinearGrid, SimpleMultiGSediment, abi
from oceanbiome.units import day


def _model(shape=(3, 3, 50), extent=(10.0, 10.0, 500.0)):
    grid = RectilinearGrid(size=shape, extent=extent, halo=(3, 3, 3), topology=("Periodic", "Periodic", "Bounded"))
    sed = SimpleMultiGSediment(grid)
    bgc = LOBSTER(grid, oxygen=Oxygen(), sediment_model=sed)
    return grid, sed, bgc


def _struct():
    grid, sed, bgc = _model()
    return sed, sed.struct(list(bgc.required_biogeochemical_tracers()))


def test_struct_resolves_the_tracers():
    sed, s = _struct()
    assert s.model == abi.OBME_SEDIMENT_SIMPLE_MULTI_G
    assert (s.nitrate, s.ammonia, s.oxygen) == (0, 1, 7)
    assert [s.sinking[q] for q in range(s.n_sinking)] == [4, 5]
    # 982 |z|^-1.548 at the bottom cell centre, -495 m
    assert math.isclose(s.sedimentation_rate, 982 * 495.0 ** -1.548, rel_tol=1e-15)
    assert O.sediment_fields(s) == 3


def test_coupled_nitrogen_is_the_remineralisation():
    _, s = _struct()
    rng = np.random.default_rng(3)
    for _ in range(200):
        u = rng.uniform(0.0, 5.0, 3)
        NO3, NH4, O2 = rng.uniform(0.1, 30.0), rng.uniform(0.01, 5.0), rng.uniform(1.0, 400.0)
        dNH4, dNO3, dO2 = O.smg_coupled(s, u, NO3, NH4, O2)
        remin = s.slow_decay_rate * u[0] + s.fast_decay_rate * u[1]
        assert math.isclose(dNH4 + dNO3, remin, rel_tol=1e-12, abs_tol=1e-300)
        assert dO2 <= 0.0


def test_field_tendencies_balance_and_decay():
    _, s = _struct()
    assert math.isclose(s.slow_fraction + s.fast_fraction, 1.0, rel_tol=0, abs_tol=1e-15)
    u = np.array([1.3, 0.7, 2.0])
    f = 0.37
    G = O.sediment_tendencies(s, u, f)
    remin = s.slow_decay_rate * u[0] + s.fast_decay_rate * u[1]
    assert math.isclose(G.sum(), f - remin, rel_tol=1e-14)
    assert math.isclose(G[2], s.refactory_fraction * f, rel_tol=1e-15)
    G0 = O.sediment_tendencies(s, u, 0.0)
    assert np.allclose(G0, [-s.slow_decay_rate * u[0], -s.fast_decay_rate * u[1], 0.0], rtol=1e-15, atol=0)


def test_fractions_where_the_logarithms_vanish():
    _, s = _struct()
    R = s.sinking_redfield
    ls, lf = s.slow_decay_rate, s.fast_decay_rate
    # k = (ls Ns + lf Nf)/(Ns + Nf) = 1/day and Cr = R (ls Ns + lf Nf) = 1/day
    ratio = (1 / day - ls) / (lf - 1 / day)          # Nf / Ns
    Ns = 1 / (day * R * (ls + lf * ratio))
    u = np.array([Ns, ratio * Ns, 0.0])
    Nr = ls * u[0] + lf * u[1]
    Cr = Nr * R
    assert math.isclose(Cr * day, 1.0, rel_tol=1e-14)
    dNH4, dNO3, dO2 = O.smg_coupled(s, u, 1.0, 1.0, 1.0)
    half = 1.0 / (s.anoxia_half_saturation + 1.0)
    pn = math.exp(s.nitrate_oxidation_params[0]) / (Nr * day) * half
    pd = math.exp(s.denitrification_params[0]) / (Cr * day) * half
    pa = math.exp(s.anoxic_params[0]) / (Cr * day)
    ps = 0.223 * s.sedimentation_rate ** 0.336
    # log(k) = log(1/day) multiplies F of the denitrification and E of the
    # anoxic fraction with log(O2) = 0: no contribution
    assert math.isclose(dNH4, (1 - pn) * Nr + 0.8 * pd * Cr, rel_tol=1e-10)
    assert math.isclose(dNO3, pn * Nr - 0.8 * pd * Cr, rel_tol=1e-10)
    assert math.isclose(dO2, -(1 - pa * ps - pd) * half * Cr - 2 * pn * Nr, rel_tol=1e-10)


def test_no_oxygen_and_an_empty_sediment():
    _, s = _struct()
    u = np.array([2.0, 1.0, 0.5])
    remin = s.slow_decay_rate * u[0] + s.fast_decay_rate * u[1]
    dNH4, dNO3, dO2 = O.smg_coupled(s, u, 10.0, 1.0, 0.0)
    assert dNH4 == remin and dNO3 == 0.0 and dO2 == 0.0
    assert O.smg_coupled(s, np.zeros(3), 10.0, 1.0, 300.0) == (0.0, 0.0, 0.0)


def _total_nitrogen(grid, names, tr, fields):
    n_names = ("NO₃", "NH₄", "P", "Z", "sPOM", "bPOM", "DOM")
    dz = np.asarray(grid.dz if hasattr(grid, "dz") else np.diff(grid.zf))
    vol = 0.0
    for n in n_names:
        a = grid.interior(tr[names.index(n)])
        vol += float(np.sum(a * dz[:, None, None]))
    return vol + float(np.sum(fields))


def test_reference_integration_conserves_nitrogen():
    grid, sed, bgc = _model()
    names = list(bgc.required_biogeochemical_tracers())
    tr = []
    init = {"NO₃": 10.0, "NH₄": 1.0, "sPOM": 1.0, "bPOM": 1.0, "O₂": 1000.0}
    for n in names:
        f = grid.new_field()
        grid.interior(f)[...] = init.get(n, 0.0)
        tr.append(f)
    cfg = bgc.config(abi.OBME_FP64)
    state = O.SedimentState(grid.Nx, grid.Ny, fields=3)
    before = _total_nitrogen(grid, names, tr, state.storage)
    O.grid_run_sediment(cfg, tr, state, 100, 1.0)
    after = _total_nitrogen(grid, names, tr, state.storage)
    assert abs(after - before) <= 0.2e-6 * abs(before), (before, after)
    assert np.all(state.storage != 0)


MUTATIONS = {
    "denitrification_nitrate_without_0.8": ("*dNO3 = pn * Nr - 0.8 * pd * Cr;", "*dNO3 = pn * Nr - pd * Cr;"),
    "fast_field_decays_with_the_slow": ("G[1] = (1 - fr) * s->fast_fraction * f - s->fast_decay_rate * u[1];",
                                        "G[1] = (1 - fr) * s->fast_fraction * f - s->fast_decay_rate * u[0];"),
}


@pytest.fixture(scope="module")
def mutant_libs():
    libs, workdir = mutants.build_many(MUTATIONS, "oracle_npd.c")
    yield libs
    shutil.rmtree(workdir, ignore_errors=True)


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistakes_fail_the_balances(name, mutant_libs):
    import ctypes as C
    L = O.load_library(mutant_libs[name])
    _, s = _struct()
    u = np.array([1.3, 0.7, 2.0])
    out = [C.c_double(), C.c_double(), C.c_double()]
    dp = C.POINTER(C.c_double)
    L.orc_smg_coupled(C.byref(s), u.ctypes.data_as(dp), 12.0, 1.5, 250.0, *[C.byref(o) for o in out])
    G = np.zeros(3)
    L.orc_sediment_tendencies(C.byref(s), u.ctypes.data_as(dp), 0.37, G.ctypes.data_as(dp))
    remin = s.slow_decay_rate * u[0] + s.fast_decay_rate * u[1]
    coupled_ok = math.isclose(out[0].value + out[1].value, remin, rel_tol=1e-12)
    field_ok = math.isclose(G.sum(), 0.37 - remin, rel_tol=1e-14)
    assert not (coupled_ok and field_ok), name
