"""The ensemble oracle (oracle.batch_run) with an InstantRemineralisation
sediment under every column (Models/Sediments/instant_remineralisation.jl,
src/Sediments), pinned by:
- a one-column ensemble against the 3-D oracle's single column with the same
  NPZD parameters, depth, temperature and sediment (grid_run_sediment, whose
  sediment functions are pinned in test_oracle_sediment.py): the ensemble
  only composes them, so the two agree to rounding;
- the reference's own sediment property (test/test_sediments.jl:37-80):
  water-column nitrogen plus sediment storage conserved to rtol 2e-7 over
  100 steps of 1 s from D = 1, every storage non-zero, here per column of a
  ragged ensemble with per-column sinking speeds;
- the sediment clock across calls (the lazy first tendency only once).
CPU only."""
import numpy as np

import oracle as O
from oceanbiome import NPZD, InstantRemineralisationSediment, RectilinearGrid, abi, workloads
from oceanbiome import TwoBandPhotosyntheticallyActiveRadiation


def _sediment_cfg(bgc, sed):
    cfg = bgc.config(abi.OBME_FP64)
    cfg.sediment = sed.struct(list(bgc.required_biogeochemical_tracers()))
    return cfg


def test_one_column_matches_the_grid_column():
    Nz, dz, T = 40, 5.0, 12.0
    grid = RectilinearGrid(size=(1, 1, Nz), z=(-Nz * dz, 0.0), halo=(3, 3, 3),
                           topology=("Periodic", "Periodic", "Bounded"))
    sed = InstantRemineralisationSediment(grid)
    bgc = NPZD(grid, light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=120.0),
               sediment_model=sed)
    rng = np.random.default_rng(4)
    cols = [rng.uniform(1.0, 10.0, Nz), rng.uniform(0.01, 0.1, Nz), rng.uniform(0.01, 0.1, Nz),
            rng.uniform(0.0, 0.5, Nz)]
    # the grid column: tracers N, P, Z, T, D
    tr = []
    for v in (cols[0], cols[1], cols[2], np.full(Nz, T), cols[3]):
        f = grid.new_field()
        grid.interior(f)[:, 0, 0] = v
        tr.append(f)
    gstate = O.SedimentState(1, 1)
    O.grid_run_sediment(bgc.config(abi.OBME_FP64), tr, gstate, 30, 60.0)
    # the one-column ensemble with the model's parameters
    params = workloads.npzd_column_defaults()[:, None].copy()
    params[19] = 120.0
    U = [c.copy() for c in cols]
    bstate = O.SedimentState(1, 1)
    O.batch_run(_sediment_cfg(bgc, sed), np.array([0, Nz]), list(params), dz, T, U, 30, 60.0, nthreads=1,
                sediment=bstate)
    for n, gi in zip(range(4), (0, 1, 2, 4)):
        np.testing.assert_allclose(U[n], grid.interior(tr[gi])[:, 0, 0], rtol=1e-13, atol=0)
    np.testing.assert_allclose(bstate.storage, gstate.storage, rtol=1e-13, atol=0)
    assert bstate.iteration == gstate.iteration == 90
    assert bstate.last_dt == gstate.last_dt


def _ragged(n=12):
    wl = workloads.npzd_columns(n_columns=n, nz_min=20, nz_max=60, dz=10.0, nsteps=1, seed=11)
    sed = InstantRemineralisationSediment()
    return wl, _sediment_cfg(wl.bgc, sed)


def _column_nitrogen(wl, U, storage):
    off, dz = wl.extra["offsets"], wl.extra["dz"]
    tot = np.array([sum(u[off[c]:off[c + 1]].sum() for u in U) * dz for c in range(len(off) - 1)])
    return tot + storage.ravel()


def test_reference_budget_per_column():
    """test/test_sediments.jl:37-80 in the ensemble: N = P = Z = 0, D = 1,
    100 steps of 1 s, every column conserves its nitrogen with its sediment."""
    wl, cfg = _ragged()
    ncell = len(wl.tracers[0])
    U = [np.zeros(ncell), np.zeros(ncell), np.zeros(ncell), np.ones(ncell)]
    state = O.SedimentState(len(wl.extra["offsets"]) - 1, 1)
    before = _column_nitrogen(wl, U, state.storage)
    O.batch_run(cfg, wl.extra["offsets"], list(wl.extra["params"]), wl.extra["dz"], wl.extra["T"], U, 100, 1.0,
                sediment=state)
    np.testing.assert_allclose(_column_nitrogen(wl, U, state.storage), before, rtol=0.2e-6)
    assert np.all(state.storage > 0)
    assert state.iteration == 300


def test_clock_across_calls_and_no_sediment():
    """Two calls of 5 steps give the run of 10 (the lazy first tendency once);
    without the sediment, the columns lose what sinks out of the bottom."""
    wl, cfg = _ragged(6)
    args = (wl.extra["offsets"], list(wl.extra["params"]), wl.extra["dz"], wl.extra["T"])
    U1 = [a.copy() for a in wl.tracers]
    s1 = O.SedimentState(6, 1)
    O.batch_run(cfg, *args, U1, 10, 60.0, sediment=s1)
    U2 = [a.copy() for a in wl.tracers]
    s2 = O.SedimentState(6, 1)
    O.batch_run(cfg, *args, U2, 5, 60.0, sediment=s2)
    O.batch_run(cfg, *args, U2, 5, 60.0, sediment=s2)
    for a, b in zip(U1, U2):
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(s1.storage, s2.storage, rtol=1e-14)
    U0 = [a.copy() for a in wl.tracers]
    O.batch_run(wl.bgc.config(abi.OBME_FP64), *args, U0, 10, 60.0)
    lost = _column_nitrogen(wl, [a.copy() for a in wl.tracers], np.zeros(6)) - _column_nitrogen(wl, U0, np.zeros(6))
    kept = _column_nitrogen(wl, U1, s1.storage) - _column_nitrogen(wl, U0, np.zeros(6))
    assert np.all(lost > 0)
    np.testing.assert_allclose(kept, lost, rtol=1e-6)
