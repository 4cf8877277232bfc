"""PISCES with an InstantRemineralisation sediment in the oracle
(oracle/pisces.py step(..., sediment=...); PISCES.jl:233,344-353 takes a
`sediment`; Models/Sediments/instant_remineralisation.jl, src/Sediments),
pinned by:
- the carbon budget: PISCES's transitions conserve carbon (to 1e-20 at the
  reference's test values, test/test_PISCES.jl:68-93), so with POC, GOC and
  CaCO3 sinking into the sediment and remineralised into DIC, the carbon of
  the water column plus the storage is kept (rtol 2e-7 over 100 steps of
  1 s, the reference's sediment test bar, test/test_sediments.jl:37-80),
  while without the sediment the same run loses carbon through the open
  bottom by far more;
- the receiver: after one short step only its bottom cells differ from the
  run without the sediment, by (1 - e) f Δt / dz;
- the sediment clock across calls (two calls of 5 steps give the run of 10).
CPU only."""
import copy

import numpy as np

import oracle as O
import pisces as OP
from oceanbiome import InstantRemineralisationSediment, workloads

CARBON = ("P", "D", "Z", "M", "DOC", "POC", "GOC", "DIC", "CaCO3")


def _case(Nx=3, Ny=2, Nz=20, seed=6):
    wl = workloads.pisces_3d(Nx, Ny, Nz, nsteps=1, seed=seed, Lz=400.0)
    m = wl.bgc
    m.sediment = InstantRemineralisationSediment(sinking_tracers=("POC", "GOC", "CaCO₃"),
                                                 remineralisation_reciever="DIC")
    g = wl.grid
    q = OP.params_from_vector(m.param_vector())
    light = dict(surface_PAR=m.surface_PAR, kw=m.kw, e=m.e, chi=m.chi, division=m.division)
    F = {n: g.interior(a).copy() for n, a in zip(OP.TRACERS, wl.tracers)}
    cc = m.carbon_chemistry
    solver = cc.solver.struct(cc.initial_pH_guess)
    sed = m.config().sediment
    return g, q, light, F, np.asarray(g.zc), np.asarray(g.zf), np.asarray(g.dz), solver, sed


def _carbon(F, dz, storage):
    return sum(np.einsum("kji,k->ji", F[n], dz) for n in CARBON) + storage


def test_carbon_budget_closes_with_the_sediment():
    g, q, light, F, zc, zf, dz, solver, sed = _case()
    F0 = copy.deepcopy(F)
    state = O.SedimentState(g.Nx, g.Ny)
    before = _carbon(F, dz, state.storage)
    OP.step(q, light, F, zc, zf, dz, 100, 1.0, 0.0, solver, nthreads=2, sediment=(sed, state))
    after = _carbon(F, dz, state.storage)
    np.testing.assert_allclose(after, before, rtol=0.2e-6)
    assert np.all(state.storage > 0)
    assert state.iteration == 300
    # without the sediment the carbon leaves through the open bottom
    OP.step(q, light, F0, zc, zf, dz, 100, 1.0, 0.0, solver, nthreads=2)
    lost = before - _carbon(F0, dz, 0.0)
    assert np.all(lost > 50 * np.abs(after - before)), (lost, after - before)
    np.testing.assert_allclose(lost, state.storage + (_carbon(F, dz, 0.0) - _carbon(F0, dz, 0.0)), rtol=1e-6)


def test_only_the_receivers_bottom_cells_differ_after_one_stage():
    g, q, light, F, zc, zf, dz, solver, sed = _case(seed=8)
    Fa, Fb = copy.deepcopy(F), copy.deepcopy(F)
    state = O.SedimentState(g.Nx, g.Ny)
    # one short step: the coupled tendency, constant to 1e-7 over it, summed
    # by the RK3 weights (γ1 + γ2 + ζ2 + γ3 + ζ3 = 1); the difference of two
    # DIC values near 2200 carries a few ulp of cancellation (~1e-12)
    dt = 1e-2
    OP.step(q, light, Fa, zc, zf, dz, 1, dt, 0.0, solver, nthreads=2, sediment=(sed, state))
    OP.step(q, light, Fb, zc, zf, dz, 1, dt, 0.0, solver, nthreads=2)
    for n in OP.WRITTEN:
        d = Fa[n] - Fb[n]
        if n == "DIC":
            assert np.all(d[0] > 0) and np.all(d[1:] == 0)
        else:
            assert np.all(d == 0), n
    f = OP.sediment_flux(q, zc, F, sed)
    expected = (1 - OP.burial_efficiency(sed, f)) * f / dz[0] * dt
    np.testing.assert_allclose(Fa["DIC"][0] - Fb["DIC"][0], expected, rtol=2e-6)


def test_clock_across_calls():
    g, q, light, F, zc, zf, dz, solver, sed = _case(seed=2, Nz=12)
    Fa, Fb = copy.deepcopy(F), copy.deepcopy(F)
    sa, sb = O.SedimentState(g.Nx, g.Ny), O.SedimentState(g.Nx, g.Ny)
    OP.step(q, light, Fa, zc, zf, dz, 10, 60.0, 0.0, solver, nthreads=2, sediment=(sed, sa))
    OP.step(q, light, Fb, zc, zf, dz, 5, 60.0, 0.0, solver, nthreads=2, sediment=(sed, sb))
    OP.step(q, light, Fb, zc, zf, dz, 5, 60.0, 300.0, solver, nthreads=2, sediment=(sed, sb))
    for n in OP.WRITTEN:
        np.testing.assert_array_equal(Fa[n], Fb[n])
    np.testing.assert_array_equal(sa.storage, sb.storage)
    assert sa.iteration == sb.iteration == 30
