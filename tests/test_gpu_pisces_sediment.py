"""GPU parity of PISCES with an InstantRemineralisation sediment
(obme_pisces_step_with_sediment; PISCES.jl:233 takes a `sediment`,
Models/Sediments/instant_remineralisation.jl, src/Sediments) against the
numpy oracle (oracle/pisces.py step(..., sediment=...)): POC, GOC and CaCO3
sinking into the sediment, remineralised into DIC.  FP64: relative L2 <=
1e-12 per tracer and for the storage; FP32-mixed (the per-stage kernels):
<= 1e-5.  Also: two calls in a row (the sediment clock), obme_pisces_step
on such a handle (the coupling alone, storage untouched), the tendency
hook's coupled term, the update-state hook's sediment steps, and the
configurations PISCES rejects."""
import copy

import numpy as np
import pytest

import oracle as O
import pisces as OP
from helpers import rel_l2
from oceanbiome import InstantRemineralisationSediment, RectilinearGrid, SimpleMultiGSediment, workloads
from oceanbiome.runtime import HIPPISCES, OBMEError

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 1e-5}


def _case(Nx=6, Ny=4, Nz=16, seed=5, sinking=("POC", "GOC", "CaCO₃"), receiver="DIC"):
    wl = workloads.pisces_3d(Nx, Ny, Nz, nsteps=1, seed=seed, medians="test", Lz=400.0)
    wl.bgc.sediment = InstantRemineralisationSediment(sinking_tracers=sinking, remineralisation_reciever=receiver)
    return wl


def _oracle(wl, calls, dt=300.0):
    g = wl.grid
    m = wl.bgc
    q = OP.params_from_vector(m.param_vector())
    light = dict(surface_PAR=m.surface_PAR, kw=m.kw, e=m.e, chi=m.chi, division=m.division)
    F = {n: g.interior(a).copy() for n, a in zip(OP.TRACERS, wl.tracers)}
    cc = m.carbon_chemistry
    solver = cc.solver.struct(cc.initial_pH_guess)
    state = O.SedimentState(g.Nx, g.Ny)
    t = 0.0
    for n in calls:
        OP.step(q, light, F, np.asarray(g.zc), np.asarray(g.zf), np.asarray(g.dz), n, dt, t, solver, nthreads=4,
                sediment=(m.config().sediment, state))
        t += n * dt
    return F, state.storage.copy()


def _device(wl, precision):
    import torch
    dev = HIPPISCES(wl.bgc, precision)
    tr = [torch.from_numpy(a).to(device="cuda", dtype=dev.dtype) for a in wl.tracers]
    return dev, tr


def _check(wl, precision, tr, storage, F, ref_storage, what):
    g = wl.grid
    errs = {n: rel_l2(g.interior(tr[i].double().cpu().numpy()), F[n]) for i, n in enumerate(OP.WRITTEN)}
    es = rel_l2(storage.cpu().numpy(), ref_storage)
    print(what, precision, "worst tracer", max(errs.values()), "storage", es, flush=True)
    assert max(errs.values()) <= TOL[precision] and es <= TOL[precision], (errs, es)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_step_with_sediment_matches_oracle(precision):
    import torch
    print("case pisces sediment", precision, flush=True)
    wl = _case()
    F, ref_s = _oracle(wl, [4])
    dev, tr = _device(wl, precision)
    dev.step(tr, 4, 300.0)
    torch.cuda.synchronize()
    assert float(dev.sediment_storage.min()) > 0
    _check(wl, precision, tr, dev.sediment_storage, F, ref_s, "one call")


def test_two_calls_fp64():
    import torch
    print("case pisces sediment two calls", flush=True)
    wl = _case(seed=7, sinking=("POC", "SFe"), receiver="Fe")
    F, ref_s = _oracle(wl, [2, 3])
    dev, tr = _device(wl, "fp64")
    dev.step(tr, 2, 300.0, t0=0.0)
    dev.step(tr, 3, 300.0, t0=600.0)
    torch.cuda.synchronize()
    _check(wl, "fp64", tr, dev.sediment_storage, F, ref_s, "two calls")


def test_coupling_alone_and_the_tendency_hook():
    """obme_pisces_step without the storage: the tracers of the oracle's run
    (they do not depend on the storage), nothing buried.  The tendency hook
    adds the coupled term into the receiver's bottom cells only."""
    import ctypes as C
    import torch
    from oceanbiome.runtime import _check as chk, _ptr_array, _stream, library
    print("case pisces sediment coupling alone", flush=True)
    wl = _case(seed=3)
    F, _ = _oracle(wl, [2])
    dev, tr = _device(wl, "fp64")
    chk(library().obme_pisces_step(dev.handle, _ptr_array(tr), 2, C.c_double(300.0), C.c_double(0.0), _stream()))
    torch.cuda.synchronize()
    assert float(dev.sediment_storage.abs().max()) == 0.0
    g = wl.grid
    worst = max(rel_l2(g.interior(tr[i].cpu().numpy()), F[n]) for i, n in enumerate(OP.WRITTEN))
    print("coupling alone worst", worst, flush=True)
    assert worst <= 1e-12
    # the hook: with and without the sediment on the same state
    wl2 = _case(seed=3)
    plain = copy.copy(wl2.bgc)
    plain.sediment = None
    dev_s, tr_s = _device(wl2, "fp64")
    dev_p = HIPPISCES(plain, "fp64")
    Gs = [torch.zeros(g.parent_shape, dtype=torch.float64, device="cuda") for _ in range(24)]
    Gp = [torch.zeros_like(x) for x in Gs]
    dev_s.update_tendencies(tr_s, Gs, t=0.0)
    dev_p.update_tendencies(tr_s, Gp, t=0.0)
    torch.cuda.synchronize()
    q = OP.params_from_vector(wl2.bgc.param_vector())
    Fi = {n: g.interior(a).copy() for n, a in zip(OP.TRACERS, wl2.tracers)}
    sed = wl2.bgc.config().sediment
    f = OP.sediment_flux(q, np.asarray(g.zc), Fi, sed)
    cpl = (1 - OP.burial_efficiency(sed, f)) * f / np.asarray(g.dz)[0]
    for i, n in enumerate(OP.WRITTEN):
        d = g.interior((Gs[i] - Gp[i]).cpu().numpy())
        if n == "DIC":
            np.testing.assert_allclose(d[0], cpl, rtol=1e-12)
            assert np.all(d[1:] == 0)
        else:
            assert np.all(d == 0), n


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_update_state_hook_sequence(precision):
    """obme_pisces_sediment_update_state call by call on fixed tracers: no
    step for a non-finite dt, the lazy first tendency, Euler on a changed dt,
    the χ-weighted step on a repeated one, against the oracle's
    sediment_time_step on the same flux."""
    import torch
    print("case pisces sediment hook", precision, flush=True)
    wl = _case(seed=4)
    g = wl.grid
    q = OP.params_from_vector(wl.bgc.param_vector())
    Fi = {n: g.interior(a).copy() for n, a in zip(OP.TRACERS, wl.tracers)}
    sed = wl.bgc.config().sediment
    f = OP.sediment_flux(q, np.asarray(g.zc), Fi, sed)
    state = O.SedimentState(g.Nx, g.Ny)
    dev, tr = _device(wl, precision)
    for dt in (float("inf"), 120.0, 120.0, 40.0, 40.0):
        dev.update_sediment_state(tr, dt)
        OP.sediment_time_step(sed, state, f, dt)
    torch.cuda.synchronize()
    es = rel_l2(dev.sediment_storage.cpu().numpy(), state.storage)
    print("hook storage", es, flush=True)
    assert es <= TOL[precision]
    assert state.iteration == 4


def test_rejected_configurations():
    print("case pisces sediment rejections", flush=True)
    wl = _case(sinking=("DOC",))
    with pytest.raises(OBMEError):
        _device(wl, "fp64")
    wl = _case()
    wl.bgc.sediment = SimpleMultiGSediment(RectilinearGrid(size=(1, 1, 16), z=(-32.0, 0.0)))
    with pytest.raises((OBMEError, ValueError)):
        _device(wl, "fp64")
