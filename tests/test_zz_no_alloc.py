"""obme_step_noalloc, the allocation-free variant of obme_step (no
counterpart in the reference; for callers that cannot allocate device
memory), against obme_step through the C ABI on the same inputs: FP64
bit-identical, FP32-mixed within 1e-5 relative L2 (its [H+] warm starts
begin at zero on every call).  Each case runs two noalloc calls in a row,
the second on a smaller grid, so stale contents of the shared static
scratch would show; a grid whose scratch exceeds the array is rejected
with a message before any device work."""
import numpy as np
import pytest

from helpers import rel_l2, to_device, to_host
from oceanbiome import workloads
from oceanbiome.runtime import HIPBiogeochemistry, OBMEError

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5


def _run(wl, prec, noalloc, T=None, S=None):
    import torch
    dev = HIPBiogeochemistry(wl.bgc, prec)
    tr = to_device(wl.tracers, dev.dtype)
    T_d = None if T is None else to_device([T], dev.dtype)[0]
    S_d = None if S is None else to_device([S], dev.dtype)[0]
    if noalloc:
        dev.step_noalloc(tr, wl.nsteps, wl.dt, T=T_d, S=S_d)
    else:
        dev.step(tr, wl.nsteps, wl.dt, T=T_d, S=S_d)
    torch.cuda.synchronize()
    return to_host(tr)


def _compare(wl, prec, T=None, S=None):
    a = _run(wl, prec, True, T, S)
    b = _run(wl, prec, False, T, S)
    g = wl.grid
    for n, (x, y) in enumerate(zip(a, b)):
        if prec == "fp64":
            assert np.array_equal(x, y), (wl.name, prec, n, float(np.max(np.abs(x - y))))
        else:
            e = rel_l2(g.interior(x), g.interior(y))
            assert e <= FP32_TOL, (wl.name, prec, n, e)
            # halos untouched
            h = np.ones(x.shape, bool)
            h[g.Hz:g.Hz + g.Nz, g.Hy:g.Hy + g.Ny, g.Hx:g.Hx + g.Nx] = False
            assert np.array_equal(x[h], wl.tracers[n][h].astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_matches_step_twice_smaller_second(prec):
    for shape in ((64, 32, 32), (16, 8, 16)):
        print("case noalloc lobster", prec, shape, flush=True)
        _compare(workloads.lobster_3d(*shape, nsteps=6), prec)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_oxygen_and_3d_temperature(prec):
    """LOBSTER + DIC/Alk + O2 with both gas exchanges (the O2 coefficients
    in the scratch too), T and S as 3-D fields (a column of surface
    constants per column), then a smaller grid."""
    for shape in ((48, 20, 24), (10, 6, 8)):
        print("case noalloc oxygen 3d", prec, shape, flush=True)
        wl = workloads.lobster_3d(*shape, nsteps=5, oxygen=True)
        g = wl.grid
        T = g.new_field()
        S = g.new_field()
        g.interior(T)[...] = wl.extra["T_profile"][:, None, None] + 0.01 * np.arange(g.Nx)[None, None, :]
        g.interior(S)[...] = 35.0 - 0.02 * np.arange(g.Ny)[None, :, None]
        _compare(wl, prec, T, S)


def test_noalloc_rejects_a_grid_beyond_the_scratch():
    """2048 x 2048 columns in FP32 need 44 B x 4 Mi > 64 MiB of scratch."""
    import torch
    print("case noalloc oversize", flush=True)
    wl = workloads.lobster_3d(2048, 2048, 1, nsteps=1)
    dev = HIPBiogeochemistry(wl.bgc, "fp32")
    tr = to_device(wl.tracers, dev.dtype)
    before = [t.clone() for t in tr]
    with pytest.raises(OBMEError, match="64 MiB"):
        dev.step_noalloc(tr, 1, wl.dt)
    torch.cuda.synchronize()
    assert all(torch.equal(x, y) for x, y in zip(tr, before))
