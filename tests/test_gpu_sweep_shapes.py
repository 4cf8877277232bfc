"""The column sweep (lane per water column) over workgroup sizes and grid
shapes: every workgroup size of 1, 2, 4 or 8 waves runs the same
arithmetic per column, so the results must be bit-identical.  The shapes
cover a partial last workgroup along i, fewer workgroups than XCDs, many
workgroups per row and columns deeper than a wave is wide; the small ones
are checked against the FP64 oracle too."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, to_device, to_host
from oceanbiome import NPZD, RectilinearGrid, TwoBandPhotosyntheticallyActiveRadiation, abi, workloads
from oceanbiome.runtime import HIPBiogeochemistry
from oceanbiome.workloads import Workload

pytestmark = pytest.mark.gpu


def _step(wl, waves, monkeypatch, nsteps):
    import torch
    monkeypatch.setenv("OBME_SWEEP_WAVES", str(waves))
    dev = HIPBiogeochemistry(wl.bgc, "fp32")
    tr = to_device(wl.tracers, dev.dtype)
    dev.step(tr, nsteps, wl.dt)
    torch.cuda.synchronize()
    return to_host(tr)


def _npzd_3d(Nx, Ny, Nz, seed=3):
    grid = RectilinearGrid(size=(Nx, Ny, Nz), extent=(100.0, 100.0, 128.0), halo=(3, 3, 3),
                           topology=("Periodic", "Periodic", "Bounded"))
    bgc = NPZD(grid, light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=100.0))
    rng = np.random.default_rng(seed)
    fields = []
    for lo, hi in ((1.0, 10.0), (0.01, 0.1), (0.01, 0.1), (8.0, 20.0), (0.0, 0.05)):
        f = grid.new_field()
        grid.interior(f)[...] = rng.uniform(lo, hi, (Nz, Ny, Nx))
        fields.append(f)
    return Workload(f"npzd_{Nx}x{Ny}x{Nz}", bgc, grid, fields, 60.0, 10)


@pytest.mark.parametrize("shape", [(45, 20, 64), (7, 3, 64), (512, 64, 64), (300, 5, 200)])
def test_sweep_waves_lobster(shape, monkeypatch):
    print("case sweep waves lobster", shape, flush=True)
    wl = workloads.lobster_3d(*shape, nsteps=4)
    a = _step(wl, 4, monkeypatch, 4)
    for waves in (1, 2, 8):
        b = _step(wl, waves, monkeypatch, 4)
        for n, (x, y) in enumerate(zip(a, b)):
            assert np.array_equal(x, y), (waves, n, float(np.max(np.abs(x - y))))
    if shape[0] * shape[1] <= 1500:      # the oracle on the small shapes
        ref = [t.copy() for t in wl.tracers]
        O.grid_run(wl.bgc.config(abi.OBME_FP64), ref, 4, wl.dt)
        g = wl.grid
        errs = [rel_l2(g.interior(x), g.interior(r)) for x, r in zip(a, ref)]
        assert max(errs) <= 1e-5, errs


def test_sweep_npzd_keeps_temperature(monkeypatch):
    print("case sweep npzd", flush=True)
    wl = _npzd_3d(40, 9, 64)
    a = _step(wl, 4, monkeypatch, 10)
    b = _step(wl, 1, monkeypatch, 10)
    for n, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x, y), (n, float(np.max(np.abs(x - y))))
    # T (tracer 3) is read-only: unchanged, halos included
    assert np.array_equal(a[3], wl.tracers[3].astype(np.float32))
    ref = [t.copy() for t in wl.tracers]
    O.grid_run(wl.bgc.config(abi.OBME_FP64), ref, 10, wl.dt)
    g = wl.grid
    errs = [rel_l2(g.interior(x), g.interior(r)) for x, r in zip(a, ref)]
    assert max(errs) <= 1e-5, errs
