"""Host-side logic: grid geometry and layout, model configuration, workload
generation.  CPU only."""
import numpy as np
import pytest

from oceanbiome import abi, workloads
from oceanbiome.grid import RectilinearGrid
from oceanbiome.models import LOBSTER, NPZD, CarbonateSystem, TwoBandPhotosyntheticallyActiveRadiation
from oceanbiome.units import day


def test_uniform_grid_geometry_and_layout():
    g = RectilinearGrid(size=(8, 4, 16), extent=(80.0, 40.0, 32.0), halo=(3, 2, 1))
    assert g.parent_shape == (16 + 2, 4 + 4, 8 + 6)
    zf = np.asarray(g.zf)
    zc = np.asarray(g.zc)
    assert zf[0] == -32.0 and zf[-1] == 0.0 and len(zf) == 17
    assert np.allclose(np.diff(zf), 2.0, rtol=0, atol=0)
    assert np.array_equal(zc, (zf[1:] + zf[:-1]) / 2)
    f = g.new_field()
    assert f.shape == g.parent_shape and f.dtype == np.float64
    g.interior(f)[...] = 1.0
    assert f.sum() == 8 * 4 * 16
    assert g.interior(f)[0, 0, 0] is not None and f[1, 2, 3] == 1.0 and f[0, 2, 3] == 0.0


def test_stretched_grid_from_faces():
    faces = np.array([-100.0, -60.0, -30.0, -10.0, 0.0])
    g = RectilinearGrid(size=(2, 2, 4), x=(0, 1), y=(0, 1), z=faces)
    assert np.array_equal(np.asarray(g.zf), faces)
    assert np.array_equal(np.asarray(g.dz), np.diff(faces))


def test_config_sinking_and_sources():
    g = RectilinearGrid(size=(4, 4, 8), extent=(1, 1, 8))
    b = LOBSTER(g, carbonate_system=CarbonateSystem(), temperature=np.linspace(4, 20, 8), salinity=35.0)
    cfg = b.config(abi.OBME_FP32_MIXED)
    assert cfg.model == abi.OBME_LOBSTER and cfg.carbonate_system == 1 and cfg.precision == abi.OBME_FP32_MIXED
    speeds = [cfg.sinking.speed[n] for n in range(9)]
    assert speeds == [0, 0, 0, 0, 3.47e-5, 200 / day, 0, 0, 0]
    assert cfg.sinking.open_bottom == 1
    assert cfg.temperature.mode == abi.OBME_FIELD_PROFILE and cfg.salinity.mode == abi.OBME_FIELD_CONSTANT
    assert cfg.temperature.profile[7] == 20.0
    n = NPZD(g, light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=50.0)).config()
    assert n.par.surface_mode == abi.OBME_SURFACE_CONSTANT and n.par.surface_value == 50.0
    assert [n.sinking.speed[q] for q in range(5)] == [0, 0.2551 / day, 0, 0, 2.7489 / day]
    assert NPZD(g).config().par.surface_mode == abi.OBME_SURFACE_DIURNAL


def test_written_tracers_and_bytes():
    wl = workloads.lobster_3d(8, 4, 16, nsteps=1)
    assert wl.n_written == 9
    assert wl.algorithmic_bytes_per_step(4) == 8 * 4 * 16 * 72
    assert wl.algorithmic_bytes_per_step(8) == 8 * 4 * 16 * 144
    n = workloads.npzd_2d()
    assert n.n_written == 4
    assert n.algorithmic_bytes_per_step(8) == 160 * 32 * (64 + 8)


def test_workloads_are_seeded():
    a = workloads.lobster_3d(6, 5, 7, nsteps=1, seed=3)
    b = workloads.lobster_3d(6, 5, 7, nsteps=1, seed=3)
    c = workloads.lobster_3d(6, 5, 7, nsteps=1, seed=4)
    assert all(np.array_equal(x, y) for x, y in zip(a.tracers, b.tracers))
    assert not np.array_equal(a.tracers[0], c.tracers[0])
    g = a.grid
    dic = g.interior(a.tracers[7])
    assert dic.min() >= 2000 and dic.max() <= 2200


def test_npzd_front_initial_state():
    wl = workloads.npzd_2d()
    g = wl.grid
    N = g.interior(wl.tracers[0])[:, 0, :]
    assert N.shape == (32, 160)
    assert 0.5 <= N.min() and N.max() <= 4.5


def test_column_ensemble():
    wl = workloads.npzd_columns(n_columns=64, nsteps=1)
    off = wl.extra["offsets"]
    nz = np.diff(off)
    assert nz.min() >= 16 and nz.max() <= 128
    assert len(wl.tracers[0]) == off[-1]
    p = wl.extra["params"]
    assert p.shape == (20, 64)
    assert np.all(p[19] >= 50) and np.all(p[19] <= 200)
