"""InstantRemineralisation sediment (Models/Sediments/instant_remineralisation.jl,
src/Sediments) in the oracle.  Pinned by: the burial efficiency's closed
forms (f = 0, f = k, f -> inf) under the reference's defaults; the flux
through the bottom face of a hand-made column (open bottom: speed times the
bottom cell, closed bottom: 0, a tracer without a drift velocity: 0); the
coupled tendency against the same model without a sediment (only the
receiver's bottom cell differs, by (1 - e) f / dz); the reference's own
sediment test property (test/test_sediments.jl:37-80: water-column nitrogen
plus sediment storage conserved to rtol 2e-7 over 100 steps of 1 s from
D = 1, storage non-zero everywhere); and the QuasiAdamsBashforth2 sequence
of the update-state hook worked out by hand (lazy first tendency, Euler on
a changed dt, (3/2 + chi) Gn - (1/2 + chi) Gm on a repeated one).  CPU only."""
import math

import numpy as np
import pytest

import oracle as O
from oceanbiome import LOBSTER, NPZD, InstantRemineralisationSediment, RectilinearGrid, abi
from oceanbiome import TwoBandPhotosyntheticallyActiveRadiation

A, B, K = 0.013, 0.53, 7.0 / 6.56


def _grid(shape=(3, 3, 50), extent=(10.0, 10.0, 500.0)):
    return RectilinearGrid(size=shape, extent=extent, halo=(3, 3, 3), topology=("Periodic", "Periodic", "Bounded"))


def _fields(grid, values):
    out = []
    for v in values:
        f = grid.new_field()
        grid.interior(f)[...] = v
        out.append(f)
    return out


def test_burial_efficiency_closed_forms():
    s = InstantRemineralisationSediment().struct(["N", "P", "Z", "T", "D"])
    assert O.burial_efficiency(s, 0.0) == A
    assert O.burial_efficiency(s, K) == pytest.approx(A + B / 4, rel=1e-15)
    assert O.burial_efficiency(s, 1e12) == pytest.approx(A + B, rel=1e-10)
    # the code's (f / (k + f))^2, not the docstring's f / (k + f)^2
    f = 3.0
    assert O.burial_efficiency(s, f) == pytest.approx(A + B * (f / (K + f)) ** 2, rel=1e-15)
    assert O.burial_efficiency(s, f) != pytest.approx(A + B * f / (K + f) ** 2, rel=1e-3)


@pytest.mark.parametrize("open_bottom", [True, False])
def test_bottom_face_flux(open_bottom):
    grid = _grid((1, 1, 8), (1.0, 1.0, 80.0))
    bgc = NPZD(grid, sediment_model=InstantRemineralisationSediment())
    bgc.detritus.open_bottom = open_bottom
    cfg = bgc.config()
    U = np.zeros((5, 8))
    U[1, 0], U[4, 0] = 2.0, 3.0          # P and D at the bottom
    U[1, 1:], U[4, 1:] = 50.0, 70.0      # the cells above do not count
    f = O.sediment_flux(cfg, U)
    if open_bottom:
        expect = (0.2551 / 86400) * 2.0 + (2.7489 / 86400) * 3.0
        assert f == pytest.approx(expect, rel=1e-15)
    else:
        assert f == 0.0
    # a listed tracer without a drift velocity (Z) adds nothing
    bgc2 = NPZD(grid, sediment_model=InstantRemineralisationSediment(sinking_tracers=("Z", "D")))
    bgc2.detritus.open_bottom = open_bottom
    U[2, 0] = 5.0
    assert O.sediment_flux(bgc2.config(), U) == pytest.approx((2.7489 / 86400) * 3.0 if open_bottom else 0.0,
                                                               rel=1e-15)


def test_coupled_tendency_is_the_only_difference():
    grid = _grid((4, 3, 20), (10.0, 10.0, 100.0))
    rng = np.random.default_rng(3)
    vals = [rng.uniform(1, 10, (20, 3, 4)), rng.uniform(0.1, 1, (20, 3, 4)), rng.uniform(0.1, 1, (20, 3, 4)),
            rng.uniform(8, 20, (20, 3, 4)), rng.uniform(0.5, 2, (20, 3, 4))]
    tr = _fields(grid, vals)
    bgc0 = NPZD(grid, light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=50.0))
    bgc1 = NPZD(grid, light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=50.0),
                sediment_model=InstantRemineralisationSediment())
    G0 = [np.zeros_like(a) for a in tr]
    G1 = [np.zeros_like(a) for a in tr]
    O.grid_tendencies(bgc0.config(), tr, G0)
    O.grid_tendencies(bgc1.config(), tr, G1)
    for n in range(5):
        d = grid.interior(G1[n]) - grid.interior(G0[n])
        if n != 0:
            assert np.all(d == 0), n
            continue
        assert np.all(d[1:] == 0)
        P0, D0 = vals[1][0], vals[4][0]
        f = (0.2551 / 86400) * P0 + (2.7489 / 86400) * D0
        e = A + B * (f / (K + f)) ** 2
        np.testing.assert_allclose(d[0], (1 - e) * f / 5.0, rtol=1e-12)
        assert np.all(d[0] > 0)


def _nitrogen(grid, tracers, names, storage):
    dz = grid.dz
    water = sum(np.einsum("kji,k->ji", grid.interior(tracers[names.index(n)]), dz) for n in ("N", "P", "Z", "D"))
    return water + storage


def test_reference_sediment_nitrogen_budget():
    """test/test_sediments.jl:37-80 with NPZD and InstantRemineralisation on
    the 3x3x50 grid (500 m deep): set D = 1, 100 steps of 1 s."""
    grid = _grid()
    bgc = NPZD(grid, sediment_model=InstantRemineralisationSediment(grid))
    names = list(bgc.required_biogeochemical_tracers())
    tr = _fields(grid, [0.0, 0.0, 0.0, 0.0, 1.0])
    state = O.SedimentState(3, 3)
    before = _nitrogen(grid, tr, names, state.storage)
    O.grid_run_sediment(bgc.config(), tr, state, 100, 1.0)
    after = _nitrogen(grid, tr, names, state.storage)
    np.testing.assert_allclose(after, before, rtol=0.2e-6)
    assert np.all(state.storage != 0)
    assert state.iteration == 300
    assert state.last_dt == (3.0 / 4.0 + -5.0 / 12.0) * 1.0


def test_lobster_budget_with_nh4_receiver():
    grid = _grid((3, 2, 30), (10.0, 10.0, 300.0))
    sed = InstantRemineralisationSediment(sinking_tracers=("sPOM", "bPOM"), remineralisation_reciever="NH₄")
    bgc = LOBSTER(grid, sediment_model=sed)
    names = list(bgc.required_biogeochemical_tracers())
    tr = _fields(grid, [0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0])
    state = O.SedimentState(3, 2)
    members = ("NO₃", "NH₄", "P", "Z", "sPOM", "bPOM", "DOM")

    def total():
        return sum(np.einsum("kji,k->ji", grid.interior(tr[names.index(n)]), grid.dz) for n in members) + state.storage

    before = total()
    O.grid_run_sediment(bgc.config(), tr, state, 100, 1.0)
    np.testing.assert_allclose(total(), before, rtol=0.2e-6)
    assert np.all(state.storage > 0)


def test_hook_sequence_by_hand():
    """update_biogeochemical_state!(model, sediment) call by call."""
    grid = _grid((2, 1, 4), (1.0, 1.0, 4.0))
    bgc = NPZD(grid, sediment_model=InstantRemineralisationSediment())
    cfg = bgc.config()
    s = cfg.sediment
    tr = _fields(grid, [1.0, 2.0, 0.5, 10.0, 3.0])
    state = O.SedimentState(2, 1)

    def ge():
        f = (0.2551 / 86400) * grid.interior(tr[1])[0, 0, 0] + (2.7489 / 86400) * grid.interior(tr[4])[0, 0, 0]
        return O.burial_efficiency(s, f) * f

    O.grid_sediment_update_state(cfg, tr, state, math.inf)        # before the first stage: no step
    assert state.iteration == 0 and np.all(state.storage == 0)
    g1 = ge()
    O.grid_sediment_update_state(cfg, tr, state, 10.0)            # iteration 0: lazy Gn, Euler
    assert state.storage[0, 0] == pytest.approx(10.0 * g1, rel=1e-15)
    grid.interior(tr[4])[...] = 6.0
    g2 = ge()
    O.grid_sediment_update_state(cfg, tr, state, 10.0)            # same dt: AB2 with Gn = Gm = g1
    st2 = 10.0 * g1 + 10.0 * (1.6 * g1 - 0.6 * g1)
    assert state.storage[0, 0] == pytest.approx(st2, rel=1e-14)
    O.grid_sediment_update_state(cfg, tr, state, 10.0)            # AB2: Gn = g2, Gm = g1
    st3 = st2 + 10.0 * (1.6 * g2 - 0.6 * g1)
    assert state.storage[0, 0] == pytest.approx(st3, rel=1e-14)
    O.grid_sediment_update_state(cfg, tr, state, 4.0)             # new dt: Euler with Gn = g2
    assert state.storage[0, 0] == pytest.approx(st3 + 4.0 * g2, rel=1e-14)
    assert state.iteration == 4 and state.last_dt == 4.0
    np.testing.assert_array_equal(state.storage[0, 0], state.storage[0, 1])


def test_host_description_of_the_sediment():
    """InstantRemineralisationSediment on the host: the defaults of
    instant_remineralisation.jl:84-98, names resolved in tracer order, bad
    names and the RungeKutta3 sediment stepper rejected."""
    names = ["NO₃", "NH₄", "P", "Z", "sPOM", "bPOM", "DOM"]
    s = InstantRemineralisationSediment(sinking_tracers=("sPOM", "bPOM"), remineralisation_reciever="NH₄").struct(names)
    assert s.model == abi.OBME_SEDIMENT_INSTANT_REMINERALISATION and s.receiver == 1
    assert s.n_sinking == 2 and list(s.sinking)[:2] == [4, 5]
    assert (s.burial_efficiency_constant1, s.burial_efficiency_constant2) == (A, B)
    assert s.burial_efficiency_half_saturation == K and s.ab2_chi == 0.1
    with pytest.raises(ValueError):
        InstantRemineralisationSediment(remineralisation_reciever="N").struct(names)
    with pytest.raises(ValueError):
        InstantRemineralisationSediment(sinking_tracers=("D",), remineralisation_reciever="NH₄").struct(names)
    with pytest.raises(NotImplementedError):
        InstantRemineralisationSediment(timestepper="RungeKutta3")
