"""GPU parity of the column ensemble with an InstantRemineralisation sediment
under every column (obme_batch_step_with_sediment; BiogeochemicalSediment,
src/Sediments, with Models/Sediments/instant_remineralisation.jl) against the
oracle's ensemble (oracle.batch_run with the sediment): ragged depths, every
rate perturbed per column.  FP64: relative L2 <= 1e-12 per tracer and for the
storage; FP32-mixed: <= 1e-5.  Also: two calls in a row (the sediment clock:
the lazy first tendency only in the first), several launches per call from a
hipGraph (the same result as one launch), obme_batch_step on such an
ensemble (the coupling alone, storage untouched), the autotuner leaving
tracers and storage alone, and the configurations the ensemble rejects."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2
from oceanbiome import InstantRemineralisationSediment, RectilinearGrid, SimpleMultiGSediment, abi, workloads
from oceanbiome.runtime import HIPBatch, OBMEError

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 1e-5}


def _ensemble(n=80, seed=9, receiver="N", sinking=("P", "D")):
    wl = workloads.npzd_columns(n_columns=n, nz_min=16, nz_max=128, dz=2.0, nsteps=1, seed=seed)
    wl.bgc.sediment = InstantRemineralisationSediment(sinking_tracers=sinking, remineralisation_reciever=receiver)
    return wl


def _args(wl):
    return wl.extra["offsets"], list(wl.extra["params"]), wl.extra["dz"], wl.extra["T"]


def _oracle(wl, calls, dt=60.0):
    U = [a.copy() for a in wl.tracers]
    st = O.SedimentState(len(wl.extra["offsets"]) - 1, 1)
    for n in calls:
        O.batch_run(wl.bgc.config(abi.OBME_FP64), *_args(wl), U, n, dt, sediment=st)
    return U, st.storage.ravel().copy()


def _device(wl, precision):
    import torch
    b = HIPBatch(wl.bgc, wl.extra["offsets"], wl.extra["params"], wl.extra["dz"], wl.extra["T"], precision)
    tr = [torch.from_numpy(a).to(device="cuda", dtype=b.dtype) for a in wl.tracers]
    return b, tr


def _check(precision, out, storage, ref, ref_storage, what):
    errs = [rel_l2(a.double().cpu().numpy(), b) for a, b in zip(out, ref)]
    es = rel_l2(storage.cpu().numpy(), ref_storage)
    print(what, precision, "tracers", errs, "storage", es, flush=True)
    assert max(errs) <= TOL[precision] and es <= TOL[precision], (errs, es)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sediment_ensemble_matches_oracle(precision):
    import torch
    print("case batch sediment", precision, flush=True)
    wl = _ensemble()
    ref, ref_s = _oracle(wl, [40])
    b, tr = _device(wl, precision)
    b.step(tr, 40, 60.0)
    torch.cuda.synchronize()
    assert float(b.sediment_storage.min()) > 0
    _check(precision, tr, b.sediment_storage, ref, ref_s, "one call")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_two_calls_and_graph_replay(precision):
    """Two calls of 15 steps against the oracle's two calls; the same with
    five launches of 3 steps per call (a hipGraph, replayed by the second)."""
    import torch
    print("case batch sediment calls", precision, flush=True)
    wl = _ensemble(n=48, seed=3, receiver="D", sinking=("D",))
    ref, ref_s = _oracle(wl, [15, 15])
    outs = []
    for spl in (0, 3):
        b, tr = _device(wl, precision)
        b.configure(1, spl)
        b.step(tr, 15, 60.0)
        b.step(tr, 15, 60.0)
        torch.cuda.synchronize()
        _check(precision, tr, b.sediment_storage, ref, ref_s, f"two calls spl={spl}")
        outs.append(([t.cpu().numpy() for t in tr], b.sediment_storage.cpu().numpy()))
    # several launches per call: the state and the FP32 residuals pass
    # between launches, so the result is the one-launch result
    for a, c in zip(outs[0][0], outs[1][0]):
        np.testing.assert_array_equal(a, c)
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_coupling_alone_without_storage():
    """obme_batch_step on an ensemble with a sediment: the receiver gets the
    coupled tendency, nothing is buried (the storage argument is absent)."""
    import ctypes as C
    import torch
    from oceanbiome.runtime import _check as chk, _ptr_array, _stream, library
    print("case batch sediment coupling alone", flush=True)
    wl = _ensemble(n=32, seed=5)
    b, tr = _device(wl, "fp64")
    chk(library().obme_batch_step(b.handle, _ptr_array(tr), 10, C.c_double(60.0), C.c_double(0.0), _stream()))
    torch.cuda.synchronize()
    assert float(b.sediment_storage.abs().max()) == 0.0
    # the tracers against the oracle's run with the sediment: they do not
    # depend on the storage, which only receives the buried part
    U, _ = _oracle(wl, [10])
    errs = [rel_l2(a.cpu().numpy(), c) for a, c in zip(tr, U)]
    print("coupling alone", errs, flush=True)
    assert max(errs) <= 1e-12


def test_autotune_leaves_tracers_and_storage():
    import torch
    print("case batch sediment autotune", flush=True)
    wl = _ensemble(n=64, seed=2)
    b, tr = _device(wl, "fp32")
    b.sediment_storage.fill_(0.25)
    before = [t.clone() for t in tr]
    b.autotune(tr, 3, 60.0)
    torch.cuda.synchronize()
    for a, c in zip(tr, before):
        assert torch.equal(a, c)
    assert float(b.sediment_storage.min()) == 0.25 and float(b.sediment_storage.max()) == 0.25
    # and the clock is untouched: a run after autotuning matches the oracle
    b.sediment_storage.zero_()
    ref, ref_s = _oracle(wl, [10])
    b.step(tr, 10, 60.0)
    torch.cuda.synchronize()
    _check("fp32", tr, b.sediment_storage, ref, ref_s, "after autotune")


def test_rejected_configurations():
    print("case batch sediment rejections", flush=True)
    wl = _ensemble(n=8, receiver="T")
    with pytest.raises(OBMEError):
        _device(wl, "fp64")
    wl = _ensemble(n=8)
    wl.bgc.sediment = SimpleMultiGSediment(RectilinearGrid(size=(1, 1, 16), z=(-32.0, 0.0)))
    with pytest.raises((OBMEError, ValueError)):
        _device(wl, "fp64")
