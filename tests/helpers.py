"""Shared helpers of the test-suite (comparison metrics, device transfer)."""
import numpy as np


def rel_l2(x, ref):
    x = np.asarray(x, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    den = np.linalg.norm(ref)
    num = np.linalg.norm(x - ref)
    return num / den if den > 0 else num


def max_rel(x, ref, floor=1e-30):
    x = np.asarray(x, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    return float(np.max(np.abs(x - ref) / np.maximum(np.abs(ref), floor)))


def to_device(arrays, dtype):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype) for a in arrays]


def to_host(tensors):
    return [t.double().cpu().numpy() for t in tensors]



def fp32_storage_floor(step_once, nsteps, F_start, F_ref, names):
    """Per-tracer relative L2 deviation from the FP64 oracle's result F_ref
    of the same FP64 oracle when its state is rounded to FP32 at the start
    and after every step (step_once(F, s) advances F in place by step s):
    the error that storing the state in FP32 alone causes.  Where the
    reference's own formulas are ill-conditioned at the inputs (PISCES's iron
    uptake factor (1 - r)/(1.05 - r) has a pole at r = 1.05,
    mixed_mondo.jl:161-181) it exceeds any fixed bar."""
    X = {k: np.asarray(v, dtype=np.float32).astype(np.float64) for k, v in F_start.items()}
    with np.errstate(all="ignore"):
        for s in range(nsteps):
            step_once(X, s)
            for k in names:
                X[k] = X[k].astype(np.float32).astype(np.float64)
    out = {}
    for k in names:
        a, b = np.asarray(X[k]), np.asarray(F_ref[k])
        ok = np.isfinite(a) & np.isfinite(b)
        out[k] = rel_l2(a[ok], b[ok]) if ok.any() else 0.0
    return out
