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
