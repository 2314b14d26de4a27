"""Helpers of the GPU parity tests."""
import numpy as np
import pytest


def device():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible to torch")
    return torch.device("cuda:0")


def to_dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device())


def rel(a, b):
    a = np.concatenate([np.ravel(x) for x in a]) if isinstance(a, (list, tuple)) else np.ravel(a)
    b = np.concatenate([np.ravel(x) for x in b]) if isinstance(b, (list, tuple)) else np.ravel(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def banner(msg):
    print("[gpu-case] " + msg, flush=True)
