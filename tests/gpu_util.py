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


def _rss_gib():
    """The process's resident memory (anonymous, file-backed, shared) and
    its peak, for the case banners."""
    keys = ("VmRSS:", "VmHWM:", "RssAnon:", "RssFile:", "RssShmem:")
    try:
        with open("/proc/self/status") as fh:
            v = {l.split(":")[0]: int(l.split()[1]) / 2.0 ** 20 for l in fh if l.startswith(keys)}
        return " [rss %.2f GiB (anon %.2f, file %.2f, shmem %.2f), peak %.2f GiB]" % (
            v["VmRSS"], v["RssAnon"], v["RssFile"], v["RssShmem"], v["VmHWM"])
    except (OSError, KeyError, ValueError):
        return ""


def banner(msg):
    print("[gpu-case] " + msg + _rss_gib(), flush=True)


def sub_reference(f, hx, hy, bc, refine=2):
    """PoissonSolver2D::solve of an hx^2-scaled right-hand side by the
    oracle, refined with the long-double residual of the same operator when
    a Dirichlet face makes it nonsingular: the discrete solution to about
    FP64 rounding (the plain oracle solve carries its own rounding floor)."""
    import oracle as O
    if not any(t in (0, 3) for t in bc):
        return O.sub_solve(f, hx, hy, bc)
    nx, ny = f.shape
    return O.reference_solution([(nx, ny, hx, hy, tuple(bc))], [], [np.asarray(f) / (hx * hx)], refine=refine)[0]
