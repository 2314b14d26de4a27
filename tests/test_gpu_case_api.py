"""The reference's validation cases written against the C++ case API
(Geometry2D / Variable2D / PoissonSolver2D / ConcatPoissonSolver2D over the C
ABI), run as compiled programs on the device and compared with the committed
fixtures.

Tolerances: rel L2 <= 1e-12 for the single-domain solver, <= 1e-10 for the
composite solver (relative GMRES tolerance 1e-13 in the programs)."""
import os

import numpy as np
import pytest

import case_programs
from gpu_util import banner

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def test_case_poisson_3x4(tmp_path):
    """case/validation/2d/poisson_solver_reference.cpp:3-95 (15 D/N combos)."""
    banner("case_poisson_3x4")
    z = np.load(os.path.join(GOLD, "poisson_3x4.npz"))
    out = tmp_path / "u.bin"
    case_programs.run("case_poisson_3x4", out)
    u = np.fromfile(out, dtype=np.float64).reshape(z["u"].shape)
    for k in range(u.shape[0]):
        assert _rel(u[k], z["u"][k]) <= 1e-12, (k, z["bc"][k])


def test_case_concat_six(tmp_path):
    """case/validation/2d/concat_solver_reference.cpp:42-108."""
    banner("case_concat_six")
    z = np.load(os.path.join(GOLD, "concat_six.npz"))
    out = tmp_path / "u.bin"
    stdout = case_programs.run("case_concat_six", out)
    assert "root T1" in stdout
    flat = np.fromfile(out, dtype=np.float64)
    pos = 0
    us, refs = [], []
    for k, (nx, ny) in enumerate(z["sizes"]):
        us.append(flat[pos:pos + nx * ny].reshape(nx, ny))
        pos += nx * ny
        refs.append(z["u%d" % k])
    assert pos == flat.size
    r = _rel(np.concatenate([u.ravel() for u in us]), np.concatenate([u.ravel() for u in refs]))
    assert r <= 1e-10, r


def test_case_cross_io(tmp_path):
    """The 32-cell mixed-boundary cross with ghost data, through
    Variable2D::set_boundary_type and buffer_map."""
    banner("case_cross_io")
    z = np.load(os.path.join(GOLD, "cross_32.npz"))
    n = 32
    sizes = [tuple(s) for s in z["sizes"]]
    bcs = [tuple(b) for b in z["bcs"]]
    parts = [z["f%d" % k].ravel() for k in range(5)]
    for d in range(5):
        for face in range(4):
            if bcs[d][face] not in (0, 1):
                continue
            key = "g%d_%d" % (d, face)
            length = sizes[d][1] if face < 2 else sizes[d][0]
            parts.append(z[key] if key in z.files else np.zeros(length))
    inp, out = tmp_path / "in.bin", tmp_path / "u.bin"
    np.concatenate(parts).astype(np.float64).tofile(inp)
    stdout = case_programs.run("case_cross_io", out, inp, n)
    assert stdout.startswith("root C"), stdout
    flat = np.fromfile(out, dtype=np.float64)
    ref = np.concatenate([z["u%d" % k].ravel() for k in range(5)])
    assert flat.size == ref.size
    r = _rel(flat, ref)
    assert r <= 1e-10, r


def test_case_manufactured_second_order():
    """Manufactured solution on an L-shaped geometry: max error O(h^2)
    through axis offsets, set_value and set_buffer (ghost centres)."""
    banner("case_manufactured")
    out = case_programs.run("case_manufactured", 32, 64, 128)
    errs = [float(line.split()[3]) for line in out.strip().splitlines()]
    assert errs[0] < 1e-3, errs
    for a, b in zip(errs, errs[1:]):
        assert 3.5 < a / b < 4.5, errs


def test_case_field2_device_mirror():
    """field2's device mirror: solves with and without host access in between
    agree bit for bit, device-current fields copy device to device, host
    writes after a solve reach the next one."""
    banner("case_field2_mirror")
    out = case_programs.run("case_field2_mirror")
    assert "mirror ok" in out, out


def test_case_cross_timing_small():
    """The case-API timing program on a 128 cross: both forms (fields
    device-resident between solves, host round trip per solve) give the same
    solution."""
    banner("case_cross_timing 128")
    out = case_programs.run("case_cross_timing", 128, 3).split()
    k = out.index("solution_norm")
    a, b = float(out[k + 1]), float(out[k + 2])
    assert abs(a - b) <= 1e-12 * abs(a), out
    assert float(out[out.index("device_resident_ms") + 1]) > 0


def test_case_projection_time_loop(tmp_path):
    """The Navier–Stokes caller's pattern through the case API: a pressure
    projection (PressureProjection2D over ConcatPoissonSolver2D) captured once
    and replayed every time step on a two-domain channel, the fields resident
    on the device; the projected velocity divergence-free to the solve's
    accuracy, inlet and wall velocities kept, interface copies equal, the
    first replay equal to the eager step, and the pressure through a
    write_binary / read_binary round trip."""
    banner("case_projection")
    out = case_programs.run("case_projection", tmp_path / "p.fpf2")
    banner(out.strip().splitlines()[0])
    assert "projection ok" in out, out
