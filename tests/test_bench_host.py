"""Host-side pieces of bench.py that its numbers rest on (CPU only)."""

import importlib.util
import os

import numpy as np
import torch

from oracle import ref_np as O

from test_oracle import assemble, random_beta

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_fp64_operator_matches_assembled_matrix():
    """The FP64 operator used for the reference pressure equals the
    scipy-assembled matrix (symmetric, zero row sums) on random data."""
    bench = _bench()
    rng = np.random.default_rng(3)
    res = (6, 5, 7)
    bx, by, bz = random_beta(res, rng, contrast=1000.0)
    p = rng.standard_normal(res)
    diag = O.compute_diag(bx, by, bz)
    got = bench.fp64_apply_A(*(torch.from_numpy(a) for a in (bx, by, bz)), torch.from_numpy(p))
    A = assemble(res, bx, by, bz)
    np.testing.assert_allclose(got.numpy().ravel(), A @ p.ravel(), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(got.numpy(), O.apply_A(p, bx, by, bz, diag), rtol=1e-13, atol=1e-13)
    # constants are in the null space (zero row sums)
    one = bench.fp64_apply_A(*(torch.from_numpy(a) for a in (bx, by, bz)), torch.ones(res, dtype=torch.float64))
    assert float(one.abs().max()) < 1e-14   # FP64 rounding only


def test_roofline_bytes():
    bench = _bench()
    n = 128 ** 3
    assert bench.smoother_bytes(n) == 20 * n          # FP32: 40 B/cell per full sweep
    assert bench.smoother_bytes(n, 2, 2) == 11 * n    # BF16 p, beta; FP32 b
    assert bench.fused_sweep_bytes(n) == 24 * n
    assert bench.fused_sweep_bytes(n, 2, 2) == 14 * n
    assert bench.p2g_bytes(10, 1) == 28 * 10 + 28
    assert bench.g2p_bytes(10, 1) == 56 * 10 + 36


def test_multi_gpu_request_exits():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "single-GPU" in r.stdout


def test_measured_traffic_records():
    """Each committed pmc traffic record is found by its own workload string
    and precision, and nothing else matches."""
    bench = _bench()
    head = bench.measured_traffic("dam break z-up 128^3 x 27 ppc, 1000:1, MG-PCG tol 1e-6", "fp32")
    assert head is not None and "k_rbgs_full level-0 sweep (128^3)" in head["kernels"]
    ad = bench.measured_traffic(
        "adaptive dam break y-up 128^3 MSBG n0=16 levels 0-2, 8 ppc, 1000:1, MG-PCG tol 1e-4", "fp32")
    m = ad["kernels"]["k_msbg_rbgs_pairs level-0 half-sweep"]
    # copy-calibrated bytes are at least the algorithmic 32 B/cell
    assert m["bytes_per_cell"] >= bench.msbg_sweep_bytes(1) and abs(m["bytes"] - m["bytes_per_cell"] * m["cells"]) < 1.0
    assert bench.measured_traffic("dam break z-up 128^3 x 27 ppc, 1000:1, MG-PCG tol 1e-6", "bf16") is None
    assert bench.measured_traffic("no such workload", "fp32") is None


def test_nested_dissection_is_a_permutation():
    bench = _bench()
    for m in (1, 5, 8, 13):
        perm = bench.nested_dissection(m)
        assert np.array_equal(np.sort(perm), np.arange(m ** 3))


def test_extraction_direct_solve_reproduces_the_full_solution():
    """The box system with Dirichlet data from the full grid's exact
    solution has that solution as its own: an independent scipy solve of the
    whole singular system (first unknown pinned, mean removed) restricted to
    the box, at a 1000:1 coefficient jump.  A dropped boundary term, a sign
    error or a transposed face array fails it."""
    from test_oracle import direct_solve
    bench = _bench()
    rng = np.random.default_rng(5)
    res = (12, 11, 13)
    bx, by, bz = random_beta(res, rng, contrast=1000.0)
    b = rng.standard_normal(res)
    b -= b.mean()
    A = assemble(res, bx, by, bz)
    p = direct_solve(A, b).reshape(res)
    for lo, m in ((3, 6), (0, 5), (6, 5)):   # interior, touching the low and the high walls
        pd, secs = bench.extraction_direct_solve(bx, by, bz, b, p, lo, m)
        box = (slice(lo, lo + m),) * 3
        assert np.abs(pd - p[box]).max() < 1e-9 * np.abs(p).max(), (lo, m)
        assert secs >= 0.0
    # with wrong Dirichlet data the box solution moves
    pd, _ = bench.extraction_direct_solve(bx, by, bz, b, p + 1.0 * (rng.random(res) > 0.5), 3, 6)
    assert np.abs(pd - p[3:9, 3:9, 3:9]).max() > 1e-3
