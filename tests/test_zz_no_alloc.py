"""pf_bin_sort_noalloc (the binning sort on the library's static scratch)
against pf_bin_sort (MI355X), through the C ABI.

Two calls in a row, the second on a smaller grid and particle set, each
compared with pf_bin_sort on the same input: keys, the sorted permutation
of the FLIP particles (stable sort: unique), the per-key ranges and the FLIP count must be equal
bit for bit.  A grid with more binning keys than the scratch holds is
refused with PF_ERR_SIZE before any device work.
"""

import sys

import numpy as np
import pytest

from conftest import gpu_available, load_pkg

pytestmark = pytest.mark.gpu

if gpu_available():
    import torch


def say(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    if not gpu_available():
        pytest.skip("no GPU")
    pkg = load_pkg()
    return sys.modules[pkg.__name__ + "._lib"]


def _particles(n, res, dx, seed):
    """Uniform positions over the box plus a few outside it; a tenth of the
    particles are not FLIP particles (pstate 1 or 2)."""
    rng = np.random.default_rng(seed)
    ext = np.array(res, np.float64) * dx
    px = (rng.random((n, 3)) * ext).astype(np.float32)
    out = rng.choice(n, max(n // 1000, 1), replace=False)
    px[out] = ((rng.random((len(out), 3)) * 1.4 - 0.2) * ext).astype(np.float32)
    st = np.zeros(n, np.int32)
    off = rng.choice(n, n // 10, replace=False)
    st[off] = rng.integers(1, 3, len(off)).astype(np.int32)
    return torch.from_numpy(px).cuda(), torch.from_numpy(st).cuda()


def _bufs(L, n, res):
    nkeys = L.load().pf_bin_nkeys(*res)
    dev = "cuda"
    return dict(keys=torch.empty(n, dtype=torch.int32, device=dev),
                rank=torch.empty(n, dtype=torch.int32, device=dev),
                perm=torch.empty(n, dtype=torch.int32, device=dev),
                start=torch.empty(nkeys + 2, dtype=torch.int32, device=dev),
                nact=torch.empty(1, dtype=torch.int32, device=dev)), nkeys


def _sort(L, name, px, st, res, dx, order, b, stable, scan_tmp=None):
    n = px.shape[0]
    args = [L.ptr(px), L.ptr(st), n, *res, dx, L.ptr(order), L.ptr(b["keys"]), L.ptr(b["rank"]),
            L.ptr(b["perm"]), L.ptr(b["start"])]
    if scan_tmp is not None:
        args.append(L.ptr(scan_tmp))
    args += [L.ptr(b["nact"]), stable, L.stream()]
    L.call(name, *args)


@pytest.mark.parametrize("stable", [1, 0])
def test_noalloc_bin_sort_equals_bin_sort(L, stable):
    lib = L.load()
    assert L.load().pf_bin_nkeys(128, 128, 128) == 2097152
    cases = [((100, 90, 120), 1.0 / 128, 3_000_000, 5, False),
             ((24, 16, 40), 1.0 / 32, 50_000, 6, True)]
    for res, dx, n, seed, with_order in cases:
        say(f"grid {res} n={n} stable={stable} order={'perm' if with_order else 'index'}")
        px, st = _particles(n, res, dx, seed)
        order = None
        if with_order:
            g = torch.Generator().manual_seed(seed)
            order = torch.randperm(n, generator=g).to(torch.int32).cuda()
        ref, nkeys = _bufs(L, n, res)
        got, _ = _bufs(L, n, res)
        for t in got.values():
            t.fill_(-7)
        scan_tmp = torch.empty(lib.pf_bin_scan_tmp_bytes(*res), dtype=torch.uint8, device="cuda")
        _sort(L, "pf_bin_sort", px, st, res, dx, order, ref, stable, scan_tmp)
        _sort(L, "pf_bin_sort_noalloc", px, st, res, dx, order, got, stable)
        torch.cuda.synchronize()
        nact = int((st == 0).sum())
        assert int(got["nact"][0]) == int(ref["nact"][0]) == nact
        assert torch.equal(got["keys"], ref["keys"])
        assert torch.equal(got["start"], ref["start"])
        assert int(got["start"][nkeys]) == nact and int(got["start"][nkeys + 1]) == n
        # the particles outside the bins (pstate != 0) close the permutation
        # in atomic-timing order in both entry points
        assert torch.equal(torch.sort(got["perm"][nact:]).values, torch.sort(ref["perm"][nact:]).values)
        if stable:
            assert torch.equal(got["perm"][:nact], ref["perm"][:nact])
        else:
            # the order within one key follows atomic timing: compare each
            # key's particles as a set (sorting by (key, index) is unique)
            def canon(p):
                p = p.to(torch.int64)
                visit = torch.arange(n, device="cuda") if order is None else order.to(torch.int64)
                key_of = torch.empty(n, dtype=torch.int64, device="cuda")
                key_of[visit] = got["keys"].to(torch.int64) & 0xFFFFFFFF
                act = p[:nact]
                return act[torch.argsort(key_of[act] * n + act)], torch.sort(p[nact:]).values
            a, b = canon(got["perm"]), canon(ref["perm"])
            assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


def test_noalloc_bin_sort_refuses_grids_beyond_its_scratch(L):
    lib = L.load()
    res = (136, 128, 128)
    assert lib.pf_bin_nkeys(*res) > 2097152
    rc = lib.pf_bin_sort_noalloc(None, None, 0, *res, 1.0 / 128, None, None, None, None,
                                 L.ptr(torch.empty(8, dtype=torch.int32, device="cuda")), None, 1, None)
    assert rc == -2
    assert b"exceed the built-in scratch" in lib.pf_last_error()
