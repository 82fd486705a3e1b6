"""FP64 numpy restatement of the MSBG Poisson stack: block grid, 2:1
operator, aggregation multigrid and flexible PCG.

TEST INFRASTRUCTURE ONLY, like ``ref_np.py``: imported by ``tests/`` and
nowhere in the product.  Each function cites the reference file:line it
restates (paths relative to the reference repository root).  Operators are
assembled once per grid as scipy sparse matrices from the per-cell face
walk, then applied as matrix products.

The red-black half-sweep updates every cell of one colour from the values
before the half-sweep (weighted Jacobi within a colour).  The reference
launches it in place; with its finest-coordinate colouring two same-colour
cells can touch across a resolution jump and then read each other's old or
new value depending on timing.  The restatement (and the HIP path) take the
old value, one of the outcomes the reference allows.

Pinning (``tests/test_oracle_msbg.py``): 2:1 grading; every finest
coordinate resolves to one cell and a level-l cell to 8^l of them; cell
counts; the assembled operator is symmetric with zero row sums and a positive
diagonal; with one level it is the 7-point Laplacian times h; a linear field
is reproduced exactly across 2:1 jumps while x^2 and x^3 leave the
h-independent per-volume truncation errors 0.5 and about 0.8 the reference
tests derive; a manufactured cosine solution is recovered; restriction is the
transpose of prolongation; the multigrid-preconditioned solver reaches the
tolerance in far fewer iterations than Jacobi; the adaptive smoother reduces
to the plain one with every block kept active (theta below every residual)
and to warm-up plus one sweep with none kept (theta above), re-activates
exactly the face neighbours of a block above theta, and as a preconditioner
reaches the plain V-cycle's solution.
"""

from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

INACTIVE = -1


def morton3(x, y, z):
    """Bit interleave of three integers below 2^21 (msbg.py:159-171)."""
    x, y, z = (np.asarray(v, np.int64) for v in (x, y, z))
    out = np.zeros(x.shape, np.int64)
    for bit in range(21):
        out |= ((x >> bit) & 1) << (3 * bit)
        out |= ((y >> bit) & 1) << (3 * bit + 1)
        out |= ((z >> bit) & 1) << (3 * bit + 2)
    return out


class Grid:
    """msbg.py:174-296 (host topology) and msbg.py:81-153 (addressing)."""

    def __init__(self, nb, n0, h0, max_level=None):
        self.nb = tuple(int(v) for v in nb)
        self.n0 = int(n0)
        self.sh0 = int(round(math.log2(self.n0)))
        self.h0 = float(h0)
        self.max_level = self.sh0 if max_level is None else int(max_level)
        self.nblk = int(np.prod(self.nb))

    def grade(self, level):
        """msbg.py:201-221: min-plus relaxation to a 2:1 graded map."""
        lv = np.asarray(level, np.int64).reshape(self.nb)
        active = lv >= 0
        w = np.where(active, lv, self.max_level + 1)
        while True:
            nxt = w.copy()
            for axis in range(3):
                for shift in (-1, 1):
                    nxt = np.minimum(nxt, np.roll(w, shift, axis=axis) + 1)
            if np.array_equal(nxt, w):
                break
            w = nxt
        return np.where(active, np.minimum(w, self.max_level), INACTIVE).ravel()

    def build(self, level, grade=True):
        """msbg.py:224-265."""
        level = np.asarray(level, np.int64).ravel()
        if grade:
            level = self.grade(level)
        self.level = level
        bx, by, bz = np.meshgrid(*[np.arange(n) for n in self.nb], indexing="ij")
        key = morton3(bx.ravel(), by.ravel(), bz.ravel())
        act = np.flatnonzero(level >= 0)
        act = act[np.argsort(key[act], kind="stable")]
        self.blk = act
        self.blev = level[act]
        per = (self.n0 >> self.blev) ** 3
        self.coff = np.concatenate(([0], np.cumsum(per)))[:-1]
        self.ncells = int(per.sum())
        self.slot = np.full(self.nblk, -1, np.int64)
        self.slot[act] = np.arange(len(act))
        self.cslot = np.repeat(np.arange(len(act)), per)
        return self.ncells

    def locate(self, f):
        """msbg.py:81-108: finest coordinate -> (addr, level), (-1, -1) outside."""
        fx, fy, fz = f
        if min(fx, fy, fz) < 0:
            return -1, -1
        b = [fx >> self.sh0, fy >> self.sh0, fz >> self.sh0]
        if b[0] >= self.nb[0] or b[1] >= self.nb[1] or b[2] >= self.nb[2]:
            return -1, -1
        s = self.slot[(b[0] * self.nb[1] + b[1]) * self.nb[2] + b[2]]
        if s < 0:
            return -1, -1
        lv = int(self.blev[s])
        n = self.n0 >> lv
        m = self.n0 - 1
        i, j, k = (fx & m) >> lv, (fy & m) >> lv, (fz & m) >> lv
        return int(self.coff[s] + (i * n + j) * n + k), lv

    def origin(self, addr):
        """msbg.py:111-141: finest coordinate of the low corner, and level."""
        s = self.cslot[addr]
        lv = int(self.blev[s])
        n = self.n0 >> lv
        r = addr - self.coff[s]
        i, r = divmod(int(r), n * n)
        j, k = divmod(r, n)
        b = int(self.blk[s])
        bz = b % self.nb[2]
        by = (b // self.nb[2]) % self.nb[1]
        bx = b // (self.nb[2] * self.nb[1])
        return ((bx << self.sh0) + (i << lv), (by << self.sh0) + (j << lv), (bz << self.sh0) + (k << lv)), lv

    def geometry(self):
        """Cell centres and volumes (msbg.py:144-153, mr_operator.py:215-220)."""
        c = np.zeros((self.ncells, 3))
        v = np.zeros(self.ncells)
        for a in range(self.ncells):
            o, lv = self.origin(a)
            ext = 1 << lv
            c[a] = (np.array(o) + 0.5 * ext) * self.h0
            v[a] = (self.h0 * ext) ** 3
        return c, v

    def neighbours(self, a):
        """The face walk of apply_row (mr_operator.py:74-160): a list of
        (neighbour, coupling factor area / dist, index of the face coefficient)."""
        o, la = self.origin(a)
        ext = 1 << la
        hA = self.h0 * ext
        out = []
        for axis in range(3):
            e = [int(axis == c) for c in range(3)]
            for side in range(2):
                f = [o[c] - e[c] for c in range(3)] if side == 0 else [o[c] + e[c] * ext for c in range(3)]
                nb, lb = self.locate(f)
                if nb < 0:
                    continue
                hB = self.h0 * (1 << lb)
                if lb >= la:
                    g = min(hA, hB) ** 2 / (0.5 * (hA + hB))
                    out.append((nb, g, (a if side == 0 else nb) * 3 + axis))
                else:
                    sub = 1 << lb
                    g = hB * hB / (0.5 * (hA + hB))
                    a1, a2 = (axis + 1) % 3, (axis + 2) % 3
                    for u in range(2):
                        for v in range(2):
                            q = list(f)
                            q[a1] += u * sub
                            q[a2] += v * sub
                            nc, _ = self.locate(q)
                            if nc >= 0:
                                out.append((nc, g, (a if side == 0 else nc) * 3 + axis))
        return out


def assemble(grid, bface):
    """Sparse A with (A p)_a = sum_f T_f (p_a - p_b), T = beta area / dist."""
    rows, cols, vals = [], [], []
    bface = np.asarray(bface, np.float64)
    for a in range(grid.ncells):
        for nb, g, fi in grid.neighbours(a):
            t = bface[fi] * g
            rows += [a, a]
            cols += [a, nb]
            vals += [t, -t]
    return sp.csr_matrix((vals, (rows, cols)), shape=(grid.ncells, grid.ncells))


def colours(grid):
    """Finest-coordinate checkerboard (mr_operator.py:205-207)."""
    out = np.zeros(grid.ncells, np.int64)
    for a in range(grid.ncells):
        o, lv = grid.origin(a)
        out[a] = ((o[0] >> lv) + (o[1] >> lv) + (o[2] >> lv)) & 1
    return out


def rbgs_half_sweep(A, diag, col, p, b, colour, omega):
    """mr_operator.py:193-212, every cell of the colour from the old values."""
    upd = (col == colour) & (diag > 0)
    r = b - A @ p
    out = p.copy()
    out[upd] = p[upd] + omega * r[upd] / diag[upd]
    return out


def coarsen_beta(cg, fg, bfine):
    """mr_multigrid.py:52-89: coarse face = mean of the 2x2 fine faces it covers."""
    out = np.zeros(cg.ncells * 3)
    for a in range(cg.ncells):
        o, lv = cg.origin(a)
        half = (1 << lv) >> 1
        for axis in range(3):
            a1, a2 = (axis + 1) % 3, (axis + 2) % 3
            acc, cnt = 0.0, 0
            for u in range(2):
                for v in range(2):
                    q = list(o)
                    q[a1] += u * half
                    q[a2] += v * half
                    nc, _ = fg.locate(q)
                    if nc >= 0:
                        acc += bfine[nc * 3 + axis]
                        cnt += 1
            out[a * 3 + axis] = acc / cnt if cnt else 0.0
    return out


def prolongation(fg, cg):
    """mr_multigrid.py:102-109 as a sparse (n_fine x n_coarse) injection;
    restriction (mr_multigrid.py:92-99) is its transpose."""
    rows = np.arange(fg.ncells)
    cols = np.array([cg.locate(fg.origin(a)[0])[0] for a in range(fg.ncells)])
    return sp.csr_matrix((np.ones(fg.ncells), (rows, cols)), shape=(fg.ncells, cg.ncells))


def block_neighbours(grid):
    """Face-neighbour slots of every active block (adaptive_relax.py:146-160)."""
    out = []
    for s in range(len(grid.blk)):
        b = int(grid.blk[s])
        bz = b % grid.nb[2]
        by = (b // grid.nb[2]) % grid.nb[1]
        bx = b // (grid.nb[2] * grid.nb[1])
        nbrs = []
        for ax in range(3):
            for sg in (-1, 1):
                q = [bx, by, bz]
                q[ax] += sg
                if all(0 <= q[c] < grid.nb[c] for c in range(3)):
                    ns = grid.slot[(q[0] * grid.nb[1] + q[1]) * grid.nb[2] + q[2]]
                    if ns >= 0:
                        nbrs.append(int(ns))
        out.append(nbrs)
    return out


def adaptive_smooth(grid, A, diag, col, p, b, omega, n_iter_max=12, theta_sigma=2.0, warmup=2, inner=1,
                    nbrs=None):
    """Algorithm 2 as adaptive_relax.py:163-241 runs it: warm-up sweeps over
    all blocks, then sweeps over the active set; per-block effective residual
    max |delta| diag; theta from the log-histogram of the first iteration's
    positive residuals; blocks above theta and their face neighbours stay
    active.  -> (p, theta, active fractions)."""
    nbrs = block_neighbours(grid) if nbrs is None else nbrs
    ns = len(grid.blk)
    active = np.ones(ns, bool)

    def sweep(p, reverse):
        delta = np.zeros(grid.ncells)
        for _ in range(inner):
            for c in ((1, 0) if reverse else (0, 1)):
                upd = (col == c) & (diag > 0) & active[grid.cslot]
                r = b - A @ p
                step = omega * r / np.where(diag > 0, diag, 1.0)
                out = p.copy()
                out[upd] = p[upd] + step[upd]
                delta[upd] = np.abs(step[upd])
                p = out
        return p, delta

    for w in range(warmup):
        p, _ = sweep(p, w % 2 == 1)
    theta, frac = None, []
    for it in range(n_iter_max):
        frac.append(active.mean())
        p, delta = sweep(p, it % 2 == 1)
        res = np.zeros(ns)
        np.maximum.at(res, grid.cslot, np.where(active[grid.cslot], delta * diag, 0.0))
        if theta is None:
            pos = res[res > 0]
            if pos.size == 0:
                break
            lg = np.log(pos)
            theta = float(np.exp(lg.mean() + theta_sigma * lg.std()))
        nxt = np.zeros(ns, bool)
        for s_ in np.flatnonzero(res > theta):
            nxt[s_] = True
            nxt[nbrs[s_]] = True
        active = nxt
    return p, theta, frac


class Multigrid:
    """mr_multigrid.py:139-246 with the red-black or the adaptive smoother."""

    def __init__(self, grid, omega=6.0 / 7.0, pre=3, post=3, coarsest_sweeps=80, min_cells=512, smoother="rbgs",
                 n_iter_max=12, theta_sigma=2.0):
        self.omega, self.pre, self.post, self.coarsest = omega, pre, post, coarsest_sweeps
        self.smoother, self.n_iter_max, self.theta_sigma = smoother, n_iter_max, theta_sigma
        self.grids = [grid]
        base = grid.level
        cap = grid.sh0
        L = 1
        while True:
            g = Grid(grid.nb, grid.n0, grid.h0, max_level=cap)
            g.build(np.where(base >= 0, np.minimum(base + L, cap), -1), grade=False)
            if g.ncells == self.grids[-1].ncells:
                break
            self.grids.append(g)
            if g.ncells <= min_cells:
                break
            L += 1
        self.P = [prolongation(self.grids[i], self.grids[i + 1]) for i in range(len(self.grids) - 1)]
        self.col = [colours(g) for g in self.grids]
        self.nbrs = [block_neighbours(g) for g in self.grids] if smoother == "adaptive" else None

    def build_hierarchy(self, bface0):
        """mr_multigrid.py:187-198."""
        self.bface = [np.asarray(bface0, np.float64)]
        for i in range(1, len(self.grids)):
            self.bface.append(coarsen_beta(self.grids[i], self.grids[i - 1], self.bface[-1]))
        self.A = [assemble(g, bf) for g, bf in zip(self.grids, self.bface)]
        self.diag = [np.asarray(A.diagonal()) for A in self.A]

    def _smooth(self, i, p, b, n, reverse=False):
        if self.smoother == "adaptive" and i < len(self.grids) - 1:
            return adaptive_smooth(self.grids[i], self.A[i], self.diag[i], self.col[i], p, b, self.omega,
                                   self.n_iter_max, self.theta_sigma, nbrs=self.nbrs[i])[0]
        for _ in range(n):
            for c in ((1, 0) if reverse else (0, 1)):
                p = rbgs_half_sweep(self.A[i], self.diag[i], self.col[i], p, b, c, self.omega)
        return p

    def _vcycle(self, i, b):
        """mr_multigrid.py:213-235."""
        n = self.grids[i].ncells
        if i == len(self.grids) - 1:
            p = self._smooth(i, np.zeros(n), b, self.coarsest)
            return p - p.mean()
        p = self._smooth(i, np.zeros(n), b, self.pre)
        r = b - self.A[i] @ p
        pc = self._vcycle(i + 1, self.P[i].T @ r)
        p = p + self.P[i] @ pc
        return self._smooth(i, p, b, self.post, reverse=True)

    def precondition(self, r):
        """mr_multigrid.py:237-243."""
        z = self._vcycle(0, np.asarray(r, np.float64))
        return z - z.mean()


def solve(grid, bface, rhs, tol=1e-4, max_iter=2000, preconditioner="mg", **mg_kwargs):
    """mr_solver.py:118-165: flexible Polak-Ribiere PCG -> (p, it, rel)."""
    A = assemble(grid, bface)
    diag = np.asarray(A.diagonal())
    if preconditioner in ("mg", "mg-adaptive"):
        if preconditioner == "mg-adaptive":
            mg_kwargs.setdefault("smoother", "adaptive")
        mg = Multigrid(grid, **mg_kwargs)
        mg.build_hierarchy(bface)
        pre = mg.precondition
    else:
        def pre(r):
            z = np.where(diag > 0, r / np.where(diag > 0, diag, 1.0), 0.0)
            return z - z.mean()
    b = np.asarray(rhs, np.float64) - np.mean(rhs)
    bnorm = np.abs(b).max()
    x = np.zeros(grid.ncells)
    if bnorm == 0.0:
        return x, 0, 0.0
    r = b.copy()
    z = pre(r)
    q = z.copy()
    rz = r @ z
    it, rel = 0, 1.0
    for it in range(1, max_iter + 1):
        Aq = A @ q
        qAq = q @ Aq
        if qAq <= 0.0:
            break
        alpha = rz / qAq
        x += alpha * q
        r_prev = r.copy()
        r -= alpha * Aq
        rel = np.abs(r).max() / bnorm
        if rel < tol:
            break
        z = pre(r)
        rz_new = r @ z
        if rz == 0.0 or rz_new <= 0.0:
            break
        beta = max((rz_new - r_prev @ z) / rz, 0.0)
        q = z + beta * q
        rz = rz_new
    return x - x.mean(), it, rel


def checker_levels(nb, max_level, pattern="slab"):
    """The refinement maps of the reference's operator tests
    (tests/test_msbg.py:20-31)."""
    bx, by, bz = np.meshgrid(*[np.arange(n) for n in nb], indexing="ij")
    if pattern == "slab":
        lv = np.where(bx < nb[0] // 2, 0, 1)
    elif pattern == "ball":
        c = np.array(nb) / 2.0
        d = np.sqrt((bx - c[0]) ** 2 + (by - c[1]) ** 2 + (bz - c[2]) ** 2)
        lv = np.clip((d / max(nb) * 4).astype(int), 0, max_level)
    else:
        lv = (bx + by + bz) % 2
    return lv.ravel().astype(np.int32)
