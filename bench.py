"""PF-FLIP step benchmark on one MI355X.

Workload: z-up dam break, 128^3 MAC grid, 27 particles per cell on a 3^3
seeded lattice with +-0.25 sub-cell jitter (56.6M particles), liquid
x < 0.4, z < 0.6, 1000:1 density contrast, FP32 particles, MG-PCG to a 1e-6
relative (inf-norm) residual, droplet stage off.  A step is one full
``Simulator.step()``: CFL dt, binning, P2G, eq. 7 / beta / gravity / RHS,
MG-PCG, projection, extrapolation, G2P + RK3 advection.

Prints one JSON line: particle-steps/s over the timed steps, the roofline of
the dominant kernel (the level-0 smoother sweep) and of P2G and G2P
(algorithmic bytes / measured average time vs 8 TB/s), the FP32 vs
BF16 MG-PCG time to 1e-6 on the last step's captured system (with each
one's pressure error against an FP64 reference solution), per-stage
milliseconds, the differentiable 64^3 dam break (adjoint vs central finite
differences), and the CPU baseline (the C++/OpenMP FP32 restatement of the
reference step, oracle/cpu_step.cpp, on the same workload: two steps on
every host thread and one step on one thread).
"""

from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def p2g_bytes(n_particles, n_cells):
    """Algorithmic bytes of one P2G: x, v (12 B each) and the mass (4 B) per
    particle; face mass x3, momentum x3 and cell mass (4 B each) per cell."""
    return 28 * n_particles + 28 * n_cells


def g2p_bytes(n_particles, n_cells):
    """Algorithmic bytes of the fused G2P + RK3 + reflection pass, which also
    leaves the next step's speed bounds: read x, v, phase and mass (32 B) and
    write x, v (24 B) per particle; read vel, dvel and vel_liq (3 x 3 faces
    x 4 B) per cell."""
    return 56 * n_particles + 36 * n_cells


def smoother_bytes(n_cells, s_p=4, s_beta=4, s_b=4):
    """Algorithmic bytes of one red-black half-sweep: every p and every face
    coefficient read once, b read and p written for half the cells."""
    return n_cells * (s_p + 3 * s_beta + s_b / 2 + s_p / 2)


def fused_sweep_bytes(n_cells, s_p=4, s_beta=4, s_b=4):
    """Algorithmic bytes of one fused full red-black sweep: p, the three face
    arrays and b read once, p written once (24 B/cell in FP32)."""
    return n_cells * (2 * s_p + 3 * s_beta + s_b)


def roofline(name, nbytes, ms):
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": name, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS, "traffic": None, "bytes_per_launch": nbytes, "avg_ms": ms}


def stream_copy_gbs(torch, nbytes=1 << 30, reps=10):
    """Achievable HBM bandwidth on this GPU: a device-to-device copy of 1 GiB
    (read + write bytes / time, HIP events)."""
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del x, y
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def measured_traffic(workload, precision):
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    (copy-calibrated) for this workload: the record of profiles/traffic*.json
    whose workload and solver precision match, or None."""
    pdir = os.path.join(ROOT, "profiles")
    try:
        names = sorted(n for n in os.listdir(pdir) if n.startswith("traffic") and n.endswith(".json"))
    except OSError:
        return None
    for n in names:
        try:
            with open(os.path.join(pdir, n)) as f:
                t = json.load(f)
        except (OSError, ValueError):
            continue
        if t.get("workload") == workload and t.get("solver_precision") == precision:
            return t
    return None


def cpu_baseline(pf, n=128, ppc=27, steps_parallel=2, steps_serial=1):
    """The C++/OpenMP restatement of the reference step (oracle/cpu_step.cpp:
    FP32, per-particle scatter with float atomics, the same MG-PCG), timed on
    this host on the benchmark workload itself (the 128^3 x 27 ppc z-up dam
    break, MG-PCG to 1e-6): ``steps_parallel`` steps with every available
    thread and ``steps_serial`` with one, from the seeded state."""
    import subprocess
    from oracle import cpu_step as cs
    from oracle import ref_np as O
    if not os.path.exists(cs.LIB):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    cfg = O.OracleConfig(res=(n, n, n), dx=1.0 / n, ppc=ppc, gravity=(0.0, 0.0, -9.81), pcg_tol=1e-6,
                         pcg_max_iter=400, droplets=False)
    pos, vel, pt = pf.scenes.dam_break(cfg, frac=(0.4, 1.0, 0.6), jitter=0.5)
    n_p = len(pos)
    out = {}
    all_threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))
    for label, threads, steps in (("parallel", all_threads, steps_parallel), ("serial", 1, steps_serial)):
        used = cs.set_threads(threads)
        sim = cs.CpuStepSim(cfg, O.derived_densities(cfg))
        sim.set_particles(pos, vel, pt)
        t0 = time.perf_counter()
        recs = [sim.step() for _ in range(steps)]
        el = time.perf_counter() - t0
        del sim
        out[label] = {"value": n_p * steps / el, "unit": "particle-steps/s", "cores": used, "kind": "port",
                      "sample": f"oracle/cpu_step.cpp (C++/OpenMP FP32 restatement of the reference step), "
                                f"dam break z-up {n}^3 x {ppc} ppc ({n_p} particles), the first {steps} "
                                f"step(s) from the seeded state in {el:.1f} s on {used} thread(s), "
                                f"MG-PCG iterations {[r['pcg_iter'] for r in recs]}"}
        print(f"[bench] cpu baseline {label}: {out[label]['sample']}", file=sys.stderr, flush=True)
    res = dict(out["parallel"])
    res["serial"] = out["serial"]
    return res


def fp64_apply_A(bx, by, bz, p):
    """A p in FP64 (the 7-point operator of pfflip/poisson.py:54-88) for
    tensors of any device: diag*p - sum_f beta_f p_nbr over interior faces,
    with diag summed from the faces in FP64 (compute_diag, poisson.py:40-51),
    so that every row sums to zero exactly as in the operator being solved."""
    diag = bx[:-1] + bx[1:] + by[:, :-1] + by[:, 1:] + bz[:, :, :-1] + bz[:, :, 1:]
    Ap = diag * p
    Ap[1:] -= bx[1:-1] * p[:-1]
    Ap[:-1] -= bx[1:-1] * p[1:]
    Ap[:, 1:] -= by[:, 1:-1] * p[:, :-1]
    Ap[:, :-1] -= by[:, 1:-1] * p[:, 1:]
    Ap[:, :, 1:] -= bz[:, :, 1:-1] * p[:, :, :-1]
    Ap[:, :, :-1] -= bz[:, :, 1:-1] * p[:, :, 1:]
    return Ap


def fp64_reference_pressure(solver, rhs, torch, tol=1e-12, max_refine=10):
    """FP64 solution of A p = b - mean(b) by iterative refinement: the
    residual is formed in FP64 and each correction comes from the FP32
    MG-PCG.  Returns (p, rel) with rel = max|b - A p| / max|b| in FP64."""
    L0 = solver.levels[0]
    bx, by, bz = (t.double() for t in (L0.bx, L0.by, L0.bz))
    b0 = rhs.double() - rhs.double().mean()
    bmax = float(b0.abs().max())
    p = torch.zeros_like(b0)
    e = torch.zeros_like(rhs)
    rel = 1.0
    for _ in range(max_refine):
        r = b0 - fp64_apply_A(bx, by, bz, p)
        rel = float(r.abs().max()) / bmax
        if rel < tol:
            break
        solver.solve(r.float(), e, tol=1e-6, max_iter=400)
        p += e.double()
        p -= p.mean()
    return p, rel


def nested_dissection(m):
    """Elimination order of an m^3 box of cells by nested dissection
    (George, "Nested dissection of a regular finite element mesh", SIAM J.
    Numer. Anal. 10, 1973): both halves of a box, then the separator plane
    between them, recursively.  For the 7-point operator it keeps the direct
    factor's fill near O(N^(4/3)), far below a general-purpose ordering."""
    import numpy as np
    idx = np.arange(m ** 3).reshape(m, m, m)
    order = []

    def rec(box):
        sz = [hi - lo for lo, hi in box]
        if max(sz) <= 4:
            order.append(idx[tuple(slice(lo, hi) for lo, hi in box)].ravel())
            return
        ax = int(np.argmax(sz))
        lo, hi = box[ax]
        mid = (lo + hi) // 2
        for half in ((lo, mid), (mid + 1, hi)):
            b = list(box)
            b[ax] = half
            rec(b)
        b = list(box)
        b[ax] = (mid, mid + 1)
        order.append(idx[tuple(slice(lo2, hi2) for lo2, hi2 in b)].ravel())

    rec([(0, m)] * 3)
    return np.concatenate(order)


def extraction_direct_solve(bx, by, bz, b0, p_ref, lo, m):
    """FP64 scipy direct solve of the pressure system restricted to the box
    of cells [lo, lo+m)^3, with the FP64 reference solution as Dirichlet
    data outside it: each box row keeps A's diagonal (the sum of its six
    face coefficients, compute_diag, pfflip/poisson.py:40-51), couples to
    box neighbours by -beta (poisson.py:54-88), and moves beta * p_ref of a
    neighbour outside the box to the right-hand side.  That removes the
    constant null space, so the solve is a plain sparse LU (SuperLU,
    nested-dissection order).  Arguments are FP64 numpy arrays of the whole
    grid; returns (p on the box, seconds)."""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    sl = (slice(lo, lo + m),) * 3
    N = m ** 3
    idx = np.arange(N).reshape(m, m, m)
    diag = (bx[lo:lo + m + 1, sl[1], sl[2]][:-1] + bx[lo:lo + m + 1, sl[1], sl[2]][1:]
            + by[sl[0], lo:lo + m + 1, sl[2]][:, :-1] + by[sl[0], lo:lo + m + 1, sl[2]][:, 1:]
            + bz[sl[0], sl[1], lo:lo + m + 1][:, :, :-1] + bz[sl[0], sl[1], lo:lo + m + 1][:, :, 1:])
    rhs = b0[sl].copy()
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [diag.ravel()]
    for ax, f in enumerate((bx, by, bz)):
        # faces of this axis from the box's low face to its high face
        fs = [sl[0], sl[1], sl[2]]
        fs[ax] = slice(lo, lo + m + 1)
        fb = f[tuple(fs)]
        inner = [slice(None)] * 3
        inner[ax] = slice(1, m)
        lo_c = [slice(None)] * 3
        lo_c[ax] = slice(0, m - 1)
        hi_c = [slice(None)] * 3
        hi_c[ax] = slice(1, m)
        w = fb[tuple(inner)].ravel()
        a, c = idx[tuple(lo_c)].ravel(), idx[tuple(hi_c)].ravel()
        rows += [a, c]
        cols += [c, a]
        vals += [-w, -w]
        # faces on the box boundary: the outside neighbour is Dirichlet data
        for side in (0, m):
            fsel = [slice(None)] * 3
            fsel[ax] = side
            csel = [slice(None)] * 3
            csel[ax] = 0 if side == 0 else m - 1
            osel = list(sl)
            o = lo - 1 if side == 0 else lo + m
            if 0 <= o < p_ref.shape[ax]:
                osel[ax] = o
                rhs[tuple(csel)] += fb[tuple(fsel)] * p_ref[tuple(osel)]
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    perm = nested_dissection(m)
    t0 = time.perf_counter()
    lu = spla.splu(A[perm][:, perm].tocsc(), permc_spec="NATURAL", options={"SymmetricMode": True})
    y = lu.solve(rhs.ravel()[perm])
    secs = time.perf_counter() - t0
    x = np.empty(N)
    x[perm] = y
    return x.reshape(m, m, m), secs


def solve_compare(pf, torch, sim, cfg, tol, extraction=64):
    """Time FP32, BF16 and FP16 MG-PCG on the same captured system (beta,
    rhs) and give each one's pressure error against an FP64 reference
    solution (iterative refinement over the whole grid) and, on a central
    box of `extraction`^3 cells, against an FP64 scipy direct solve with the
    reference as Dirichlet data (0: no extraction)."""
    out = {}
    beta = [b.clone() for b in sim.beta]
    rhs = sim.rhs.clone()
    p_ref = None
    sols = {}
    for prec in ("fp32", "bf16", "fp16"):
        s = pf.MGPCG(sim.res, cfg.device, omega=cfg.mg_omega, pre=cfg.mg_pre_sweeps,
                     post=cfg.mg_post_sweeps, coarsest=cfg.mg_coarsest_sweeps,
                     min_dim=cfg.mg_min_dim, precision=prec)
        p = torch.zeros_like(rhs)
        s.build_hierarchy(*beta)
        s.solve(rhs, p, tol=tol, max_iter=400)          # warm-up
        reps, times = 3, []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            it, rel = s.solve(rhs, p, tol=tol, max_iter=400)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        if p_ref is None:
            p_ref, ref_rel = fp64_reference_pressure(s, rhs, torch)
            out["fp64_reference"] = {"method": "FP64-residual iterative refinement",
                                     "rel_residual": ref_rel}
        d = p.double() - p_ref
        d -= d.mean()
        p_err = float(d.abs().max() / p_ref.abs().max())
        out[prec] = {"ms": 1e3 * min(times), "iters": it, "rel": rel, "p_err_vs_fp64": p_err}
        sols[prec] = p.double().cpu().numpy()
        del s, d
    if extraction:
        n = sim.res[0]
        lo = n // 2 - extraction // 2
        b0 = rhs.double() - rhs.double().mean()
        f64 = [t.double().cpu().numpy() for t in beta]
        pr = p_ref.cpu().numpy()
        pd, secs = extraction_direct_solve(*f64, b0.cpu().numpy(), pr, lo, extraction)
        box = (slice(lo, lo + extraction),) * 3
        scale = float(abs(pr).max())
        out["fp64_direct_extraction"] = {
            "box": [lo, lo + extraction], "cells": extraction ** 3,
            "method": "scipy.sparse.linalg.splu (SuperLU, nested-dissection order), FP64, Dirichlet "
                      "data from the FP64 reference outside the box",
            "seconds": secs,
            "reference_vs_direct": float(abs(pr[box] - pd).max()) / scale}
        for prec, sol in sols.items():
            out[prec]["p_err_vs_direct"] = float(abs(sol[box] - pd).max()) / scale
    return out


def adjoint_check(pf, torch, n=64, ppc=8, n_steps=16, n_dirs=8, eps=1e-3, tol=1e-8,
                  eps_sweep=(2e-3, 5e-4)):
    """Differentiable dam break: 64^3, 8 ppc, 16 steps, FP32.  Control: a
    seeded Gaussian-random initial liquid velocity field (sigma 0.1 after a
    4-cell Gaussian smoothing), sampled trilinearly at the liquid particles.
    Loss: squared distance of the final liquid centre of mass to
    (0.6, 0.5, 0.2).  Hand-written adjoint vs central finite differences
    (eps = 1e-3) along n_dirs seeded smooth directions, dt held fixed, MG-PCG
    (forward and adjoint) to `tol` so that the solver's stopping point does
    not move the loss by more than the finite-difference signal.  The
    tags-fixed differences are also taken at each step of `eps_sweep`: the
    step map is only piecewise smooth (clamps, trilinear kinks at cell
    faces), so a central difference carries an error that shrinks with eps,
    which the sweep shows apart from the adjoint's own."""
    import numpy as np
    import scipy.ndimage as ndi
    import torch.nn.functional as F

    cfg = pf.Config(res=(n, n, n), dx=1.0 / n, ppc=ppc, gravity=(0.0, 0.0, -9.81), droplets=False,
                    pcg_tol=tol, pcg_max_iter=1000)
    pos, vel, pt = pf.scenes.dam_break(cfg, frac=(0.4, 1.0, 0.6), jitter=0.5)
    sim = pf.Simulator(cfg)
    sim.set_particles(pos, 0.0 * vel, pt)
    rng = np.random.default_rng(cfg.seed)

    def smooth_field():
        f = np.stack([ndi.gaussian_filter(rng.standard_normal((n, n, n)), sigma=4.0, mode="wrap")
                      for _ in range(3)])
        return f

    u = smooth_field()
    u0 = torch.from_numpy((0.1 / u.std() * u).astype(np.float32)).cuda()
    x0 = sim.px.clone()
    N = x0.shape[0]
    liq = (sim.ptype == 1).float()[:, None]
    L = float(n * cfg.dx)
    grid = (2.0 * x0 / L - 1.0)[:, [2, 1, 0]].reshape(1, N, 1, 1, 3)

    def sample(field):
        out = F.grid_sample(field.unsqueeze(0), grid, mode="bilinear", padding_mode="border",
                            align_corners=False)
        return (out.reshape(3, N).t() * liq).contiguous()

    target = torch.tensor([0.6, 0.5, 0.2], dtype=torch.float64, device="cuda")
    lmask = sim.ptype == 1

    def loss_of(xT):
        return ((xT[lmask].double().mean(0) - target) ** 2).sum()

    # base run: the time steps and liquid tags of every step
    v0 = sample(u0)
    sim.px.copy_(x0)
    sim.pv.copy_(v0)
    dts, tags = [], []
    for _ in range(n_steps):
        dts.append(sim.step()["dt"])
        tags.append(sim.is_liquid.clone())

    def run(v, frozen):
        sim.px.copy_(x0)
        sim.pv.copy_(v)
        for k in range(n_steps):
            sim.step(dts[k], tags=tags[k] if frozen else None)
        return float(loss_of(sim.px))

    adj = pf.StepAdjoint(sim, tol=tol, max_iter=1000)
    uu = u0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xT, _ = pf.rollout(sim, x0, sample(uu), n_steps, dts=dts, adjoint=adj)
    loss = loss_of(xT)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    iters = [r["pcg_iter"] for r in sim.stats[-n_steps:]]
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    g = uu.grad.detach().double()
    rows = []
    for _ in range(n_dirs):
        dvec = smooth_field()
        d = torch.from_numpy((dvec / np.abs(dvec).max()).astype(np.float32)).cuda()
        ad = float((g * d.double()).sum())
        with torch.no_grad():
            row = {"adjoint": ad}
            for name, frozen in (("fd_tags_fixed", True), ("fd_free", False)):
                fd = (run(sample(u0 + eps * d), frozen) - run(sample(u0 - eps * d), frozen)) / (2.0 * eps)
                row[name] = fd
                row["rel_" + name] = abs(fd - ad) / max(abs(fd), 1e-30)
            for e in eps_sweep:
                row[f"fd_tags_fixed@{e:g}"] = (run(sample(u0 + e * d), True)
                                               - run(sample(u0 - e * d), True)) / (2.0 * e)
        rows.append(row)
    def vec_err(name):
        fd = np.array([r[name] for r in rows])
        ad = np.array([r["adjoint"] for r in rows])
        return float(np.linalg.norm(fd - ad) / np.linalg.norm(fd))

    return {"workload": f"differentiable dam break {n}^3 x {ppc} ppc, {n_steps} steps, FP32",
            "particles": N, "loss": float(loss.detach()), "forward_ms": 1e3 * (t1 - t0),
            "backward_ms": 1e3 * (t2 - t1), "eps": eps, "pcg_tol": tol, "pcg_iters": iters,
            "rel_err_tags_fixed": vec_err("fd_tags_fixed"), "rel_err_free": vec_err("fd_free"),
            "rel_err_tags_fixed_by_eps": {f"{e:g}": vec_err(f"fd_tags_fixed@{e:g}") if e != eps
                                          else vec_err("fd_tags_fixed")
                                          for e in sorted(set(eps_sweep) | {eps}, reverse=True)},
            "max_rel_err_tags_fixed": max(r["rel_fd_tags_fixed"] for r in rows),
            "max_rel_err_free": max(r["rel_fd_free"] for r in rows),
            "note": "the adjoint holds the liquid tags fixed; fd_tags_fixed differences the "
                    "simulation along that branch (tags replayed from the base run), fd_free "
                    "lets the eq. (7) threshold flip tags between the two runs",
            "directions": rows}


def p2g_kernel_ms(torch, sim, reps=5):
    """Average milliseconds of the P2G launches alone (k_max_count, k_p2g,
    k_p2g_gather through the current bins), HIP events on the stream they
    run on around back-to-back replays: the host enqueues faster than the
    GPU runs them, so no host time falls inside.  The replays rewrite the
    same grid sums from the same particles."""
    sim._splat()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        sim._splat()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def g2p_kernel_ms(pf, torch, sim, dt, reps=5):
    """Average milliseconds of one fused G2P + RK3 + reflection launch
    through the bins (k_g2p_tiled, leaving the speed bounds as in the step),
    HIP events around the launch alone.  Each replay starts from the same
    particle state (restored by a copy outside the event pair) and the state
    after the step is put back afterwards."""
    import ctypes
    lib = pf._lib
    x1, v1 = sim.px.clone(), sim.pv.clone()
    x0, v0 = getattr(sim, "_bench_x0"), getattr(sim, "_bench_v0")
    bins = (lib.ptr(sim.perm), lib.ptr(sim.cell_start), lib.ptr(sim._n_active))
    a = sim._g2p_args(lib.GA_G2P | lib.GA_ADVECT | lib.GA_REFLECT, dt)
    total = 0.0
    for r in range(reps + 1):
        sim.px.copy_(x0)
        sim.pv.copy_(v0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        lib.call("pf_g2p_advect_binned", ctypes.byref(a), *bins, lib.ptr(sim.pmass),
                 lib.ptr(sim._bounds), lib.stream())
        e1.record()
        torch.cuda.synchronize()
        if r > 0:
            total += e0.elapsed_time(e1)
    sim.px.copy_(x1)
    sim.pv.copy_(v1)
    sim._bounds_token = None
    torch.cuda.synchronize()
    return total / reps


def particle_kernel_rooflines(pf, torch, sim, n_p, n_c, dt):
    """P2G and G2P rooflines from kernel-only event timings (see
    p2g_kernel_ms / g2p_kernel_ms)."""
    return [roofline("p2g (k_max_count + k_p2g + k_p2g_gather)", p2g_bytes(n_p, n_c), p2g_kernel_ms(torch, sim)),
            roofline("k_g2p_tiled (G2P + RK3 + reflect + bounds)", g2p_bytes(n_p, n_c),
                     g2p_kernel_ms(pf, torch, sim, dt))]


def stage_rooflines(pf, torch, sim, n_p, n_c, s16, dt):
    """Per-kernel GB/s (algorithmic bytes / measured time) of the step's
    dominant kernels: the level-0 sweep (HIP events via pf_mg_time_smoother),
    P2G and G2P (kernel-only HIP events around replays)."""
    out = []
    if sim.solver.small_from > 0:
        ms = pf.poisson.time_smoother(sim.solver, 0, reps=10)
        out.append(roofline(f"k_rbgs_full level-0 sweep ({sim.res[0]}^3)",
                            fused_sweep_bytes(n_c, s16, s16), ms))
    return out + particle_kernel_rooflines(pf, torch, sim, n_p, n_c, dt)


def timed_steps(torch, sim, steps, label="step"):
    """Run and time ``steps`` steps: per-step wall times from the step
    records, the whole region bracketed by device synchronisations.  After
    it, untimed, the particle state is kept and one more step is run, so the
    kernel replays (P2G, G2P) have a step's bins and grid fields together
    with the particle state that step started from; its dt is
    ``sim._bench_dt``."""
    import gc
    # a full collection of the interpreter's ~2e5 tracked objects takes tens
    # of milliseconds and lands inside whatever step allocates the object
    # that triggers it; freeze the set-up's objects so collections stay cheap
    gc.collect()
    gc.freeze()
    sim.stage_ms = {}
    recs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step_ms = []
    for k in range(steps):
        ts = time.perf_counter()
        recs.append(sim.step())        # ends with a device read, so it has finished
        step_ms.append(1e3 * (time.perf_counter() - ts))
        print(f"[bench] {label} {k}: pcg_iter={recs[-1]['pcg_iter']} dt={recs[-1]['dt']:.4g} "
              f"wall={step_ms[-1]:.2f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stages = {k: v / steps for k, v in sim.stage_ms.items()}
    sim.stage_ms = None
    # the replays' step (the bench's own work, outside the timed region)
    sim._bench_x0, sim._bench_v0 = sim.px.clone(), sim.pv.clone()
    sim._bench_dt = sim.step()["dt"]
    torch.cuda.synchronize()
    return recs, elapsed, stages, step_ms


def step_time_summary(step_ms):
    import statistics
    return {"ms_per_step_median": statistics.median(step_ms),
            "ms_per_step_min": min(step_ms), "ms_per_step_max": max(step_ms),
            "step_ms": [round(t, 3) for t in step_ms]}


def run_rt128(pf, torch, args):
    """Rayleigh-Taylor instability at 128^3 x 27 ppc, 1000:1, heavy over light
    at half height with a seeded 8-mode interface perturbation (amplitude
    0.02, wavenumbers 1-4), 60 steps, mixed-precision MG-PCG (FP16 iterates
    and coefficients; BF16's 8-bit mantissa can stall this system's FPCG just
    above 1e-6) to 1e-6.
    alpha_phi = 0.5 and CFL = 0.5 as the reference's own Rayleigh-Taylor
    run uses (scripts/run.py:42-43).  Reports per-kernel GB/s and the
    V-cycles per step (one per FPCG iteration plus the first)."""
    n, steps = 128, args.rt_steps
    cfg = pf.Config(res=(n, n, n), dx=1.0 / n, ppc=27, gravity=(0.0, 0.0, -9.81), droplets=False,
                    pcg_tol=1e-6, pcg_max_iter=400, solver_precision="fp16", alpha_phi=0.5, cfl=0.5)
    pos, vel, pt = pf.scenes.rayleigh_taylor_modes(cfg)
    sim = pf.Simulator(cfg)
    sim.set_particles(pos, vel, pt)
    n_p, n_c = len(pos), n ** 3
    del pos, vel, pt
    recs, elapsed, stages, step_ms = timed_steps(torch, sim, steps, "rt step")
    kernels = stage_rooflines(pf, torch, sim, n_p, n_c, 2, sim._bench_dt)
    return {"metric": "Rayleigh-Taylor 128^3 x 27 ppc particle-steps/s, 60 steps, FP16 MG-PCG",
            "value": n_p * steps / elapsed, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
            "ms_per_step": 1e3 * elapsed / steps, **step_time_summary(step_ms),
            "higher_is_better": True, "dtype": "f32",
            "data": "synthetic (seeded jittered lattice)",
            "config": {"workload": f"rayleigh-taylor z-up {n}^3 x 27 ppc, 8 modes amplitude 0.02, "
                                   f"1000:1, alpha_phi 0.5, CFL 0.5, FP16 MG-PCG tol 1e-6",
                       "particles": n_p},
            "vcycles_per_step": [r["pcg_iter"] + 1 for r in recs],
            "roofline_kernels": kernels, "stages_ms": stages}


def run_surge256(pf, torch, args):
    """Surge on a column at 256^3 x 27 ppc: liquid slab of depth 0.3 moving at
    +2.0 along x in its top 0.1, a solid column of width 0.1 at x = 0.7 cut
    from the grid (particles inside dropped), mixed-precision MG-PCG (BF16)
    to 1e-6; timed forward steps, then one adjoint step from a checkpoint.
    Reports the HBM footprint, particle-steps/s and roofline fractions."""
    n = 256
    cfg = pf.Config(res=(n, n, n), dx=1.0 / n, ppc=27, gravity=(0.0, 0.0, -9.81), droplets=False,
                    pcg_tol=1e-6, pcg_max_iter=400, solver_precision="bf16")
    t_seed = time.perf_counter()
    pos, vel, pt, solid = pf.scenes.surge_column(cfg)
    print(f"[bench] seeded {len(pos)} particles in {time.perf_counter() - t_seed:.1f} s",
          file=sys.stderr, flush=True)
    torch.cuda.reset_peak_memory_stats()
    sim = pf.Simulator(cfg)
    sim.set_solids(solid)
    sim.set_particles(pos, vel, pt)
    n_p, n_c = len(pos), n ** 3
    del pos, vel, pt
    for _ in range(args.warmup):
        sim.step()
    recs, elapsed, stages, step_ms = timed_steps(torch, sim, args.steps, "surge step")
    kernels = stage_rooflines(pf, torch, sim, n_p, n_c, 2, sim._bench_dt)
    del sim._bench_x0, sim._bench_v0
    fwd_peak = torch.cuda.max_memory_allocated()
    # one adjoint step with checkpointing: loss = liquid centre of mass
    x0, v0 = sim.px.clone(), sim.pv.clone().requires_grad_(True)
    lmask = sim.ptype == 1
    adj = pf.StepAdjoint(sim)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    xT, _ = pf.rollout(sim, x0, v0, 1, adjoint=adj)
    loss = (xT[lmask].double().mean(0) ** 2).sum()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    grad_norm = float(v0.grad.double().norm())
    return {"metric": "surge 256^3 x 27 ppc particle-steps/s, BF16 MG-PCG, + one adjoint step",
            "value": n_p * args.steps / elapsed, "unit": "particle-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps,
            **step_time_summary(step_ms),
            "higher_is_better": True, "dtype": "f32", "data": "synthetic (seeded jittered lattice)",
            "config": {"workload": f"surge slab 0.3 (+2.0 in top 0.1) on a solid column, {n}^3 x 27 ppc, "
                                   f"1000:1, BF16 MG-PCG tol 1e-6", "particles": n_p,
                       "solid_cells": int(solid.sum())},
            "hbm_bytes_forward_peak": fwd_peak, "hbm_bytes_with_adjoint_peak": torch.cuda.max_memory_allocated(),
            "pcg_iters": [r["pcg_iter"] for r in recs], "roofline_kernels": kernels, "stages_ms": stages,
            "adjoint_step": {"forward_ms": 1e3 * (t2 - t1), "backward_ms": 1e3 * (t3 - t2),
                             "adjoint_pcg": list(adj.last_solve), "grad_norm": grad_norm}}


def msbg_sweep_bytes(n_cells):
    """Algorithmic bytes of one MSBG red-black half-sweep: p read and written,
    b, the diagonal, the three owned face coefficients and the cell's block
    slot read once per cell (32 B/cell); block tables are a few KB."""
    return 32 * n_cells


def run_adaptive128(pf, torch, args):
    """The adaptive path: a two-phase dam break (y up, liquid x < 0.4,
    y < 0.6) on a 128^3 MSBG (8^3 blocks of 16^3 cells, levels 0-2), 8 ppc
    seeded at level 0, alpha_phi 0.25 (the measured-safe noise margin at 8
    ppc), beta_bw 0.5, MG-PCG to the reference's 1e-4; particles refine and
    coarsen every step.  Warm-up steps let the resampling settle."""
    n = 128
    cfg = pf.Config(res=(n, n, n), dx=1.0 / n, ppc=8, alpha_phi=0.25, droplets=False, pcg_tol=1e-4,
                    pcg_max_iter=400)
    sim = pf.AdaptiveSimulator(cfg, (8, 8, 8), n0=16, max_level=2, beta_bw=0.5)
    pos, vel, pt = pf.scenes.dam_break(cfg)
    sim.set_particles(pos, vel, pt)
    n0 = len(pos)
    del pos, vel, pt
    for w in range(args.warmup):
        r = sim.step()
        print(f"[bench] warmup {w}: particles {r['particles']} cells {r['cells']} cg {r['cg_iter']}",
              file=sys.stderr, flush=True)
    recs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        recs.append(sim.step())
        print(f"[bench] step {k}: particles {recs[-1]['particles']} cells {recs[-1]['cells']} "
              f"cg {recs[-1]['cg_iter']} wall {recs[-1]['wall_s'] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # particles entering each timed step (the last record holds the count after it)
    n_in = [sim.stats[-args.steps - 1]["particles"] if len(sim.stats) > args.steps else n0] + \
        [r["particles"] for r in recs[:-1]]
    # dominant kernel: the level-0 half-sweep of the MSBG V-cycle, timed with
    # HIP events on the stream the kernels run on
    lib = pf._lib
    lv = sim.solver.mg.levels[0]
    g = lv.grid
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record()
    for i in range(reps):
        lib.call("pf_msbg_rbgs", g.ref(), lib.ptr(lv.p), lib.ptr(lv.p2), lib.ptr(lv.b), lib.ptr(lv.bface),
                 lib.ptr(lv.diag), i & 1, 6.0 / 7.0, lib.stream())
    e1.record()
    torch.cuda.synchronize()
    sweep_ms = e0.elapsed_time(e1) / reps
    rl = roofline(f"k_msbg_rbgs_pairs level-0 half-sweep ({g.ncells} cells)", msbg_sweep_bytes(g.ncells), sweep_ms)
    workload = "adaptive dam break y-up 128^3 MSBG n0=16 levels 0-2, 8 ppc, 1000:1, MG-PCG tol 1e-4"
    tr = measured_traffic(workload, "fp32")
    if tr is not None:
        # HBM bytes per cell of the level-0 half-sweep, measured on this
        # workload's grid; the sweep streams its arrays once, so the figure
        # scales with the cell count
        m = tr["kernels"]["k_msbg_rbgs_pairs level-0 half-sweep"]
        rl["traffic"] = m["bytes_per_cell"] * g.ncells
        rl["traffic_source"] = tr["source"]
    return {
        "metric": "adaptive PF-FLIP step particle-steps/s, 128^3 MSBG (levels 0-2)",
        "value": sum(n_in) / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded jittered lattice, both phases)",
        "config": {"workload": workload, "particles_seeded": n0,
                   "particles": [r["particles"] for r in recs], "cells": [r["cells"] for r in recs],
                   "compression": recs[-1]["compression"], "levels": recs[-1]["levels"],
                   "parallelism": "single GPU"},
        "pcg_iters": [r["cg_iter"] for r in recs],
        "roofline": rl,
        "cpu_baseline": None if args.skip_cpu else adaptive_cpu_baseline(pf),
    }


def adaptive_cpu_baseline(pf, steps=3):
    """The FP64 numpy restatement of the adaptive step (oracle/ref_mr_sim.py)
    on one core, on a bounded sample of the same workload: the dam break on
    a 16^3 MSBG (2^3 blocks of 8^3, levels 0-2), 8 ppc, both phases."""
    import numpy as np
    from oracle import ref_mr_sim as R

    n = 16
    cfg = pf.Config(res=(n, n, n), dx=1.0 / n, ppc=8, alpha_phi=0.25, droplets=False, pcg_tol=1e-4,
                    pcg_max_iter=400)
    pos, vel, pt = pf.scenes.dam_break(cfg)
    st = R.State(cfg, (2, 2, 2), 8, 2, 0.5, pos, vel, pt)
    t0 = time.perf_counter()
    done = 0
    for _ in range(steps):
        done += len(st.x)
        R.step(st)
    el = time.perf_counter() - t0
    return {"value": done / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle/ref_mr_sim.py FP64 adaptive step, dam break 16^3 MSBG x 8 ppc "
                      f"({len(pos)} particles), {steps} steps in {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--res", type=int, default=128)
    ap.add_argument("--ppc", type=int, default=27)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16", "fp16"])
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-solve-compare", action="store_true")
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--workload", default="dambreak128",
                    choices=["dambreak128", "rt128", "surge256", "adaptive128"],
                    help="dambreak128: the headline benchmark (default); rt128 / surge256 / adaptive128: the "
                         "Rayleigh-Taylor, large surge and adaptive (MSBG) configurations (their own JSON line)")
    ap.add_argument("--rt-steps", type=int, default=60)
    ap.add_argument("--extraction", type=int, default=64,
                    help="edge of the central box whose FP64 scipy direct solve checks the solvers' "
                         "pressure (0: none; the 64^3 box takes about a minute of host time)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 512 MiB device-to-device copy three times (the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "only single-GPU runs are supported: the step runs on one "
                                   "GPU (a multi-GPU version would shard the grid into x-slabs "
                                   "with halo exchange)", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch
    if args.copy_calibration:
        nbytes = 512 << 20
        x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read": nbytes, "bytes_written": nbytes,
                                               "launches": 3}}), flush=True)
        return
    pf = importlib.import_module("implementation-of-adaptive-phase-field-flip_amd")
    if args.workload == "rt128":
        print(json.dumps(run_rt128(pf, torch, args)), flush=True)
        return
    if args.workload == "surge256":
        print(json.dumps(run_surge256(pf, torch, args)), flush=True)
        return
    if args.workload == "adaptive128":
        print(json.dumps(run_adaptive128(pf, torch, args)), flush=True)
        return

    n = args.res
    cfg = pf.Config(res=(n, n, n), dx=1.0 / n, ppc=args.ppc, gravity=(0.0, 0.0, -9.81),
                    droplets=False, pcg_tol=1e-6, pcg_max_iter=400,
                    solver_precision=args.precision)
    t_seed = time.perf_counter()
    pos, vel, pt = pf.scenes.dam_break(cfg, frac=(0.4, 1.0, 0.6), jitter=0.5)
    print(f"[bench] seeded {len(pos)} particles in {time.perf_counter() - t_seed:.1f} s",
          file=sys.stderr, flush=True)
    sim = pf.Simulator(cfg)
    sim.set_particles(pos, vel, pt)
    n_p, n_c = len(pos), n ** 3
    del pos, vel, pt
    for w in range(args.warmup):
        rec = sim.step()
        print(f"[bench] warmup {w}: {rec}", file=sys.stderr, flush=True)

    recs, elapsed, stages, step_ms = timed_steps(torch, sim, args.steps)

    # the dominant kernel: the level-0 red-black sweep of the V-cycle (one
    # fused kernel per sweep), timed with HIP events on the solver's stream
    # right after the timed steps; P2G and G2P from kernel-only replays
    s16 = 2 if args.precision != "fp32" else 4
    kernels = stage_rooflines(pf, torch, sim, n_p, n_c, s16, sim._bench_dt)
    del sim._bench_x0, sim._bench_v0
    workload = f"dam break z-up {n}^3 x {args.ppc} ppc, 1000:1, MG-PCG tol 1e-6"
    copy_gbs = stream_copy_gbs(torch)
    for k in kernels:
        k["frac_of_copy"] = k["achieved"] / copy_gbs
    tr = measured_traffic(workload, args.precision)
    if tr is not None:
        for k in kernels:
            m = tr["kernels"].get(k["kernel"])
            if m is not None:
                k["traffic"] = m["bytes"]
                k["traffic_source"] = tr["source"]
    result = {
        "metric": "PF-FLIP step particle-steps/s at 128^3/27ppc; MG-PCG ms to 1e-6, mixed vs FP32",
        "value": n_p * args.steps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        **step_time_summary(step_ms),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered lattice)",
        "config": {"workload": workload,
                   "grid": n, "ppc": args.ppc, "particles": n_p, "pcg_tol": 1e-6,
                   "solver_precision": args.precision, "parallelism": "single GPU"},
        "roofline": kernels[0],
        "roofline_kernels": kernels[1:],
        "stream_copy_gbs": copy_gbs,
        "pcg_iters": [r["pcg_iter"] for r in recs],
        "stages_ms": stages,
    }
    if not args.skip_solve_compare:
        result["mgpcg_1e-6"] = solve_compare(pf, torch, sim, cfg, 1e-6, args.extraction)
    if not args.skip_adjoint:
        del sim
        torch.cuda.empty_cache()
        result["adjoint_64"] = adjoint_check(pf, torch)
    if not args.skip_cpu:
        result["cpu_baseline"] = cpu_baseline(pf, n, args.ppc)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
