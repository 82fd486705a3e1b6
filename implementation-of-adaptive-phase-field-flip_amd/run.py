"""Run a PF-FLIP scene and write frames (drop-in for scripts/run.py).

    python -m implementation-of-adaptive-phase-field-flip_amd.run --scene dam_break --res 64 --frames 120

Each frame advances the simulation to ``(frame + 1) / fps`` with CFL steps,
the last one shortened to land on the frame time (run.py:54-73); then the
frame is written (``io.save_frame``), optionally a quick-look PNG, and at
the end ``stats.json`` with every step's record.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

from . import io, scenes
from .config import Config


def advance_to(sim, target: float, eps: float = 1e-9):
    """Steps until ``sim.time`` reaches ``target`` (run.py:56-61) -> (substeps, last record)."""
    nsub, rec = 0, None
    while sim.time < target - eps:
        dt = min(sim.dt_from_cfl(), target - sim.time)
        rec = sim.step(dt)
        nsub += 1
    return nsub, rec


def make_config(args) -> Config:
    r = args.res
    cfg = Config(res=(r, r, r), dx=1.0 / r, cfl=args.cfl, ppc=args.ppc, device=args.device, dt_max=1.0 / args.fps)
    if args.scene == "rayleigh_taylor":  # paper 8.2
        cfg.alpha_phi, cfg.cfl = 0.5, 0.5
    return cfg


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--scene", default="dam_break", choices=sorted(scenes.SCENES))
    ap.add_argument("--res", type=int, default=64)
    ap.add_argument("--frames", type=int, default=60)
    ap.add_argument("--fps", type=float, default=24.0)
    ap.add_argument("--cfl", type=float, default=1.0)
    ap.add_argument("--ppc", type=int, default=8)
    ap.add_argument("--out", default="out")
    ap.add_argument("--device", default="cuda:0")
    ap.add_argument("--png", action="store_true", help="also write quick-look PNGs")
    ap.add_argument("--npz-stride", type=int, default=4)
    args = ap.parse_args(argv)

    from .sim import Simulator

    cfg = make_config(args)
    sim = Simulator(cfg)
    out = scenes.SCENES[args.scene](cfg)
    pos, vel, ptype = out[:3]
    sim.set_particles(pos, vel, ptype)
    if len(out) > 3:  # scenes with a solid mask
        sim.set_solids(out[3])
    outdir = os.path.join(args.out, f"{args.scene}_{args.res}")
    os.makedirs(outdir, exist_ok=True)
    n_liq = int((ptype == 1).sum())
    print(f"[run] {args.scene} on {cfg.res[0]}x{cfg.res[1]}x{cfg.res[2]}: {len(pos)} particles, "
          f"{n_liq} liquid; writing {args.frames} frames to {outdir}", flush=True)
    frame_dt = 1.0 / args.fps
    t0 = time.perf_counter()
    for frame in range(args.frames):
        nsub, rec = advance_to(sim, (frame + 1) * frame_dt)
        io.save_frame(sim, os.path.join(outdir, f"frame_{frame:04d}.npz"), stride=args.npz_stride)
        if args.png:
            io.render_slices(sim, os.path.join(outdir, f"frame_{frame:04d}.png"))
        if rec is not None:
            print(f"[run] frame {frame + 1}/{args.frames} at t = {sim.time:.4f}: {nsub} steps, "
                  f"last solve {rec['pcg_iter']} iterations, max |div u| {rec['max_div']:.3g}, "
                  f"{time.perf_counter() - t0:.1f} s elapsed", flush=True)
    with open(os.path.join(outdir, "stats.json"), "w") as f:
        json.dump(sim.stats, f)
    print(f"[run] {len(sim.stats)} steps in {time.perf_counter() - t0:.1f} s; records in "
          f"{os.path.join(outdir, 'stats.json')}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
