"""Frame output and quick-look rendering (drop-in for pfflip/io.py).

``save_frame`` writes the reference's compressed ``.npz`` frame schema
(io.py:10-29), the format downstream tools read:

=========  =========  ======================================================
key        dtype      content
=========  =========  ======================================================
``t``      float32    simulation time
``step``   int32      step index
``dx``     float32    cell size
``phi``    float16    cell-centred phase field, shape ``res``
``px``     float16    every ``stride``-th particle position, ``(n, 3)``
``ptype``  int8       their types (1 = liquid)
=========  =========  ======================================================

The arrays come off the device with one copy each; writing is host work.
"""

from __future__ import annotations

import os

import numpy as np


def _host(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def save_frame(sim, path: str, *, particles: bool = True, stride: int = 1):
    """Dump one frame as a compressed ``.npz`` (io.py:10-29)."""
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    data = {
        "t": np.float32(sim.time),
        "step": np.int32(sim.step_index),
        "dx": np.float32(sim.cfg.dx),
        "phi": _host(sim.phi_cell).astype(np.float16),
    }
    if particles:
        n = getattr(sim, "n_particles", None)
        px, pt = _host(sim.px), _host(sim.ptype)
        if n is not None:
            px, pt = px[:n], pt[:n]
        data["px"] = px[::stride].astype(np.float16)
        data["ptype"] = pt[::stride].astype(np.int8)
    np.savez_compressed(path, **data)


def load_frame(path: str) -> dict:
    """Read a frame written by ``save_frame`` (no pickled objects)."""
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def render_slices(sim, path: str, *, axis: int = 2, dpi: int = 110):
    """Quick-look PNG: a phase-field slice and the particles of a slab
    through the domain centre (io.py:32-79)."""
    import matplotlib

    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    phi = _host(sim.phi_cell)
    sl = [slice(None)] * 3
    sl[axis] = phi.shape[axis] // 2
    img = phi[tuple(sl)].T
    n = getattr(sim, "n_particles", None)
    px, pt = _host(sim.px), _host(sim.ptype)
    if n is not None:
        px, pt = px[:n], pt[:n]
    sz = sim.cfg.domain_size
    band = np.abs(px[:, axis] - 0.5 * sz[axis]) < 1.5 * sim.cfg.dx
    other = [a for a in range(3) if a != axis]
    fig, axes = plt.subplots(1, 2, figsize=(11, 5), constrained_layout=True)
    im = axes[0].imshow(img, origin="lower", extent=[0, sz[other[0]], 0, sz[other[1]]], vmin=0, vmax=1,
                        cmap="Blues")
    axes[0].set_title(f"phase field $\\phi$   t = {sim.time:.3f}s")
    fig.colorbar(im, ax=axes[0], fraction=0.046)
    liq, air = band & (pt == 1), band & (pt == 0)
    axes[1].scatter(px[air, other[0]], px[air, other[1]], s=0.3, c="#cccccc", linewidths=0, label="air")
    axes[1].scatter(px[liq, other[0]], px[liq, other[1]], s=0.4, c="#1f6fb4", linewidths=0, label="liquid")
    axes[1].set_xlim(0, sz[other[0]])
    axes[1].set_ylim(0, sz[other[1]])
    axes[1].set_aspect("equal")
    axes[1].set_title(f"particles (slab)   step {sim.step_index}")
    axes[1].legend(loc="upper right", markerscale=12, framealpha=0.9)
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    fig.savefig(path, dpi=dpi)
    plt.close(fig)
