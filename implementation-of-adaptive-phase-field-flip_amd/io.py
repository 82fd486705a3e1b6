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
    """Quick-look PNG of the current state, viewed along ``axis``.

    Left: the cell phase field on the grid plane through the domain centre,
    with the phi = 0.5 contour that separates liquid-tagged from air-tagged
    cells.  Right: where the liquid particles are, as the number of liquid
    particles per grid column along the viewing axis, divided by the count a
    full column holds (ppc * cells along the axis): 1 means a column full of
    liquid.  Both panels share the same in-plane axes.
    """
    import matplotlib

    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    phi = _host(sim.phi_cell).astype(np.float32)
    h, v = [a for a in range(3) if a != axis]
    plane = np.take(phi, phi.shape[axis] // 2, axis=axis)      # axes (h, v)
    n = getattr(sim, "n_particles", None)
    px, pt = _host(sim.px), _host(sim.ptype)
    if n is not None:
        px, pt = px[:n], pt[:n]
    liq = px[pt == 1]
    res = phi.shape
    size = sim.cfg.domain_size
    counts, _, _ = np.histogram2d(liq[:, h], liq[:, v], bins=(res[h], res[v]),
                                  range=((0.0, size[h]), (0.0, size[v])))
    fill = counts / float(max(sim.cfg.ppc * res[axis], 1))
    names = "xyz"
    extent = (0.0, size[h], 0.0, size[v])

    fig, (ax_phi, ax_fill) = plt.subplots(1, 2, figsize=(10, 4.4), sharey=True)
    a = ax_phi.imshow(plane.T, origin="lower", extent=extent, cmap="viridis", vmin=0.0, vmax=1.0,
                      interpolation="nearest")
    gh = (np.arange(res[h]) + 0.5) * size[h] / res[h]
    gv = (np.arange(res[v]) + 0.5) * size[v] / res[v]
    if plane.min() < 0.5 < plane.max():
        ax_phi.contour(gh, gv, plane.T, levels=[0.5], colors="w", linewidths=0.8)
    ax_phi.set_title(f"phi, {names[axis]} = mid-plane (white: phi = 0.5)")
    b = ax_fill.imshow(fill.T, origin="lower", extent=extent, cmap="cividis", vmin=0.0,
                       vmax=max(1.0, float(fill.max())), interpolation="nearest")
    ax_fill.set_title(f"liquid particles per column / full column, along {names[axis]}")
    for ax in (ax_phi, ax_fill):
        ax.set_xlabel(names[h])
    ax_phi.set_ylabel(names[v])
    fig.colorbar(a, ax=ax_phi, shrink=0.85)
    fig.colorbar(b, ax=ax_fill, shrink=0.85)
    fig.suptitle(f"step {sim.step_index}, t = {sim.time:.4f}")
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    fig.savefig(path, dpi=dpi)
    plt.close(fig)
