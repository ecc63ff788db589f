"""Snapshot pictures of a grain state: DEMPort's ``snapshot`` / ``plot_p4p`` (utils.jl:80-124).

The reference draws every grain as a sphere of its radius, coloured by id, in a 1200 x 800 3-D
axis with data aspect, azimuth 7.3 rad, elevation 0.189 rad and perspectiveness 0.5 (GLMakie
``meshscatter``, utils.jl:80-94), and saves ``snapshot_<step>.png``.  Here the same view is drawn
with matplotlib's 3-D scatter: each grain is a disc whose diameter is its sphere's diameter at the
axis scale, and perspectiveness p maps to a field of view of p x 90 degrees.  The picture is for
looking at, not for comparison: the projection and shading are matplotlib's, not GLMakie's.

matplotlib is imported only when a picture is drawn; without it these functions raise ImportError.
"""
from __future__ import annotations

import math
import os
from typing import List, Sequence, Union

import numpy as np

FIGSIZE_PX = (1200, 800)   # utils.jl:93
AZIMUTH = 7.3              # rad, utils.jl:90
ELEVATION = 0.189          # rad, utils.jl:91
PERSPECTIVENESS = 0.5      # utils.jl:92
_DPI = 100


def render(k: np.ndarray, r: np.ndarray, pid: np.ndarray, path: str) -> str:
    """Draw grains at positions ``k`` (n x 3) with radii ``r`` coloured by ``pid`` into ``path``."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    k = np.asarray(k, dtype=np.float64).reshape(-1, 3)
    r = np.asarray(r, dtype=np.float64).reshape(-1)
    pid = np.asarray(pid).reshape(-1)
    fig = plt.figure(figsize=(FIGSIZE_PX[0] / _DPI, FIGSIZE_PX[1] / _DPI), dpi=_DPI)
    ax = fig.add_subplot(projection="3d")
    if len(k):
        lo = (k - r[:, None]).min(axis=0)
        hi = (k + r[:, None]).max(axis=0)
    else:
        lo, hi = np.zeros(3), np.ones(3)
    span = np.maximum(hi - lo, 1e-12)
    ax.set_xlim(lo[0], hi[0])
    ax.set_ylim(lo[1], hi[1])
    ax.set_zlim(lo[2], hi[2])
    ax.set_box_aspect(tuple(span))  # aspect = :data
    ax.view_init(elev=math.degrees(ELEVATION), azim=math.degrees(AZIMUTH))
    fov = PERSPECTIVENESS * math.pi / 2
    ax.set_proj_type("persp", focal_length=1.0 / math.tan(fov / 2))
    ax.set_xlabel("x")
    ax.set_ylabel("y")
    ax.set_zlabel("z")
    if len(k):
        # marker area in pt^2: the sphere's diameter scaled by the axis box's width in points
        box_pt = 0.5 * min(FIGSIZE_PX) * 72.0 / _DPI
        d_pt = 2.0 * r / span.max() * box_pt
        ax.scatter(k[:, 0], k[:, 1], k[:, 2], s=np.maximum(d_pt, 0.5) ** 2, c=pid,
                   cmap="viridis", depthshade=True, linewidths=0)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    fig.savefig(path, dpi=_DPI)
    plt.close(fig)
    return path


def snapshot(grains: np.ndarray, step: int, out_dir: str = ".") -> str:
    """utils.jl:96-106: ``snapshot_<step>.png`` of a grain array (GRAIN_DTYPE rows)."""
    return render(grains["k"], grains["r"], grains["id"],
                  os.path.join(out_dir, f"snapshot_{step}.png"))


def plot_p4p(p4p: Union[str, Sequence], out_dir: str = ".") -> List[str]:
    """utils.jl:108-124: one ``snapshot_<i>.png`` per frame (1-based) of a p4p file or of the
    frames ``read_p4p`` returns; radii from the volumes, r = cbrt(V / (4/3 pi))."""
    from .p4io import read_p4p
    records = read_p4p(p4p) if isinstance(p4p, str) else p4p
    out = []
    for i, (_t, g) in enumerate(records, start=1):
        r = np.cbrt(g["V"] / (4.0 / 3.0 * math.pi))
        out.append(render(g["k"], r, g["id"], os.path.join(out_dir, f"snapshot_{i}.png")))
    return out
