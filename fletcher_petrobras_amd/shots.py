"""Multi-shot acquisition setup: shot sub-grids cut from one parent model.

Each shot owns an aperture window of the parent interior; its absorbing zone and halo are
built the reference way around that window (the window's edge values extended outward, then
the random-velocity boundary of boundary.c:7-98 with its own rand() stream, srand(1+shot)).
"""
import numpy as np

from . import BORD, grid_shape, index, random_boundary


def parent_interior(model, absorb, bord=BORD):
    """Interior part (no halo, no absorbing zone) of a full model."""
    b = bord + absorb
    return [m[b:m.shape[0] - b, b:m.shape[1] - b, b:m.shape[2] - b] for m in model]


def shot_model(interior, x0, y0, nx, ny, absorb, seed, bord=BORD):
    """Model arrays of the shot whose interior is interior[:, y0:y0+ny, x0:x0+nx] (full depth)."""
    nz = interior[0].shape[0]
    pad = bord + absorb
    out = []
    for a in interior:
        w = a[:, y0:y0 + ny, x0:x0 + nx]
        out.append(np.ascontiguousarray(np.pad(w, pad, mode="edge"), dtype=np.float32))
    random_boundary(out, (nx, ny, nz), absorb, seed=seed, bord=bord)
    return out


def make_shots(interior, nshots, absorb, nmin, nmax, src_depth, rec_depth, seed=32, bord=BORD):
    """nshots shots with horizontal apertures uniform in [nmin, nmax] placed uniformly in the
    parent; source at the sub-grid centre in x, y and at array depth src_depth; receivers on
    plane rec_depth (-1: none).  Returns dicts for fletcher_petrobras_amd.Batch."""
    rng = np.random.default_rng(seed)
    nz, py, px = interior[0].shape
    shots = []
    for s in range(nshots):
        nx = int(rng.integers(nmin, nmax + 1))
        ny = int(rng.integers(nmin, nmax + 1))
        x0 = int(rng.integers(0, px - nx + 1))
        y0 = int(rng.integers(0, py - ny + 1))
        m = shot_model(interior, x0, y0, nx, ny, absorb, seed=1 + s, bord=bord)
        shape = grid_shape(nx, ny, nz, absorb, bord)
        shots.append({"model": m, "isource": index(shape, shape[2] // 2, shape[1] // 2, src_depth),
                      "rec_iz": rec_depth, "window": (x0, y0, nx, ny), "shape": shape})
    return shots
