"""Shared builders and independent checkers for the tests.

Nothing here calls the product's kernels: tiny hand-made worlds, numpy restatements of the
neighbour predicate and of the map order (checked against the oracle), and comparison helpers.
"""
import numpy as np
from numpy.lib import recfunctions

from alien_amd.abi import CELL_DTYPE, MOTION_COLLISION, MOTION_FLUID, PARTICLE_DTYPE, default_params


def cells(n):
    c = np.zeros(n, dtype=CELL_DTYPE)
    c["id"] = np.arange(1, n + 1, dtype=np.uint64)
    c["energy"] = 100.0
    c["stiffness"] = 1.0
    c["maxConnections"] = 6
    c["connections"]["cellIndex"] = -1
    return c


def particles(n, first_id=10_000):
    p = np.zeros(n, dtype=PARTICLE_DTYPE)
    p["id"] = np.arange(first_id, first_id + n, dtype=np.uint64)
    return p


def bond(c, a, b, rest, angle_a=360.0, angle_b=360.0):
    """Adds the bond a<->b at the end of both bond lists (angles given per side)."""
    for x, y, ang in ((a, b, angle_a), (b, a, angle_b)):
        k = int(c["numConnections"][x])
        c["connections"][x, k]["cellIndex"] = y
        c["connections"][x, k]["distance"] = rest
        c["connections"][x, k]["angleFromPrevious"] = ang
        c["numConnections"][x] = k + 1


def params(motion=MOTION_COLLISION, **kw):
    return default_params(motionType=motion, **kw)


def f32_len(dx, dy):
    dx = np.float32(dx)
    dy = np.float32(dy)
    return np.sqrt(np.float32(dx * dx) + np.float32(dy * dy), dtype=np.float32)


def torus_delta32(a, b, size):
    """a - b on the torus in FP32 for wrapped coordinates: the period is subtracted from the
    operand beyond it first (exact), then the difference is rounded once."""
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    s = np.float32(size)
    d = (a - b).astype(np.float32)
    half = s / np.float32(2)
    out = np.where(d > half, (a - s) - b, d)
    out = np.where(d < -half, a - (b - s), out)
    return out.astype(np.float32)


def brute_force_neighbors(c, size, p):
    """All (i, j) with j a neighbour of i, by the predicate of the neighbour pass, O(N^2).

    Fluid: |torus(p_i - p_j)| <= 2h.  Collision: |p_j - p_i| <= r with the uncorrected difference
    (CellProcessor.cuh:296-301).  (The detached rule never applies here: cells set through the
    boundary are never detached.)  The radius is at most one bin edge, so every such pair lies in
    the 3x3 stencil of bins."""
    W, H = size
    pos = c["pos"].astype(np.float32)
    n = len(c)
    out = {}
    for i in range(n):
        if p.motionType == MOTION_FLUID:
            dx = torus_delta32(pos[i, 0], pos[:, 0], W)
            dy = torus_delta32(pos[i, 1], pos[:, 1], H)
            d = np.sqrt(dx * dx + dy * dy, dtype=np.float32)
            ok = d <= np.float32(p.smoothingLength) * np.float32(2)
        else:
            dx = (pos[:, 0] - pos[i, 0]).astype(np.float32)
            dy = (pos[:, 1] - pos[i, 1]).astype(np.float32)
            d = np.sqrt(dx * dx + dy * dy, dtype=np.float32)
            ok = d <= np.float32(p.cellMaxCollisionDistance)
        ok[i] = False
        out[int(c["id"][i])] = set(int(x) for x in c["id"][ok])
    return out


def bin_geometry(size, p):
    """Bin edge and counts: edge = max(2, ceil(radius)), the last bin absorbs the remainder."""
    import math

    r = p.smoothingLength * 2.0 if p.motionType == MOTION_FLUID else p.cellMaxCollisionDistance
    s = max(2, int(math.ceil(np.float32(r))))
    return s, size[0] // s, size[1] // s


def bins_of(pos, size, p):
    s, nbx, nby = bin_geometry(size, p)
    sx = np.floor(pos[:, 0]).astype(np.int64) % size[0]
    sy = np.floor(pos[:, 1]).astype(np.int64) % size[1]
    bx = np.minimum(sx // s, nbx - 1)
    by = np.minimum(sy // s, nby - 1)
    return by * nbx + bx


def angle_sums(c):
    """Per bonded cell, the sum of angleFromPrevious over its bonds."""
    n = c["numConnections"].astype(np.int64)
    ang = c["connections"]["angleFromPrevious"].astype(np.float64)
    mask = np.arange(6)[None, :] < n[:, None]
    return (ang * mask).sum(axis=1)[n > 0]


def bonds_symmetric(c):
    """Every bond a->b has the matching b->a with the same rest length."""
    n = c["numConnections"]
    for a in range(len(c)):
        for k in range(int(n[a])):
            b = int(c["connections"][a, k]["cellIndex"])
            rest = c["connections"][a, k]["distance"]
            back = [m for m in range(int(n[b])) if int(c["connections"][b, m]["cellIndex"]) == a]
            if len(back) != 1 or c["connections"][b, back[0]]["distance"] != rest:
                return False
    return True


def normwise_rel(a, b):
    """max |a - b| / max |b| over all components."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.abs(b).max()
    if scale == 0:
        return float(np.abs(a - b).max()) if a.size else 0.0
    return float(np.abs(a - b).max() / scale)


def torus_diff(a, b, size):
    """Componentwise torus difference of two position arrays (float64)."""
    d = np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64)
    for k, s in enumerate(size):
        d[:, k] = (d[:, k] + s / 2.0) % s - s / 2.0
    return d


def packed_bytes(a):
    """The record bytes without the alignment padding (which a copy may leave uninitialised)."""
    return recfunctions.repack_fields(np.ascontiguousarray(a)).tobytes()
