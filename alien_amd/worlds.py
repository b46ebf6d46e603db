"""Seeded synthetic worlds for the physics path (host side, numpy).

The soft bodies follow ALIEN's createHex geometry (source/EngineInterface/DescriptionEditService.cpp:
40-81): hexagonal lattice at spacing 1.0, bonds between lattice neighbours at their generated
distance, maxConnections 6, and angleFromPrevious bookkeeping where a cell's bond angles run in
increasing angleOfVector order and sum to 360 degrees (Descriptions.cpp:273-356).  A patch of n
cells is the first n points of the lattice in spiral order (ring, then polar angle).

Patches are placed on a randomly chosen subset of a square slot grid with jitter, which keeps
bodies at least one unit apart without sequential rejection sampling.

Workloads:
  plain_soft_body      torus 512^2, 800 patches of U{7..37} cells, 10k particles     (collision)
  large_fluid          torus 4096^2, 4M free cells, 500k particles                    (fluid, h=0.65)
  large_soft_body      torus 4096^2, 40k bodies of U{20..100} cells + free cells      (collision)
  batched_worlds       256 worlds, side logU[256,2048], density U(0.02,0.3)           (collision)
"""
import math

import numpy as np

from .abi import CELL_DTYPE, MAX_CELL_BONDS, MOTION_COLLISION, MOTION_FLUID, PARTICLE_DTYPE, default_params

SEED_BASE = 0xA11E0000

_AXIAL_DIRS = ((1, 0), (-1, 0), (0, 1), (0, -1), (1, -1), (-1, 1))
_TEMPLATE_RING = 7


def _hex_template():
    R = _TEMPLATE_RING
    pts = []
    for q in range(-R, R + 1):
        for r in range(-R, R + 1):
            s = -q - r
            ring = max(abs(q), abs(r), abs(s))
            if ring <= R:
                x = q + r / 2.0
                y = r * math.sqrt(3.0) / 2.0
                ang = math.atan2(y, x) % (2 * math.pi)
                pts.append((ring, ang, q, r, x, y))
    pts.sort(key=lambda t: (t[0], t[1]))
    index = {(p[2], p[3]): i for i, p in enumerate(pts)}
    xy = np.array([(p[4], p[5]) for p in pts], dtype=np.float64)
    ring = np.array([p[0] for p in pts], dtype=np.int64)
    adj = np.full((len(pts), 6), -1, dtype=np.int64)
    for i, p in enumerate(pts):
        for d, (dq, dr) in enumerate(_AXIAL_DIRS):
            j = index.get((p[2] + dq, p[3] + dr))
            if j is not None:
                adj[i, d] = j
    return xy, ring, adj


_TPL_XY, _TPL_RING, _TPL_ADJ = _hex_template()


def patch_extent(n):
    """Half-width of the bounding box of the first n spiral lattice points."""
    return float(np.abs(_TPL_XY[:n]).max())


def angle_of_vector(dx, dy):
    """Math::angleOfVector in float64 (Math.cuh:140-155): 0 deg = (0,-1), increasing clockwise
    on screen."""
    dx = np.asarray(dx, dtype=np.float64)
    dy = np.asarray(dy, dtype=np.float64)
    ln = np.sqrt(dx * dx + dy * dy)
    with np.errstate(invalid="ignore", divide="ignore"):
        nvy = np.clip(-dy / ln, -1.0, 1.0)
    a = np.degrees(np.arcsin(nvy))
    out = np.where(dx >= 0, 90.0 - a, a + 270.0)
    return np.where(ln < 1e-5, 0.0, out)


def _torus_delta(d, size):
    return d - size * np.round(d / size)


def _assign_bond_angles(cells, world):
    """Orders each cell's bonds by angleOfVector and fills angleFromPrevious (sum 360)."""
    n = len(cells)
    W, H = world
    k = cells["numConnections"].astype(np.int64)
    tgt = cells["connections"]["cellIndex"].astype(np.int64)
    valid = np.arange(MAX_CELL_BONDS)[None, :] < k[:, None]
    safe = np.where(valid, tgt, np.arange(n)[:, None])
    px = cells["pos"][:, 0].astype(np.float64)
    py = cells["pos"][:, 1].astype(np.float64)
    dx = _torus_delta(px[safe] - px[:, None], W)
    dy = _torus_delta(py[safe] - py[:, None], H)
    ang = np.where(valid, angle_of_vector(dx, dy), np.inf)
    dist = np.sqrt(dx * dx + dy * dy)
    order = np.argsort(ang, axis=1, kind="stable")
    ang_s = np.take_along_axis(ang, order, axis=1)
    tgt_s = np.take_along_axis(safe, order, axis=1)
    dist_s = np.take_along_axis(dist, order, axis=1)
    prev = np.roll(ang_s, 1, axis=1)
    with np.errstate(invalid="ignore"):  # unused slots hold inf - inf; masked below
        gaps = ang_s - prev
    last = np.take_along_axis(ang_s, np.maximum(k - 1, 0)[:, None], axis=1)[:, 0]
    gaps[:, 0] = 360.0 - (last - ang_s[:, 0])
    gaps = np.where(np.arange(MAX_CELL_BONDS)[None, :] < k[:, None], gaps, 0.0)
    conn = cells["connections"]
    conn["cellIndex"] = np.where(valid, tgt_s, 0).astype(np.int32)
    conn["distance"] = np.where(valid, dist_s, 0.0).astype(np.float32)
    conn["angleFromPrevious"] = gaps.astype(np.float32)
    cells["connections"] = conn


def _place_slots(rng, world, num, slot):
    W, H = world
    nsx, nsy = W // slot, H // slot
    total = nsx * nsy
    if num > total:
        num = total
    chosen = rng.choice(total, size=num, replace=False)
    return chosen % nsx, chosen // nsx, num


def soft_bodies(rng, world, sizes, vel_sigma, stiffness=None, omega_sigma=0.0, energy=100.0):
    """Hex patches of the given sizes on a jittered slot grid.  Returns (cells, slot_size,
    occupied slot mask)."""
    W, H = world
    sizes = np.asarray(sizes, dtype=np.int64)
    ext = max(patch_extent(int(sizes.max())), 0.5)
    slot = int(math.ceil(2 * ext + 2.0))
    sx, sy, num = _place_slots(rng, world, len(sizes), slot)
    sizes = sizes[:num]
    jitter = slot - (2 * ext + 1.0)
    cx = sx * slot + ext + 0.5 + rng.uniform(0.0, jitter, size=num)
    cy = sy * slot + ext + 0.5 + rng.uniform(0.0, jitter, size=num)
    vb = rng.normal(0.0, vel_sigma, size=(num, 2))
    om = rng.normal(0.0, omega_sigma, size=num) if omega_sigma > 0 else np.zeros(num)
    stiff = np.ones(num) if stiffness is None else np.asarray(stiffness(num), dtype=np.float64)
    total = int(sizes.sum())
    cells = np.zeros(total, dtype=CELL_DTYPE)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    for n in np.unique(sizes):
        sel = np.nonzero(sizes == n)[0]
        base = starts[sel][:, None] + np.arange(n)[None, :]
        tpl = _TPL_XY[:n]
        x = cx[sel][:, None] + tpl[None, :, 0]
        y = cy[sel][:, None] + tpl[None, :, 1]
        vx = vb[sel, 0][:, None] - om[sel][:, None] * tpl[None, :, 1]
        vy = vb[sel, 1][:, None] + om[sel][:, None] * tpl[None, :, 0]
        flat = base.ravel()
        cells["pos"][flat, 0] = np.mod(x, W).ravel().astype(np.float32)
        cells["pos"][flat, 1] = np.mod(y, H).ravel().astype(np.float32)
        cells["vel"][flat, 0] = vx.ravel().astype(np.float32)
        cells["vel"][flat, 1] = vy.ravel().astype(np.float32)
        cells["stiffness"][flat] = np.repeat(stiff[sel], n).astype(np.float32)
        adj = _TPL_ADJ[:n]
        ok = (adj >= 0) & (adj < n)
        cnt = ok.sum(axis=1)
        # compact the template neighbours into the first cnt slots
        order = np.argsort(~ok, axis=1, kind="stable")
        adjc = np.take_along_axis(np.where(ok, adj, 0), order, axis=1)
        tgt = starts[sel][:, None, None] + adjc[None, :, :]
        conn = cells["connections"][flat]
        conn["cellIndex"] = tgt.reshape(-1, 6).astype(np.int32)
        cells["connections"][flat] = conn
        cells["numConnections"][flat] = np.tile(cnt, len(sel)).astype(np.uint8)
    cells["energy"] = np.float32(energy)
    cells["maxConnections"] = 6
    _assign_bond_angles(cells, world)
    occupied = np.zeros((H // slot, W // slot), dtype=bool)
    occupied[sy[:num], sx[:num]] = True
    return cells, slot, occupied


def free_cells(rng, world, count, vel_sigma, energy=100.0, max_connections=0, exclude=None):
    W, H = world
    pos = np.empty((0, 2))
    need = count
    while need > 0:
        cand = rng.uniform(0.0, 1.0, size=(int(need * 1.3) + 16, 2)) * np.array([W, H])
        if exclude is not None:
            slot, occ = exclude
            ix = np.minimum((cand[:, 0] // slot).astype(np.int64), occ.shape[1] - 1)
            iy = np.minimum((cand[:, 1] // slot).astype(np.int64), occ.shape[0] - 1)
            inside = (cand[:, 0] < occ.shape[1] * slot) & (cand[:, 1] < occ.shape[0] * slot)
            cand = cand[~(inside & occ[iy, ix])]
        pos = np.concatenate([pos, cand[:need]])
        need = count - len(pos)
    cells = np.zeros(count, dtype=CELL_DTYPE)
    cells["pos"] = pos.astype(np.float32)
    cells["vel"] = rng.normal(0.0, vel_sigma, size=(count, 2)).astype(np.float32)
    cells["energy"] = np.float32(energy)
    cells["stiffness"] = np.float32(1.0)
    cells["maxConnections"] = max_connections
    return cells


def particles(rng, world, count, vel_sigma):
    W, H = world
    p = np.zeros(count, dtype=PARTICLE_DTYPE)
    p["pos"] = (rng.uniform(0.0, 1.0, size=(count, 2)) * np.array([W, H])).astype(np.float32)
    p["vel"] = rng.normal(0.0, vel_sigma, size=(count, 2)).astype(np.float32)
    p["energy"] = rng.uniform(1.0, 5.0, size=count).astype(np.float32)
    return p


def concat_cells(parts):
    """Concatenates cell arrays, shifting bond indices."""
    out = np.concatenate(parts)
    off = 0
    for part in parts:
        n = len(part)
        sl = slice(off, off + n)
        k = out["numConnections"][sl].astype(np.int64)
        valid = np.arange(MAX_CELL_BONDS)[None, :] < k[:, None]
        ci = out["connections"]["cellIndex"][sl]
        conn = out["connections"][sl]
        conn["cellIndex"] = np.where(valid, ci + off, 0)
        out["connections"][sl] = conn
        off += n
    return out


def assign_ids(cells, parts, first=1):
    cells["id"] = np.arange(first, first + len(cells), dtype=np.uint64)
    parts["id"] = np.arange(first + len(cells), first + len(cells) + len(parts), dtype=np.uint64)


class World:
    """A generated world: geometry, parameters and transfer arrays."""

    def __init__(self, name, size, params, cells, parts, steps, seed):
        self.name = name
        self.size = size
        self.params = params
        self.cells = cells
        self.particles = parts
        self.steps = steps
        self.seed = seed

    @property
    def num_cells(self):
        return len(self.cells)

    def __repr__(self):
        return (f"World({self.name}, {self.size[0]}x{self.size[1]}, cells={len(self.cells)}, "
                f"particles={len(self.particles)}, steps={self.steps})")


def plain_soft_body(seed=SEED_BASE + 0, size=(512, 512), patches=800, num_particles=10_000, steps=100):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(7, 38, size=patches)
    cells, _, _ = soft_bodies(rng, size, sizes, vel_sigma=0.05)
    parts = particles(rng, size, num_particles, vel_sigma=0.05)
    assign_ids(cells, parts)
    params = default_params(motionType=MOTION_COLLISION, cellMaxCollisionDistance=1.3, friction=0.001)
    return World("plain_soft_body", size, params, cells, parts, steps, seed)


def large_fluid(seed=SEED_BASE + 1, size=(4096, 4096), num_cells=4_000_000, num_particles=500_000, steps=1000):
    rng = np.random.default_rng(seed)
    cells = free_cells(rng, size, num_cells, vel_sigma=0.1)
    parts = particles(rng, size, num_particles, vel_sigma=0.1)
    assign_ids(cells, parts)
    params = default_params(motionType=MOTION_FLUID, smoothingLength=0.65, pressureStrength=0.1,
                            viscosityStrength=0.1)
    return World("large_fluid", size, params, cells, parts, steps, seed)


def large_soft_body(seed=SEED_BASE + 2, size=(4096, 4096), bodies=40_000, free_density=0.06, steps=1000,
                    num_particles=0):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(20, 101, size=bodies)
    body, slot, occ = soft_bodies(rng, size, sizes, vel_sigma=0.2, omega_sigma=0.01,
                                  stiffness=lambda n: rng.uniform(0.5, 1.0, size=n))
    free = free_cells(rng, size, int(round(free_density * size[0] * size[1])), vel_sigma=0.2,
                      exclude=(slot, occ))
    cells = concat_cells([body, free])
    parts = particles(rng, size, num_particles, vel_sigma=0.1)
    assign_ids(cells, parts)
    params = default_params(motionType=MOTION_COLLISION, cellMaxCollisionDistance=1.3)
    return World("large_soft_body", size, params, cells, parts, steps, seed)


def batch_world(rng, side, density, steps, seed=0):
    size = (int(side), int(side))
    target = density * side * side
    sizes = []
    total = 0
    while total < target:
        n = int(rng.integers(7, 38))
        sizes.append(n)
        total += n
    cells, _, _ = soft_bodies(rng, size, sizes, vel_sigma=0.05)
    num_particles = int(round(10_000 * side * side / (512.0 * 512.0)))
    parts = particles(rng, size, num_particles, vel_sigma=0.05)
    assign_ids(cells, parts)
    params = default_params(motionType=MOTION_COLLISION, cellMaxCollisionDistance=1.3, friction=0.001)
    return World(f"batch_{side}", size, params, cells, parts, steps, seed)


def batched_world_specs(seed=SEED_BASE + 3, count=256, side_range=(256, 2048), density_range=(0.02, 0.3)):
    """(side, density, child seed) of every world of the batched workload: sides log-uniform,
    densities uniform, each world filled from its own child seed (so worlds can be built in any
    order or in parallel)."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(side_range[0]), math.log(side_range[1])
    sides = np.exp(rng.uniform(lo, hi, size=count)).astype(np.int64)
    dens = rng.uniform(density_range[0], density_range[1], size=count)
    children = np.random.SeedSequence(seed).spawn(count)
    return [(int(s), float(d), c) for s, d, c in zip(sides, dens, children)]


def batched_world(spec, steps=500):
    side, dens, child = spec
    return batch_world(np.random.default_rng(child), side, dens, steps, seed=int(child.entropy))


def batched_worlds(seed=SEED_BASE + 3, count=256, side_range=(256, 2048), density_range=(0.02, 0.3), steps=500,
                   processes=1):
    specs = batched_world_specs(seed, count, side_range, density_range)
    if processes <= 1:
        return [batched_world(sp, steps) for sp in specs]
    import functools
    import multiprocessing

    with multiprocessing.get_context("fork").Pool(processes) as pool:
        return pool.map(functools.partial(batched_world, steps=steps), specs, chunksize=1)


def serving_job_specs(seed=SEED_BASE + 4, rate=20.0, duration=600.0, side_range=(256, 2048),
                      density_range=(0.02, 0.3), step_range=(100, 10_000), max_jobs=None):
    """Poisson arrivals: list of (arrival_time, side, density, steps, child seed)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    specs = []
    lo, hi = math.log(side_range[0]), math.log(side_range[1])
    slo, shi = math.log(step_range[0]), math.log(step_range[1])
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration or (max_jobs is not None and len(specs) >= max_jobs):
            break
        side = int(math.exp(rng.uniform(lo, hi)))
        dens = rng.uniform(*density_range)
        steps = int(math.exp(rng.uniform(slo, shi)))
        specs.append((t, side, dens, steps))
    children = np.random.SeedSequence(seed).spawn(len(specs))
    return [sp + (ch,) for sp, ch in zip(specs, children)]


def serving_job(spec):
    t, side, dens, steps, child = spec
    return t, batch_world(np.random.default_rng(child), side, dens, steps, seed=int(child.entropy))


def serving_jobs(seed=SEED_BASE + 4, rate=20.0, duration=600.0, side_range=(256, 2048),
                 density_range=(0.02, 0.3), step_range=(100, 10_000), max_jobs=None, processes=1):
    """Poisson arrivals: list of (arrival_time, World)."""
    specs = serving_job_specs(seed, rate, duration, side_range, density_range, step_range, max_jobs)
    if processes <= 1:
        return [serving_job(sp) for sp in specs]
    import multiprocessing

    with multiprocessing.get_context("fork").Pool(processes) as pool:
        return pool.map(serving_job, specs, chunksize=1)


def random_mix(seed, size=(64, 64), num_free=300, patches=6, num_particles=200, num_barrier=20,
               motion=MOTION_COLLISION, vel_sigma=0.3, free_max_connections=6):
    """Small world that exercises every branch: bodies, free cells that can fuse, barrier cells,
    fast cells and particles."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(3, 20, size=patches)
    body, slot, occ = soft_bodies(rng, size, sizes, vel_sigma=vel_sigma, omega_sigma=0.02,
                                  stiffness=lambda n: rng.uniform(0.5, 1.0, size=n))
    free = free_cells(rng, size, num_free, vel_sigma=vel_sigma, max_connections=free_max_connections,
                      exclude=(slot, occ))
    bar = free_cells(rng, size, num_barrier, vel_sigma=0.0, exclude=(slot, occ))
    bar["barrier"] = 1
    bar["vel"] = (rng.normal(0.0, 0.05, size=(num_barrier, 2))).astype(np.float32)
    cells = concat_cells([body, free, bar])
    cells["color"] = rng.integers(0, 7, size=len(cells)).astype(np.uint8)
    parts = particles(rng, size, num_particles, vel_sigma=0.5)
    parts["color"] = rng.integers(0, 7, size=len(parts)).astype(np.uint8)
    assign_ids(cells, parts)
    params = default_params(motionType=motion, smoothingLength=0.65)
    return World(f"mix_{seed}", size, params, cells, parts, 50, seed)


def crowded_bins(seed, motion=MOTION_COLLISION, size=(96, 96), background=400):
    """Free cells crowded into a few bins so that the neighbour pass takes each of its slower
    paths: one bin of 140 cells (a staged bin row above 128 entries), three vertically stacked
    bins of 95 cells (three consecutive rows above 256 entries together), two stacked bins of 70
    cells (a two-row stage above 128 entries whose single rows fit) and a 3x3 block of bins with 9
    cells each (cells with more than 64 candidates), over a sparse background.  Bins are 2 units
    wide at these parameters (interaction radius 1.3)."""
    rng = np.random.default_rng(seed)
    params = default_params(motionType=motion, smoothingLength=0.65)

    def in_bin(bx, by, n):
        p = np.empty((n, 2))
        p[:, 0] = 2.0 * bx + rng.uniform(0.05, 1.95, n)
        p[:, 1] = 2.0 * by + rng.uniform(0.05, 1.95, n)
        return p

    groups = [in_bin(10, 10, 140), in_bin(30, 20, 95), in_bin(30, 21, 95), in_bin(30, 22, 95),
              in_bin(20, 34, 70), in_bin(20, 35, 70)]
    groups += [in_bin(5 + dx, 40 + dy, 9) for dx in range(3) for dy in range(3)]
    crowd_pos = np.concatenate(groups)
    crowd = np.zeros(len(crowd_pos), dtype=CELL_DTYPE)
    crowd["pos"] = crowd_pos.astype(np.float32)
    crowd["vel"] = rng.normal(0.0, 0.05, size=(len(crowd), 2)).astype(np.float32)
    crowd["energy"] = np.float32(100.0)
    crowd["stiffness"] = np.float32(1.0)
    free = free_cells(rng, size, background, vel_sigma=0.2)
    cells = concat_cells([crowd, free])
    parts = particles(rng, size, 50, vel_sigma=0.5)
    assign_ids(cells, parts)
    return World(f"crowded_{seed}", size, params, cells, parts, 4, seed)
