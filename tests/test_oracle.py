"""Pins the CPU oracle (the checker of the HIP path) against what the mathematics and ALIEN's own
tests fix, independently of the oracle's code:

  closed forms       two-cell repulsion (CellProcessor.cuh:329-340), Hooke springs
                     (CellProcessor.cuh:434), SPH kernel values (CellProcessor.cuh:92-119 = the M4
                     cubic spline of Monaghan with sigma = 1/pi), two-cell pressure, free-cell Verlet
  known answers      MuscleTests.cpp:31-65 (bonded pair at rest stays at rest),
                     DescriptionHelperTests.cpp:39-48 (angle sum 360 degrees),
                     DataTransferTests.cpp:19-148 (exact set/get round trip),
                     CellConnectionTests.cpp:19-37 (energy conservation)
  brute force        O(N^2) neighbour sets, the row-major stable map order
  invariants         momentum of symmetric pair forces, energy of transfers, FP32 vs FP64
"""
import math

import numpy as np
import pytest

from alien_amd import abi, worlds
from alien_amd.abi import MOTION_COLLISION, MOTION_FLUID
from oracle.pyoracle import OracleWorld, from_world, lib

from . import helpers as H


def run(size, p, c, parts=None, steps=1, precision=64):
    o = OracleWorld(size, p, precision=precision)
    o.set_data(c, H.particles(0) if parts is None else parts)
    o.step(steps)
    return o.get_data()


# --------------------------------------------------------------------------------------------
# scalar helpers


@pytest.mark.parametrize("precision,tol", [(64, 1e-12), (32, 2e-7)])
def test_sph_kernel_closed_form(oracle_lib, precision, tol):
    import ctypes

    L = lib()
    L.oracle_sph_kernel.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]

    def k(q):
        w, dw = ctypes.c_double(), ctypes.c_double()
        L.oracle_sph_kernel(precision, q, ctypes.byref(w), ctypes.byref(dw))
        return w.value, dw.value

    # Monaghan's M4 spline, sigma = 1/pi: f = 1 - 1.5q^2 + 0.75q^3 (q<1), 0.25(2-q)^3 (q<2)
    def m4(q):
        if q < 1:
            return (1 - 1.5 * q * q + 0.75 * q ** 3) / math.pi, (-3 * q + 2.25 * q * q) / math.pi
        if q < 2:
            return 0.25 * (2 - q) ** 3 / math.pi, -0.75 * (2 - q) ** 2 / math.pi
        return 0.0, 0.0

    for q in (0.0, 0.5, 1.0, 1.5, 2.0, 0.25, 0.999, 1.7, 3.0):
        w, dw = k(q)
        ew, edw = m4(q)
        assert abs(w - ew) <= tol * max(1.0, abs(ew)), (q, w, ew)
        assert abs(dw - edw) <= tol * max(1.0, abs(edw)), (q, dw, edw)
    # W' is the derivative of W (central differences), and W is continuous at q = 1 and q = 2
    if precision == 64:
        for q in (0.3, 0.8, 1.2, 1.9):
            e = 1e-6
            fd = (k(q + e)[0] - k(q - e)[0]) / (2 * e)
            assert abs(fd - k(q)[1]) < 1e-8
        assert abs(k(1 - 1e-12)[0] - k(1.0)[0]) < 1e-10
        assert abs(k(2 - 1e-9)[0]) < 1e-10


@pytest.mark.parametrize("precision,tol", [(64, 1e-9), (32, 3e-5)])
def test_angle_of_vector(oracle_lib, precision, tol):
    """Math.cuh:140-155: 0 degrees at (0,-1), 90 at (1,0), clockwise on screen: atan2(x, -y)."""
    import ctypes

    L = lib()
    L.oracle_angle_of_vector.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.oracle_angle_of_vector.restype = ctypes.c_double
    L.oracle_subtract_angle.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.oracle_subtract_angle.restype = ctypes.c_double
    rng = np.random.default_rng(5)
    for x, y in list(rng.normal(size=(500, 2)) * 3) + [(1, 0), (0, 1), (-1, 0), (0.3, -1)]:
        if precision == 32:
            x, y = float(np.float32(x)), float(np.float32(y))
        a = L.oracle_angle_of_vector(precision, x, y)
        e = math.degrees(math.atan2(x, -y)) % 360.0
        d = abs(a - e)
        assert min(d, 360 - d) <= tol * 360, (x, y, a, e)
    assert L.oracle_angle_of_vector(precision, 0.0, 0.0) == 0.0
    assert L.oracle_subtract_angle(precision, 10.0, 350.0) == 20.0
    assert L.oracle_subtract_angle(precision, 350.0, 10.0) == 340.0


# --------------------------------------------------------------------------------------------
# one-step closed forms


@pytest.mark.parametrize("d", [0.5, 0.9, 1.25])
def test_two_cell_repulsion(oracle_lib, d):
    """Each cell gets F = n (r - d) 0.08 from its own visit and again from the partner's visit
    (CellProcessor.cuh:329-340), so after one step |v| = 2 (1.3 - d) 0.08 pointing apart."""
    c = H.cells(2)
    c["pos"] = [[50.0, 50.0], [50.0 + d, 50.0]]
    d = float(c["pos"][1, 0]) - 50.0  # the distance after rounding to FP32
    p = H.params(MOTION_COLLISION, friction=0.0)
    out, _ = run((100, 100), p, c)
    v = 2 * (1.3 - d) * 0.08
    np.testing.assert_allclose(out["vel"][0], [-v, 0.0], atol=1e-7)
    np.testing.assert_allclose(out["vel"][1], [v, 0.0], atol=1e-7)
    np.testing.assert_allclose(out["pos"][0], [50.0 - v, 50.0], atol=1e-5)
    np.testing.assert_allclose(out["pos"][1], [50.0 + d + v, 50.0], atol=1e-5)


def test_two_cell_repulsion_across_seam_is_invisible_in_collision_mode(oracle_lib):
    """The collision predicate uses the uncorrected difference (CellProcessor.cuh:296-301), so two
    cells touching across the torus seam do not interact."""
    c = H.cells(2)
    c["pos"] = [[99.8, 50.0], [0.3, 50.0]]
    p = H.params(MOTION_COLLISION, friction=0.0)
    out, _ = run((100, 100), p, c)
    np.testing.assert_array_equal(out["vel"], 0.0)


def test_overlap_push(oracle_lib):
    """Below cellMinDistance the position is pushed by delta * cellMinDistance / 5 in the
    neighbour pass (CellProcessor.cuh:303-306), before the velocity update."""
    d = 0.2
    c = H.cells(2)
    c["pos"] = [[50.0, 50.0], [50.0 + d, 50.0]]
    d = float(c["pos"][1, 0]) - 50.0  # the distance after rounding to FP32
    p = H.params(MOTION_COLLISION, friction=0.0)
    out, _ = run((100, 100), p, c)
    v = 2 * (1.3 - d) * 0.08
    push = d * 0.3 / 5
    np.testing.assert_allclose(out["vel"][0], [-v, 0.0], atol=1e-7)
    np.testing.assert_allclose(out["pos"][0], [50.0 - push - v, 50.0], atol=1e-5)


def test_fast_approach_uses_velocity_branch(oracle_lib):
    """|v_i| > 0.5 and approaching: F = delta (dv . delta) / (-2 (d^2 + 0.25)) per visit."""
    d = 1.0
    c = H.cells(2)
    c["pos"] = [[50.0, 50.0], [50.0 + d, 50.0]]
    d = float(c["pos"][1, 0]) - 50.0  # the distance after rounding to FP32
    c["vel"] = [[0.55, 0.0], [0.0, 0.0]]
    c["maxConnections"] = 0  # no fusion
    p = H.params(MOTION_COLLISION, friction=0.0)
    out, _ = run((100, 100), p, c)
    # cell 0 visits 1: delta = p0 - p1 = (-d, 0), dv = v0 - v1 = (0.55, 0), approaching:
    # f = delta (dv . delta) / (-2 (d^2 + 0.25)) = (-0.55 d^2 / (2 (d^2 + 0.25)), 0)
    f_own = -0.55 * d * d / (2 * (d * d + 0.25))
    # cell 1 visits 0 (|v_1| = 0: repulsion branch) and adds -f to cell 0, f = (+(r - d) 0.08, 0)
    f_partner = -(1.3 - d) * 0.08
    v0 = 0.55 + f_own + f_partner
    np.testing.assert_allclose(out["vel"][0, 0], v0, atol=1e-7)


def test_hooke_spring(oracle_lib):
    """Bonded pair stretched to d from rest d0: F = n (d - d0)(k1^2 + k2^2) / 6
    (CellProcessor.cuh:434), Verlet x' = x + F/2, v' = (F + F')/2 (CellProcessor.cuh:525-559)."""
    d, d0, k1, k2 = 1.2, 1.0, 0.9, 0.6
    c = H.cells(2)
    c["pos"] = [[50.0, 50.0], [50.0 + d, 50.0]]
    d = float(c["pos"][1, 0]) - 50.0  # the distance after rounding to FP32
    c["stiffness"] = [k1, k2]
    H.bond(c, 0, 1, d0)
    p = H.params(MOTION_COLLISION, friction=0.0)
    out, _ = run((100, 100), p, c)
    kk = (k1 * k1 + k2 * k2) / 6
    f0 = (d - d0) * kk
    d1 = d - f0  # each end moves f0 / 2 toward the other
    f1 = (d1 - d0) * kk
    np.testing.assert_allclose(out["pos"][0], [50.0 + f0 / 2, 50.0], atol=1e-6)
    np.testing.assert_allclose(out["pos"][1], [50.0 + d - f0 / 2, 50.0], atol=1e-6)
    np.testing.assert_allclose(out["vel"][0], [(f0 + f1) / 2, 0.0], atol=1e-7)
    np.testing.assert_allclose(out["vel"][1], [-(f0 + f1) / 2, 0.0], atol=1e-7)


@pytest.mark.parametrize("precision", [32, 64])
def test_bonded_pair_at_rest_does_nothing(oracle_lib, precision):
    """MuscleTests.cpp:31-65 (doNothing): bond at rest length, no velocity, no friction."""
    c = H.cells(2)
    c["pos"] = [[10.0, 10.0], [11.0, 10.0]]
    H.bond(c, 0, 1, 1.0)
    p = H.params(MOTION_COLLISION, friction=0.0)
    out, _ = run((100, 100), p, c, precision=precision)
    np.testing.assert_array_equal(out["vel"], 0.0)
    assert abs(float(out["pos"][1, 0] - out["pos"][0, 0]) - 1.0) < 1e-6


def test_free_cell_verlet_is_exact_in_fp32(oracle_lib):
    """A lone cell: x' = x + v dt, v' = v (1 - friction), both exactly rounded FP32 operations."""
    rng = np.random.default_rng(3)
    c = H.cells(1)
    for _ in range(20):
        x = rng.uniform(0, 100, size=2).astype(np.float32)
        v = rng.normal(0, 0.5, size=2).astype(np.float32)
        c["pos"][0] = x
        c["vel"][0] = v
        p = H.params(MOTION_COLLISION, friction=0.001)
        out, _ = run((100, 100), p, c, precision=32)
        ex = (x + v * np.float32(1.0)).astype(np.float32) % np.float32(100)
        ev = (v * (np.float32(1) - np.float32(0.001))).astype(np.float32)
        np.testing.assert_array_equal(out["pos"][0], ex)
        np.testing.assert_array_equal(out["vel"][0], ev)


def test_velocity_clamp(oracle_lib):
    c = H.cells(1)
    c["pos"][0] = [10.0, 10.0]
    c["vel"][0] = [3.0, 4.0]
    p = H.params(MOTION_COLLISION, friction=0.0)
    out, _ = run((100, 100), p, c)
    np.testing.assert_allclose(out["vel"][0], [1.2, 1.6], atol=1e-9)


def test_two_cell_pressure(oracle_lib):
    """Fluid, both densities 1 after set: F_i = 0.1 * (-delta/d) * (1/1 + 1/1) * W'(d/h) / h^3,
    viscosity 0 for equal velocities (CellProcessor.cuh:201-224, 277-278)."""
    h, d = 0.65, 0.8
    c = H.cells(2)
    c["pos"] = [[50.0, 50.0], [50.0 + d, 50.0]]
    d = float(c["pos"][1, 0]) - 50.0  # the distance after rounding to FP32
    c["maxConnections"] = 0
    p = H.params(MOTION_FLUID, friction=0.0, smoothingLength=h)
    out, _ = run((100, 100), p, c)
    q = d / h
    dw = (-3 * q + 2.25 * q * q) / math.pi if q < 1 else -0.75 * (2 - q) ** 2 / math.pi
    f = 0.1 * 2 * dw / h ** 3  # force on cell 0 along +x (W' < 0: pushes cell 0 to -x)
    np.testing.assert_allclose(out["vel"][0], [f, 0.0], atol=1e-9)
    np.testing.assert_allclose(out["vel"][1], [-f, 0.0], atol=1e-9)


def test_particle_absorption_transfers_energy(oracle_lib):
    """A particle moving into a cell's slot gives its energy to that cell and dies
    (RadiationProcessor.cuh:134-165, radiationAbsorption 1)."""
    c = H.cells(1)
    c["pos"][0] = [20.5, 20.5]
    c["energy"][0] = 100.0
    pa = H.particles(1)
    pa["pos"][0] = [19.5, 20.5]
    pa["vel"][0] = [0.5, 0.0]
    pa["energy"][0] = 3.25
    p = H.params(MOTION_COLLISION, friction=0.0)
    out, parts = run((100, 100), p, c, parts=pa, precision=32)
    assert len(parts) == 0
    assert out["energy"][0] == np.float32(103.25)


# --------------------------------------------------------------------------------------------
# boundary and map


def test_set_get_round_trip_is_exact(oracle_lib):
    """DataTransferTests.cpp:19-148: what goes in comes out, ordered by id."""
    w = worlds.random_mix(11, motion=MOTION_COLLISION)
    o = from_world(w)
    gc, gp = o.get_data()
    oc = np.sort(w.cells, order="id")
    op = np.sort(w.particles, order="id")
    assert H.packed_bytes(gc) == H.packed_bytes(oc)
    assert H.packed_bytes(gp) == H.packed_bytes(op)


@pytest.mark.parametrize("motion", [MOTION_FLUID, MOTION_COLLISION])
def test_neighbor_sets_brute_force(oracle_lib, motion):
    w = worlds.random_mix(21 + motion, motion=motion)
    o = from_world(w)
    off, ids = o.neighbors()
    c = np.sort(w.cells, order="id")
    bf = H.brute_force_neighbors(c, w.size, w.params)
    got = {int(c["id"][i]): set(int(x) for x in ids[off[i]:off[i + 1]]) for i in range(len(c))}
    assert got == bf


@pytest.mark.parametrize("motion", [MOTION_FLUID, MOTION_COLLISION])
def test_map_order_row_major_stable(oracle_lib, motion):
    """Cells sorted by bin (row-major), ties by the previous logical order, which right after a
    set call is the order of the input array."""
    w = worlds.random_mix(31 + motion, motion=motion)
    o = from_world(w)
    bins, ids, (s, nbx, nby) = o.get_map()
    assert (s, nbx, nby) == H.bin_geometry(w.size, w.params)
    c = np.sort(w.cells, order="id")
    eb = H.bins_of(c["pos"], w.size, w.params)
    order = np.lexsort((c["id"], eb))
    np.testing.assert_array_equal(ids, c["id"][order])
    np.testing.assert_array_equal(bins, eb[order])
    # a shuffled input gives the same bins, with ties in the shuffled input order
    perm = np.random.default_rng(1).permutation(len(w.cells))
    o2 = OracleWorld(w.size, w.params)
    shuffled = w.cells[perm].copy()
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    conn = shuffled["connections"]
    idx = conn["cellIndex"]
    mask = np.arange(6)[None, :] < shuffled["numConnections"][:, None]
    idx[mask] = inv[idx[mask]]
    conn["cellIndex"] = idx
    shuffled["connections"] = conn
    o2.set_data(shuffled, w.particles)
    bins2, ids2, _ = o2.get_map()
    eb2 = H.bins_of(shuffled["pos"], w.size, w.params)
    order2 = np.lexsort((np.arange(len(shuffled)), eb2))
    np.testing.assert_array_equal(ids2, shuffled["id"][order2])
    np.testing.assert_array_equal(bins2, bins)


# --------------------------------------------------------------------------------------------
# invariants over many steps


def test_generated_bodies_are_consistent():
    w = worlds.random_mix(41)
    assert H.bonds_symmetric(w.cells)
    np.testing.assert_allclose(H.angle_sums(w.cells), 360.0, atol=1e-3)


@pytest.mark.parametrize("motion", [MOTION_FLUID, MOTION_COLLISION])
def test_topology_invariants_over_steps(oracle_lib, motion):
    """Fusion, bond breaking and deletions keep every bond list symmetric and every angle sum at
    360 degrees (DescriptionHelperTests.cpp:39-48, DebugKernels.cu:48-74)."""
    w = worlds.random_mix(51 + motion, motion=motion)
    o = from_world(w, precision=32)
    n0 = int(w.cells["numConnections"].sum())
    o.step(60)
    c, _ = o.get_data()
    assert H.bonds_symmetric(c)
    np.testing.assert_allclose(H.angle_sums(c), 360.0, atol=2e-3)
    assert int(c["numConnections"].sum()) != n0  # the mix really changed its topology


@pytest.mark.parametrize("motion", [MOTION_FLUID, MOTION_COLLISION])
def test_energy_conservation(oracle_lib, motion):
    """CellConnectionTests.cpp:19-37 / IntegrationTestFramework.cpp:30-40: cell + particle energy
    only moves between the two (absorption, particle fusion)."""
    w = worlds.random_mix(61 + motion, motion=motion)
    o = from_world(w, precision=64)
    e0 = o.totals()["energy"]
    np0 = o.counts()[1]
    o.step(100)
    t = o.totals()
    assert abs(t["energy"] - e0) <= 1e-12 * e0
    assert o.counts()[1] < np0  # particles were absorbed or merged


def test_momentum_conservation_collision(oracle_lib):
    """Pair forces are applied as +f to one cell and -f to the other and springs/angle forces are
    balanced, so with no friction, no clamp, no barriers and no particles sum(v) is constant."""
    w = worlds.random_mix(71, motion=MOTION_COLLISION, num_barrier=0, num_particles=0)
    p = w.params.copy()
    p.friction = 0.0
    p.cellMaxVelocity = 1e9
    o = OracleWorld(w.size, p, precision=64)
    o.set_data(w.cells, w.particles)
    m0 = np.array(o.totals()["momentum"])
    o.step(100)
    m1 = np.array(o.totals()["momentum"])
    scale = np.abs(w.cells["vel"].astype(np.float64)).sum()
    assert np.abs(m1 - m0).max() <= 1e-12 * scale


def matched_particles(a, b):
    """Particles present in both results, minus those whose energy differs: a merge or absorption
    decided differently at a threshold (a discrete event, not a rounding difference)."""
    common = np.intersect1d(a["id"], b["id"])
    a = a[np.isin(a["id"], common)]
    b = b[np.isin(b["id"], common)]
    same = np.abs(a["energy"].astype(np.float64) - b["energy"]) <= 1e-5 * np.abs(b["energy"].astype(np.float64))
    return a[same], b[same], len(common) - int(same.sum())


@pytest.mark.parametrize("name", ["mix_collision", "mix_fluid", "plain_soft_body", "fluid_256k"])
def test_fp32_single_step_within_1e5_of_fp64(oracle_lib, name):
    """One step of the FP32 build agrees with the FP64 build within 1e-5, norm-wise
    (max |x32 - x64| / max |x64|) for cell positions, cell velocities and particle velocities."""
    w = {
        "mix_collision": lambda: worlds.random_mix(91, motion=MOTION_COLLISION),
        "mix_fluid": lambda: worlds.random_mix(92, motion=MOTION_FLUID),
        "plain_soft_body": lambda: worlds.plain_soft_body(),
        "fluid_256k": lambda: worlds.large_fluid(num_cells=256_000, num_particles=32_000, size=(1024, 1024)),
    }[name]()
    o32 = from_world(w, precision=32, threads=8)
    o64 = from_world(w, precision=64, threads=8)
    o32.step(1)
    o64.step(1)
    c32, p32 = o32.get_data()
    c64, p64 = o64.get_data()
    assert c32["id"].tolist() == c64["id"].tolist()
    assert np.abs(H.torus_diff(c32["pos"], c64["pos"], w.size)).max() <= 1e-5 * np.abs(c64["pos"]).max()
    assert H.normwise_rel(c32["vel"], c64["vel"]) <= 1e-5
    a, b, flips = matched_particles(p32, p64)
    assert flips <= max(1, len(p64) // 1000)
    if len(b):
        assert np.abs(H.torus_diff(a["pos"], b["pos"], w.size)).max() <= 1e-5 * np.abs(b["pos"]).max()
        assert H.normwise_rel(a["vel"], b["vel"]) <= 1e-5


def test_fp32_drift_1000_steps(oracle_lib):
    """Over 1000 steps the FP32 build holds total energy and momentum within 1e-3 (relative) of
    the FP64 build.  Collision mode without barriers and without the velocity clamp: then the
    only momentum change is friction, M(t) = (1 - f)^t M(0), which both builds must follow."""
    w = worlds.random_mix(81, motion=MOTION_COLLISION, num_barrier=0)
    p = w.params.copy()
    p.cellMaxVelocity = 1e9
    res = {}
    for prec in (32, 64):
        o = OracleWorld(w.size, p, precision=prec, threads=4)
        o.set_data(w.cells, w.particles)
        o.step(1000)
        res[prec] = o.totals()
    t32, t64 = res[32], res[64]
    m0 = w.cells["vel"].astype(np.float64).sum(axis=0)
    expect = m0 * (1.0 - np.float32(p.friction)) ** 1000
    m32, m64 = np.array(t32["momentum"]), np.array(t64["momentum"])
    assert np.abs(m64 - expect).max() <= 1e-3 * np.abs(expect).max()
    assert np.abs(m32 - m64).max() <= 1e-3 * np.abs(m64).max()
    assert abs(t32["energy"] - t64["energy"]) <= 1e-3 * t64["energy"]
    e0 = float(w.cells["energy"].astype(np.float64).sum() + w.particles["energy"].astype(np.float64).sum())
    assert abs(t64["energy"] - e0) <= 1e-9 * e0


def test_asymmetric_bonds_are_rejected(oracle_lib):
    c = H.cells(3)
    c["pos"] = [[1.0, 1.0], [2.0, 1.0], [3.0, 1.0]]
    c["numConnections"][0] = 1
    c["connections"][0, 0]["cellIndex"] = 1  # 0 -> 1 without 1 -> 0
    o = OracleWorld((16, 16), H.params())
    with pytest.raises(RuntimeError, match="symmetric"):
        o.set_data(c, H.particles(0))


def test_particle_merge(oracle_lib):
    """Two particles in one 1x1 slot closer than sqrt(0.5): the one that does not own the slot
    merges into the owner with energy-weighted velocity (RadiationProcessor.cuh:78-96).  The slot
    owner is the lowest-index particle whose pre-move position lies in it."""
    pa = H.particles(2)
    pa["pos"] = [[10.2, 10.2], [10.6, 10.6]]
    pa["vel"] = [[0.1, 0.0], [0.0, 0.1]]
    pa["energy"] = [2.0, 3.0]
    out_c, out_p = run((100, 100), H.params(MOTION_COLLISION), H.cells(0), parts=pa)
    assert len(out_p) == 1 and out_p["id"][0] == pa["id"][0]
    assert out_p["energy"][0] == np.float32(5.0)
    np.testing.assert_allclose(out_p["vel"][0], [0.1 * 0.4, 0.1 * 0.6], atol=1e-7)
    np.testing.assert_allclose(out_p["pos"][0], [10.3, 10.2], atol=1e-5)


def test_particle_reflects_off_barrier(oracle_lib):
    """A particle heading into a barrier cell's slot: v' = vr - 2 r (vr.r) / max(0.1, |r|^2) + v_c
    with r = particle - cell (RadiationProcessor.cuh:98-106); no energy moves."""
    c = H.cells(1)
    c["pos"][0] = [20.5, 20.5]
    c["barrier"] = 1
    pa = H.particles(1)
    pa["pos"][0] = [19.5, 20.5]
    pa["vel"][0] = [0.6, 0.0]
    pa["energy"][0] = 2.5
    out_c, out_p = run((100, 100), H.params(MOTION_COLLISION), c, parts=pa)
    rx = 20.1 - 20.5
    vr = 0.6
    dot = vr * rx
    v = vr - rx * 2 * dot / max(0.1, rx * rx)
    np.testing.assert_allclose(out_p["vel"][0], [v, 0.0], atol=1e-5)
    assert out_p["energy"][0] == np.float32(2.5)
    assert out_c["energy"][0] == np.float32(100.0)


def test_fusion_of_two_fast_cells(oracle_lib):
    """Two free cells closing at |dv| >= cellFusionVelocity (0.6) get a bond at the end of the
    step (CellProcessor.cuh:343-360, CellConnectionProcessor.cuh:211-330) whose rest length is
    their distance then, with angleFromPrevious 360 on both sides."""
    c = H.cells(2)
    c["pos"] = [[30.0, 30.0], [31.0, 30.0]]
    c["vel"] = [[0.4, 0.0], [-0.4, 0.0]]
    p = H.params(MOTION_COLLISION, friction=0.0)
    out, _ = run((100, 100), p, c)
    v = 0.4 - 2 * (1.3 - 1.0) * 0.08  # repulsion from both visits (|v| < 0.5: no fast branch)
    d = 1.0 - 2 * v
    assert out["numConnections"].tolist() == [1, 1]
    assert out["connections"][0, 0]["cellIndex"] == 1 and out["connections"][1, 0]["cellIndex"] == 0
    np.testing.assert_allclose(out["connections"]["distance"][:, 0], [d, d], atol=1e-5)
    np.testing.assert_array_equal(out["connections"]["angleFromPrevious"][:, 0], [360.0, 360.0])


def test_no_fusion_below_fusion_velocity(oracle_lib):
    c = H.cells(2)
    c["pos"] = [[30.0, 30.0], [31.0, 30.0]]
    c["vel"] = [[0.25, 0.0], [-0.25, 0.0]]
    out, _ = run((100, 100), H.params(MOTION_COLLISION, friction=0.0), c)
    assert out["numConnections"].tolist() == [0, 0]


def test_overstretched_bond_breaks(oracle_lib):
    """A bond longer than cellMaxBindingDistance (3.6) after the position update is deleted on
    both sides (CellProcessor.cuh:493-523)."""
    c = H.cells(3)
    c["pos"] = [[10.0, 10.0], [13.7, 10.0], [20.0, 10.0]]
    H.bond(c, 0, 1, 3.7)
    H.bond(c, 1, 2, 6.3, angle_a=180.0, angle_b=360.0)
    c["connections"][1, 0]["angleFromPrevious"] = 180.0
    out, _ = run((100, 100), H.params(MOTION_COLLISION, friction=0.0), c)
    assert out["numConnections"].tolist() == [0, 0, 0]


def test_angle_force_closed_form(oracle_lib):
    """Cell 1 with bonds to the left (angle 270) and right (angle 90) neighbours, reference angle
    150 for the first bond against the actual 180: strength = 30 / 2000 (stiffness 1); force1 =
    cw(n0) * strength to cell 0, force2 = ccw(n_prev) * strength to cell 2, both negated because
    reference < actual, and -(force1 + force2) on cell 1 (CellProcessor.cuh:450-485).  Springs are
    at rest, so after one step each cell has moved by its force / 2."""
    c = H.cells(3)
    c["pos"] = [[49.0, 50.0], [50.0, 50.0], [51.0, 50.0]]
    H.bond(c, 1, 0, 1.0, angle_a=150.0, angle_b=360.0)
    H.bond(c, 1, 2, 1.0, angle_a=210.0, angle_b=360.0)
    out, _ = run((100, 100), H.params(MOTION_COLLISION, friction=0.0), c)
    s = 30.0 / 2000.0
    np.testing.assert_allclose(out["pos"][:, 1], [50.0 + s / 2, 50.0 - s, 50.0 + s / 2], atol=1e-5)
    np.testing.assert_allclose(out["pos"][:, 0], [49.0, 50.0, 51.0], atol=1e-5)
    # without angle forces (timestep 1: only every third step) nothing moves
    o = OracleWorld((100, 100), H.params(MOTION_COLLISION, friction=0.0), precision=64)
    o.set_data(c, H.particles(0))
    o.timestep = 1
    o.step(1)
    out1, _ = o.get_data()
    np.testing.assert_allclose(out1["pos"], c["pos"], atol=1e-6)


def test_statistics_by_colour_closed_form():
    """Per-colour counts and energies (StatisticsKernels.cu:11-45) on a hand-made world."""
    from oracle.pyoracle import statistics

    cells = np.zeros(3, dtype=abi.CELL_DTYPE)
    cells["color"] = [0, 2, 2]
    cells["energy"] = [1.5, 2.0, 0.25]
    parts = np.zeros(2, dtype=abi.PARTICLE_DTYPE)
    parts["color"] = [2, 6]
    parts["energy"] = [3.0, 0.5]
    st = statistics(cells, parts)
    assert st["numCells"] == [1, 0, 2, 0, 0, 0, 0]
    assert st["numParticles"] == [0, 0, 1, 0, 0, 0, 1]
    assert st["totalEnergy"] == [1.5, 0.0, 5.25, 0.0, 0.0, 0.0, 0.5]
