"""Closed-form checks of the oracle terms that the whole-step tests leave open.

Each expectation below is written from the reference formulas (file:line given per test) in
float64 numpy, independently of the oracle's code, on worlds small enough that every term can be
followed by hand.  They are chosen so that a plausible slip in the oracle changes the result:

  SPH density sum        CellProcessor.cuh:183-186   (self term W(0)/h^2 included, barriers excluded)
  SPH pressure factor    CellProcessor.cuh:201-211   (1/rho_i + 1/rho_j, last step's densities)
  SPH viscosity          CellProcessor.cuh:215-217   ((v_i - v_j)/rho_j * d * W'/(d^2 + 0.25))
  collision barrier      CellProcessor.cuh:318-339   (a barrier's own visit counts twice)
  fluid barrier          CellProcessor.cuh:234-274   (lone barrier: r = cell - barrier; bonded wall:
                                                      r = rotated edge between the two partners)
  fusion insertion       CellConnectionProcessor.cuh:346-395  (angle fraction times the reference
                                                      angle of the gap, remainder to the next bond)
  binding-energy decay   CellProcessor.cuh:786-790   (energy above cellMaxBindingEnergy drops all bonds)

The SPH kernel is Monaghan's M4 spline with sigma = 1/pi (pinned against CellProcessor.cuh:92-119
in test_oracle.py), and angles are atan2(x, -y) in degrees (Math.cuh:140-155, also pinned there).
"""
import math

import numpy as np
import pytest

from alien_amd.abi import MOTION_COLLISION, MOTION_FLUID
from oracle.pyoracle import OracleWorld

from . import helpers as H


def m4_w(q):
    if q < 1:
        return (1 - 1.5 * q * q + 0.75 * q ** 3) / math.pi
    if q < 2:
        return 0.25 * (2 - q) ** 3 / math.pi
    return 0.0


def m4_dw(q):
    if q < 1:
        return (-3 * q + 2.25 * q * q) / math.pi
    if q < 2:
        return -0.75 * (2 - q) ** 2 / math.pi
    return 0.0


def angle_deg(v):
    return math.degrees(math.atan2(v[0], -v[1])) % 360.0


def run(size, p, c, parts=None, steps=1, precision=64, timestep=0):
    o = OracleWorld(size, p, precision=precision)
    o.set_data(c, H.particles(0) if parts is None else parts)
    o.timestep = timestep
    o.step(steps)
    return o


# --------------------------------------------------------------------------------------------
# SPH: density sum, pressure factor and viscosity over two steps (unequal densities in step 2)


def fluid_line(vel):
    """Three free cells in a row (A-B 0.9 apart, B-C 1.0, A-C 1.9 > 2h): the densities that
    step 1 leaves are all different (B has two neighbours, A and C one each, at other distances)."""
    c = H.cells(3)
    c["pos"] = [[5.0, 5.0], [5.9, 5.0], [6.9, 5.0]]
    c["vel"] = vel
    c["maxConnections"] = 0
    return c


def sph_step(pos, vel, rho, h, pressure, viscosity):
    """One step of the SPH pass for unbonded free cells, no friction, no clamp reached:
    returns (pos', vel', rho') with rho' the density sums of this step."""
    n = len(pos)
    F = np.zeros((n, 2))
    rho_new = np.zeros(n)
    for i in range(n):
        Fp = np.zeros(2)
        Fv = np.zeros(2)
        for j in range(n):
            d_vec = pos[i] - pos[j]
            d = float(np.hypot(*d_vec))
            if d > 2 * h:
                continue
            rho_new[i] += m4_w(d / h) / h ** 2  # includes the cell itself (d = 0)
            if i == j:
                continue
            kd = m4_dw(d / h) / h ** 3
            Fp += d_vec / (-d) * (1.0 / rho[i] + 1.0 / rho[j]) * kd
            Fv += (vel[i] - vel[j]) / rho[j] * d * kd / (d * d + 0.25)
        F[i] = Fp * pressure + Fv * viscosity
    vel2 = vel + F
    return pos + vel2, vel2, rho_new


@pytest.mark.parametrize("pressure,viscosity", [(0.01, 0.0), (0.0, 0.05), (0.01, 0.05)])
def test_sph_two_steps_closed_form(oracle_lib, pressure, viscosity):
    """Step 1 runs with the densities reset to 1 (set data); step 2 with the density sums of
    step 1.  Pressure alone pins the density sum through 1/rho_i + 1/rho_j; viscosity alone pins
    the 1/rho_j weight and the d^2 + 0.25 softening (unequal velocities and densities)."""
    h = 0.65
    vel = [[0.02, 0.01], [-0.01, 0.015], [0.005, -0.02]]
    c = fluid_line(vel)
    p = H.params(MOTION_FLUID, friction=0.0, smoothingLength=h, pressureStrength=pressure,
                 viscosityStrength=viscosity)
    pos0 = c["pos"].astype(np.float64)
    vel0 = c["vel"].astype(np.float64)
    pos1, vel1, rho1 = sph_step(pos0, vel0, np.ones(3), h, pressure, viscosity)
    pos2, vel2, _ = sph_step(pos1, vel1, rho1, h, pressure, viscosity)
    # the three densities really differ, and step 2 really depends on them
    assert len({round(r, 6) for r in rho1}) == 3
    o = run((32, 32), p, c, steps=1)
    out1, _ = o.get_data()
    np.testing.assert_allclose(out1["vel"], vel1, rtol=0, atol=2e-8)
    np.testing.assert_allclose(out1["pos"], pos1, rtol=0, atol=2e-6)
    o.step(1)
    out2, _ = o.get_data()
    np.testing.assert_allclose(out2["vel"], vel2, rtol=0, atol=2e-8)
    np.testing.assert_allclose(out2["pos"], pos2, rtol=0, atol=2e-6)
    # a density sum without the self term, or weighted by 1/rho_i, would be far off
    _, vel2_wrong, _ = sph_step(pos1, vel1, rho1 * 2.0, h, pressure, viscosity)
    assert np.abs(vel2_wrong - vel2).max() > 100 * 2e-8


# --------------------------------------------------------------------------------------------
# collision mode: barrier factor


@pytest.mark.parametrize("fast", [False, True])
def test_collision_barrier_counts_twice(oracle_lib, fast):
    """A free cell next to a barrier cell gets its own visit's force (factor 1) and the negated
    force of the barrier's visit (factor 2, CellProcessor.cuh:322): in the repulsion branch
    3 (r - d) 0.08 in total.  Moving fast toward the barrier, its own visit takes the velocity
    branch instead while the barrier (at rest) still repels it with factor 2."""
    c = H.cells(2)
    c["pos"] = [[10.0, 10.0], [10.9, 10.0]]
    c["barrier"] = [1, 0]
    v_free = -0.6 if fast else 0.0
    c["vel"] = [[0.0, 0.0], [v_free, 0.0]]
    d = float(np.float32(10.9)) - 10.0
    p = H.params(MOTION_COLLISION, friction=0.0)
    out, _ = run((32, 32), p, c).get_data()
    rep = (1.3 - d) * 0.08
    if fast:
        # own visit: delta = (d, 0), dv = (v, 0): delta (dv . delta) / (-2 (d^2 + 0.25))
        own = d * (v_free * d) / (-2 * (d * d + 0.25))
        expect = v_free + own + 2 * rep
    else:
        expect = 3 * rep
    np.testing.assert_allclose(out["vel"][1], [expect, 0.0], atol=1e-7)
    np.testing.assert_array_equal(out["vel"][0], [0.0, 0.0])  # barriers do not move
    np.testing.assert_array_equal(out["pos"][0], [10.0, 10.0])


# --------------------------------------------------------------------------------------------
# fluid mode: barrier reflection


def reflection(vel_cell, r, dist):
    """CellProcessor.cuh:265-271 for a barrier at rest: F = (vr - 2 r (vr.r)/max(0.05, |r|^2)
    - v_cell) / max(0.05, dist) if vr.r < 0."""
    vr = np.asarray(vel_cell, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    dot = float(vr @ r)
    if dot >= 0:
        return np.zeros(2)
    return (vr - r * 2 * dot / max(0.05, float(r @ r)) - vr) / max(0.05, dist)


def test_fluid_lone_barrier_reflection(oracle_lib):
    """A lone barrier (no bonds): r is the direction from the barrier to the cell.  The barrier is
    not part of the density sum, so the cell has no pressure or viscosity partner."""
    h = 0.65
    c = H.cells(2)
    c["pos"] = [[10.0, 10.0], [10.3, 10.8]]
    c["barrier"] = [1, 0]
    c["vel"] = [[0.0, 0.0], [0.1, -0.3]]
    c["maxConnections"] = 0
    p = H.params(MOTION_FLUID, friction=0.0, smoothingLength=h)
    out, _ = run((32, 32), p, c).get_data()
    r = np.array([float(np.float32(10.3)) - 10.0, float(np.float32(10.8)) - 10.0])
    F = reflection([0.1, -0.3], r, float(np.hypot(*r)))
    np.testing.assert_allclose(out["vel"][1], np.array([0.1, -0.3]) + F, atol=1e-7)
    assert F[1] > 0.5  # a real reflection, not a rounding effect


def test_fluid_bonded_wall_reflection(oracle_lib):
    """A wall of three bonded barriers; the closest one has two bonds, so r is the edge between
    the two partners that bracket the direction to the cell, rotated a quarter counterclockwise
    ((x, y) -> (y, -x), Math.cuh:164-169): here (-2, 0) -> (0, 2), the wall normal, whatever the
    cell's offset along the wall."""
    h = 0.65
    c = H.cells(4)
    c["pos"] = [[9.0, 10.0], [10.0, 10.0], [11.0, 10.0], [10.3, 10.8]]
    c["barrier"] = [1, 1, 1, 0]
    c["vel"][3] = [0.1, -0.3]
    c["maxConnections"] = [6, 6, 6, 0]
    # wall cell 1 lists its partners as [0, 2]: bond angles 180/180 (partners at 270 and 90 degrees)
    H.bond(c, 1, 0, 1.0, angle_a=180.0, angle_b=360.0)
    H.bond(c, 1, 2, 1.0, angle_a=180.0, angle_b=360.0)
    p = H.params(MOTION_FLUID, friction=0.0, smoothingLength=h)
    out, _ = run((32, 32), p, c, timestep=1).get_data()
    d = np.array([float(np.float32(10.3)) - 10.0, float(np.float32(10.8)) - 10.0])
    F = reflection([0.1, -0.3], np.array([0.0, 2.0]), float(np.hypot(*d)))
    np.testing.assert_allclose(out["vel"][3], np.array([0.1, -0.3]) + F, atol=1e-7)
    # the lone-barrier rule (r = cell - barrier) would give another velocity
    F_lone = reflection([0.1, -0.3], d, float(np.hypot(*d)))
    assert np.abs(F - F_lone).max() > 1e-2


# --------------------------------------------------------------------------------------------
# fusion into a cell that already has two or three bonds


@pytest.mark.parametrize("case", ["two_bonds", "three_bonds"])
def test_fusion_insertion_angle_fraction(oracle_lib, case):
    """A fast cell F fuses into X between X's partners P (before) and N (after) in clockwise
    order.  The new bond's angleFromPrevious is refAngle * (angle(F) - angle(P)) / (angle(N) -
    angle(P)), refAngle being the stored angle of the gap's bond, and that bond keeps the rest
    (CellConnectionProcessor.cuh:367-395).  The stored angles differ from the geometric ones on
    purpose (no angle forces at timestep 1), and F comes in off the gap's bisector, so a half
    split of the gap, or the geometric gap instead of refAngle, gives other angles."""
    X = np.array([20.0, 20.0])

    def at(deg, r=1.0):
        a = math.radians(deg)
        return X + r * np.array([math.sin(a), -math.cos(a)])

    if case == "two_bonds":
        partners = [at(0.0), at(180.0)]
        stored = [250.0, 110.0]  # X's bond list [P0, P1]; the gap P0 -> P1 holds 110
        f_deg, gap = 45.0, 1  # F between P0 (0 deg) and P1 (180 deg): the bond at index 1
    else:
        partners = [at(0.0), at(120.0), at(240.0)]
        stored = [100.0, 140.0, 120.0]
        f_deg, gap = 30.0, 1
    n = len(partners)
    c = H.cells(n + 2)
    c["pos"][0] = X
    for k, q in enumerate(partners):
        c["pos"][1 + k] = q
    f_dir = at(f_deg) - X
    c["pos"][n + 1] = X + 1.0 * f_dir
    c["vel"][n + 1] = -0.8 * f_dir
    c["maxConnections"][1:n + 1] = 1  # the partners are full: only X and F can fuse
    for k in range(n):
        H.bond(c, 0, 1 + k, 1.0, angle_a=stored[k], angle_b=360.0)
    p = H.params(MOTION_COLLISION, friction=0.0)
    o = run((64, 64), p, c, timestep=1)
    out, _ = o.get_data()
    pos = out["pos"].astype(np.float64)
    conn = out["connections"][0]
    assert int(out["numConnections"][0]) == n + 1
    targets = [int(conn[k]["cellIndex"]) for k in range(n + 1)]
    # the new bond sits at the gap's index, the later bonds move up by one
    expect_targets = [1 + k for k in range(gap)] + [n + 1] + [1 + k for k in range(gap, n)]
    assert targets == expect_targets
    # the angle bookkeeping with the end-of-step positions
    ang = lambda q: angle_deg(q - pos[0])
    a_prev, a_next, a_new = ang(pos[gap]), ang(pos[1 + gap]), ang(pos[n + 1])
    frac = ((a_new - a_prev) % 360.0) / ((a_next - a_prev) % 360.0)
    afp = stored[gap] * frac
    angles = [float(conn[k]["angleFromPrevious"]) for k in range(n + 1)]
    expect = list(stored[:gap]) + [afp, stored[gap] - afp] + list(stored[gap + 1:])
    np.testing.assert_allclose(angles, expect, atol=2e-3)
    assert abs(afp - 0.5 * stored[gap]) > 5.0  # not the half split
    np.testing.assert_allclose(sum(angles), 360.0, atol=1e-3)
    np.testing.assert_allclose(float(conn[gap]["distance"]), np.hypot(*(pos[n + 1] - pos[0])), atol=1e-5)


# --------------------------------------------------------------------------------------------
# binding-energy decay


@pytest.mark.parametrize("energy,kept", [(200.0, False), (120.0, True)])
def test_binding_energy_decay(oracle_lib, energy, kept):
    """A cell whose energy exceeds cellMaxBindingEnergy loses all its bonds at the end of the
    step, and its partners lose the bonds to it (CellProcessor.cuh:786-790 schedules
    deleteAllConnections)."""
    c = H.cells(3)
    c["pos"] = [[10.0, 10.0], [11.0, 10.0], [12.0, 10.0]]
    H.bond(c, 0, 1, 1.0)
    H.bond(c, 1, 2, 1.0, angle_a=180.0, angle_b=360.0)
    c["connections"][1, 0]["angleFromPrevious"] = 180.0
    c["energy"][0] = energy
    p = H.params(MOTION_COLLISION, friction=0.0, cellMaxBindingEnergy=150.0)
    out, _ = run((32, 32), p, c, timestep=1).get_data()
    if kept:
        assert out["numConnections"].tolist() == [1, 2, 1]
    else:
        # cell 0 drops its bond, cell 1 drops the bond to 0 and keeps the one to 2
        assert out["numConnections"].tolist() == [0, 1, 1]
        assert int(out["connections"][1, 0]["cellIndex"]) == 2
        assert float(out["connections"][1, 0]["angleFromPrevious"]) == 360.0


# --------------------------------------------------------------------------------------------
# the remaining helpers and rules


@pytest.mark.parametrize("precision", [32, 64])
def test_segment_crossing_against_orientation_predicate(oracle_lib, precision):
    """Math::crossing (Math.cuh:363-395) on random segments in general position agrees with the
    textbook test: the segments cross where both segment parameters of the intersection of the
    two lines lie strictly inside (0, 1).  Cases within 1e-3 of an end or nearly parallel are
    skipped (the reference's own tolerances decide those), and segments sharing an endpoint never
    cross."""
    import ctypes

    L = oracle_lib
    D2 = ctypes.c_double * 2
    L.oracle_crossing.argtypes = [ctypes.c_int, D2, D2, D2, D2]
    L.oracle_crossing.restype = ctypes.c_int
    rng = np.random.default_rng(17)
    checked = crossed = 0
    for _ in range(4000):
        s, e, os_, oe = (rng.uniform(-2, 2, size=2) for _ in range(4))
        if precision == 32:
            s, e, os_, oe = (np.float32(v).astype(np.float64) for v in (s, e, os_, oe))
        v1, v2 = e - s, oe - os_
        den = v1[0] * v2[1] - v1[1] * v2[0]
        if abs(den) < 1e-2:
            continue
        w = os_ - s
        t = (w[0] * v2[1] - w[1] * v2[0]) / den  # along (s, e)
        u = (w[0] * v1[1] - w[1] * v1[0]) / den  # along (os, oe)
        if min(abs(t), abs(1 - t), abs(u), abs(1 - u)) < 1e-3:
            continue
        expect = 0 < t < 1 and 0 < u < 1
        got = L.oracle_crossing(precision, D2(*s), D2(*e), D2(*os_), D2(*oe))
        assert bool(got) == expect, (s, e, os_, oe, t, u)
        checked += 1
        crossed += expect
    assert checked > 3000 and 300 < crossed < checked - 300
    a, b, c = D2(0.0, 0.0), D2(1.0, 1.0), D2(1.0, 0.0)
    assert L.oracle_crossing(precision, a, b, a, c) == 0  # shared endpoint


def test_fluid_pair_across_the_seam_equals_the_pair_inside(oracle_lib):
    """Fluid mode takes torus differences (CellProcessor.cuh:168-170): a pair 0.75 apart across
    the seam moves exactly like the same pair in the middle of the world."""
    h = 0.65
    p = H.params(MOTION_FLUID, friction=0.0, smoothingLength=h)
    res = []
    for xa, xb in ((0.25, 31.5), (16.25, 15.5)):
        c = H.cells(2)
        c["pos"] = [[xa, 10.0], [xb, 10.0]]
        c["vel"] = [[0.01, 0.02], [-0.03, 0.0]]
        c["maxConnections"] = 0
        out, _ = run((32, 32), p, c).get_data()
        res.append(out)
    np.testing.assert_array_equal(res[0]["vel"], res[1]["vel"])
    assert abs(res[0]["vel"][0, 0]) > 0.05  # the pair really interacted
    shift = np.array([16.0, 0.0])
    np.testing.assert_allclose(H.torus_diff(res[0]["pos"] + shift, res[1]["pos"], (32, 32)), 0.0, atol=1e-5)


def test_partial_absorption_and_last_absorbed_cell(oracle_lib):
    """radiationAbsorption 0.25 moves a quarter of the particle's energy into the cell; the
    particle remembers the cell and does not give to it again next step although it still heads
    into its slot (RadiationProcessor.cuh:108-165)."""
    c = H.cells(1)
    c["pos"][0] = [20.5, 20.5]
    c["color"][0] = 3
    pa = H.particles(1)
    pa["pos"][0] = [19.25, 20.5]
    pa["vel"][0] = [0.5, 0.0]
    pa["energy"][0] = 3.0
    p = H.params(MOTION_COLLISION, friction=0.0)
    p.radiationAbsorption[3] = 0.25
    o = run((64, 64), p, c, parts=pa, steps=1)
    out, parts = o.get_data()
    assert float(out["energy"][0]) == 100.75 and float(parts["energy"][0]) == 2.25
    o.step(1)  # the particle at 20.25 still targets slot 20 (20.25 + 0.5)
    out, parts = o.get_data()
    assert float(out["energy"][0]) == 100.75 and float(parts["energy"][0]) == 2.25
    o.step(1)  # now at 20.75: target 21.25, an empty slot
    out, parts = o.get_data()
    assert float(out["energy"][0]) == 100.75 and len(parts) == 1


def test_tiny_particle_gives_everything(oracle_lib):
    """A particle below 0.01 energy gives all of it, whatever the absorption factor, and dies
    (RadiationProcessor.cuh:149-158)."""
    c = H.cells(1)
    c["pos"][0] = [20.5, 20.5]
    pa = H.particles(1)
    pa["pos"][0] = [19.25, 20.5]
    pa["vel"][0] = [0.5, 0.0]
    pa["energy"][0] = 0.0078125
    p = H.params(MOTION_COLLISION, friction=0.0)
    for k in range(7):
        p.radiationAbsorption[k] = 0.25
    out, parts = run((64, 64), p, c, parts=pa, precision=32).get_data()
    assert len(parts) == 0
    assert out["energy"][0] == np.float32(100.0) + np.float32(0.0078125)


@pytest.mark.parametrize("closed", [True, False])
def test_triangle_suppresses_angle_forces(oracle_lib, closed):
    """Angle forces act only where the two partners of consecutive bonds are not bonded to each
    other (CellProcessor.cuh:446-458).  An equilateral triangle with wrong stored angles, springs
    at rest, stays put at an angle-force step; opened at one edge, it moves."""
    r3 = math.sqrt(3.0) / 2
    c = H.cells(3)
    c["pos"] = [[10.0, 10.0], [11.0, 10.0], [10.5, 10.0 + r3]]
    pos = c["pos"].astype(np.float64)
    dist = lambda a, b: float(np.hypot(*(pos[a] - pos[b])))
    # cell 0: partners 1 (90 deg) then 2 (150 deg): the true gap is 60, stored 100 / 260
    H.bond(c, 0, 1, dist(0, 1), angle_a=260.0, angle_b=360.0)
    H.bond(c, 0, 2, dist(0, 2), angle_a=100.0, angle_b=360.0)
    if closed:
        H.bond(c, 1, 2, dist(1, 2), angle_a=300.0, angle_b=60.0)
    p = H.params(MOTION_COLLISION, friction=0.0)
    out, _ = run((32, 32), p, c, timestep=0).get_data()
    moved = np.abs(out["pos"].astype(np.float64) - pos).max()
    if closed:
        assert moved < 1e-5
    else:
        assert moved > 5e-3


def test_fusion_into_cell_with_one_bond(oracle_lib):
    """With one bond, the new bond's angleFromPrevious is the clockwise angle from the old partner
    to the new one, and the old bond gets 360 minus that (CellConnectionProcessor.cuh:313-337)."""
    X = np.array([20.0, 20.0])
    c = H.cells(3)
    c["pos"] = [X, X + [0.0, -1.0], X + 1.0 * np.array([math.sin(math.radians(70)), -math.cos(math.radians(70))])]
    c["vel"][2] = -0.8 * (c["pos"][2] - X)
    c["maxConnections"][1] = 1
    H.bond(c, 0, 1, 1.0)
    o = run((64, 64), H.params(MOTION_COLLISION, friction=0.0), c, timestep=1)
    out, _ = o.get_data()
    pos = out["pos"].astype(np.float64)
    assert int(out["numConnections"][0]) == 2
    conn = out["connections"][0]
    assert [int(conn[0]["cellIndex"]), int(conn[1]["cellIndex"])] == [1, 2]
    diff = (angle_deg(pos[2] - pos[0]) - angle_deg(pos[1] - pos[0])) % 360.0
    assert 50 < diff < 90
    np.testing.assert_allclose([conn[0]["angleFromPrevious"], conn[1]["angleFromPrevious"]], [360.0 - diff, diff],
                               atol=2e-3)


@pytest.mark.parametrize("victim,expect", [(1, [240.0, 120.0]), (2, [100.0, 260.0]), (3, [220.0, 140.0])])
def test_deleted_bond_angle_goes_to_the_next_bond(oracle_lib, victim, expect):
    """deleteConnectionOneWay (CellConnectionProcessor.cuh:429-448): the deleted bond's
    angleFromPrevious is added to the bond after it (the first bond when the last is deleted).
    Cell 0 has partners 1, 2, 3 with stored angles 100, 140, 120; the victim partner's energy is
    above cellMaxBindingEnergy, so it drops all its bonds."""
    X = np.array([20.0, 20.0])
    at = lambda deg: X + np.array([math.sin(math.radians(deg)), -math.cos(math.radians(deg))])
    c = H.cells(4)
    c["pos"] = [X, at(0.0), at(120.0), at(240.0)]
    for k, a in enumerate((100.0, 140.0, 120.0)):
        H.bond(c, 0, 1 + k, 1.0, angle_a=a, angle_b=360.0)
    c["energy"][victim] = 500.0
    p = H.params(MOTION_COLLISION, friction=0.0, cellMaxBindingEnergy=150.0)
    out, _ = run((64, 64), p, c, timestep=1).get_data()
    conn = out["connections"][0]
    assert int(out["numConnections"][0]) == 2 and int(out["numConnections"][victim]) == 0
    remaining = [k for k in (1, 2, 3) if k != victim]
    assert [int(conn[0]["cellIndex"]), int(conn[1]["cellIndex"])] == remaining
    np.testing.assert_array_equal([conn[0]["angleFromPrevious"], conn[1]["angleFromPrevious"]], expect)
