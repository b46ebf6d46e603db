"""Closed-form checks of the oracle's inner friction, force decay and living states.

  counter-based draws   Philox4x32-10 against the Random123 known-answer vectors, restated here in
                        numpy (independently of the oracle's C++), then used to predict exactly
                        which cells the force decay selects
  inner friction        CellProcessor.cuh:678-705 with the Jacobi schedule of rule R11: each bond
                        relaxes each of its cells twice per pass toward the partner's start-of-pass
                        velocity, v <- v (1 - f) + (v + w)/2 f; every third step only
                        (SimulationKernelsLauncher.cu:32, 69-71); barrier partners skipped
  force decay           CellProcessor.cuh:367-386: |F| of the neighbour pass above cellMaxForce and
                        a draw below cellMaxForceDecayProb delete all bonds of the cell
  living states         CellProcessor.cuh:493-523 (Detaching marks of overstretched bonds),
                        561-594 (aging), 596-676 (transitions), 775-858 (decay to Dying), under
                        cellDeathConsequences = None
"""
import numpy as np
import pytest

from alien_amd.abi import MOTION_COLLISION
from oracle.pyoracle import OracleWorld, lib

from . import helpers as H

READY, UNDER_CONSTRUCTION, ACTIVATING, DETACHING, REVIVING, DYING = range(6)  # CellFunctionConstants.h:24-33

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85


def philox4x32_10(counter, key):
    """Philox4x32-10 (Salmon et al., SC 2011) on python ints."""
    c = [int(x) & 0xFFFFFFFF for x in counter]
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    for _ in range(10):
        p0, p1 = M0 * c[0], M1 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k0) & 0xFFFFFFFF, p1 & 0xFFFFFFFF, ((p0 >> 32) ^ c[3] ^ k1) & 0xFFFFFFFF,
             p0 & 0xFFFFFFFF]
        k0, k1 = (k0 + W0) & 0xFFFFFFFF, (k1 + W1) & 0xFFFFFFFF
    return c


def draw(seed, timestep, cell_id, purpose):
    c = philox4x32_10([cell_id & 0xFFFFFFFF, cell_id >> 32, timestep & 0xFFFFFFFF, purpose],
                      [seed & 0xFFFFFFFF, seed >> 32])
    return (c[0] >> 8) / 16777216.0


# Random123 kat_vectors for philox4x32_10: (counter, key, expected output)
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


def test_philox_known_answers(oracle_lib):
    import ctypes

    for counter, key, expect in KAT:
        assert philox4x32_10(counter, key) == expect
        c = (ctypes.c_uint32 * 4)(*counter)
        lib().oracle_philox4x32_10(c, key[0], key[1])
        assert list(c) == expect
    rng = np.random.default_rng(4)
    for _ in range(200):
        seed, ts, cid = (int(x) for x in rng.integers(0, 2**63, size=3))
        assert lib().oracle_uniform_draw(seed, ts, cid, 1) == draw(seed, ts, cid, 1)
    u = np.array([draw(7, t, 1000 + t, 1) for t in range(20000)])
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01


# --------------------------------------------------------------------------------------------
# inner friction


def relax(v, w, f):
    avg = (v + w) / 2.0
    return v * (1.0 - f) + avg * f


@pytest.mark.parametrize("timestep", [0, 1, 3])
def test_inner_friction_two_cells(oracle_lib, timestep):
    """Two bonded cells with stiffness 0 (no connection forces): every third step each relaxes twice
    toward the other's start-of-pass velocity, v_a' = v_a a^2 + v_b (1 - a^2) with a = 1 - f/2;
    the pair's momentum is unchanged; on other steps nothing happens."""
    f = 0.3
    c = H.cells(2)
    c["pos"] = [[10.0, 10.0], [11.0, 10.0]]
    c["vel"] = [[0.01, 0.05], [0.01, -0.03]]
    c["stiffness"] = 0.0
    H.bond(c, 0, 1, 1.0)
    p = H.params(MOTION_COLLISION, friction=0.0, innerFriction=f)
    o = OracleWorld((32, 32), p, precision=64)
    o.set_data(c, H.particles(0))
    o.timestep = timestep
    o.step(1)
    out, _ = o.get_data()
    v = c["vel"].astype(np.float64)
    if timestep % 3 == 0:
        va = relax(relax(v[0], v[1], f), v[1], f)
        vb = relax(relax(v[1], v[0], f), v[0], f)
        a2 = (1 - f / 2) ** 2
        np.testing.assert_allclose(va, a2 * v[0] + (1 - a2) * v[1], atol=1e-15)
    else:
        va, vb = v[0], v[1]
    np.testing.assert_allclose(out["vel"][0], va, atol=1e-8)
    np.testing.assert_allclose(out["vel"][1], vb, atol=1e-8)
    np.testing.assert_allclose(out["vel"].astype(np.float64).sum(axis=0), v.sum(axis=0), atol=1e-8)


def test_inner_friction_chain_order_and_barriers(oracle_lib):
    """A chain A-B-C plus a barrier D bonded to B: B relaxes toward A's and C's start-of-pass
    velocities in bond order, twice (own visits, then the partners' visits), skipping D; the
    barrier's velocity is untouched; friction is applied after inner friction."""
    f, fr = 0.3, 0.01
    c = H.cells(4)
    c["pos"] = [[10.0, 10.0], [11.0, 10.0], [12.0, 10.0], [11.0, 11.0]]
    c["vel"] = [[0.02, 0.0], [-0.01, 0.01], [0.0, -0.02], [0.03, 0.03]]
    c["stiffness"] = 0.0
    c["barrier"][3] = 1
    H.bond(c, 1, 0, 1.0, angle_a=180.0)
    H.bond(c, 1, 2, 1.0, angle_a=90.0)
    H.bond(c, 1, 3, 1.0, angle_a=90.0)
    p = H.params(MOTION_COLLISION, friction=fr, innerFriction=f, cellMaxCollisionDistance=0.5)
    o = OracleWorld((32, 32), p, precision=64)
    o.set_data(c, H.particles(0))
    o.step(1)
    out, _ = o.get_data()
    v0 = c["vel"].astype(np.float64)
    vb = v0[1]
    for _visit in range(2):
        for j in (0, 2):  # bond order of B: A, C, (D is a barrier)
            vb = relax(vb, v0[j], f)
    va = relax(relax(v0[0], v0[1], f), v0[1], f)
    vc = relax(relax(v0[2], v0[1], f), v0[1], f)
    np.testing.assert_allclose(out["vel"][:3], np.array([va, vb, vc]) * (1 - fr), atol=1e-8)
    np.testing.assert_array_equal(out["vel"][3], c["vel"][3])


# --------------------------------------------------------------------------------------------
# force decay


def pushed_pairs(n, gap):
    """n bonded pairs (A at x, B at x + 1) far apart, each with a free cell F `gap` to the left of
    A: A feels F's repulsion (B is bonded to A, so A and B exert no collision force), B feels none."""
    c = H.cells(3 * n)
    for k in range(n):
        x, y = 10.0 + 8.0 * (k % 60), 10.0 + 8.0 * (k // 60)
        c["pos"][3 * k:3 * k + 3] = [[x, y], [x + 1.0, y], [x - gap, y]]
        H.bond(c, 3 * k, 3 * k + 1, 1.0)
        c["maxConnections"][3 * k + 2] = 0
    return c


@pytest.mark.parametrize("max_force,expect_any", [(0.0, True), (0.063, True), (0.07, False)])
def test_force_decay_selects_the_predicted_cells(oracle_lib, max_force, expect_any):
    """With |F| above cellMaxForce, a cell's bonds break exactly when its draw
    (seed, timestep, id, purpose 1) is below cellMaxForceDecayProb; above |F| nothing breaks."""
    n, prob, seed, gap = 1200, 0.3, 0xBEEF, 0.9
    c = pushed_pairs(n, gap)
    force = (1.3 - gap) * 0.08  # each visit's repulsion; A gets 2x from F's and its own visit
    assert 2 * force > 0.0 and 2 * force < 0.065
    p = H.params(MOTION_COLLISION, friction=0.0, cellMaxForceDecayProb=prob)
    for k in range(7):
        p.cellMaxForce[k] = max_force
    o = OracleWorld((512, 512), p, precision=64, seed=seed)
    o.set_data(c, H.particles(0))
    o.timestep = 4
    o.step(1)
    out, _ = o.get_data()
    broken = out["numConnections"][0::3] == 0
    if not expect_any:
        assert not broken.any()
        return
    ids = c["id"][0::3].astype(np.uint64)
    predicted = np.array([draw(seed, 4, int(i), 1) < prob for i in ids])
    np.testing.assert_array_equal(broken, predicted)
    assert abs(predicted.mean() - prob) < 0.05
    np.testing.assert_array_equal(out["numConnections"][1::3] == 0, predicted)  # partners drop the bond too


# --------------------------------------------------------------------------------------------
# living states


def states_after(c, steps, p=None, parts=None, timestep=1):
    o = OracleWorld((64, 64), p or H.params(MOTION_COLLISION, friction=0.0), precision=64)
    o.set_data(c, H.particles(0) if parts is None else parts)
    o.timestep = timestep
    res = []
    for _ in range(steps):
        o.step(1)
        out, _ = o.get_data()
        res.append(out)
    return res


def chain(n, x0=10.0):
    c = H.cells(n)
    c["pos"] = [[x0 + k, 20.0] for k in range(n)]
    for k in range(n - 1):
        H.bond(c, k, k + 1, 1.0, angle_a=180.0 if k > 0 else 360.0)
    return c


def test_low_energy_cell_dies_and_detaches_its_partners(oracle_lib):
    """A cell below cellMinEnergy becomes Dying at the end of the step and stays so; its partners
    are Detaching for that step and Ready after the next (CellProcessor.cuh:836-850, :642-648)."""
    c = chain(3)
    c["energy"][0] = 40.0
    r = states_after(c, 3)
    assert r[0]["livingState"].tolist() == [DYING, DETACHING, READY]
    assert r[1]["livingState"].tolist() == [DYING, READY, READY]
    assert r[2]["livingState"].tolist() == [DYING, READY, READY]


def test_absorbed_energy_counts_before_the_decay_check(oracle_lib):
    """The particle absorbed in the step lifts the cell above cellMinEnergy before decay runs."""
    c = H.cells(1)
    c["pos"][0] = [20.5, 20.5]
    c["energy"][0] = 40.0
    pa = H.particles(1)
    pa["pos"][0] = [19.25, 20.5]
    pa["vel"][0] = [0.5, 0.0]
    pa["energy"][0] = 30.0
    r = states_after(c, 1, parts=pa)
    assert r[0]["livingState"][0] == READY and r[0]["energy"][0] == 70.0
    assert states_after(c, 1)[0]["livingState"][0] == DYING


def test_dying_partner_stays_dying(oracle_lib):
    """A partner that is already Dying is not turned into Detaching by a newly dying cell."""
    c = chain(2)
    c["energy"][0] = 40.0
    c["livingState"][1] = DYING
    r = states_after(c, 2)
    assert r[0]["livingState"].tolist() == [DYING, DYING]
    assert r[1]["livingState"].tolist() == [DYING, DYING]


def test_activation_spreads_through_its_creature(oracle_lib):
    """UnderConstruction becomes Activating next to an Activating cell of the same creature, and
    Activating becomes Ready one step later: the activation walks along a creature's bonds one
    cell per step and does not cross into another creature (CellProcessor.cuh:622-636)."""
    c = chain(4)
    c["livingState"] = [ACTIVATING, UNDER_CONSTRUCTION, UNDER_CONSTRUCTION, UNDER_CONSTRUCTION]
    c["creatureId"] = [7, 7, 7, 8]
    r = states_after(c, 4)
    assert r[0]["livingState"].tolist() == [READY, ACTIVATING, UNDER_CONSTRUCTION, UNDER_CONSTRUCTION]
    assert r[1]["livingState"].tolist() == [READY, READY, ACTIVATING, UNDER_CONSTRUCTION]
    assert r[2]["livingState"].tolist() == [READY, READY, READY, UNDER_CONSTRUCTION]
    assert r[3]["livingState"].tolist() == [READY, READY, READY, UNDER_CONSTRUCTION]


def test_aging_and_activation_time(oracle_lib):
    """Every step ages non-barrier cells by one and counts a Ready cell's activationTime down to 0
    (an UnderConstruction cell keeps its count); barrier cells do not age and are Ready; Reviving
    and Detaching become Ready (CellProcessor.cuh:561-594, :630-645)."""
    c = H.cells(5)
    c["pos"] = [[10.0, 10.0], [20.0, 10.0], [30.0, 10.0], [40.0, 10.0], [50.0, 10.0]]
    c["age"] = [3, 0, 9, 1, 1]
    c["activationTime"] = [5, 5, 2, 0, 0]
    c["livingState"] = [READY, UNDER_CONSTRUCTION, REVIVING, DETACHING, DETACHING]
    c["barrier"][4] = 1
    r = states_after(c, 3)
    out = r[2]
    assert out["age"].tolist() == [6, 3, 12, 4, 1]
    assert out["activationTime"].tolist() == [2, 5, 0, 0, 0]  # the Reviving cell counted 2 steps as Ready
    assert out["livingState"].tolist() == [READY, UNDER_CONSTRUCTION, READY, READY, READY]
    assert r[0]["activationTime"].tolist() == [4, 5, 2, 0, 0]  # Reviving at aging time: no count


def test_overstretched_bond_detaches_its_partners(oracle_lib):
    """checkConnections sets all partners of a cell with an overstretched bond to Detaching before
    aging (no activation count that step) and the transition turns them Ready in the same step, a
    Dying cell included (CellProcessor.cuh:515-520).  Chain D-A-B-C with only B-C overstretched:
    B and C drop all their bonds, A (a partner of B) loses its count, D keeps counting."""
    c = H.cells(4)
    c["pos"] = [[7.0, 10.0], [8.0, 10.0], [9.0, 10.0], [12.7, 10.0]]
    H.bond(c, 0, 1, 1.0)
    H.bond(c, 1, 2, 1.0, angle_a=180.0)
    H.bond(c, 2, 3, 3.7, angle_a=180.0)
    c["livingState"][2] = DYING
    c["activationTime"] = 4
    out = states_after(c, 1)[0]
    assert out["numConnections"].tolist() == [1, 1, 0, 0]
    assert out["livingState"].tolist() == [READY, READY, READY, READY]
    assert out["activationTime"].tolist() == [3, 4, 4, 4]
