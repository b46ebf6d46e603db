"""Closed-form checks of the oracle's cell radiation (CellProcessor::radiation,
CellProcessor.cuh:725-773, and RadiationProcessor::createEnergyParticle, RadiationProcessor.cuh:
248-364, without radiation sources or external energy).

The draws are predicted with the numpy Philox4x32-10 of test_oracle_living_friction (itself pinned
by the Random123 known-answer vectors), so each check names the exact cells that radiate and the
exact energies they lose:

  which cells radiate   draw(seed, timestep, id, 2) < radiationProb, barrier cells never
  factor                highRadiationFactor above highRadiationMinCellEnergy (strictly) plus
                        radiationCellAgeStrength above radiationMinCellAge (strictly)
  energy                loss = 2 * (E * factor / radiationProb) * draw(..., 3), in that order of
                        FP32 operations; at most E - 1; nothing if E <= 1; a loss below NEAR_ZERO
                        leaves the cell but creates no particle
  particle              velocity v = cell velocity * multiplier + (sin a, -cos a) * perturbation
                        with a = 360 draw(..., 4) degrees; created at cell + 1.5 v/|v| - v and
                        moved by v in the same step; id 2^63 | timestep << 32 | cell index (R18)
  age                   a cell ages by one per step (its age at the step is the age set plus the
                        steps before)
Radiation absorption is switched off (0 for every colour) where a particle must survive its step.
"""
import math

import numpy as np
import pytest

from alien_amd.abi import MOTION_COLLISION
from oracle.pyoracle import OracleWorld

from . import helpers as H
from .test_oracle_living_friction import draw

SEED = 77
CREATED = 1 << 63


def rad_params(age_strength=0.0, prob=1.0, high_factor=0.0, high_min=500.0, min_age=0, absorb=0.0, **kw):
    p = H.params(MOTION_COLLISION, cellMaxCollisionDistance=1.3, friction=0.0, **kw)
    p.radiationProb = prob
    for c in range(7):
        p.radiationCellAgeStrength[c] = age_strength
        p.highRadiationFactor[c] = high_factor
        p.highRadiationMinCellEnergy[c] = high_min
        p.radiationMinCellAge[c] = min_age
        p.radiationAbsorption[c] = absorb
    return p


def lattice(n_side, spacing=4.0, energy=100.0, age=5):
    c = H.cells(n_side * n_side)
    k = np.arange(n_side * n_side)
    c["pos"][:, 0] = 2.25 + spacing * (k % n_side)
    c["pos"][:, 1] = 2.75 + spacing * (k // n_side)
    c["energy"] = energy
    c["age"] = age
    c["color"] = k % 7
    return c


def loss32(e, factor, prob, u):
    f32 = np.float32
    loss = f32(e) * f32(factor)
    loss = f32(loss / f32(prob))
    return f32(f32(f32(2.0) * loss) * f32(u))


def run(cells, params, steps=1, precision=32, size=(64, 64)):
    o = OracleWorld(size, params, precision=precision, seed=SEED)
    o.set_data(cells, H.particles(0))
    o.step(steps)
    return o.get_data()


@pytest.mark.parametrize("precision", [32, 64])
def test_single_cell_emission_closed_form(oracle_lib, precision):
    c = H.cells(1)
    c["pos"][0] = (10.3, 20.7)
    c["vel"][0] = (0.2, -0.1)
    c["energy"] = 100.0
    c["age"] = 5
    c["color"] = 2
    p = rad_params(age_strength=0.01, radiationVelocityMultiplier=0.75, radiationVelocityPerturbation=0.5)
    cells, parts = run(c, p, precision=precision)
    assert len(parts) == 1
    q = parts[0]
    cid = int(c["id"][0])
    u1, u2 = draw(SEED, 0, cid, 3), draw(SEED, 0, cid, 4)
    loss = float(loss32(100.0, np.float32(0.01), np.float32(1.0), u1)) if precision == 32 else \
        2.0 * (100.0 * float(np.float32(0.01)) / 1.0) * u1
    assert int(q["id"]) == CREATED | 0  # timestep 0, cell index 0
    assert q["color"] == 2
    if precision == 32:
        assert q["energy"] == np.float32(loss) and cells["energy"][0] == np.float32(np.float32(100.0) - np.float32(loss))
    else:
        assert math.isclose(float(q["energy"]), loss, rel_tol=1e-6)
        assert math.isclose(float(cells["energy"][0]), 100.0 - loss, rel_tol=1e-6)
    a = math.radians(360.0 * u2)
    v = (0.2 * 0.75 + 0.5 * math.sin(a), -0.1 * 0.75 - 0.5 * math.cos(a))
    tol = 1e-6 if precision == 32 else 1e-7
    assert abs(float(q["vel"][0]) - v[0]) < tol and abs(float(q["vel"][1]) - v[1]) < tol
    # created 1.5 ahead and one velocity back, then moved by its velocity: 1.5 ahead of the cell
    n = math.hypot(*v)
    expect = (10.3 + 1.5 * v[0] / n, 20.7 + 1.5 * v[1] / n)
    assert abs(float(q["pos"][0]) - expect[0]) < 1e-5 and abs(float(q["pos"][1]) - expect[1]) < 1e-5


def test_which_cells_radiate_and_how_much(oracle_lib):
    prob, strength = 0.3, 0.002
    c = lattice(12)
    p = rad_params(age_strength=strength, prob=prob)
    cells, parts = run(c, p, size=(48, 48))
    ids = c["id"].astype(np.uint64)
    expect = [i for i in range(len(c)) if draw(SEED, 0, int(ids[i]), 2) < np.float32(prob)]
    assert 20 < len(expect) < 70
    got = sorted(int(q["id"]) & 0xFFFFFFFF for q in parts)
    assert got == expect
    assert all(int(q["id"]) >> 32 == 1 << 31 for q in parts)  # created at timestep 0
    by_index = {int(q["id"]) & 0xFFFFFFFF: q for q in parts}
    final = {int(x["id"]): float(x["energy"]) for x in cells}
    for i in range(len(c)):
        cid = int(ids[i])
        if i in by_index:
            loss = loss32(100.0, np.float32(strength), np.float32(prob), draw(SEED, 0, cid, 3))
            assert by_index[i]["energy"] == loss
            assert by_index[i]["color"] == c["color"][i]
            assert final[cid] == np.float32(np.float32(100.0) - loss)
        else:
            assert final[cid] == np.float32(100.0)


def test_conditions_on_the_cell(oracle_lib):
    c = lattice(3)
    # 0: barrier; 1: energy 1 (not above 1); 2: energy 1.5, huge factor (loss capped at 0.5);
    # 3: age 5 = radiationMinCellAge (no age factor), energy 400 (no high radiation);
    # 4: energy 600 above highRadiationMinCellEnergy, age 5 (high factor only);
    # 5: age 6 above the minimum, energy 400 (age factor only); 6..8: reference cells
    c["barrier"][0] = 1
    c["energy"][1] = 1.0
    c["energy"][2] = 1.5
    c["age"] = 6
    c["age"][[3, 4]] = 5
    c["energy"][3] = 400.0
    c["energy"][4] = 600.0
    c["energy"][5] = 400.0
    c["color"] = 0
    p = rad_params(age_strength=0.001, prob=1.0, high_factor=0.01, high_min=500.0, min_age=5)
    cells, parts = run(c, p, size=(48, 48))
    byid = {int(x["id"]): x for x in cells}
    emitted = {int(q["id"]) & 0xFFFFFFFF: q for q in parts}
    assert 0 not in emitted and 1 not in emitted and 3 not in emitted
    assert byid[int(c["id"][0])]["energy"] == np.float32(100.0)
    assert byid[int(c["id"][1])]["energy"] == np.float32(1.0)
    assert byid[int(c["id"][3])]["energy"] == np.float32(400.0)
    # the cap: huge loss (2 * 1.5 * factor / 1 * u) limited to E - 1
    p2 = rad_params(age_strength=10.0, prob=1.0, min_age=5)
    cells2, parts2 = run(c, p2, size=(48, 48))
    q2 = {int(q["id"]) & 0xFFFFFFFF: q for q in parts2}
    assert q2[2]["energy"] == np.float32(0.5)
    assert {int(x["id"]): float(x["energy"]) for x in cells2}[int(c["id"][2])] == 1.0
    # high factor only for cell 4, age factor only for cell 5 (and the reference cells)
    for i, factor in ((4, 0.01), (5, 0.001), (6, 0.001)):
        cid = int(c["id"][i])
        loss = loss32(float(c["energy"][i]), np.float32(factor), np.float32(1.0), draw(SEED, 0, cid, 3))
        assert emitted[i]["energy"] == loss, i
    # both factors add: energy 600, age 6
    c3 = c.copy()
    c3["age"][4] = 6
    cells3, parts3 = run(c3, p, size=(48, 48))
    q3 = {int(q["id"]) & 0xFFFFFFFF: q for q in parts3}
    f = np.float32(np.float32(0.01) + np.float32(0.001))
    assert q3[4]["energy"] == loss32(600.0, f, np.float32(1.0), draw(SEED, 0, int(c["id"][4]), 3))


def test_tiny_loss_creates_no_particle(oracle_lib):
    c = lattice(2)
    p = rad_params(age_strength=1e-9, prob=1.0)
    cells, parts = run(c, p, size=(48, 48))
    assert len(parts) == 0
    for x, cid in zip(cells, c["id"]):
        loss = loss32(100.0, np.float32(1e-9), np.float32(1.0), draw(SEED, 0, int(cid), 3))
        assert 0 < loss < 1e-5
        assert x["energy"] == np.float32(np.float32(100.0) - loss)


def test_age_grows_by_one_per_step(oracle_lib):
    c = lattice(2, age=0)
    p = rad_params(age_strength=0.001, prob=1.0, min_age=1)
    # ages 0 then 1: not above the minimum 1 in steps 0 and 1; age 2 in step 2
    _, parts = run(c, p, steps=2, size=(48, 48))
    assert len(parts) == 0
    _, parts = run(c, p, steps=3, size=(48, 48))
    assert sorted(int(q["id"]) for q in parts) == [CREATED | (2 << 32) | i for i in range(4)]


def test_radiated_particles_are_absorbed_like_any_other(oracle_lib):
    # with absorption on, energy only moves between cells and particles (no loss below NEAR_ZERO
    # at these sizes), so the total is conserved over many steps
    c = lattice(10, spacing=2.0)
    p = rad_params(age_strength=0.005, prob=0.2, absorb=1.0)
    o = OracleWorld((48, 48), p, precision=64, seed=SEED)
    o.set_data(c, H.particles(0))
    e0 = o.totals()
    o.step(60)
    e1 = o.totals()
    assert e1["particleEnergy"] > 0
    total0 = e0["cellEnergy"] + e0["particleEnergy"]
    total1 = e1["cellEnergy"] + e1["particleEnergy"]
    assert math.isclose(total0, total1, rel_tol=1e-9)


# --------------------------------------------------------------------------------------------
# particle splitting (RadiationProcessor.cuh:172-204): after absorption, with probability 0.01
# (draw(seed, timestep, particle id, 5)), a particle above particleSplitEnergy of its colour halves
# its energy; the halves go 1/5 of the unit vector of angle 360 draw(..., 6) either side of its
# (moved) position, with velocities v -/+ u radiationVelocityPerturbation / (E/2 + 1)

SPLIT = 1 << 62


def split_params(threshold, **kw):
    p = rad_params(**kw)
    for c in range(7):
        p.particleSplitEnergy[c] = threshold
    return p


def first_split_step(pid, limit=5000):
    return next(t for t in range(limit) if draw(SEED, t, pid, 5) < np.float32(0.01))


@pytest.mark.parametrize("precision", [32, 64])
def test_split_closed_form(oracle_lib, precision):
    q = H.particles(1)
    q["pos"][0] = (20.25, 30.75)
    q["vel"][0] = (0.3, -0.2)
    q["energy"] = 10.0
    q["color"] = 3
    pid = int(q["id"][0])
    t = first_split_step(pid)
    o = OracleWorld((64, 64), split_params(5.0), precision=precision, seed=SEED)
    o.set_data(H.cells(0), q)
    o.timestep = t
    o.step(1)
    _, parts = o.get_data()
    assert len(parts) == 2
    byid = {int(x["id"]): x for x in parts}
    half = byid[CREATED | SPLIT | (t << 32) | 0]
    orig = byid[pid]
    a = math.radians(360.0 * draw(SEED, t, pid, 6))
    u = (math.sin(a), -math.cos(a))
    moved = (20.25 + 0.3, 30.75 - 0.2)
    k = 0.5 / (5.0 + 1.0)
    tol = 4e-6  # (both builds return float32 transfer objects)
    for x, sign in ((orig, -1.0), (half, 1.0)):
        assert float(x["energy"]) == 5.0 and x["color"] == 3
        assert abs(float(x["pos"][0]) - (moved[0] + sign * u[0] / 5)) < tol
        assert abs(float(x["pos"][1]) - (moved[1] + sign * u[1] / 5)) < tol
        assert abs(float(x["vel"][0]) - (0.3 + sign * u[0] * k)) < tol
        assert abs(float(x["vel"][1]) - (-0.2 + sign * u[1] * k)) < tol


def test_which_particles_split(oracle_lib):
    n = 30
    q = H.particles(n * n)
    k = np.arange(n * n)
    q["pos"][:, 0] = 1.5 + 4.0 * (k % n)
    q["pos"][:, 1] = 1.5 + 4.0 * (k // n)
    q["energy"] = np.where(k % 3 == 0, 2.0, 8.0)  # every third below the threshold 2.0 (strict)
    p = split_params(2.0)
    for t in (0, 7):
        o = OracleWorld((128, 128), p, seed=SEED)
        o.set_data(H.cells(0), q)
        o.timestep = t
        o.step(1)
        _, parts = o.get_data()
        expect = sorted(int(q["id"][i]) for i in k
                        if q["energy"][i] > 2.0 and draw(SEED, t, int(q["id"][i]), 5) < np.float32(0.01))
        # the particles that split kept half their energy; each half is a new particle
        halved = sorted(int(x["id"]) for x in parts if not int(x["id"]) & CREATED and x["energy"] == np.float32(4.0))
        halves = [x for x in parts if int(x["id"]) & SPLIT]
        assert halved == expect and len(halves) == len(expect) and len(parts) == n * n + len(expect)
        assert all(x["energy"] == np.float32(4.0) for x in halves)
    assert len(expect) > 0


def test_split_and_radiation_together(oracle_lib):
    # a particle that splits at step t among radiating cells: one half with the split id of step t
    # (the halves, 0.4 apart, fuse again in the next step), and the radiated particles with theirs
    c = lattice(2)
    q = H.particles(1)
    q["pos"][0] = (30.5, 30.5)
    q["energy"] = 10.0
    pid = int(q["id"][0])
    t = first_split_step(pid)
    p = split_params(5.0, age_strength=0.001, prob=1.0)
    o = OracleWorld((48, 48), p, seed=SEED)
    o.set_data(c, q)
    o.timestep = t
    o.step(1)
    _, parts = o.get_data()
    ids = [int(x["id"]) for x in parts]
    # the particle is map entry 4, behind the four radiated ones (bins in row order)
    assert [i for i in ids if i & SPLIT] == [CREATED | SPLIT | (t << 32) | 4]
    radiated = [i for i in ids if i & CREATED and not i & SPLIT]
    assert sorted(radiated) == [CREATED | (t << 32) | k for k in range(4)]
