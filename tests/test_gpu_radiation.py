"""Parity of the HIP path with the oracle for cell radiation (CellProcessor.cuh:725-773): the
emission kernel, the staging slots that bring the emitted particles into the step's particle map,
the ids of created particles, the growth of a world's particle room, and the paths that switch
radiation on (through the C ABI; needs an MI355X).  Bar: cells and particles byte-identical to
the FP32 oracle after every compared step."""
import numpy as np
import pytest

from alien_amd import engine, worlds
from alien_amd.abi import MOTION_COLLISION, MOTION_FLUID
from oracle.pyoracle import OracleWorld, from_world

from . import helpers as H
from .test_gpu_living_friction import living_mix, run_checkpoints
from .test_gpu_parity import CPU_THREADS, assert_identical, say

pytestmark = pytest.mark.gpu

CREATED = 1 << 63


def radiating(w, strength=0.002, prob=0.2, absorb=None, high=None):
    p = w.params.copy()
    p.radiationProb = prob
    for c in range(7):
        p.radiationCellAgeStrength[c] = strength
        if absorb is not None:
            p.radiationAbsorption[c] = absorb
        if high is not None:
            p.highRadiationFactor[c] = high[0]
            p.highRadiationMinCellEnergy[c] = high[1]
    return worlds.World(w.name + "_rad", w.size, p, w.cells, w.particles, w.steps, w.seed)


@pytest.mark.parametrize("motion", [MOTION_FLUID, MOTION_COLLISION])
def test_radiation_bit_exact(motion):
    w = radiating(worlds.random_mix(401 + motion, motion=motion, num_particles=300), high=(0.01, 120.0))
    say(f"radiation: {w.name} motion={motion} cells={len(w.cells)}")
    sim, _ = run_checkpoints(w, [1, 2, 5, 20, 40])
    parts = sim.get_data()[1]
    assert (parts["id"] >= CREATED).sum() > 50 and (parts["id"] < CREATED).any()


def test_particle_room_grows():
    """Every cell radiates every step and nothing absorbs: the live particles outgrow the room the
    world was built with (8 steps of emissions) and the particle arrays grow between steps."""
    w = radiating(worlds.random_mix(411, motion=MOTION_COLLISION, num_particles=0), strength=0.001, prob=1.0,
                  absorb=0.0)
    say(f"particle room: {w.name} cells={len(w.cells)}")
    sim, _ = run_checkpoints(w, [3, 9, 21, 33])
    assert len(sim.get_data()[1]) > 5 * len(w.cells)


def test_created_particles_survive_get_and_set_data():
    w = radiating(worlds.random_mix(421, motion=MOTION_FLUID, num_particles=100))
    say(f"round trip: {w.name}")
    sim = engine.Simulation.from_world(w)
    sim.step(12)
    cells, parts = sim.get_data()
    assert (parts["id"] >= CREATED).any()
    again = engine.Simulation(w.size, w.params, cells, parts, timestep=12, seed=w.seed)
    ref = OracleWorld(w.size, w.params, seed=w.seed, threads=CPU_THREADS)
    ref.set_data(cells, parts)
    ref.timestep = 12
    again.step(10)
    ref.step(10)
    assert_identical(again.get_data(), ref.get_data(), "after get_data / set_data")


def test_set_params_starts_radiation():
    """Radiation switched on by set_params: a world without tracked living states is rebuilt
    around the new parameters, a world with them keeps its set and gets its staging area and
    particle room before the next step."""
    for w in (worlds.random_mix(431, motion=MOTION_FLUID, num_particles=50), living_mix(432, MOTION_COLLISION)):
        say(f"set_params: {w.name}")
        sim = engine.Simulation.from_world(w)
        ref = from_world(w, threads=CPU_THREADS)
        sim.step(3)
        ref.step(3)
        p = radiating(w, strength=0.003, prob=0.3).params
        sim.set_params(p)
        ref.set_params(p)
        sim.step(7)
        ref.step(7)
        got = sim.get_data()
        assert_identical(got, ref.get_data(), f"{w.name} after radiation starts")
        assert (got[1]["id"] >= CREATED).any()


@pytest.mark.parametrize("graph", [True, False])
def test_batch_with_radiation_equals_alone(graph):
    ws = [radiating(worlds.random_mix(441, motion=MOTION_FLUID, num_particles=80)), worlds.random_mix(442),
          radiating(living_mix(443, MOTION_COLLISION), strength=0.004, prob=0.5, absorb=0.0)]
    b = engine.Batch()
    for w in ws:
        b.add(w)
    b.set_use_graph(graph)
    b.step(6)
    b.step(7)
    for i, w in enumerate(ws):
        ref = from_world(w, threads=CPU_THREADS)
        ref.step(13)
        assert_identical(b.get_data(i), ref.get_data(), f"batch world {i} ({w.name}) graph={graph}")


def test_server_with_radiation_equals_alone():
    """Radiating and splitting jobs through the job server (set rebuilds by device copies as jobs
    come and go, split halves waiting between steps carried along, worlds frozen between chunks):
    each snapshot equal to the oracle's run of the job alone."""
    ws = [radiating(worlds.random_mix(451, motion=MOTION_FLUID, num_particles=60), strength=0.003, prob=0.3),
          worlds.random_mix(452),
          radiating(living_mix(453, MOTION_COLLISION), strength=0.002, prob=0.5, absorb=0.0),
          splitting(worlds.random_mix(454, motion=MOTION_COLLISION, num_particles=500), 1.5)]
    steps = [21, 6, 14, 26]
    srv = engine.Server(budget_bytes=1 << 30, chunk_steps=4)
    jobs = [srv.submit(w.size, w.params, w.cells, w.particles, s, seed=w.seed) for w, s in zip(ws, steps)]
    for j in jobs:
        srv.wait(j, timeout=120)
    for j, w, s in zip(jobs, ws, steps):
        ref = from_world(w, threads=CPU_THREADS)
        ref.step(s)
        assert_identical(srv.snapshot(j), ref.get_data(), f"server job {w.name}")
    srv.close()


def splitting(w, threshold):
    p = w.params.copy()
    for c in range(7):
        p.particleSplitEnergy[c] = threshold
    return worlds.World(w.name + "_split", w.size, p, w.cells, w.particles, w.steps, w.seed)


@pytest.mark.parametrize("motion", [MOTION_FLUID, MOTION_COLLISION])
def test_splitting_bit_exact(motion):
    """Particles above the split energy split with probability 0.01 per step (RadiationProcessor.cuh:
    172-204): the halves, staged at the end of a step, enter the next map; GPU equal to the oracle."""
    w = splitting(worlds.random_mix(461 + motion, motion=motion, num_particles=600), 1.5)
    say(f"splitting: {w.name} motion={motion}")
    sim, _ = run_checkpoints(w, [1, 2, 7, 25, 45])
    ids = sim.get_data()[1]["id"].astype(np.uint64)
    assert ((ids & np.uint64(1 << 62)) != 0).sum() > 0


def sparse_lattice(n=32, spacing=6.0):
    """Cells on a sparse lattice, slow and aged: the particles they emit spread out and rarely fuse."""
    c = H.cells(n * n)
    k = np.arange(n * n)
    c["pos"][:, 0] = 1.5 + spacing * (k % n)
    c["pos"][:, 1] = 1.5 + spacing * (k // n)
    c["vel"] = np.random.default_rng(5).normal(0, 0.05, size=(n * n, 2)).astype(np.float32)
    c["age"] = 3
    p = H.params(MOTION_COLLISION, cellMaxCollisionDistance=1.3)
    return worlds.World("sparse_lattice", (int(n * spacing), int(n * spacing)), p, c, H.particles(0), 10, 9)


def test_splitting_with_radiation_and_room_growth():
    """Every cell radiates every step, nothing absorbs, particles split above a low energy: both
    kinds of created particles, the per-step room check, and array growth while split halves wait
    in their staging slots for the next step."""
    base = sparse_lattice()
    w = splitting(radiating(base, strength=0.01, prob=1.0, absorb=0.0), 0.05)
    say(f"splitting with radiation: {w.name} cells={len(w.cells)}")
    sim, _ = run_checkpoints(w, [1, 3, 10, 16])
    parts = sim.get_data()[1]
    ids = parts["id"].astype(np.uint64)
    split = (ids & np.uint64(1 << 62)) != 0
    assert split.sum() > 100 and ((ids >= np.uint64(CREATED)) & ~split).sum() > 1000
    # the room the world was built with, 2 * 8 * cells, was outgrown: the arrays grew
    assert 2 * (len(parts) + len(w.cells)) > 2 * 8 * len(w.cells)


def test_living_states_without_bond_work():
    """A radiating world whose cells have no bonds and no bond slots (no bond work): its living
    states (every state, activation counts, decay to Dying as radiation drains cells below
    cellMinEnergy) advance in one pass per cell, equal to the oracle's."""
    base = sparse_lattice(n=24, spacing=4.0)
    c = base.cells.copy()
    rng = np.random.default_rng(481)
    c["maxConnections"] = 0
    c["livingState"] = rng.integers(0, 6, size=len(c)).astype(np.uint8)
    c["activationTime"] = np.where(rng.random(len(c)) < 0.4, rng.integers(1, 6, size=len(c)), 0).astype(np.uint32)
    c["energy"] = rng.uniform(45.0, 60.0, size=len(c)).astype(np.float32)
    w = radiating(worlds.World("unbonded_living", base.size, base.params, c, base.particles, 10, base.seed),
                  strength=0.02, prob=0.5)
    say(f"living states without bond work: cells={len(c)}")
    sim, _ = run_checkpoints(w, [1, 2, 6, 15])
    st = sim.get_data()[0]["livingState"]
    assert (st == 5).sum() > len(c) // 4 and (st == 0).any()
