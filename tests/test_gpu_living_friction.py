"""Parity of the HIP path with the oracle for inner friction, the force-driven bond decay (with the
counter-based random draws) and the living states, through the C ABI (needs an MI355X).  Bar:
cells (including living state, activation count and age), particles and bonds byte-identical to
the FP32 oracle after every compared step."""
import numpy as np
import pytest

from alien_amd import abi, engine, worlds
from alien_amd.abi import MOTION_COLLISION, MOTION_FLUID
from oracle.pyoracle import from_world

from .test_gpu_parity import CPU_THREADS, assert_identical, say

pytestmark = pytest.mark.gpu


def with_params(w, **kw):
    p = w.params.copy()
    for k, v in kw.items():
        if k in ("cellMaxForce", "cellMinEnergy"):
            for c in range(abi.MAX_COLORS):
                getattr(p, k)[c] = v
        else:
            setattr(p, k, v)
    return worlds.World(w.name, w.size, p, w.cells, w.particles, w.steps, w.seed)


def living_mix(seed, motion):
    """A random mix whose cells carry every living state, activation counts, two creatures per
    body, and energies below cellMinEnergy (some of them get particles)."""
    w = worlds.random_mix(seed, motion=motion, num_particles=400)
    rng = np.random.default_rng(seed + 1000)
    c = w.cells.copy()
    n = len(c)
    pick = rng.random(n) < 0.35
    c["livingState"][pick] = rng.integers(0, 6, size=int(pick.sum())).astype(np.uint8)
    c["activationTime"] = np.where(rng.random(n) < 0.3, rng.integers(1, 6, size=n), 0).astype(np.uint32)
    c["creatureId"] = rng.integers(1, 3, size=n).astype(np.uint32)
    low = rng.random(n) < 0.2
    c["energy"][low] = rng.uniform(20.0, 49.0, size=int(low.sum())).astype(np.float32)
    return worlds.World(f"living_{seed}", w.size, w.params, c, w.particles, 10, w.seed)


def run_checkpoints(w, checkpoints):
    sim = engine.Simulation.from_world(w)
    ref = from_world(w, threads=CPU_THREADS)
    done = 0
    for s in checkpoints:
        sim.step(s - done)
        ref.step(s - done)
        done = s
        assert sim.timestep == ref.timestep == s
        assert_identical(sim.get_data(), ref.get_data(), f"{w.name} step {s}")
    return sim, ref


@pytest.mark.parametrize("motion", [MOTION_FLUID, MOTION_COLLISION])
def test_inner_friction_bit_exact(motion):
    w = with_params(worlds.random_mix(301 + motion, motion=motion), innerFriction=0.3)
    say(f"inner friction: {w.name} motion={motion}")
    run_checkpoints(w, [1, 3, 4, 12])


@pytest.mark.parametrize("motion", [MOTION_FLUID, MOTION_COLLISION])
def test_force_decay_bit_exact(motion):
    """cellMaxForce low enough that many bonded cells qualify: bonds break by the random draws, on
    the GPU exactly as in the oracle (same seed, same ids, same steps)."""
    base = worlds.random_mix(311 + motion, motion=motion)
    w = with_params(base, cellMaxForceDecayProb=0.5, cellMaxForce=0.02)
    say(f"force decay: {w.name} motion={motion}")
    sim, ref = run_checkpoints(w, [1, 2, 6])
    plain = from_world(base)
    plain.step(6)
    bonds = int(sim.get_data()[0]["numConnections"].sum())
    assert bonds < int(plain.get_data()[0]["numConnections"].sum())  # the decay really broke bonds


def test_alien_default_parameters_bit_exact():
    """ALIEN's defaults (inner friction 0.3, force decay 0.2 above 0.8, cell radiation with
    probability 0.03 and age strength 0.00002) on the 512x512 soft-body world."""
    base = worlds.plain_soft_body()
    p = abi.alien_default_params(motionType=MOTION_COLLISION, cellMaxCollisionDistance=1.3, friction=0.001)
    w = worlds.World("plain_soft_body_alien_defaults", base.size, p, base.cells, base.particles, 30, base.seed)
    say("ALIEN default parameters on the plain soft-body world")
    run_checkpoints(w, [1, 3, 30])


@pytest.mark.parametrize("motion", [MOTION_FLUID, MOTION_COLLISION])
def test_living_states_bit_exact(motion):
    w = living_mix(321 + motion, motion)
    say(f"living states: {w.name} motion={motion}")
    sim, _ = run_checkpoints(w, [1, 2, 5, 10])
    st = sim.get_data()[0]["livingState"]
    assert (st == 5).any() and (st == 0).any()  # cells died and others are Ready


def test_min_energy_raised_mid_run():
    """A world without tracked living states gets them when cellMinEnergy rises above its cells'
    energies: every non-barrier cell becomes Dying, as in the oracle."""
    w = worlds.random_mix(331, motion=MOTION_COLLISION)
    sim = engine.Simulation.from_world(w)
    ref = from_world(w, threads=CPU_THREADS)
    sim.step(3)
    ref.step(3)
    p = with_params(w, cellMinEnergy=120.0).params
    sim.set_params(p)
    ref.set_params(p)
    sim.step(3)
    ref.step(3)
    got = sim.get_data()
    assert_identical(got, ref.get_data(), "after raising cellMinEnergy")
    c = got[0]
    assert (c["livingState"][c["barrier"] == 0] == 5).all()


@pytest.mark.parametrize("graph", [True, False])
def test_batch_with_new_physics_equals_alone(graph):
    ws = [living_mix(341, MOTION_COLLISION), with_params(worlds.random_mix(342), innerFriction=0.3),
          with_params(worlds.random_mix(343, motion=MOTION_FLUID), cellMaxForceDecayProb=0.5, cellMaxForce=0.02),
          worlds.random_mix(344), living_mix(345, MOTION_FLUID)]
    b = engine.Batch()
    for w in ws:
        b.add(w)
    b.set_use_graph(graph)
    b.step(4)
    b.step(5)
    for i, w in enumerate(ws):
        ref = from_world(w, threads=CPU_THREADS)
        ref.step(9)
        assert_identical(b.get_data(i), ref.get_data(), f"batch world {i} ({w.name}) graph={graph}")


def test_server_with_new_physics_equals_alone():
    ws = [living_mix(351, MOTION_COLLISION), with_params(worlds.random_mix(352), innerFriction=0.3,
                                                         cellMaxForceDecayProb=0.5, cellMaxForce=0.02)]
    steps = [17, 9]
    srv = engine.Server(budget_bytes=1 << 30, chunk_steps=4)
    jobs = [srv.submit(w.size, w.params, w.cells, w.particles, s, seed=w.seed) for w, s in zip(ws, steps)]
    for j in jobs:
        srv.wait(j, timeout=120)
    for j, w, s in zip(jobs, ws, steps):
        ref = from_world(w, threads=CPU_THREADS)
        ref.step(s)
        assert_identical(srv.snapshot(j), ref.get_data(), f"server job {w.name}")
    srv.close()
