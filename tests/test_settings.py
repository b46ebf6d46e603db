"""ALIEN's .settings.json (AuxiliaryDataParserService.cpp:1100-1137, every value a string under
nested keys): alien_settings_read on the settings file ALIEN wrote itself
(tests/golden/autosave.settings.json), checked field by field against Python's own parse of the
same file, and write -> read round trips.  Host code only (no GPU)."""
import json
import math
import os

import numpy as np
import pytest

from alien_amd import abi, engine

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "autosave.settings.json")


def f32(x):
    return float(np.float32(float(x)))


def test_read_alien_settings_file():
    text = open(FIXTURE).read()
    raw = json.loads(text)
    g, sp = raw["general"], raw["simulation parameters"]
    size, p, timestep, unsim = engine.read_settings(text)
    assert size == (int(g["world size"]["x"]), int(g["world size"]["y"])) == (2000, 1000)
    assert timestep == int(g["time step"]) == 85540
    assert p.motionType == int(sp["motion"]["type"]) == abi.MOTION_FLUID
    assert p.timestepSize == f32(sp["time step size"])
    assert p.smoothingLength == f32(sp["fluid"]["smoothing length"])
    assert p.pressureStrength == f32(sp["fluid"]["pressure strength"])
    assert p.viscosityStrength == f32(sp["fluid"]["viscosity strength"])
    assert p.friction == f32(sp["friction"])
    cell = sp["cell"]
    assert p.cellMaxVelocity == f32(cell["max velocity"])
    assert p.cellMinDistance == f32(cell["min distance"])
    assert p.cellFusionVelocity == f32(cell["fusion velocity"])
    assert p.cellMaxBindingEnergy == f32(cell["max binding energy"])
    assert p.cellMaxForceDecayProb == f32(cell["max force decay probability"])
    for k in range(7):
        assert p.cellMaxBindingDistance[k] == f32(cell[f"max binding distance[{k}]"])
        assert p.cellMaxForce[k] == f32(cell[f"max force[{k}]"])
        assert p.cellMinEnergy[k] == f32(cell[f"min energy[{k}]"])
        assert p.radiationAbsorption[k] == f32(sp["radiation"][f"absorption[{k}]"])
    assert p.innerFriction == f32(0.3)  # not part of ALIEN's settings files: the default stays
    # cell radiation (AuxiliaryDataParserService.cpp:203-268), simulated
    rad, high = sp["radiation"], sp["high radiation"]
    assert p.radiationProb == f32(rad["probability"]) == f32(0.03)
    assert p.radiationVelocityMultiplier == f32(rad["velocity multiplier"])
    assert p.radiationVelocityPerturbation == f32(rad["velocity perturbation"])
    for k in range(7):
        assert p.radiationCellAgeStrength[k] == f32(rad[f"factor[{k}]"])
        assert p.highRadiationFactor[k] == f32(high[f"factor[{k}]"])
        assert p.highRadiationMinCellEnergy[k] == f32(high[f"min cell energy[{k}]"])
        assert p.radiationMinCellAge[k] == int(float(rad.get(f"min cell age[{k}]", "0")))
    assert p.radiationCellAgeStrength[1] == f32(0.000093) and p.highRadiationFactor[0] == f32(0.000478)
    # the file switches on zones (3 spots), particle sources and cell death; its absorption
    # penalties are 0
    assert int(sp["spots"]["num spots"]) == 3
    assert all(float(v) == 0.0 for key, v in rad.items() if "penalty" in key)
    expect = abi.SETTINGS_ZONES | abi.SETTINGS_CELL_DEATH
    if int(sp["particle sources"]["num sources"]) > 0:
        expect |= abi.SETTINGS_RADIATION_SOURCES
    assert unsim == expect


@pytest.mark.parametrize("motion", [abi.MOTION_FLUID, abi.MOTION_COLLISION])
def test_write_read_round_trip(motion):
    p = abi.alien_default_params(motionType=motion, smoothingLength=0.65, cellMaxCollisionDistance=1.5,
                                 friction=0.0025, cellFusionVelocity=0.5)
    for k in range(7):
        p.radiationAbsorption[k] = 0.1 * (k + 1)
        p.cellMinEnergy[k] = 40.0 + k
        p.radiationMinCellAge[k] = 10 * k
        p.highRadiationFactor[k] = 0.001 * k
        p.highRadiationMinCellEnergy[k] = 300.0 + k
        p.radiationCellAgeStrength[k] = 0.00001 * (k + 1)
    p.radiationProb = 0.05
    p.radiationVelocityMultiplier = 0.75
    p.radiationVelocityPerturbation = 0.25
    text = engine.write_settings((640, 480), p, timestep=1234)
    tree = json.loads(text)  # valid JSON, every leaf a string like ALIEN's
    assert tree["general"]["world size"]["x"] == "640"
    assert all(isinstance(v, str) for v in tree["simulation parameters"]["cell"].values())
    size, q, t, unsim = engine.read_settings(text)
    assert size == (640, 480) and t == 1234 and unsim == 0
    for name, _ in abi.Params._fields_:
        if name == "innerFriction":
            continue  # not stored in settings files
        a, b = getattr(p, name), getattr(q, name)
        if motion == abi.MOTION_FLUID and name in ("cellMaxCollisionDistance", "cellRepulsionStrength"):
            continue  # the other motion's values are not written (as in ALIEN)
        if motion == abi.MOTION_COLLISION and name in ("smoothingLength", "pressureStrength", "viscosityStrength"):
            continue
        if isinstance(a, float):
            assert math.isclose(a, b, rel_tol=1e-7) or (math.isinf(a) and b > 3e38), name
        elif isinstance(a, int):
            assert a == b, name
        else:
            np.testing.assert_allclose(list(a), list(b), rtol=1e-7, err_msg=name)


def test_bad_settings_are_rejected():
    with pytest.raises(engine.AlienError) as e:
        engine.read_settings('{"general": {"world size": {"x": "12"')
    assert e.value.code == abi.E_ARG
    with pytest.raises(engine.AlienError) as e:
        engine.read_settings('{"simulation parameters": {"friction": "much"}}')
    assert e.value.code == abi.E_ARG
    size, p, _, _ = engine.read_settings("{}")  # everything at ALIEN's defaults
    assert size == (1000, 500) and p.innerFriction == f32(0.3) and p.cellMaxForceDecayProb == f32(0.2)
