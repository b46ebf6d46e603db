"""Python mirror of the C ABI data layouts in include/alien_hip.h.

The numpy dtypes below have exactly the C layout of alien_cell_to / alien_particle_to (ALIEN's
CellTO / ParticleTO, source/EngineGpuKernels/TOs.cuh:10-19 and :156-192), so a structured numpy
array can be handed to the library as a plain pointer.  tests/test_abi_layout.py checks every
offset against a C compiler's view of the header.
"""
import ctypes

import numpy as np

MAX_CELL_BONDS = 6
MAX_COLORS = 7
MOTION_FLUID = 0
MOTION_COLLISION = 1

OK = 0
E_ARG = -1
E_OOM = -2
E_CAPACITY = -3
E_HIP = -4
E_STATE = -5
E_BUDGET = -6
E_NOTFOUND = -7

JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, 3

# processes a settings file may switch on that the physics path does not simulate
SETTINGS_ZONES = 0x1
SETTINGS_RADIATION_SOURCES = 0x2
SETTINGS_RADIATION = 0x4
SETTINGS_CELL_DEATH = 0x8
SETTINGS_PARTICLE_SPLIT = 0x10
SETTINGS_RIGIDITY = 0x20
SETTINGS_PARTICLE_TRANSFORMATION = 0x40
SETTINGS_ABSORPTION_CONTROL = 0x80

CONNECTION_DTYPE = np.dtype(
    [("cellIndex", "<i4"), ("distance", "<f4"), ("angleFromPrevious", "<f4")], align=True)

CELL_DTYPE = np.dtype(
    [
        ("id", "<u8"),
        ("connections", CONNECTION_DTYPE, (MAX_CELL_BONDS,)),
        ("pos", "<f4", (2,)),
        ("vel", "<f4", (2,)),
        ("energy", "<f4"),
        ("stiffness", "<f4"),
        ("color", "u1"),
        ("maxConnections", "u1"),
        ("numConnections", "u1"),
        ("barrier", "u1"),
        ("age", "<u4"),
        ("livingState", "u1"),
        ("cellFunction", "u1"),
        ("ancestorMutationId", "u1"),
        ("selected", "u1"),
        ("creatureId", "<u4"),
        ("mutationId", "<u4"),
        ("genomeComplexity", "<f4"),
        ("activationTime", "<u4"),
    ],
    align=True,
)

PARTICLE_DTYPE = np.dtype(
    [
        ("id", "<u8"),
        ("energy", "<f4"),
        ("pos", "<f4", (2,)),
        ("vel", "<f4", (2,)),
        ("color", "u1"),
        ("selected", "u1"),
        ("pad_", "u1", (2,)),
    ],
    align=True,
)


class Params(ctypes.Structure):
    """alien_params: physics subset of SimulationParameters (SimulationParameters.h:18-206)."""

    _fields_ = [
        ("timestepSize", ctypes.c_float),
        ("motionType", ctypes.c_int32),
        ("smoothingLength", ctypes.c_float),
        ("viscosityStrength", ctypes.c_float),
        ("pressureStrength", ctypes.c_float),
        ("cellMaxCollisionDistance", ctypes.c_float),
        ("cellRepulsionStrength", ctypes.c_float),
        ("friction", ctypes.c_float),
        ("innerFriction", ctypes.c_float),
        ("cellMaxVelocity", ctypes.c_float),
        ("cellMinDistance", ctypes.c_float),
        ("cellFusionVelocity", ctypes.c_float),
        ("cellMaxBindingEnergy", ctypes.c_float),
        ("cellMaxForceDecayProb", ctypes.c_float),
        ("cellMaxBindingDistance", ctypes.c_float * MAX_COLORS),
        ("radiationAbsorption", ctypes.c_float * MAX_COLORS),
        ("cellMaxForce", ctypes.c_float * MAX_COLORS),
        ("cellMinEnergy", ctypes.c_float * MAX_COLORS),
        ("radiationProb", ctypes.c_float),
        ("radiationVelocityMultiplier", ctypes.c_float),
        ("radiationVelocityPerturbation", ctypes.c_float),
        ("radiationMinCellAge", ctypes.c_int32 * MAX_COLORS),
        ("highRadiationFactor", ctypes.c_float * MAX_COLORS),
        ("highRadiationMinCellEnergy", ctypes.c_float * MAX_COLORS),
        ("radiationCellAgeStrength", ctypes.c_float * MAX_COLORS),
        ("particleSplitEnergy", ctypes.c_float * MAX_COLORS),
    ]

    def copy(self):
        other = Params()
        ctypes.pointer(other)[0] = self
        return other


def alien_default_params(**overrides):
    """ALIEN's defaults for the physics subset (Motion.h, SimulationParameters.h,
    SimulationParametersZoneValues.h), as alien_params_default() fills them: inner friction 0.3,
    force decay probability 0.2 above cellMaxForce 0.8, cellMinEnergy 50, cell radiation with
    probability 0.03 and age strength 0.00002."""
    p = default_params(innerFriction=0.3, cellMaxForceDecayProb=0.2)
    for c in range(MAX_COLORS):
        p.radiationCellAgeStrength[c] = 0.00002
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def default_params(**overrides):
    """The parameter set of the benchmark workloads and most tests: ALIEN's defaults for the
    physics subset except inner friction 0, force decay probability 0 and no cell radiation (age
    strength and high-radiation factor 0; the workloads are specified without these processes, and
    the tests of them set them explicitly)."""
    p = Params()
    p.timestepSize = 1.0
    p.motionType = MOTION_FLUID
    p.smoothingLength = 0.8
    p.viscosityStrength = 0.1
    p.pressureStrength = 0.1
    p.cellMaxCollisionDistance = 1.3
    p.cellRepulsionStrength = 0.08
    p.friction = 0.001
    p.innerFriction = 0.0
    p.cellMaxVelocity = 2.0
    p.cellMinDistance = 0.3
    p.cellFusionVelocity = 0.6
    p.cellMaxBindingEnergy = float("inf")
    p.cellMaxForceDecayProb = 0.0
    for c in range(MAX_COLORS):
        p.cellMaxBindingDistance[c] = 3.6
        p.radiationAbsorption[c] = 1.0
        p.cellMaxForce[c] = 0.8
        p.cellMinEnergy[c] = 50.0
        p.radiationMinCellAge[c] = 0
        p.highRadiationFactor[c] = 0.0
        p.highRadiationMinCellEnergy[c] = 500.0
        p.radiationCellAgeStrength[c] = 0.0
        p.particleSplitEnergy[c] = float("inf")
    p.radiationProb = 0.03
    p.radiationVelocityMultiplier = 1.0
    p.radiationVelocityPerturbation = 0.5
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


class WorldDesc(ctypes.Structure):
    _fields_ = [
        ("worldSizeX", ctypes.c_int32),
        ("worldSizeY", ctypes.c_int32),
        ("cellCapacity", ctypes.c_uint64),
        ("particleCapacity", ctypes.c_uint64),
        ("randomSeed", ctypes.c_uint64),
    ]


class Statistics(ctypes.Structure):
    """alien_statistics: per-colour counts and energies (TimestepStatistics subset)."""
    _fields_ = [
        ("numCells", ctypes.c_int32 * 7),
        ("numParticles", ctypes.c_int32 * 7),
        ("totalEnergy", ctypes.c_double * 7),
        ("timestep", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {
            "numCells": list(self.numCells),
            "numParticles": list(self.numParticles),
            "totalEnergy": list(self.totalEnergy),
            "timestep": int(self.timestep),
        }


class Totals(ctypes.Structure):
    _fields_ = [
        ("cellEnergy", ctypes.c_double),
        ("particleEnergy", ctypes.c_double),
        ("momentum", ctypes.c_double * 2),
        ("kinetic", ctypes.c_double),
        ("numCells", ctypes.c_uint64),
        ("numParticles", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {
            "cellEnergy": self.cellEnergy,
            "particleEnergy": self.particleEnergy,
            "energy": self.cellEnergy + self.particleEnergy,
            "momentum": (self.momentum[0], self.momentum[1]),
            "kinetic": self.kinetic,
            "numCells": self.numCells,
            "numParticles": self.numParticles,
        }


class JobInfo(ctypes.Structure):
    _fields_ = [
        ("state", ctypes.c_int32),
        ("stepsDone", ctypes.c_uint64),
        ("stepsRequested", ctypes.c_uint64),
        ("numCells", ctypes.c_uint64),
        ("numParticles", ctypes.c_uint64),
        ("submitTime", ctypes.c_double),
        ("startTime", ctypes.c_double),
        ("finishTime", ctypes.c_double),
        ("deviceBytes", ctypes.c_uint64),
    ]


def empty_cells(n):
    return np.zeros(n, dtype=CELL_DTYPE)


def empty_particles(n):
    return np.zeros(n, dtype=PARTICLE_DTYPE)
