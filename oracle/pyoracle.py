"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker or the
timed CPU baseline; the product (alien_amd/) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

from alien_amd.abi import CELL_DTYPE, PARTICLE_DTYPE, Params, Totals

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [i32, i32, i32, ctypes.POINTER(Params)]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_set_threads.argtypes = [vp, i32]
        L.oracle_set_seed.argtypes = [vp, u64]
        L.oracle_uniform_draw.argtypes = [u64, u64, u64, ctypes.c_uint32]
        L.oracle_uniform_draw.restype = ctypes.c_float
        L.oracle_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_set_params.argtypes = [vp, ctypes.POINTER(Params)]
        L.oracle_set_data.argtypes = [vp, vp, u64, vp, u64]
        L.oracle_get_counts.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.oracle_get_data.argtypes = [vp, vp, u64, ctypes.POINTER(u64), vp, u64, ctypes.POINTER(u64)]
        L.oracle_step.argtypes = [vp, u64]
        L.oracle_get_timestep.argtypes = [vp]
        L.oracle_get_timestep.restype = u64
        L.oracle_set_timestep.argtypes = [vp, u64]
        L.oracle_get_totals.argtypes = [vp, ctypes.POINTER(Totals)]
        L.oracle_get_map.argtypes = [vp, vp, vp, u64, ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
        L.oracle_neighbors.argtypes = [vp, vp, u64, vp, u64, ctypes.POINTER(u64)]
        L.oracle_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())


class OracleWorld:
    def __init__(self, size, params, precision=32, threads=1, seed=0):
        self._h = lib().oracle_create(precision, int(size[0]), int(size[1]), ctypes.byref(params))
        if not self._h:
            raise RuntimeError(lib().oracle_last_error().decode())
        self.size = size
        self.precision = precision
        lib().oracle_set_threads(self._h, threads)
        lib().oracle_set_seed(self._h, int(seed))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    def set_threads(self, n):
        lib().oracle_set_threads(self._h, n)

    def set_params(self, params):
        _check(lib().oracle_set_params(self._h, ctypes.byref(params)))

    def set_data(self, cells, parts):
        cells = np.ascontiguousarray(cells, dtype=CELL_DTYPE)
        parts = np.ascontiguousarray(parts, dtype=PARTICLE_DTYPE)
        _check(lib().oracle_set_data(self._h, cells.ctypes.data, len(cells), parts.ctypes.data, len(parts)))

    def counts(self):
        n, p = ctypes.c_uint64(), ctypes.c_uint64()
        lib().oracle_get_counts(self._h, ctypes.byref(n), ctypes.byref(p))
        return n.value, p.value

    def get_data(self):
        n, p = self.counts()
        cells = np.zeros(max(n, 1), dtype=CELL_DTYPE)
        parts = np.zeros(max(p, 1), dtype=PARTICLE_DTYPE)
        on, op = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().oracle_get_data(self._h, cells.ctypes.data, len(cells), ctypes.byref(on), parts.ctypes.data,
                                     len(parts), ctypes.byref(op)))
        return cells[: on.value], parts[: op.value]

    def step(self, n=1):
        _check(lib().oracle_step(self._h, n))

    @property
    def timestep(self):
        return lib().oracle_get_timestep(self._h)

    @timestep.setter
    def timestep(self, t):
        lib().oracle_set_timestep(self._h, t)

    def totals(self):
        t = Totals()
        lib().oracle_get_totals(self._h, ctypes.byref(t))
        return t.as_dict()

    def get_map(self):
        n, _ = self.counts()
        bins = np.zeros(max(n, 1), dtype=np.int32)
        ids = np.zeros(max(n, 1), dtype=np.uint64)
        on = ctypes.c_uint64()
        bs, bx, by = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(lib().oracle_get_map(self._h, bins.ctypes.data, ids.ctypes.data, len(bins), ctypes.byref(on),
                                    ctypes.byref(bs), ctypes.byref(bx), ctypes.byref(by)))
        return bins[: on.value], ids[: on.value], (bs.value, bx.value, by.value)

    def neighbors(self, max_ids=None):
        n, _ = self.counts()
        cap = max_ids or max(64 * n, 1)
        off = np.zeros(n + 1, dtype=np.uint64)
        ids = np.zeros(cap, dtype=np.uint64)
        m = ctypes.c_uint64()
        _check(lib().oracle_neighbors(self._h, off.ctypes.data, len(off), ids.ctypes.data, cap, ctypes.byref(m)))
        return off, ids[: m.value]


def statistics(cells, particles):
    """Per-colour counts and energies of a world's data (StatisticsKernels.cu:11-45,
    cudaUpdateTimestepStatistics_substep2: numCells and energy per cell colour, numParticles and
    energy per particle colour), energies summed in double."""
    n_cells = np.bincount(cells["color"].astype(np.int64) % 7, minlength=7)
    n_parts = np.bincount(particles["color"].astype(np.int64) % 7, minlength=7)
    energy = np.bincount(cells["color"].astype(np.int64) % 7, weights=cells["energy"].astype(np.float64), minlength=7)
    energy = energy + np.bincount(particles["color"].astype(np.int64) % 7,
                                  weights=particles["energy"].astype(np.float64), minlength=7)
    return {"numCells": n_cells.tolist(), "numParticles": n_parts.tolist(), "totalEnergy": energy.tolist()}


def from_world(world, precision=32, threads=1):
    o = OracleWorld(world.size, world.params, precision=precision, threads=threads, seed=world.seed)
    o.set_data(world.cells, world.particles)
    return o
