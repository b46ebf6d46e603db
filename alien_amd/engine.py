"""ctypes binding of libalien_hip.so (include/alien_hip.h).

The classes mirror the physics subset of ALIEN's facade (source/EngineInterface/SimulationFacade.h:
13-110): newSimulation -> Simulation(...), setSimulationData / getSimulationData, calcTimesteps,
get/setCurrentTimestep, get/setSimulationParameters, getWorldSize, getGpuName.  There is no CPU
fallback: importing works anywhere, but every constructor raises if the HIP library is missing or
no GPU is visible.
"""
import ctypes
import os

import numpy as np

from .abi import CELL_DTYPE, PARTICLE_DTYPE, JobInfo, Params, Statistics, Totals, WorldDesc

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libalien_hip.so")
_lib = None


class KernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 40), ("launches", ctypes.c_uint64), ("totalMs", ctypes.c_double)]


def _kernel_times(getter, handle):
    cap = 64
    buf = (KernelTime * cap)()
    n = ctypes.c_int32()
    check(getter(handle, buf, cap, ctypes.byref(n)))
    return {buf[i].name.decode(): (buf[i].launches, buf[i].totalMs) for i in range(min(n.value, cap))}


class AlienError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


class AlienMemoryError(AlienError):
    """CudaMemoryAllocationException analog (Base/Exceptions.h:15-20)."""


def lib():
    """Loads the HIP library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32
        pu64 = ctypes.POINTER(u64)
        L.alien_last_error.restype = ctypes.c_char_p
        L.alien_abi_version.restype = ctypes.c_int
        L.alien_get_device_name.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        L.alien_params_default.argtypes = [ctypes.POINTER(Params)]
        L.alien_params_default.restype = None
        L.alien_world_create.argtypes = [ctypes.POINTER(WorldDesc), ctypes.POINTER(Params), ctypes.POINTER(vp)]
        L.alien_world_destroy.argtypes = [vp]
        L.alien_set_params.argtypes = [vp, ctypes.POINTER(Params)]
        L.alien_get_params.argtypes = [vp, ctypes.POINTER(Params)]
        L.alien_set_data.argtypes = [vp, vp, u64, vp, u64]
        L.alien_get_data.argtypes = [vp, vp, u64, pu64, vp, u64, pu64]
        L.alien_get_counts.argtypes = [vp, pu64, pu64]
        L.alien_step.argtypes = [vp, u64]
        L.alien_get_timestep.argtypes = [vp]
        L.alien_get_timestep.restype = u64
        L.alien_set_timestep.argtypes = [vp, u64]
        L.alien_get_totals.argtypes = [vp, ctypes.POINTER(Totals)]
        L.alien_get_statistics.argtypes = [vp, ctypes.POINTER(Statistics)]
        L.alien_get_statistics_noalloc.argtypes = [vp, ctypes.POINTER(Statistics)]
        L.alien_clear.argtypes = [vp]
        L.alien_debug_get_map.argtypes = [vp, vp, vp, u64, pu64, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                          ctypes.POINTER(i32)]
        L.alien_debug_neighbors.argtypes = [vp, vp, u64, vp, u64, pu64]
        L.alien_batch_create.argtypes = [ctypes.POINTER(vp)]
        L.alien_batch_destroy.argtypes = [vp]
        L.alien_batch_add_world.argtypes = [vp, ctypes.POINTER(WorldDesc), ctypes.POINTER(Params), vp, u64, vp, u64,
                                            ctypes.POINTER(i32)]
        L.alien_batch_num_worlds.argtypes = [vp]
        L.alien_batch_step.argtypes = [vp, u64]
        L.alien_batch_get_data.argtypes = [vp, i32, vp, u64, pu64, vp, u64, pu64]
        L.alien_batch_get_counts.argtypes = [vp, i32, pu64, pu64]
        L.alien_batch_get_timestep.argtypes = [vp, i32]
        L.alien_batch_get_timestep.restype = u64
        L.alien_batch_set_block_size.argtypes = [vp, i32]
        L.alien_batch_set_use_graph.argtypes = [vp, i32]
        L.alien_batch_autotune.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                           ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_double)]
        L.alien_set_kernel_timing.argtypes = [vp, i32]
        L.alien_get_kernel_timing.argtypes = [vp, ctypes.POINTER(KernelTime), i32, ctypes.POINTER(i32)]
        L.alien_get_path_counts.argtypes = [vp, pu64, pu64]
        L.alien_batch_set_kernel_timing.argtypes = [vp, i32]
        L.alien_batch_get_kernel_timing.argtypes = [vp, ctypes.POINTER(KernelTime), i32, ctypes.POINTER(i32)]
        L.alien_batch_get_path_counts.argtypes = [vp, pu64, pu64]
        L.alien_server_create.argtypes = [u64, ctypes.c_uint32, ctypes.POINTER(vp)]
        L.alien_server_destroy.argtypes = [vp]
        L.alien_server_submit.argtypes = [vp, ctypes.POINTER(WorldDesc), ctypes.POINTER(Params), vp, u64, vp, u64, u64,
                                          pu64]
        L.alien_server_poll.argtypes = [vp, u64, ctypes.POINTER(JobInfo)]
        L.alien_server_wait.argtypes = [vp, u64, ctypes.c_double]
        L.alien_server_get_snapshot.argtypes = [vp, u64, vp, u64, pu64, vp, u64, pu64]
        L.alien_server_release.argtypes = [vp, u64]
        L.alien_settings_read.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(WorldDesc),
                                          ctypes.POINTER(Params), pu64, ctypes.POINTER(ctypes.c_uint32)]
        L.alien_settings_write.argtypes = [ctypes.POINTER(WorldDesc), ctypes.POINTER(Params), u64, ctypes.c_char_p,
                                           ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.alien_server_estimate_bytes.argtypes = [u64, u64, i32, i32]
        L.alien_server_estimate_bytes.restype = u64
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().alien_last_error().decode()
        if rc == -2:
            raise AlienMemoryError(rc, msg)
        raise AlienError(rc, msg)


def read_settings(text):
    """ALIEN's .settings.json text -> (world size, alien_params, timestep, unsimulated bits)."""
    raw = text.encode() if isinstance(text, str) else bytes(text)
    d, p = WorldDesc(), Params()
    t, u = ctypes.c_uint64(), ctypes.c_uint32()
    check(lib().alien_settings_read(raw, len(raw), ctypes.byref(d), ctypes.byref(p), ctypes.byref(t), ctypes.byref(u)))
    return (d.worldSizeX, d.worldSizeY), p, t.value, u.value


def write_settings(size, params, timestep=0):
    """The physics subset as ALIEN's .settings.json text."""
    d = _desc(size)
    n = ctypes.c_size_t()
    rc = lib().alien_settings_write(ctypes.byref(d), ctypes.byref(params), int(timestep), None, 0, ctypes.byref(n))
    if rc not in (0, -3):
        check(rc)
    buf = ctypes.create_string_buffer(n.value + 1)
    check(lib().alien_settings_write(ctypes.byref(d), ctypes.byref(params), int(timestep), buf, len(buf), ctypes.byref(n)))
    return buf.value.decode()


def gpu_name():
    buf = ctypes.create_string_buffer(256)
    check(lib().alien_get_device_name(buf, 256))
    return buf.value.decode()


def _cells(a):
    a = np.ascontiguousarray(a, dtype=CELL_DTYPE)
    return a, (a.ctypes.data if len(a) else None)


def _parts(a):
    a = np.ascontiguousarray(a, dtype=PARTICLE_DTYPE)
    return a, (a.ctypes.data if len(a) else None)


def _desc(size, seed=0):
    d = WorldDesc()
    d.worldSizeX, d.worldSizeY = int(size[0]), int(size[1])
    d.randomSeed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return d


def _read(fn_counts, fn_data):
    n, p = ctypes.c_uint64(), ctypes.c_uint64()
    check(fn_counts(ctypes.byref(n), ctypes.byref(p)))
    while True:
        cells = np.zeros(max(n.value, 1), dtype=CELL_DTYPE)
        parts = np.zeros(max(p.value, 1), dtype=PARTICLE_DTYPE)
        on, op = ctypes.c_uint64(), ctypes.c_uint64()
        rc = fn_data(cells.ctypes.data, len(cells), ctypes.byref(on), parts.ctypes.data, len(parts), ctypes.byref(op))
        if rc == -3 and (on.value > len(cells) or op.value > len(parts)):
            n, p = on, op
            continue
        check(rc)
        return cells[: on.value], parts[: op.value]


class Simulation:
    """One world on the GPU (ALIEN's _SimulationFacade physics subset over the C ABI)."""

    def __init__(self, size, params, cells=None, particles=None, timestep=0, seed=0):
        self._h = None
        h = ctypes.c_void_p()
        self._size = (int(size[0]), int(size[1]))
        check(lib().alien_world_create(ctypes.byref(_desc(size, seed)), ctypes.byref(params), ctypes.byref(h)))
        self._h = h
        if timestep:
            self.set_timestep(timestep)
        if cells is not None:
            self.set_data(cells, particles if particles is not None else np.zeros(0, PARTICLE_DTYPE))

    @classmethod
    def from_world(cls, world):
        return cls(world.size, world.params, world.cells, world.particles, seed=world.seed)

    def close(self):
        if self._h:
            lib().alien_world_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def world_size(self):
        return self._size

    def set_params(self, params):
        check(lib().alien_set_params(self._h, ctypes.byref(params)))

    def get_params(self):
        p = Params()
        check(lib().alien_get_params(self._h, ctypes.byref(p)))
        return p

    def set_data(self, cells, particles):
        c, cp = _cells(cells)
        p, pp = _parts(particles)
        check(lib().alien_set_data(self._h, cp, len(c), pp, len(p)))

    def get_data(self):
        L = lib()
        return _read(lambda n, p: L.alien_get_counts(self._h, n, p),
                     lambda c, cc, n, p, pc, np_: L.alien_get_data(self._h, c, cc, n, p, pc, np_))

    def counts(self):
        n, p = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().alien_get_counts(self._h, ctypes.byref(n), ctypes.byref(p)))
        return n.value, p.value

    def step(self, n=1):
        check(lib().alien_step(self._h, int(n)))

    calc_timesteps = step

    @property
    def timestep(self):
        return lib().alien_get_timestep(self._h)

    def set_timestep(self, t):
        check(lib().alien_set_timestep(self._h, int(t)))

    def totals(self):
        t = Totals()
        check(lib().alien_get_totals(self._h, ctypes.byref(t)))
        return t.as_dict()

    def statistics(self):
        """Per-colour cell and particle counts and energies, reduced on the device."""
        s = Statistics()
        check(lib().alien_get_statistics(self._h, ctypes.byref(s)))
        return s.as_dict()

    def clear(self):
        check(lib().alien_clear(self._h))

    def set_kernel_timing(self, on):
        check(lib().alien_set_kernel_timing(self._h, 1 if on else 0))

    def kernel_times(self):
        """{kernel name: (launches, total ms)} since the last call (kernel timing mode)."""
        return _kernel_times(lib().alien_get_kernel_timing, self._h)

    def path_counts(self):
        """(chunks on the global-memory path, cells on the single-cell path) since set_data."""
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().alien_get_path_counts(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def get_map(self):
        n, _ = self.counts()
        bins = np.zeros(max(n, 1), dtype=np.int32)
        ids = np.zeros(max(n, 1), dtype=np.uint64)
        on = ctypes.c_uint64()
        bs, bx, by = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().alien_debug_get_map(self._h, bins.ctypes.data, ids.ctypes.data, len(bins), ctypes.byref(on),
                                        ctypes.byref(bs), ctypes.byref(bx), ctypes.byref(by)))
        return bins[: on.value], ids[: on.value], (bs.value, bx.value, by.value)

    def neighbors(self):
        n, _ = self.counts()
        off = np.zeros(n + 1, dtype=np.uint64)
        cap = max(64 * n, 1)
        while True:
            ids = np.zeros(cap, dtype=np.uint64)
            m = ctypes.c_uint64()
            rc = lib().alien_debug_neighbors(self._h, off.ctypes.data, len(off), ids.ctypes.data, cap, ctypes.byref(m))
            if rc == -3 and m.value > cap:
                cap = int(m.value)
                continue
            check(rc)
            return off, ids[: m.value]


class Batch:
    """Many independent worlds advanced together by one captured graph per call."""

    def __init__(self):
        self._h = None
        h = ctypes.c_void_p()
        check(lib().alien_batch_create(ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().alien_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_world(self, size, params, cells, particles, seed=0):
        c, cp = _cells(cells)
        p, pp = _parts(particles)
        idx = ctypes.c_int32()
        check(lib().alien_batch_add_world(self._h, ctypes.byref(_desc(size, seed)), ctypes.byref(params), cp, len(c), pp,
                                          len(p), ctypes.byref(idx)))
        return idx.value

    def add(self, world):
        return self.add_world(world.size, world.params, world.cells, world.particles, seed=world.seed)

    def __len__(self):
        return lib().alien_batch_num_worlds(self._h)

    def step(self, n=1):
        check(lib().alien_batch_step(self._h, int(n)))

    def get_data(self, w):
        L = lib()
        return _read(lambda n, p: L.alien_batch_get_counts(self._h, w, n, p),
                     lambda c, cc, n, p, pc, np_: L.alien_batch_get_data(self._h, w, c, cc, n, p, pc, np_))

    def counts(self, w):
        n, p = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().alien_batch_get_counts(self._h, w, ctypes.byref(n), ctypes.byref(p)))
        return n.value, p.value

    def timestep(self, w):
        return lib().alien_batch_get_timestep(self._h, w)

    def set_block_size(self, threads):
        check(lib().alien_batch_set_block_size(self._h, int(threads)))

    def set_use_graph(self, on):
        check(lib().alien_batch_set_use_graph(self._h, 1 if on else 0))

    def set_kernel_timing(self, on):
        check(lib().alien_batch_set_kernel_timing(self._h, 1 if on else 0))

    def autotune(self, trial_steps=20):
        """Picks the fastest launch shape on a device copy; returns (block, strip cap, chunk rows
        (0: the default rule), ms/step)."""
        blk, strip, rows, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        check(lib().alien_batch_autotune(self._h, int(trial_steps), ctypes.byref(blk), ctypes.byref(strip),
                                         ctypes.byref(rows), ctypes.byref(ms)))
        return blk.value, strip.value, rows.value, ms.value

    def kernel_times(self):
        return _kernel_times(lib().alien_batch_get_kernel_timing, self._h)

    def path_counts(self):
        """(chunks on the global-memory path, cells on the single-cell path) since the last rebuild."""
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().alien_batch_get_path_counts(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


class Server:
    """Continuous-batching job server keeping every admitted world resident on the GPU."""

    def __init__(self, budget_bytes, chunk_steps=50):
        self._h = None
        h = ctypes.c_void_p()
        check(lib().alien_server_create(int(budget_bytes), int(chunk_steps), ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().alien_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, size, params, cells, particles, steps, seed=0):
        c, cp = _cells(cells)
        p, pp = _parts(particles)
        jid = ctypes.c_uint64()
        check(lib().alien_server_submit(self._h, ctypes.byref(_desc(size, seed)), ctypes.byref(params), cp, len(c), pp,
                                        len(p), int(steps), ctypes.byref(jid)))
        return jid.value

    def poll(self, job):
        info = JobInfo()
        check(lib().alien_server_poll(self._h, int(job), ctypes.byref(info)))
        return info

    def wait(self, job, timeout=600.0):
        check(lib().alien_server_wait(self._h, int(job), float(timeout)))

    def snapshot(self, job):
        L = lib()
        info = self.poll(job)
        cells = np.zeros(max(info.numCells, 1), dtype=CELL_DTYPE)
        parts = np.zeros(max(info.numParticles, 1), dtype=PARTICLE_DTYPE)
        on, op = ctypes.c_uint64(), ctypes.c_uint64()
        check(L.alien_server_get_snapshot(self._h, int(job), cells.ctypes.data, len(cells), ctypes.byref(on),
                                          parts.ctypes.data, len(parts), ctypes.byref(op)))
        return cells[: on.value], parts[: op.value]

    def release(self, job):
        check(lib().alien_server_release(self._h, int(job)))

    @staticmethod
    def estimate_bytes(num_cells, num_particles, size):
        return lib().alien_server_estimate_bytes(int(num_cells), int(num_particles), int(size[0]), int(size[1]))
