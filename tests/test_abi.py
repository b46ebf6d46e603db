"""The C ABI (include/alien_hip.h) without a GPU: the library loads and exports every declared
entry point, the Python mirrors of the structs match the C layout that gcc computes from the
header, and the pure host calls behave as documented.  No compute call is made here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from alien_amd import abi, engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "alien_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(alien_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_facade_surface():
    names = declared_functions()
    for required in ("alien_world_create", "alien_world_destroy", "alien_set_params", "alien_get_params",
                     "alien_set_data", "alien_get_data", "alien_step", "alien_get_timestep", "alien_set_timestep",
                     "alien_get_totals", "alien_clear", "alien_last_error", "alien_batch_create", "alien_batch_step",
                     "alien_server_create", "alien_server_submit", "alien_server_get_snapshot"):
        assert required in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(engine.LIB_PATH), "build the HIP library first (__graft_entry__.build())"
    L = ctypes.CDLL(engine.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_library_exports_nothing_undeclared():
    out = subprocess.run(["nm", "-D", "--defined-only", engine.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = sorted(set(re.findall(r"\b(alien_[a-z0-9_]+)\b", out)))
    assert exported == declared_functions()


STRUCTS = {
    "alien_cell_to": abi.CELL_DTYPE,
    "alien_particle_to": abi.PARTICLE_DTYPE,
    "alien_connection_to": abi.CONNECTION_DTYPE,
}
CTYPES = {
    "alien_params": abi.Params,
    "alien_world_desc": abi.WorldDesc,
    "alien_totals": abi.Totals,
    "alien_statistics": abi.Statistics,
    "alien_job_info": abi.JobInfo,
    "alien_kernel_time": engine.KernelTime,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    """sizeof and offsetof of every mirrored struct, as gcc lays them out from the header."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for name, dt in STRUCTS.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for f in dt.names:
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    for name, ct in CTYPES.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for f, _ in ct._fields_:
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines.append("return 0; }")
    d = tmp_path_factory.mktemp("layout")
    src = d / "layout.c"
    src.write_text("\n".join(lines))
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return {k: int(v) for k, v in (ln.rsplit(" ", 1) for ln in out.splitlines())}


def test_numpy_records_match_c(c_layout):
    for name, dt in STRUCTS.items():
        assert c_layout[f"{name} size"] == dt.itemsize, name
        for f in dt.names:
            assert c_layout[f"{name}.{f}"] == dt.fields[f][1], (name, f)


def test_ctypes_structs_match_c(c_layout):
    for name, ct in CTYPES.items():
        assert c_layout[f"{name} size"] == ctypes.sizeof(ct), name
        for f, _ in ct._fields_:
            assert c_layout[f"{name}.{f}"] == getattr(ct, f).offset, (name, f)


def test_host_only_calls():
    L = engine.lib()
    assert L.alien_abi_version() >= 1
    p = abi.Params()
    L.alien_params_default(ctypes.byref(p))
    ref = abi.alien_default_params()
    for f, _ in abi.Params._fields_:
        a, b = getattr(p, f), getattr(ref, f)
        if isinstance(a, (int, float)):
            assert a == b, f
        else:
            assert list(a) == list(b), f


def test_product_has_no_cpu_fallback(monkeypatch, tmp_path):
    """Without the HIP library the product raises instead of computing anything elsewhere."""
    monkeypatch.setattr(engine, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(engine, "_lib", None)
    with pytest.raises(RuntimeError, match="HIP library not built"):
        engine.Simulation((64, 64), abi.default_params())


def test_product_does_not_use_the_oracle():
    pattern = re.compile(r"import\s+oracle|from\s+oracle|pyoracle|liboracle|alien_oracle|oracle_capi")
    for dirpath, _, files in os.walk(os.path.join(ROOT, "alien_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hpp", ".hip", ".h")) or f == "Makefile":
                text = open(os.path.join(dirpath, f), errors="replace").read()
                assert not pattern.search(text), os.path.join(dirpath, f)


def test_cell_record_round_trips_through_numpy():
    c = np.zeros(3, dtype=abi.CELL_DTYPE)
    c["connections"]["cellIndex"][1, 2] = 7
    raw = c.tobytes()
    assert len(raw) == 3 * abi.CELL_DTYPE.itemsize
    assert np.frombuffer(raw, dtype=abi.CELL_DTYPE)["connections"]["cellIndex"][1, 2] == 7


MAX_SET_CELLS = (2**31 - 1) // 6  # bond planes are addressed as b * cap + x in 32-bit ints


def _desc(size=64):
    d = abi.WorldDesc()
    d.worldSizeX = d.worldSizeY = size
    return d


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_set_size_limit_is_checked_before_any_read():
    """A world set holds at most (2^31 - 1) / 6 cells: larger requests fail with ALIEN_E_CAPACITY
    before the transfer array is read (here a one-cell buffer with a count of 400M), for one
    batch world, for a batch's running total, and for a serving job.  No device call is made."""
    L = engine.lib()
    one = np.zeros(1, dtype=abi.CELL_DTYPE)
    nopart = np.zeros(1, dtype=abi.PARTICLE_DTYPE)
    d, p = _desc(), abi.default_params(motionType=abi.MOTION_COLLISION)
    b = ctypes.c_void_p()
    assert L.alien_batch_create(ctypes.byref(b)) == abi.OK
    idx = ctypes.c_int32()
    add = lambda cells, n: L.alien_batch_add_world(b, ctypes.byref(d), ctypes.byref(p), _ptr(cells), ctypes.c_uint64(n),
                                                   _ptr(nopart), ctypes.c_uint64(0), ctypes.byref(idx))
    assert add(one, 400_000_000) == abi.E_CAPACITY
    assert b"cells in one world set" in L.alien_last_error()
    assert add(one, MAX_SET_CELLS + 1) == abi.E_CAPACITY
    small = np.zeros(200, dtype=abi.CELL_DTYPE)
    small["pos"] = np.random.default_rng(0).uniform(0, 64, size=(200, 2))
    assert add(small, 200) == abi.OK
    assert add(one, MAX_SET_CELLS - 100) == abi.E_CAPACITY  # 200 + that exceeds the limit
    assert L.alien_batch_num_worlds(b) == 1
    assert L.alien_batch_destroy(b) == abi.OK

    s = ctypes.c_void_p()
    assert L.alien_server_create(ctypes.c_uint64(1 << 40), 50, ctypes.byref(s)) == abi.OK
    jid = ctypes.c_uint64()
    rc = L.alien_server_submit(s, ctypes.byref(d), ctypes.byref(p), _ptr(one), ctypes.c_uint64(400_000_000),
                               _ptr(nopart), ctypes.c_uint64(0), ctypes.c_uint64(10), ctypes.byref(jid))
    assert rc == abi.E_CAPACITY
    assert L.alien_server_destroy(s) == abi.OK


def test_estimator_at_the_set_limit():
    """The admission estimate grows linearly with cells and particles and counts every per-cell
    array at the largest set (a 400M-cell request would need more than 160 GB)."""
    L = engine.lib()
    L.alien_server_estimate_bytes.restype = ctypes.c_uint64
    L.alien_server_estimate_bytes.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32]
    e0 = L.alien_server_estimate_bytes(0, 0, 4096, 4096)
    e1 = L.alien_server_estimate_bytes(100_000_000, 0, 4096, 4096)
    e4 = L.alien_server_estimate_bytes(400_000_000, 0, 4096, 4096)
    assert e4 - e0 == 4 * (e1 - e0)
    assert (e1 - e0) / 100_000_000 > 400  # bytes per cell: two hot copies with six bond planes, scratch
    assert e4 > 160e9


def test_server_admission_counts_the_rebuild_peak():
    """The serving controller's admission rule: while a rebuild builds the new resident set, the old
    one still holds its memory, so old + kept + admitted must fit the budget; admission is FIFO and
    stops at the first job that does not fit."""
    L = ctypes.CDLL(engine.LIB_PATH)
    f = L.alien_server_plan_admission
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64]

    def plan(budget, old, kept, queued):
        arr = (ctypes.c_uint64 * max(1, len(queued)))(*queued)
        return f(budget, old, kept, arr, len(queued))

    GB = 1 << 30
    assert plan(200 * GB, 0, 0, [50 * GB, 60 * GB, 100 * GB]) == 2      # 50 + 60 fit, + 100 does not
    assert plan(200 * GB, 0, 0, [50 * GB, 60 * GB, 90 * GB]) == 3       # exactly the budget
    # a 120 GB resident set: rebuilding it with one more job needs 120 (old) + 120 (kept) + the job
    assert plan(200 * GB, 120 * GB, 120 * GB, [10 * GB]) == 0
    assert plan(300 * GB, 120 * GB, 120 * GB, [10 * GB, 60 * GB]) == 1
    # FIFO: a small job behind one that does not fit waits
    assert plan(200 * GB, 0, 0, [250 * GB, 1 * GB]) == 0
    assert plan(200 * GB, 0, 0, []) == 0
    # no overflow past 2^64
    assert plan(2**64 - 1, 2**63, 2**62, [2**63]) == 0


@pytest.mark.parametrize("what", ["cell", "particle"])
def test_colour_above_the_palette_is_rejected(what):
    """ALIEN has colours 0..6 (MAX_COLORS = 7); a transfer object with colour 7 or more is refused
    with ALIEN_E_ARG instead of being read as colour % 7."""
    L = engine.lib()
    cells = np.zeros(2, dtype=abi.CELL_DTYPE)
    cells["pos"] = [[1.0, 1.0], [5.0, 5.0]]
    parts = np.zeros(1, dtype=abi.PARTICLE_DTYPE)
    parts["pos"] = [2.0, 2.0]
    (cells if what == "cell" else parts)["color"][0] = 9
    b = ctypes.c_void_p()
    assert L.alien_batch_create(ctypes.byref(b)) == abi.OK
    d, p = _desc(), abi.default_params(motionType=abi.MOTION_COLLISION)
    rc = L.alien_batch_add_world(b, ctypes.byref(d), ctypes.byref(p), _ptr(cells), ctypes.c_uint64(2), _ptr(parts),
                                 ctypes.c_uint64(1), None)
    assert rc == abi.E_ARG and b"color" in L.alien_last_error()
    (cells if what == "cell" else parts)["color"][0] = 6
    rc = L.alien_batch_add_world(b, ctypes.byref(d), ctypes.byref(p), _ptr(cells), ctypes.c_uint64(2), _ptr(parts),
                                 ctypes.c_uint64(1), None)
    assert rc == abi.OK
    assert L.alien_batch_destroy(b) == abi.OK
