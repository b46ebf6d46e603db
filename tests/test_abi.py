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
    ref = abi.default_params()
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
