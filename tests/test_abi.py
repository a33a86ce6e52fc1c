"""CPU tests of the C-ABI library: it loads, exports every symbol the header
declares, and its host-side argument checks follow the reference's
preconditions.  No compute call is made here."""

import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "onedal_amd", "kmeans_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:km_status|const char\*)\s+(km_\w+)\s*\(", text, re.M)))


def test_header_declares_expected_entry_points():
    from onedal_amd.engine import EXPORTED_SYMBOLS

    assert declared_functions() == sorted(EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    from onedal_amd.engine import load_library
    from onedal_amd.build import LIB_PATH

    lib = load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    for name in declared_functions():
        assert name in exported, name


def test_null_context_is_rejected():
    from onedal_amd.engine import load_library, KM_INVALID_ARGUMENT

    lib = load_library()
    assert lib.km_context_destroy(None) == KM_INVALID_ARGUMENT
    assert lib.km_context_set_option(None, 1, 0) == KM_INVALID_ARGUMENT
    assert lib.km_lloyd_train_f32(None, None, 0, 0, None, 0, 0, 0.0, 0, None, None, None,
                                  None) == KM_INVALID_ARGUMENT


def test_library_does_not_link_the_oracle():
    from onedal_amd.build import LIB_PATH

    out = subprocess.run(["nm", "-D", LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    assert "oracle_" not in out
    ldd = subprocess.run(["ldd", LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    assert "oracle" not in ldd


def test_cpp_api_header_compiles_and_validates():
    """The oneapi::dal::kmeans C++ surface: descriptor setters and argument checks
    throw the reference's exception types with its messages (no device needed)."""
    src = os.path.join(ROOT, "tests", "cpp", "test_api_host.cpp")
    if not os.path.exists(src):
        pytest.skip("C++ API test source not present")
    exe = os.path.join(ROOT, "tests", "cpp", "build", "test_api_host")
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    lib_dir = os.path.join(ROOT, "onedal_amd", "lib")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"), src,
                    "-L", lib_dir, "-lkmeans_hip", "-Wl,-rpath," + lib_dir, "-o", exe], check=True)
    res = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert res.returncode == 0, res.stdout


def test_struct_layouts_match_the_header(tmp_path):
    """Every struct the binding passes across the ABI has the header's size and
    field offsets (gcc on the header itself prints them)."""
    from onedal_amd import engine as E

    structs = {"km_stats": E.KmStats, "km_batch_instance_f32": E.KmBatchInstance,
               "km_server_config": E.KmServerConfig, "km_job_result": E.KmJobResult,
               "km_server_stats": E.KmServerStats}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "onedal_amd/kmeans_hip.h"',
             'int main(void) {']
    for cname, cls in structs.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for field, _ in cls._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, field, cname, field))
    lines += ['return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            name, field, value = line.split()
            got[(name, field)] = int(value)
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, (cname, field)
