"""The public header as a foreign-function binding sees it (CPU only).

include/nigiri_amd.h must compile on its own as C99 and C++17 (a cgo or JNI
binding includes it from C), and the ctypes mirror in nigiri_amd/__init__.py
must agree with it byte for byte: the size of every mirrored struct and the
offset of every field, as gcc lays them out for this platform.
"""
import ctypes as C
import os
import subprocess

import pytest

import nigiri_amd as na

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nigiri_amd.h")

MIRRORS = [
    (na.TtView, "nigiri_tt_view"),
    (na.Query, "nigiri_query"),
    (na.Journey, "nigiri_journey"),
    (na.Leg, "nigiri_leg"),
    (na.Result, "nigiri_result"),
    (na.AlgoInit, "nigiri_algo_init"),
    (na.EngineStats, "nigiri_engine_stats"),
    (na.EngineConfig, "nigiri_engine_config"),
    (na.TuneTrial, "nigiri_tune_trial"),
]


def c_field(name):
    return "from" if name == "from_" else name  # a Python keyword in the mirror


@pytest.mark.parametrize("compiler,std", [("gcc", "c99"), ("g++", "c++17")])
def test_header_compiles_on_its_own(tmp_path, compiler, std):
    src = tmp_path / ("h.c" if compiler == "gcc" else "h.cc")
    src.write_text(f'#include "{HEADER}"\nint main(void) {{ return nigiri_amd_version() == 0; }}\n')
    subprocess.run([compiler, f"-std={std}", "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only",
                    str(src)], check=True, capture_output=True)


def test_ctypes_mirror_matches_the_header_layout(tmp_path):
    lines = [f'#include "{HEADER}"', "#include <stddef.h>", "#include <stdio.h>", "int main(void) {"]
    for cls, cname in MIRRORS:
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            fc = c_field(f[0])
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {fc}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in filter(None, out):
        cname, field, value = ln.split()
        got[(cname, field)] = int(value)
    for cls, cname in MIRRORS:
        assert got[(cname, "sizeof")] == C.sizeof(cls), cname
        for f in cls._fields_:
            assert got[(cname, f[0])] == getattr(cls, f[0]).offset, (cname, f[0])
