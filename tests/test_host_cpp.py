"""The C++20 host surface (ugsdr_amd/host/ugsdr_hip/*.hpp): the CRTP bases and
Hip* backends compile against include/ugsdr_hip.h (CPU), and the C++ test
program passes the reference's unit KATs and the real_8 acquisition KAT on
the GPU through both the per-operator and the fused engine, and finds the
strong satellites of the packed BbpDdc recording through the fused engine."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "test_host_api.cpp")
BIN = os.path.join(ROOT, "_build", "test_host_api")
HOST = os.path.join(ROOT, "ugsdr_amd", "host")


def build_binary():
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    headers = [os.path.join(HOST, "ugsdr_hip", f) for f in os.listdir(os.path.join(HOST, "ugsdr_hip"))]
    lib = os.path.join(ROOT, "ugsdr_amd", "libugsdr_hip.so")
    deps = [SRC, lib, *headers]
    if os.path.exists(BIN) and all(os.path.getmtime(BIN) >= os.path.getmtime(d) for d in deps):
        return
    subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", "-Werror", "-I" + HOST, SRC,
                    "-L" + os.path.join(ROOT, "ugsdr_amd"), "-lugsdr_hip",
                    "-Wl,-rpath," + os.path.join(ROOT, "ugsdr_amd"), "-o", BIN], check=True)


def test_host_surface_compiles():
    build_binary()
    assert os.path.exists(BIN)


def test_reference_bases_mode_compiles():
    """with UGSDR_HIP_REFERENCE_BASES the headers expect the reference's bases:
    supplying them from our own bases.hpp (as a stand-in include) must still compile"""
    code = ('#include "ugsdr_hip/bases.hpp"\n#define UGSDR_HIP_REFERENCE_BASES\n#include "ugsdr_hip/backends.hpp"\n'
            "int main() { return 0; }\n")
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-I" + HOST, "-x", "c++", "-"], input=code,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_host_api_on_gpu():
    print("case C++ host API (unit KATs + real_8 per-operator vs fused)", flush=True)
    build_binary()
    r = subprocess.run([BIN, os.path.join(ROOT, "tests", "golden", "real_8_first4ms.bin"),
                        os.path.join(ROOT, "tests", "golden", "bbp_ddc_gps_L1_first4ms.bin")], capture_output=True,
                       text=True, timeout=600)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "host API ok" in r.stdout
