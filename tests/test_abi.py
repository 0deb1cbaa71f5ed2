"""CPU checks of the C ABI library: it loads, exports every entry point that
include/ugsdr_hip.h declares, and its host-only functions (PRN codes, rate
rules) agree with the oracle.  No GPU compute is called here."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import ugsdr_amd as U
from ugsdr_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ugsdr_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w]+\s*\*?\s*(ugsdr_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = U.lib()
    decl = declared_functions()
    assert len(decl) >= 30
    for name in decl:
        assert hasattr(lib, name), name
        assert name in U.SIGNATURES, f"{name} has no ctypes signature"
    out = subprocess.run(["nm", "-D", "--defined-only", U.HIP_LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(ugsdr_\w+)$", out, flags=re.M))
    assert set(decl) <= exported
    assert U.lib().ugsdr_hip_version().decode().endswith("gfx950")


def test_library_is_gfx950_code_object():
    """the offload bundle of the library built here targets gfx950 only"""
    blob = open(U.HIP_LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    for other in (b"--gfx942", b"--gfx90a", b"--gfx1100"):
        assert other not in blob


def test_codes_match_oracle(oracle):
    for sv in range(37):
        np.testing.assert_array_equal(U.gps_l1ca(sv), oracle.gps_l1ca(sv))
    for sv in range(32):
        np.testing.assert_array_equal(U.gps_l5i(sv), oracle.gps_l5i(sv))
        np.testing.assert_array_equal(U.gps_l5q(sv), oracle.gps_l5q(sv))
    np.testing.assert_array_equal(U.glonass_of(), oracle.glonass_of())
    c = U.gps_l1ca(4)
    np.testing.assert_array_equal(U.boc(c), oracle.boc_expand(c))
    with pytest.raises(U.UgsdrError, match="ARG"):
        U.gps_l1ca(37)
    with pytest.raises(U.UgsdrError, match="ARG"):
        U.gps_l5i(32)


def test_rate_rules_match_oracle(oracle):
    for fs in (4.092e6, 16.368e6, 20.46e6, 38.192e6, 79.5e6, 53e6, 4e6, 8.184e6):
        for target in (2.048e6, 8.192e6, 20.46e6, 40.96e6):
            assert U.adjust_sampling_rate(fs, target) == oracle.adjust_sampling_rate(fs, target), (fs, target)
    for n, a, b in [(152768, 9548000, 38192000), (12, 4, 6), (1000, 8192000, 4000000), (5, 3, 7)]:
        assert U.resample_len(n, a, b) == len(oracle.resample(np.ones(n, np.complex64), a, b))


def test_doppler_bins_follow_reference_loop():
    """f = -range; f <= range; f += step (fse.hpp:250-252)"""
    d = U.doppler_bins(5000, 500)
    assert len(d) == 21 and d[0] == -5000 and d[-1] == 5000
    assert len(U.doppler_bins(5000, 250)) == 41
    assert len(U.doppler_bins(6000, 125)) == 97
    assert len(U.doppler_bins(6000, 200)) == 61


def test_context_without_gpu_fails_cleanly():
    try:
        import torch
        has_gpu = torch.cuda.device_count() > 0
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(U.UgsdrError):
        U.Context(0)


def test_siggen_deterministic_and_calibrated():
    code = U.gps_l1ca(0)
    sat = W.Sat(0, code, 1.023e6, W.L1_HZ, 100.25, 1500.0, 0.1, 45.0)
    a = W.synthesize([sat], 4.092e6, 40920, 8.0, 7, 32, threads=1)
    b = W.synthesize([sat], 4.092e6, 40920, 8.0, 7, 32, threads=5)
    np.testing.assert_array_equal(a, b)
    noise = W.synthesize([], 4.092e6, 200000, 8.0, 7, 32)
    assert abs(noise.std() - 8.0) < 0.05 and abs(noise.mean()) < 0.05
    # carrier-to-noise density of the planted satellite: A^2 / (2 sigma^2 / fs)
    x = a[0::2] + 1j * a[1::2]
    n = np.arange(40920)
    chip = np.floor(n * 1.023e6 * (1 + 1500.0 / W.L1_HZ) / 4.092e6 - 100.25).astype(int) % 1023
    ref = code[chip] * np.exp(2j * np.pi * (1500.0 * n / 4.092e6 + 0.1))
    amp = np.vdot(ref, x).real / len(n)
    cn0 = 10 * np.log10(amp ** 2 * 4.092e6 / (2 * 64.0))
    assert abs(cn0 - 45.0) < 0.6
    q = W.synthesize([sat], 4.092e6, 4092, 8.0, 7, 8)
    assert q.dtype == np.int8 and np.abs(q).max() <= 127


def test_workloads_shapes():
    w = W.l1ca_1ms()
    assert w.n_samples == 4092 and w.cells == 32 * 21 * 4092 and w.iq.dtype == np.int8
    assert len(w.truth()) == 6
    reps = w.replicas()
    assert reps.shape == (32, 4092)
    np.testing.assert_array_equal(reps[3][::4], U.gps_l1ca(3))


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in (r.stdout + r.stderr)
