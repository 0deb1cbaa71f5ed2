"""Per-operator drop-ins (the Hip* CRTP backends' C ABI) vs the oracle and the
reference's unit KATs (test/basic_tests.cpp)."""
import numpy as np
import pytest

import ugsdr_amd as U

pytestmark = pytest.mark.gpu


def test_mixer(ctx, oracle):
    print("case ops/mixer", flush=True)
    ones = np.ones(1000, np.complex64)
    y = ctx.mix_inplace(ones, 1e6, 5e5, 0.0)   # test/basic_tests.cpp:321-350
    assert np.array_equal(y.real, 1.0 - 2.0 * (np.arange(1000) & 1))
    assert np.max(np.abs(y.imag)) < 1e-6
    rng = np.random.default_rng(5)
    x = (rng.normal(size=20000) + 1j * rng.normal(size=20000)).astype(np.complex64)
    for f, ph in [(-1250.0, 0.0), (3000.0, 1.0), (-16.368e6 / 3, -7.0)]:
        np.testing.assert_allclose(ctx.mix(x, 16.368e6, f, ph), oracle.mix(x, 16.368e6, f, ph), atol=2e-5)
    np.testing.assert_array_equal(ctx.mix_inplace(x, 4e6, 0.0, 1.0), x)   # mixer.hpp:36-37


TABLE_MIXER_CASES = [
    (4.092e6, 1.25e3, 0.0), (4.092e6, -3.7e3, 1.0), (40.92e6, 5.0e3, -2.5), (16.368e6, 4.1234e6, 9.0),
    (2.0e6, 1.0e6, 0.0), (4.2e6, 0.0, 1.0), (1.0e6, 1.3e6, 0.3), (33.25e6, -8.4205e6, 6.283185307179586),
]


@pytest.mark.parametrize("fs,f,phase", TABLE_MIXER_CASES)
def test_table_mixer_bit_exact(ctx, oracle, fs, f, phase):
    """HipTableMixer against the oracle's TableMixer, itself bit-exact with the
    reference's own TableMixer (tests/test_ref_pins.py): every output bit equal"""
    print(f"case ops/table_mixer fs={fs} f={f} phase={phase}", flush=True)
    rng = np.random.default_rng(int(fs) % 997)
    x = (100 * (rng.normal(size=70001) + 1j * rng.normal(size=70001))).astype(np.complex64)
    got = ctx.table_mix(x, fs, f, phase)
    np.testing.assert_array_equal(got.view(np.uint32), oracle.table_mix(x, fs, f, phase).view(np.uint32))
    from oracle import refsrc
    if refsrc.available():
        np.testing.assert_array_equal(got.view(np.uint32), refsrc.table_mix(x, fs, f, phase).view(np.uint32))


def test_table_mixer_member_and_device(ctx, oracle):
    """member Translate (mixer.hpp:53-60) over five 1 ms calls, and the device
    memory form, equal the oracle bit for bit"""
    from conftest import DeviceArray
    print("case ops/table_mixer member + device", flush=True)
    fs, f, phase, n = 4.092e6, 2.375e3, 0.25, 4092
    rng = np.random.default_rng(11)
    x = (30 * (rng.normal(size=5 * n) + 1j * rng.normal(size=5 * n))).astype(np.complex64)
    ph_g = ph_o = phase
    for c in range(5):
        g, ph_g = ctx.table_mixer_member(x[c * n:(c + 1) * n], fs, f, ph_g)
        o, ph_o = oracle.table_mixer_member(x[c * n:(c + 1) * n], fs, f, ph_o)
        np.testing.assert_array_equal(g.view(np.uint32), o.view(np.uint32))
        assert ph_g == ph_o
    d = DeviceArray(x)
    try:
        ctx.table_mix_device(d.ptr.value, len(x), fs, -f, 1.5)
        np.testing.assert_array_equal(d.get().view(np.uint32), oracle.table_mix(x, fs, -f, 1.5).view(np.uint32))
    finally:
        d.close()
    with pytest.raises(U.UgsdrError):
        ctx.table_mix(x[:8], fs, -3 * fs, 0.0)


@pytest.mark.parametrize("n", [511, 1000, 4092, 16368, 38192])
def test_dft(ctx, n):
    print(f"case ops/dft n={n}", flush=True)
    rng = np.random.default_rng(n)
    x = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    ref = np.fft.fft(x.astype(np.complex128))
    assert np.max(np.abs(ctx.dft(x) - ref)) < 2e-6 * np.abs(ref).max()
    refi = np.fft.ifft(x.astype(np.complex128))
    assert np.max(np.abs(ctx.dft(x, True) - refi)) < 2e-6 * np.abs(refi).max()


def test_dft_refuses_large_prime_factor(ctx):
    """a stage's cost is n times its radix: lengths with a prime factor over
    4096 are refused instead of running a quadratic stage"""
    with pytest.raises(U.UgsdrError):
        ctx.dft(np.zeros(4099, np.complex64))
    assert ctx.dft(np.ones(4093, np.complex64))[0] == pytest.approx(4093, rel=1e-5)


def test_dft_tone_kat(ctx):
    x = np.exp(2j * np.pi * np.arange(1000) * 0.1).astype(np.complex64)   # basic_tests.cpp:482-517
    y = ctx.dft(x)
    assert abs(y[100].real - 1000) < 5e-3 and abs(y[100].imag) < 5e-3


def test_matched_filter_kat(ctx, oracle):
    """basic_tests.cpp:368-379 (GLONASS m-sequence: 511 / -1)"""
    print("case ops/matched filter", flush=True)
    g = U.glonass_of().astype(np.float32)
    y = ctx.mf_filter_code(g.astype(np.complex64), g)
    assert abs(y[0].real - 511) < 5e-3 and np.max(np.abs(y[1:].real + 1)) < 5e-3
    spec = ctx.mf_prepare(g)
    np.testing.assert_allclose(spec, oracle.mf_prepare(g), atol=1e-3)
    rng = np.random.default_rng(9)
    x = (rng.normal(size=511) + 1j * rng.normal(size=511)).astype(np.complex64)
    np.testing.assert_allclose(ctx.mf_filter(x, spec), oracle.mf_process(x, oracle.mf_prepare(g)), atol=1e-4)


def test_math_ops(ctx, oracle):
    print("case ops/math", flush=True)
    a = ctx.abs(np.full(1000, 1 - 1j, np.complex64))   # basic_tests.cpp:408-433
    assert np.all(np.abs(a - np.sqrt(2)) < 1e-2)
    v, i = ctx.max_index(np.arange(1000, dtype=np.float32))   # :528-553
    assert (v, i) == (999.0, 999)
    assert ctx.max_index(np.array([1, 5, 5, 2], np.float32))[1] == 1
    rng = np.random.default_rng(0)
    m, s = ctx.mean_std(rng.normal(0, 1, 1000000).astype(np.float32))   # :564-596
    assert abs(m) < 0.01 and abs(s - 1) < 0.01
    x = rng.normal(size=5000).astype(np.float32)
    m, s = ctx.mean_std(x)
    mo, so = oracle.mean_std(x)
    assert abs(m - mo) < 1e-6 and abs(s - so) < 1e-6 * so
    np.testing.assert_array_equal(ctx.reshape_sum(np.ones(10000, np.float32), 1000), np.full(1000, 10.0))   # :607-630
    z = (rng.normal(size=400) + 1j * rng.normal(size=400)).astype(np.complex64)
    np.testing.assert_allclose(ctx.reshape_sum(z, 100), z.reshape(4, 100).sum(axis=0), atol=1e-5)
    with pytest.raises(U.UgsdrError, match="SIZE"):
        ctx.reshape_sum(np.ones(10, np.float32), 3)


def test_resample_ops(ctx, oracle):
    print("case ops/resample", flush=True)
    slope = (np.arange(1000, dtype=np.float32) * (1 + 1j)).astype(np.complex64)
    np.testing.assert_allclose(ctx.resample(slope, 100, 1000), oracle.accumulate(slope, 10), atol=1e-4)
    np.testing.assert_array_equal(ctx.upsample(np.arange(1000, dtype=np.float32), 10000),
                                  np.repeat(np.arange(1000, dtype=np.float32), 10))
    rng = np.random.default_rng(2)
    x = (rng.normal(size=152768) + 1j * rng.normal(size=152768)).astype(np.complex64)
    np.testing.assert_allclose(ctx.resample(x, 9548000, 38192000), oracle.resample(x, 9548000, 38192000),
                               atol=1e-5)
    np.testing.assert_allclose(ctx.resample(x[:1200], 4, 6), oracle.resample(x[:1200], 4, 6), atol=1e-5)


def test_correlator(ctx, oracle):
    print("case ops/correlator", flush=True)
    c = U.gps_l1ca(0).astype(np.float32)
    assert ctx.correlate(c.astype(np.complex64), c) == complex(1023, 0)   # basic_tests.cpp:72-86
    rng = np.random.default_rng(4)
    x = (rng.normal(size=40920) + 1j * rng.normal(size=40920)).astype(np.complex64)
    code = np.sign(rng.normal(size=40920)).astype(np.float32)
    r = ctx.correlate(x, code)
    o = oracle.correlate(x, code)
    assert abs(r - o) < 1e-4 * np.sqrt(40920)
    with pytest.raises(U.UgsdrError, match="SIZE"):   # correlator.hpp:30-31
        ctx.correlate(x[:10], code)
