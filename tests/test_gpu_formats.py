"""Recorded sample formats on the GPU, through the C ABI.

* ugsdr_hip_unpack against the oracle's decoders (BbpDdc, Nt1065 channels 1-4)
  and numpy (CI8, CI16, CF32, R8): bit-exact;
* a broken BbpDdc ms counter is refused before any device work;
* acquisition straight from the BbpDdc recording (tests/golden, 4 ms at
  33.25 MHz, decoded inside the front kernel, resampled to 4.2 MHz) against the
  oracle's decode -> resample -> FP64 acquisition, with the parity rules of
  tests/parity.py, and PRNs 9 and 23 detected;
* acquisition from packed Nt1065 bytes equals acquisition from the same
  samples handed over as complex float (the front kernel sees identical
  values, so every result field matches exactly);
* int16 / float I/Q read in place by the class transforms equal int8 I/Q
  decoded by the front kernel: results, peaks and surfaces bit for bit.
"""
import os

import numpy as np
import pytest

import ugsdr_amd as U
from ugsdr_amd import workloads as W

from parity import compare
from test_formats import BBP_FS, BBP_L1, bbp_recording_search

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _bbp_stream(rng, spm, n_ep, first=1000):
    eb = 16 * ((spm + 16 + 31) // 32)
    raw = rng.integers(0, 256, n_ep * eb, dtype=np.uint8)
    for e in range(n_ep):
        raw[e * eb: e * eb + 8] = np.frombuffer(np.uint64(first + e).tobytes(), np.uint8)
    return raw


@pytest.mark.parametrize("fmt", [U.FMT_BBP_DDC, U.FMT_NT1065_1, U.FMT_NT1065_2, U.FMT_NT1065_3, U.FMT_NT1065_4,
                                 U.FMT_CI8, U.FMT_CI16, U.FMT_CF32, U.FMT_R8])
def test_unpack_matches_oracle(ctx, oracle, fmt):
    rng = np.random.default_rng(100 + fmt)
    say(f"case unpack fmt={fmt}")
    if fmt == U.FMT_BBP_DDC:
        fs, n = 4.0e6, 3 * 4000 + 1234          # a partial last epoch
        raw = _bbp_stream(rng, 4000, 4)
        ref = oracle.unpack_bbp(raw, n, fs)
    elif fmt in (U.FMT_NT1065_1, U.FMT_NT1065_2, U.FMT_NT1065_3, U.FMT_NT1065_4):
        fs, n = 79.5e6, 100003
        raw = rng.integers(0, 256, n, dtype=np.uint8)
        ref = oracle.unpack_nt1065(raw, n, 2 * (U.FMT_NT1065_4 - fmt))
    elif fmt == U.FMT_CF32:
        fs, n = 1e6, 5000
        ref = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
        raw = ref.view(np.float32)
    elif fmt == U.FMT_R8:
        fs, n = 1e6, 5000
        raw = rng.integers(-128, 128, n, dtype=np.int8)
        ref = raw.astype(np.float32).astype(np.complex64)
    else:
        fs, n = 1e6, 5000
        dt = np.int8 if fmt == U.FMT_CI8 else np.int16
        raw = rng.integers(np.iinfo(dt).min, np.iinfo(dt).max, 2 * n, dtype=dt)
        ref = (raw[0::2].astype(np.float32) + 1j * raw[1::2].astype(np.float32)).astype(np.complex64)
    got = ctx.unpack(raw, fmt, n, fs)
    assert np.array_equal(got, ref)


def test_unpack_bbp_rejects_counter_gap(ctx):
    rng = np.random.default_rng(3)
    raw = _bbp_stream(rng, 4000, 3)
    eb = U.format_bytes(U.FMT_BBP_DDC, 4000, 4e6)
    raw[2 * eb] += 1                          # epoch 2 skips a millisecond
    with pytest.raises(U.UgsdrError) as e:
        ctx.unpack(raw, U.FMT_BBP_DDC, 3 * 4000, 4e6)
    assert e.value.code == -1


def test_acquire_bbp_recording_vs_oracle(ctx, oracle):
    raw = np.fromfile(BBP_L1, np.uint8)
    spm_in = int(BBP_FS / 1e3)
    fs_acq = 4.2e6
    N = int(fs_acq / 1e3)
    say("case BBP recording: 4 ms @ 33.25 MHz -> 4.2 MHz, 32 PRN x 41 bins x 4 blocks")
    ref, dop = bbp_recording_search(oracle)
    keys = [ctx.replica_put(U.gps_l1ca(s), 1, N) for s in range(32)]
    try:
        gpu, _, _ = ctx.acquire(raw, U.FMT_BBP_DDC, n_samples=4 * spm_in, fs_in=BBP_FS, fs_acq=fs_acq, n_fft=N,
                                n_blocks=4, surface_len=N, doppler=dop, keys=keys)
    finally:
        for k in keys:
            ctx.replica_drop(k)
    fails, notes = compare(gpu, ref)
    for n in notes:
        say("  note: " + n)
    assert not fails, fails
    det = {p for p, r in enumerate(gpu) if r["detected"]}
    say(f"  detected sv {sorted(det)}")
    assert {8, 22} <= det


def test_acquire_nt1065_packed_equals_decoded(ctx, oracle):
    """a GPS L1 signal quantised to 2 bits in channel 3 of Nt1065 bytes, other
    channels random; packed and pre-decoded inputs give identical results"""
    w = W.l1ca_1ms()
    rng = np.random.default_rng(11)
    x = w.iq[0::2].astype(np.float32)                 # real part as an IF-free real signal
    q = np.clip(np.round(x / (np.std(x) + 1e-9)), -2, 1)   # -> {-2,-1,0,1} -> values {-3,-1,1,3}
    sign = (q >= 0).astype(np.uint8)
    mag = ((q == -2) | (q == 1)).astype(np.uint8)
    raw = rng.integers(0, 256, x.size, dtype=np.uint8)
    raw = (raw & ~np.uint8(0x0C)) | (sign << 2) | (mag << 3)
    dec = oracle.unpack_nt1065(raw, raw.size, 2)
    say("case NT1065 channel 3 packed vs decoded, L1 C/A 4092 x 21 bins")
    keys = [ctx.replica_put(U.gps_l1ca(s), 1, w.n_fft) for s in range(32)]
    kw = dict(fs_in=w.fs, fs_acq=w.fs, n_fft=w.n_fft, n_blocks=1, surface_len=w.n_fft, doppler=w.doppler, keys=keys)
    try:
        a, _, _ = ctx.acquire(raw, U.FMT_NT1065_3, n_samples=raw.size, **kw)
        b, _, _ = ctx.acquire(dec, U.FMT_CF32, n_samples=dec.size, **kw)
    finally:
        for k in keys:
            ctx.replica_drop(k)
    assert a == b


@pytest.mark.parametrize("make", [W.l1ca_1ms, W.l1ca_4ms], ids=["L1CA_1MS", "L1CA_4MS_fold"])
def test_in_place_reads_equal_front_kernel(ctx, make):
    """int16 and float I/Q at the acquisition rate with no tone are read in
    place by the class transforms (no front kernel); int8 I/Q goes through the
    front kernel.  The same int8 values in the three formats decode to the same
    floats, so results, peaks and surfaces match exactly; the entry point that
    allocates nothing matches them to the parity rules."""
    w = make()
    say(f"case in-place reads vs front kernel, {w.name} n_fft={w.n_fft}")
    x8 = w.iq.astype(np.int8)
    x16 = x8.astype(np.int16)
    xf = x8.astype(np.float32)
    keys = [ctx.replica_put(c, w.repeats, w.n_fft) for c in w.codes]
    kw = dict(fs_in=w.fs, fs_acq=w.fs, n_fft=w.n_fft, n_blocks=w.n_blocks, surface_len=w.surface_len,
              doppler=w.doppler, keys=keys, reshape=w.reshape)
    try:
        ref = ctx.acquire(x8, U.FMT_CI8, want_surfaces=True, want_peaks=True, **kw)
        for x, fmt in ((x16, U.FMT_CI16), (xf, U.FMT_CF32)):
            got = ctx.acquire(x, fmt, want_surfaces=True, want_peaks=True, **kw)
            assert got[0] == ref[0], fmt
            for g, r in zip(got[1:], ref[1:]):
                np.testing.assert_array_equal(g, r)
        # the entry point without allocation runs the length's default variant
        # (no autotuned choice), so its sums may round differently: parity rules
        a, _, _ = ctx.acquire_noalloc(x16, U.FMT_CI16, **kw)
        fails, _ = compare(a, ref[0])
        assert not fails, fails
    finally:
        for k in keys:
            ctx.replica_drop(k)
