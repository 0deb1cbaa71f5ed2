"""Anti-jamming front end on the GPU (ugsdr_hip_stft / _istft /
_jse_process) against the oracle's restatement of ShortTimeFourierTransform,
JammingDetector and JammingSuppressionEngine.

Bar: STFT cells and overlap-added samples within 2e-6 of the largest
magnitude (FP32 DFTs against FP64 ones rounded to FP32); detections equal in
type, bin and frequency, J/N within 1e-4 relative; suppressed samples within
1e-4 of the input's largest magnitude.  The reference's own STFT KATs run on
the GPU transform too."""
import numpy as np
import pytest

import ugsdr_amd as U
from oracle import refcpu as R

from conftest import DeviceArray
from test_jse import chirp, sine

pytestmark = pytest.mark.gpu

WINDOWS = [U.STFT_BLACKMAN_HARRIS, U.STFT_HANN]


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("window", WINDOWS)
@pytest.mark.parametrize("n", [128, 4092, 4095, 40920])
def test_stft_matches_oracle(ctx, oracle, window, n):
    rng = np.random.default_rng(n)
    x = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    say(f"case stft window {window} n {n}")
    got = ctx.stft(x, window)
    ref = R.stft(x, window)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= 2e-6 * np.max(np.abs(ref))


@pytest.mark.parametrize("ncol", [1, 7, 1275])
def test_istft_matches_oracle(ctx, oracle, ncol):
    rng = np.random.default_rng(ncol)
    cols = (rng.normal(size=(ncol, 128)) + 1j * rng.normal(size=(ncol, 128))).astype(np.complex64)
    say(f"case istft {ncol} columns")
    got = ctx.istft(cols)
    ref = R.istft(cols)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= 2e-6 * np.max(np.abs(ref))


@pytest.mark.parametrize("window", WINDOWS)
def test_stft_reference_kats_on_gpu(ctx, window):
    fs = 50e6
    cols = ctx.stft(sine(1e6, fs), window)
    peaks = np.argmax(np.abs(cols), axis=1) * fs / 128
    assert np.mean(np.abs(peaks - 1e6) <= fs / 128 / 2) >= 0.95
    freqs = np.argmax(np.abs(ctx.stft(chirp(1e6, 10e6, 20e-6, fs), window)), axis=1) * fs / 128
    occupied = (10e6 - 1e6) / (fs / 128)
    for f in np.unique(freqs):
        share = np.mean(freqs == f) * 100
        assert share >= 100 / occupied / 2 or share < 1.0


def scenes():
    fs, n = 4.092e6, 4 * 4092
    k = np.arange(n)
    rng = np.random.default_rng(3)
    cw = (0.5 * (rng.normal(size=n) + 1j * rng.normal(size=n))
          + 6.0 * np.exp(2j * np.pi * (96 * fs / 128 / 8 * k / fs + 0.1))).astype(np.complex64)
    rng = np.random.default_rng(8)
    burst = (k >= 8000) & (k < 8640)
    notch = (0.5 * (rng.normal(size=n) + 1j * rng.normal(size=n)) + 6.0 * np.exp(2j * np.pi * 12 * k / 128)
             + burst * 40.0 * np.exp(2j * np.pi * 70 * k / 128)).astype(np.complex64)
    rng = np.random.default_rng(12)
    sweep = (0.5 * (rng.normal(size=n) + 1j * rng.normal(size=n))
             + 4.0 * np.tile(chirp(-1.5e6, 1.5e6, 4e-3, fs), 4)[:n]).astype(np.complex64)
    clean = (0.5 * (rng.normal(size=n) + 1j * rng.normal(size=n))).astype(np.complex64)
    return fs, {"cw": cw, "notch": notch, "sweep": sweep, "clean": clean}


@pytest.mark.parametrize("name", ["cw", "notch", "sweep", "clean"])
@pytest.mark.parametrize("window", WINDOWS)
def test_jse_matches_oracle(ctx, oracle, name, window):
    fs, sc = scenes()
    x = sc[name]
    say(f"case jse scene {name} window {window}")
    y, det = ctx.jse_process(x, fs, window)
    ry, rdet = R.jse_process(x, fs, window)
    say(f"  detections {[(d[0], d[1]) for d in det]}")
    assert [(d[0], d[1], d[2]) for d in det] == [(d[0], d[1], d[2]) for d in rdet]
    for a, b in zip(det, rdet):
        assert a[3] == pytest.approx(b[3], rel=1e-4)
    assert np.max(np.abs(y - ry)) <= 1e-4 * np.max(np.abs(x))
    if name == "clean":
        assert det == [] and np.array_equal(y, x)
    # the same in place on device memory
    buf = DeviceArray(x.copy())
    try:
        _, det_d = ctx.jse_process(None, fs, window, device_ptr=buf.ptr.value, n=len(x))
        assert det_d == det
        assert np.array_equal(buf.get(), y)
    finally:
        buf.close()


def test_jse_restores_acquisition(ctx):
    fs, spm, n_ms = 4.092e6, 4092, 4
    n = spm * n_ms
    rng = np.random.default_rng(3)
    sv, f0, tau = 7, 1250.0, 2000.0
    code = U.gps_l1ca(sv).astype(np.float64)
    k = np.arange(n)
    chip = np.floor((k - tau) * 1.023e6 / fs).astype(np.int64) % 1023
    x = (0.05 * code[chip] * np.exp(2j * np.pi * f0 * k / fs) + 0.5 * (rng.normal(size=n) + 1j * rng.normal(size=n))
         + 6.0 * np.exp(2j * np.pi * (96 * fs / 128 / 8 * k / fs + 0.1))).astype(np.complex64)
    y, det = ctx.jse_process(x, fs)
    key = ctx.replica_put(U.gps_l1ca(sv), 1, spm)
    try:
        kw = dict(fs_in=fs, fs_acq=fs, n_fft=spm, n_blocks=n_ms, surface_len=spm, doppler=U.doppler_bins(5000, 250),
                  keys=[key])
        before, _, _ = ctx.acquire(x, U.FMT_CF32, **kw)
        after, _, _ = ctx.acquire(y, U.FMT_CF32, **kw)
    finally:
        ctx.replica_drop(key)
    say(f"case jse then acquisition: before snr {before[0]['snr']:.2f}, after {after[0]['snr']:.2f}")
    assert not before[0]["detected"]
    assert after[0]["detected"] and abs(after[0]["index"] - tau) <= 1 and after[0]["doppler"] == f0


def test_jse_refusals(ctx):
    with pytest.raises(U.UgsdrError):
        ctx.jse_process(np.zeros(100, np.complex64), 4e6)
    with pytest.raises(U.UgsdrError):
        ctx.jse_process(np.zeros(1000, np.complex64), 4e6, window=7)
