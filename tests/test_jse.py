"""The anti-jamming front end's oracle (STFT, JammingDetector,
JammingSuppressionEngine restated in oracle/refcpu.py), pinned by the
reference's own STFT tests, closed forms and hand-made tables:

* the reference's STFT KATs (test/basic_tests.cpp:664-730): a 1 MHz tone at
  50 MHz peaks within half a bin in >= 95 % of the columns, a 1-10 MHz chirp
  visits its bins, for both windows;
* an on-bin tone's cell magnitude is A sum(w) in every column (closed form);
* the periodic Hann window at hop 32 overlap-adds to 2, so the inverse of the
  forward STFT is 2 x inside the covered span (closed form);
* both NullRegion variants on hand cases (wrap at both ends, and the
  engine's variant on a spectrum of at most 256 bins, where its loop bounds
  do not wrap);
* AnalyzeStftData on hand-made candidate tables: one frequency in >= 90 % of
  the columns (alone: narrowband and, by the size_t(size * 0.9) rule, also
  wideband; beside a rare one: narrowband only); candidates spread over many
  frequencies: wideband;
* end to end: a CW jammer 30 dB over the noise is found at its bin and
  removed (> 30 dB less power at its frequency), and a GPS L1 C/A signal
  that the jammer hides from the acquisition is found again afterwards.
"""
import numpy as np
import pytest

import ugsdr_amd as U
from oracle import refcpu as R


def sine(f, fs):
    n = int(fs / 1e3)
    return np.exp(2j * np.pi * np.arange(n) * f / fs).astype(np.complex64)


def chirp(f0, f1, period, fs):
    n = int(fs / 1e3)
    k = (f1 - f0) / period
    spc = int(np.ceil(period * fs))
    t = (np.arange(n) % spc) / fs
    return np.exp(2j * np.pi * (k * t * t / 2 + f0 * t)).astype(np.complex64)


@pytest.mark.parametrize("kind", [R.WINDOW_BLACKMAN_HARRIS, R.WINDOW_HANN])
def test_stft_reference_kats(oracle, kind):
    fs = 50e6
    cols = R.stft(sine(1e6, fs), kind)
    peaks = np.argmax(np.abs(cols), axis=1) * fs / 128
    near = np.abs(peaks - 1e6) <= fs / 128 / 2
    assert np.mean(near) >= 0.95
    cols = R.stft(chirp(1e6, 10e6, 20e-6, fs), kind)
    freqs = np.argmax(np.abs(cols), axis=1) * fs / 128
    step = fs / 128
    occupied = (10e6 - 1e6) / step
    for f in np.unique(freqs):
        share = np.mean(freqs == f) * 100
        assert share >= 100 / occupied / 2 or share < 1.0


@pytest.mark.parametrize("kind", [R.WINDOW_BLACKMAN_HARRIS, R.WINDOW_HANN])
def test_stft_on_bin_tone_closed_form(oracle, kind):
    fs, k0, amp = 4.096e6, 37, 1.7
    x = (amp * np.exp(2j * np.pi * k0 * np.arange(4096) / 128)).astype(np.complex64)
    cols = R.stft(x, kind)
    assert cols.shape == ((4096 - 96) // 32, 128)
    w = R.stft_window(kind).astype(np.float64)
    assert np.allclose(np.abs(cols[:, k0]), amp * w.sum(), rtol=1e-5)


def test_hann_overlap_add_closed_form(oracle):
    rng = np.random.default_rng(1)
    x = (rng.normal(size=2048) + 1j * rng.normal(size=2048)).astype(np.complex64)
    y = R.istft(R.stft(x, R.WINDOW_HANN))
    assert len(y) == 128 + 32 * ((2048 - 96) // 32 - 1)
    assert np.allclose(y[96:len(y) - 96], 2 * x[96:len(y) - 96], atol=1e-5)


def test_null_region_variants(oracle):
    d = np.ones(16, np.float32)
    R.null_region_detector(1, 3, d)
    assert list(np.flatnonzero(d == 0)) == [0, 1, 2, 3, 4, 14, 15]
    d = np.ones(16, np.float32)
    R.null_region_detector(14, 3, d)
    assert list(np.flatnonzero(d == 0)) == [0, 1, 11, 12, 13, 14, 15]
    d = np.ones(600, np.float32)
    R.null_region_jse(5, 128, d)
    z = np.flatnonzero(d == 0)
    assert len(z) == 257 and z[0] == 0 and z[133] == 133 and z[134] == 600 - 123
    # a spectrum of 256 or fewer bins: start = 50 - 128 + 200 = 122 does not
    # exceed finish = 178, so the loop zeroes [122, 178], away from bin 50
    d = np.ones(200, np.float32)
    R.null_region_jse(50, 128, d)
    assert list(np.flatnonzero(d == 0)) == list(range(122, 179))


def _col(*cands):
    return [(b, b * 1e3, jn) for b, jn in cands]


def test_analyze_hand_tables(oracle):
    one = [_col((10, 20.0))] * 50
    det = R.jse_analyze(one)
    assert [d[0] for d in det] == [R.NARROWBAND, R.WIDEBAND] and det[0][1] == 10 and det[0][3] == pytest.approx(20.0)
    mostly = [_col((10, 20.0))] * 48 + [_col((3, 9.0))] * 2
    det = R.jse_analyze(mostly)
    assert det == [(R.NARROWBAND, 10, 10e3, pytest.approx(20.0 * 48 / 50))]
    spread = [_col((b % 40, 7.0)) for b in range(200)]
    det = R.jse_analyze(spread)
    assert [d[0] for d in det] == [R.WIDEBAND]
    assert R.jse_analyze([[] for _ in range(10)]) == []
    # the mean count is an integer division: 1.9 candidates a column -> one round
    two = [_col((10, 30.0), (50, 8.0))] * 9 + [_col((10, 30.0))]
    assert len(R.jse_analyze(two)) == 2


def test_cw_jammer_removed_and_acquisition_recovers(oracle):
    fs, spm, n_ms = 4.092e6, 4092, 4
    n = spm * n_ms
    rng = np.random.default_rng(3)
    sv, f0, tau = 7, 1250.0, 2000.0
    code = R.gps_l1ca(sv).astype(np.float64)
    k = np.arange(n)
    chip = np.floor((k - tau) * 1.023e6 / fs).astype(np.int64) % 1023
    sig = 0.05 * code[chip] * np.exp(2j * np.pi * f0 * k / fs)
    noise = 0.5 * (rng.normal(size=n) + 1j * rng.normal(size=n))
    fj = 96 * fs / 128 / 8                         # on an STFT bin
    jam = 6.0 * np.exp(2j * np.pi * (fj * k / fs + 0.1))
    x = (sig + noise + jam).astype(np.complex64)
    y, det = R.jse_process(x, fs, R.WINDOW_BLACKMAN_HARRIS)
    # every column's strongest bin is the jammer's: a one-entry map, reported
    # narrowband and wideband, so the STFT excision runs
    assert [(d[0], d[1]) for d in det] == [(R.NARROWBAND, 12), (R.WIDEBAND, 12)]
    assert det[0][2] == pytest.approx(fj)
    tone = np.exp(-2j * np.pi * fj * k / fs)
    p_before, p_after = abs(np.mean(x * tone)) ** 2, abs(np.mean(y * tone)) ** 2
    assert 10 * np.log10(p_before / p_after) > 30

    rep = R.repeat_upsample(R.gps_l1ca(sv), 1, spm)[None, :]
    kw = dict(n_fft=spm, n_blocks=n_ms, surface_len=spm, fs_in=fs, fs_acq=fs, doppler=U.doppler_bins(5000, 250))
    before, _, _ = R.acquire(x, rep, **kw)
    after, _, _ = R.acquire(y, rep, **kw)
    assert not before[0]["detected"]
    assert after[0]["detected"] and abs(after[0]["index"] - tau) <= 1 and after[0]["doppler"] == f0


def test_narrowband_notch_path(oracle):
    """a CW jammer that tops 96 % of the columns and a stronger burst at
    another frequency that tops the rest: a two-entry map with one entry over
    0.9, so only a narrowband detection, and the whole-signal notch (+-128
    bins around the CW) removes the CW and leaves the burst"""
    fs, n = 4.092e6, 4 * 4092
    rng = np.random.default_rng(8)
    k = np.arange(n)
    f1, f2 = 12 * fs / 128, 70 * fs / 128
    x = 0.5 * (rng.normal(size=n) + 1j * rng.normal(size=n)) + 6.0 * np.exp(2j * np.pi * f1 * k / fs)
    burst = (k >= 8000) & (k < 8000 + 20 * 32)
    x = (x + burst * 40.0 * np.exp(2j * np.pi * f2 * k / fs)).astype(np.complex64)
    y, det = R.jse_process(x, fs, R.WINDOW_BLACKMAN_HARRIS)
    assert [(d[0], d[1]) for d in det] == [(R.NARROWBAND, 12)]
    p = lambda v, f, m: abs(np.mean((v * np.exp(-2j * np.pi * f * k / fs))[m])) ** 2
    everywhere = np.ones(n, bool)
    assert 10 * np.log10(p(x, f1, everywhere) / p(y, f1, everywhere)) > 30
    assert 10 * np.log10(p(x, f2, burst) / p(y, f2, burst)) < 0.5
