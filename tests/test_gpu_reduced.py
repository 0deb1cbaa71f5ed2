"""Reduced-precision path (16-bit storage, matrix-core DFT stages) through the
C ABI, against numpy's FP64 FFT, the FP64 oracle and the FP32 GPU path.

Stated error bounds (written into the asserts below):
  DFT operator, relative RMS error vs FP64:  FP16 <= 1e-3, BF16 <= 8e-3
  acquisition surface, ||S - S64|| / ||S64||: FP16 <= 1e-3, BF16 <= 8e-3
  peak-to-noise (level / (mean + sigma)) vs FP32, planted satellites: <= 0.1 dB
  code-phase index and Doppler bin of planted satellites: equal to FP32's
"""
import numpy as np
import pytest

import ugsdr_amd as U
from ugsdr_amd import workloads as W

from parity import iq_to_complex

pytestmark = pytest.mark.gpu

DFT_TOL = {U.PREC_FP16: 1e-3, U.PREC_BF16: 8e-3}
SURF_TOL = {U.PREC_FP16: 1e-3, U.PREC_BF16: 8e-3}
PNR_DB = 0.1
NAME = {U.PREC_FP16: "fp16", U.PREC_BF16: "bf16"}


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("prec", [U.PREC_FP16, U.PREC_BF16])
@pytest.mark.parametrize("n", [60, 112, 120, 180, 240, 341])
def test_dft_reduced_vs_fp64(ctx, n, prec):
    """every two-stage matrix-core plan, both directions, random input"""
    rng = np.random.default_rng(n)
    x = (rng.normal(size=(37, n)) + 1j * rng.normal(size=(37, n))).astype(np.complex64)
    say(f"case dft_reduced n={n} {NAME[prec]}")
    for inverse in (False, True):
        got = ctx.dft_reduced(x, prec, inverse)
        ref = np.fft.ifft(x.astype(np.complex128), axis=1) if inverse else np.fft.fft(x.astype(np.complex128), axis=1)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        say(f"  inverse={inverse} rel rms err {err:.2e}")
        assert err <= DFT_TOL[prec], (n, prec, inverse, err)
    # a tone lands in one bin
    t = np.exp(2j * np.pi * 7 * np.arange(n) / n).astype(np.complex64)
    spec = ctx.dft_reduced(t, prec)[0]
    assert int(np.argmax(np.abs(spec))) == 7


def test_dft_reduced_errors(ctx):
    with pytest.raises(U.UgsdrError):
        ctx.dft_reduced(np.zeros(100, np.complex64), U.PREC_FP16)
    with pytest.raises(U.UgsdrError):
        ctx.dft_reduced(np.zeros(240, np.complex64), U.PREC_FP32)


def _boc_subset():
    w = W.boc_4ms()
    planted = sorted(s.prn for s in w.sats)
    absent = [p for p in range(len(w.codes)) if p not in planted][:2]
    return w, planted[:4] + absent, planted[:4]


@pytest.mark.parametrize("prec", [U.PREC_FP16, U.PREC_BF16])
def test_reduced_boc_vs_fp64_and_fp32(ctx, oracle, prec):
    """BOC(1,1) 4 ms x 5 blocks, N = 81840, all 97 bins, 4 planted + 2 absent PRNs"""
    w, prns, planted = _boc_subset()
    say(f"case reduced BOC_4MS {NAME[prec]} PRNs {prns}")
    k32 = [ctx.replica_put(w.codes[p], 1, w.n_fft) for p in prns]
    kr = [ctx.replica_put(w.codes[p], 1, w.n_fft, prec) for p in prns]
    kw = dict(fs_in=w.fs, fs_acq=w.fs, n_fft=w.n_fft, n_blocks=w.n_blocks, surface_len=w.surface_len,
              doppler=w.doppler, reshape=w.reshape)
    g32, s32, _ = ctx.acquire(w.iq, w.fmt, want_surfaces=True, keys=k32, **kw)
    gr, sr, _ = ctx.acquire(w.iq, w.fmt, want_surfaces=True, keys=kr, precision=prec, **kw)
    x = iq_to_complex(w.iq)
    reps = w.replicas()[prns]
    ref, s64, _ = oracle.acquire(x, reps, n_fft=w.n_fft, n_blocks=w.n_blocks, surface_len=w.surface_len,
                                 fs_in=w.fs, fs_acq=w.fs, doppler=w.doppler, reshape=w.reshape, precision=64,
                                 want_surfaces=True)
    err_r = np.linalg.norm(sr - s64) / np.linalg.norm(s64)
    err_32 = np.linalg.norm(s32 - s64) / np.linalg.norm(s64)
    say(f"  surface rel err vs FP64: reduced {err_r:.2e}, FP32 {err_32:.2e}")
    assert err_r <= SURF_TOL[prec]
    for i, p in enumerate(prns):
        d_db = 20 * np.log10(gr[i]["snr"] / g32[i]["snr"])
        d64 = 20 * np.log10(gr[i]["snr"] / ref[i]["snr"])
        say(f"  PRN {p}: snr fp32 {g32[i]['snr']:.4f} reduced {gr[i]['snr']:.4f} fp64 {ref[i]['snr']:.4f} "
            f"delta {d_db:+.4f} dB (vs fp64 {d64:+.4f} dB) idx {gr[i]['index']}/{g32[i]['index']} "
            f"bin {gr[i]['doppler_bin']}/{g32[i]['doppler_bin']}")
        if p in planted:
            assert abs(d_db) <= PNR_DB, (p, d_db)
            assert abs(d64) <= PNR_DB, (p, d64)
            assert gr[i]["index"] == g32[i]["index"] == ref[i]["index"]
            assert gr[i]["doppler_bin"] == g32[i]["doppler_bin"] == ref[i]["doppler_bin"]
    for k in k32 + kr:
        ctx.replica_drop(k)


@pytest.mark.parametrize("prec", [U.PREC_FP16, U.PREC_BF16])
@pytest.mark.parametrize("n_fft", [20460, 38192, 40920, 61380])
def test_reduced_fourstep_lengths_vs_fp64(ctx, oracle, prec, n_fft):
    """the other matrix-core plans (N1 = 60, 112, 120, 180 times 341): L5-like
    1-3 ms at 20.46 MHz and the real_8 recording's 38192, 2 blocks, 9 bins,
    3 PRNs with one planted"""
    rng = np.random.default_rng(n_fft + prec)
    fs = 20.46e6 if n_fft % 20460 == 0 else 9.548e6
    n_chips = 10230 * (n_fft // 20460) if n_fft % 20460 == 0 else 4092
    dop = U.doppler_bins(2000, 500)
    codes = [np.where(rng.random(n_chips) < 0.5, -1, 1).astype(np.int8) for _ in range(3)]
    reps = np.stack([oracle.repeat_upsample(c, 1, n_fft) for c in codes])
    t = np.arange(2 * n_fft)
    x = (rng.normal(0, 20, 2 * n_fft) + 1j * rng.normal(0, 20, 2 * n_fft)).astype(np.complex64)
    x += (2.0 * np.roll(np.tile(reps[1], 2), 4321) * np.exp(2j * np.pi * 1000 * t / fs)).astype(np.complex64)
    iq = np.ascontiguousarray(x).view(np.float32)
    say(f"case reduced N={n_fft} {NAME[prec]}")
    kw = dict(fs_in=fs, fs_acq=fs, n_fft=n_fft, n_blocks=2, surface_len=n_fft, doppler=dop)
    k32 = [ctx.replica_put(c, 1, n_fft) for c in codes]
    kr = [ctx.replica_put(c, 1, n_fft, prec) for c in codes]
    g32, s32, _ = ctx.acquire(iq, U.FMT_CF32, want_surfaces=True, keys=k32, **kw)
    gr, sr, _ = ctx.acquire(iq, U.FMT_CF32, want_surfaces=True, keys=kr, precision=prec, **kw)
    ref, s64, _ = oracle.acquire(x, reps, n_fft=n_fft, n_blocks=2, surface_len=n_fft, fs_in=fs, fs_acq=fs,
                                 doppler=dop, precision=64, want_surfaces=True)
    err_r = np.linalg.norm(sr - s64) / np.linalg.norm(s64)
    say(f"  surface rel err vs FP64: {err_r:.2e}; FP32 {np.linalg.norm(s32 - s64) / np.linalg.norm(s64):.2e}")
    assert err_r <= SURF_TOL[prec]
    d_db = 20 * np.log10(gr[1]["snr"] / g32[1]["snr"])
    say(f"  planted: idx {gr[1]['index']} bin {gr[1]['doppler_bin']} PNR delta vs FP32 {d_db:+.4f} dB")
    assert abs(d_db) <= PNR_DB
    assert gr[1]["index"] == g32[1]["index"] == ref[1]["index"] == 4321
    assert gr[1]["doppler_bin"] == g32[1]["doppler_bin"] == ref[1]["doppler_bin"]
    for k in k32 + kr:
        ctx.replica_drop(k)


@pytest.mark.parametrize("prec", [U.PREC_FP16, U.PREC_BF16])
@pytest.mark.parametrize("n1,samples_per_chip,threads", [(24, 10, 512), (24, 10, 1024), (12, 5, 512),
                                                         (12, 4, 1024)])
def test_reduced_upsampled_code_vs_fp64(ctx, oracle, prec, n1, samples_per_chip, threads):
    """the upsampled-code form of the reduced search (codes of N1 * 341
    chips held for U samples, U phase transforms of N1 * 341 points per row
    on the matrix cores): 3 blocks, 7 bins, 3 PRNs with one planted, against
    the FP64 oracle's n_fft-point search and the FP32 GPU search, in both
    workgroup shapes"""
    n_chips = n1 * 341
    n_fft = n_chips * samples_per_chip
    fs = n_fft * 250.0   # 4 ms blocks
    rng = np.random.default_rng(n_fft + prec + threads)
    dop = U.doppler_bins(750, 250)
    codes = [np.where(rng.random(n_chips) < 0.5, -1, 1).astype(np.int8) for _ in range(3)]
    reps = np.stack([oracle.repeat_upsample(c, 1, n_fft) for c in codes])
    nb, delay = 3, 12345 % n_fft
    t = np.arange(nb * n_fft)
    x = (rng.normal(0, 20, nb * n_fft) + 1j * rng.normal(0, 20, nb * n_fft)).astype(np.complex64)
    x += (1.5 * np.roll(np.tile(reps[2], nb), delay) * np.exp(2j * np.pi * 500 * t / fs)).astype(np.complex64)
    iq = np.ascontiguousarray(x).view(np.float32)
    say(f"case reduced upsampled code N1={n1} U={samples_per_chip} N={n_fft} {NAME[prec]} threads={threads}")
    kw = dict(fs_in=fs, fs_acq=fs, n_fft=n_fft, n_blocks=nb, surface_len=n_fft, doppler=dop)
    k32 = [ctx.replica_put(c, 1, n_fft) for c in codes]
    kr = [ctx.replica_put(c, 1, n_fft, prec) for c in codes]
    ctx.set_threads(n_fft, threads)
    try:
        g32, s32, _ = ctx.acquire(iq, U.FMT_CF32, want_surfaces=True, keys=k32, **kw)
        gr, sr, _ = ctx.acquire(iq, U.FMT_CF32, want_surfaces=True, keys=kr, precision=prec, **kw)
        form = ctx.last_search_form()
    finally:
        ctx.set_threads(n_fft, 0)
    assert form["samples_per_chip"] == samples_per_chip and form["n_transform"] == n_fft, form
    ref, s64, _ = oracle.acquire(x, reps, n_fft=n_fft, n_blocks=nb, surface_len=n_fft, fs_in=fs, fs_acq=fs,
                                 doppler=dop, precision=64, want_surfaces=True)
    err_r = np.linalg.norm(sr - s64) / np.linalg.norm(s64)
    say(f"  surface rel err vs FP64: {err_r:.2e}; FP32 {np.linalg.norm(s32 - s64) / np.linalg.norm(s64):.2e}")
    assert err_r <= SURF_TOL[prec]
    for i in range(3):
        d64 = 20 * np.log10(gr[i]["snr"] / ref[i]["snr"])
        say(f"  PRN {i}: idx {gr[i]['index']}/{ref[i]['index']} bin {gr[i]['doppler_bin']}/{ref[i]['doppler_bin']} "
            f"PNR vs FP64 {d64:+.4f} dB")
    d_db = 20 * np.log10(gr[2]["snr"] / g32[2]["snr"])
    assert abs(d_db) <= PNR_DB, d_db
    assert gr[2]["index"] == g32[2]["index"] == ref[2]["index"] == delay
    assert gr[2]["doppler_bin"] == g32[2]["doppler_bin"] == ref[2]["doppler_bin"]
    assert gr[2]["detected"] and ref[2]["detected"]
    for k in k32 + kr:
        ctx.replica_drop(k)


def test_reduced_rejects_unsupported(ctx):
    """lengths without a matrix-core plan and folded surfaces report UNSUPPORTED"""
    w = W.l1ca_1ms()
    key = ctx.replica_put(w.codes[0], w.repeats, w.n_fft, U.PREC_FP16)
    with pytest.raises(U.UgsdrError) as e:
        ctx.acquire(w.iq, w.fmt, fs_in=w.fs, fs_acq=w.fs, n_fft=w.n_fft, n_blocks=1, surface_len=w.surface_len,
                    doppler=w.doppler, keys=[key], precision=U.PREC_FP16)
    assert e.value.code == -4
    ctx.replica_drop(key)
