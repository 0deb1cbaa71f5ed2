"""Correlator bank (ugsdr_hip_correlate_bank) vs the oracle.

Tolerance: FP32 I/Q accumulators against FP64 oracle sums, |err| <= 1e-5 of
sum |x| over the block (the magnitude scale of any accumulator of the block).
"""
import numpy as np
import pytest

import ugsdr_amd as U
from ugsdr_amd import workloads as W

from parity import iq_to_complex

pytestmark = pytest.mark.gpu


def test_bank_nco_vs_oracle(ctx, oracle):
    print("case bank NCO 6 ch x 5 taps x 3 blocks @ 40.92 MHz", flush=True)
    w = W.bank_l5(n_ch=6, n_blocks=3)
    out, states = ctx.correlate_bank(w.iq, U.FMT_CI16, spm=w.spm, n_blocks=w.n_blocks, fs=w.fs,
                                     channels=w.channels, codes=w.codes, taps_chips=w.taps, mode=U.BANK_NCO)
    x = iq_to_complex(w.iq)
    ref = oracle.bank_nco(x, w.spm, w.n_blocks, w.fs, w.channels, w.codes, w.taps)
    scale = np.abs(x).reshape(w.n_blocks, w.spm).sum(axis=1)
    err = np.abs(out - ref).max(axis=(1, 2))
    assert np.all(err <= 1e-5 * scale), (err, scale)
    # locked channels: prompt dominates, taps fall off symmetrically
    p = np.abs(out[:, :, 2])
    assert np.all(p > np.abs(out[:, :, 0])) and np.all(p > np.abs(out[:, :, 4]))
    # states advanced by n_blocks ms of FP64 NCO
    for (f, th, r, tau), (f2, th2, r2, tau2) in zip(w.channels, states):
        span = w.n_blocks * w.spm / w.fs
        assert abs(((th + f * span) % 1.0) - th2) < 1e-9
        assert abs(((tau + r * span) % 10230) - tau2) < 1e-6
    # continuing from the returned states equals one longer call
    out2, _ = ctx.correlate_bank(w.iq[2 * w.spm * 2:], U.FMT_CI16, spm=w.spm, n_blocks=1, fs=w.fs,
                                 channels=[(c[0], c[1] + c[0] * 2e-3, c[2], (c[3] + c[2] * 2e-3) % 10230)
                                           for c in w.channels],
                                 codes=w.codes, taps_chips=w.taps, mode=U.BANK_NCO)
    assert np.max(np.abs(out2[0] - out[2])) <= 1e-5 * scale[2]


def _bank_case(ctx, oracle, w, fmt, taps, channels=None, codes=None):
    channels = w.channels if channels is None else channels
    codes = w.codes if codes is None else codes
    x = iq_to_complex(w.iq)
    if fmt == U.FMT_CI8:
        iq = np.clip(np.round(w.iq / 16), -127, 127).astype(np.int8)
        x = iq_to_complex(iq)
    elif fmt == U.FMT_CF32:
        iq = x.view(np.float32)
    else:
        iq = w.iq
    out, _ = ctx.correlate_bank(iq, fmt, spm=w.spm, n_blocks=w.n_blocks, fs=w.fs, channels=channels, codes=codes,
                                taps_chips=taps, mode=U.BANK_NCO)
    ref = oracle.bank_nco(x, w.spm, w.n_blocks, w.fs, channels, codes, taps)
    scale = np.abs(x).reshape(w.n_blocks, w.spm).sum(axis=1)
    err = np.abs(out - ref).max(axis=(1, 2))
    assert np.all(err <= 1e-5 * scale), (err, scale)


@pytest.mark.parametrize("fmt,taps", [
    (U.FMT_CI16, [0.25, 0.0, -0.25]),              # quarter grid, descending, three taps
    (U.FMT_CI8, [-0.5, -0.25, 0.0, 0.25, 0.5]),     # quarter grid, int8 samples
    (U.FMT_CF32, [-0.25, 0.0, 0.25]),               # quarter grid, float samples
    (U.FMT_CI16, [-0.3, 0.0, 0.7, 1.1]),            # off the quarter grid: the multi-channel kernel
    (U.FMT_CI16, [-1.0, -0.5, 0.0, 0.5, 1.0, 0.125, -0.125, 2.0]),   # eight taps
])
def test_bank_nco_tap_sets_and_formats(ctx, oracle, fmt, taps):
    """every NCO kernel shape against the oracle: taps a quarter chip apart in
    either direction (the quarter-table kernel) and arbitrary taps (the
    multi-channel kernel), 7 channels (a partial channel group), 3 blocks"""
    print(f"case bank NCO fmt={fmt} taps={taps}", flush=True)
    w = W.bank_l5(n_ch=7, n_blocks=3, seed=0xC41)
    _bank_case(ctx, oracle, w, fmt, np.array(taps))


def test_bank_nco_code_shorter_than_block(ctx, oracle):
    """a code that repeats within a block (1023 chips at 2.046 Mchip/s over
    1 ms blocks) runs the one-channel kernel"""
    print("case bank NCO short code", flush=True)
    w = W.bank_l5(n_ch=3, n_blocks=2, seed=0xC42)
    rng = np.random.default_rng(5)
    codes = np.where(rng.random((3, 1023)) < 0.5, -1, 1).astype(np.int8)
    chans = [(c[0], c[1], 2.046e6, c[3] % 1023) for c in w.channels]
    _bank_case(ctx, oracle, w, U.FMT_CI16, np.array([-0.5, 0.0, 0.5]), channels=chans, codes=codes)


def test_bank_ref_split_vs_oracle(ctx, oracle):
    """Tracker::GetEpl / CorrelateSplit semantics, E/P/L at +-0.25 chip"""
    print("case bank ref_split 4 ch x 3 taps x 3 blocks", flush=True)
    fs, spm = 4.092e6, 4092
    codes = np.stack([U.gps_l1ca(s) for s in (0, 5, 11, 20)])
    rng = np.random.default_rng(12)
    sats = [W.Sat(i, codes[i], 1.023e6, W.L1_HZ, rng.uniform(0, 1023), rng.uniform(-3000, 3000),
                  rng.uniform(0, 1), 50.0) for i in range(4)]
    iq = W.synthesize(sats, fs, 3 * spm, 8.0, 77, 16)
    chans = []
    for s in sats:
        code_phase = (s.code_phase * 4.0) % spm + 0.37   # samples, non-integer
        chans.append((s.doppler, rng.uniform(0, 6.28), 0.0, code_phase))
    taps = [0.25, 0.0, -0.25]
    out, states = ctx.correlate_bank(iq, U.FMT_CI16, spm=spm, n_blocks=3, fs=fs, channels=chans, codes=codes,
                                     taps_chips=taps, mode=U.BANK_REF_SPLIT, samples_per_chip=4.0,
                                     code_period_ms=1.0)
    x = iq_to_complex(iq)
    for c, (f, ph, _, cp) in enumerate(chans):
        code3 = oracle.repeat_upsample(codes[c], 3, 3 * spm)
        for b in range(3):
            xb = x[b * spm:(b + 1) * spm].copy()
            ref = oracle.epl(xb, fs, f, ph + b * 2 * np.pi * f / 1000.0, code3, 1.0, cp, 4.0, taps)
            scale = np.abs(xb).sum()
            assert np.max(np.abs(out[b, c] - ref)) <= 1e-5 * scale, (c, b, out[b, c], ref)
        assert abs(states[c][1] - (ph + 3 * 2 * np.pi * f / 1000.0)) < 1e-9


def test_bank_errors(ctx):
    w = W.bank_l5(n_ch=2, n_blocks=1)
    with pytest.raises(U.UgsdrError, match="ARG"):
        ctx.correlate_bank(w.iq, U.FMT_CI16, spm=w.spm, n_blocks=1, fs=w.fs, channels=w.channels, codes=w.codes,
                           taps_chips=np.zeros(9), mode=U.BANK_NCO)
