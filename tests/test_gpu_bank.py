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
