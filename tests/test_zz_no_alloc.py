"""ugsdr_hip_acquire_noalloc (the entry point that allocates no device memory)
against ugsdr_hip_acquire on the same requests, through the C ABI, to the
parity bar of tests/parity.py (same detections, equal index and bin outside
near-ties, level and SNR within 1e-3): two calls in a row, the second with a
smaller input, then a hypotheses request, a device-memory input and the
limit errors."""
import numpy as np
import pytest

import ugsdr_amd as U
from ugsdr_amd import workloads as W

from parity import compare

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def run_both(ctx, w, keys, peaks=False, **kw):
    """peaks keep the full surfaces on the device (the selected bin's row is
    copied from them), so only small searches ask for them"""
    args = dict(fs_in=w.fs, fs_acq=w.fs, n_fft=w.n_fft, n_blocks=w.n_blocks, surface_len=w.surface_len,
                doppler=w.doppler, keys=keys)
    args.update(kw)
    ref, rs, rp = ctx.acquire(w.iq, w.fmt, want_peaks=peaks, **args)
    got, gs, gp = ctx.acquire_noalloc(w.iq, w.fmt, want_peaks=peaks, **args)
    return ref, got, rp, gp


def assert_same(ref, got, rp, gp, tag):
    fails, notes = compare(got, ref)
    say(f"{tag}: {len(notes)} notes {notes}")
    assert not fails, fails
    if rp is None:
        return
    for p, r in enumerate(ref):
        if r["doppler_bin"] >= 0 and got[p]["doppler_bin"] == r["doppler_bin"]:
            assert np.max(np.abs(gp[p] - rp[p])) <= 2e-5 * max(1e-30, float(np.max(np.abs(rp[p]))))


def test_noalloc_matches_acquire_twice_then_smaller(ctx):
    w = W.l1ca_wide()
    say("case noalloc L1CA_WIDE 32 PRN x 41 bins x 10 blocks, N=16368")
    keys = [ctx.replica_put(c, w.repeats, w.n_fft) for c in w.codes]
    ref, got, rp, gp = run_both(ctx, w, keys)
    assert_same(ref, got, rp, gp, "noalloc/wide")
    assert sum(r["detected"] for r in got) >= 5
    for k in keys:
        ctx.replica_drop(k)
    # second call, smaller input: 1 ms at 4.092 MHz, 12 PRNs
    w2 = W.l1ca_1ms()
    say("case noalloc L1CA_1MS 12 PRN x 21 bins, N=4092 (second, smaller call)")
    keys2 = [ctx.replica_put(c, w2.repeats, w2.n_fft) for c in w2.codes[:12]]
    ref2, got2, rp2, gp2 = run_both(ctx, w2, keys2, peaks=True)
    assert_same(ref2, got2, rp2, gp2, "noalloc/1ms")
    for k in keys2:
        ctx.replica_drop(k)


def test_noalloc_four_step_and_hypotheses(ctx):
    """N = 81840 (four-step, fewer workgroups when the workspace is short),
    6 PRNs x 97 bins x 5 blocks; then the same replicas as 3 satellites x 2
    hypotheses"""
    w = W.boc_4ms()
    prns = sorted(s.prn for s in w.sats)[:6]
    say(f"case noalloc BOC_4MS PRNs {prns} N=81840")
    keys = [ctx.replica_put(w.codes[p], 1, w.n_fft) for p in prns]
    ref, got, rp, gp = run_both(ctx, w, keys)
    assert_same(ref, got, rp, gp, "noalloc/boc")
    ref_h, _, _ = ctx.acquire(w.iq, w.fmt, fs_in=w.fs, fs_acq=w.fs, n_fft=w.n_fft, n_blocks=w.n_blocks,
                              surface_len=w.surface_len, doppler=w.doppler, keys=keys, hypotheses=2)
    got_h, _, _ = ctx.acquire_noalloc(w.iq, w.fmt, fs_in=w.fs, fs_acq=w.fs, n_fft=w.n_fft, n_blocks=w.n_blocks,
                                      surface_len=w.surface_len, doppler=w.doppler, keys=keys, hypotheses=2)
    assert len(got_h) == 3
    fails, _ = compare(got_h, ref_h)
    assert not fails, fails
    for k in keys:
        ctx.replica_drop(k)


def test_noalloc_limits(ctx):
    w = W.l1ca_wide()
    keys = [ctx.replica_put(c, w.repeats, w.n_fft) for c in w.codes]
    say("case noalloc limits: full surfaces beyond the workspace, reduced precision")
    for want in ("surfaces", "peaks"):            # 32 x 41 x 16368 floats = 86 MB of surfaces
        with pytest.raises(U.UgsdrError) as e:
            ctx.acquire_noalloc(w.iq, w.fmt, want_surfaces=want == "surfaces", want_peaks=want == "peaks",
                                fs_in=w.fs, fs_acq=w.fs, n_fft=w.n_fft, n_blocks=w.n_blocks,
                                surface_len=w.surface_len, doppler=w.doppler, keys=keys)
        assert e.value.code == -5
    for k in keys:
        ctx.replica_drop(k)
    kr = ctx.replica_put(w.codes[0], w.repeats, w.n_fft, U.PREC_FP16)
    with pytest.raises(U.UgsdrError) as e:
        ctx.acquire_noalloc(w.iq, w.fmt, fs_in=w.fs, fs_acq=w.fs, n_fft=w.n_fft, n_blocks=w.n_blocks,
                            surface_len=w.surface_len, doppler=w.doppler, keys=[kr], precision=U.PREC_FP16)
    assert e.value.code == -4
    ctx.replica_drop(kr)
