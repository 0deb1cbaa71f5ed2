"""Fused acquisition (ugsdr_hip_acquire) vs the FP64 oracle, through the C ABI.

Parity bar (tests/parity.py): same detected PRNs, bit-exact surface index and
Doppler bin outside near-ties, level / metric within 1e-3 relative; full
surfaces within 2e-5 of the oracle's peak.
"""
import os

import numpy as np
import pytest

import ugsdr_amd as U
from ugsdr_amd import workloads as W

from parity import compare, iq_to_complex

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg):
    print(msg, flush=True)


def put_replicas(ctx, w, precision=U.PREC_FP32):
    return [ctx.replica_put(c, w.repeats, w.n_fft, precision) for c in w.codes]


def oracle_run(oracle, w, prns=None, dop=None, reshape=None, coherent=True, surfaces=False):
    x = iq_to_complex(w.iq)
    reps = w.replicas()
    if prns is not None:
        reps = reps[prns]
    return oracle.acquire(x, reps, n_fft=w.n_fft, n_blocks=w.n_blocks, surface_len=w.surface_len, fs_in=w.fs,
                          fs_acq=w.fs, doppler=w.doppler if dop is None else dop,
                          reshape=w.reshape if reshape is None else reshape, coherent=coherent, precision=64,
                          want_surfaces=surfaces)


def gpu_run(ctx, w, keys, dop=None, reshape=None, coherent=True, surfaces=False, peaks=False):
    return ctx.acquire(w.iq, w.fmt, want_surfaces=surfaces, want_peaks=peaks, fs_in=w.fs, fs_acq=w.fs,
                       n_fft=w.n_fft, n_blocks=w.n_blocks, surface_len=w.surface_len,
                       doppler=w.doppler if dop is None else dop, keys=keys,
                       reshape=w.reshape if reshape is None else reshape, coherent=coherent)


def check_truth(w, res, tol_idx=1.5):
    """every planted satellite the oracle detects sits at the planted delay / Doppler"""
    step = w.doppler[1] - w.doppler[0]
    for prn, idx, dop, cn0 in w.truth():
        r = res[prn]
        if not r["detected"]:
            continue
        d = abs(r["index"] - idx)
        d = min(d, w.surface_len - d)
        assert d <= tol_idx, (prn, r["index"], idx)
        assert abs(r["doppler"] - dop) <= step, (prn, r["doppler"], dop)


def assert_parity(gpu, ref, tag):
    fails, notes = compare(gpu, ref)
    say(f"{tag}: {len(notes)} notes {notes}")
    assert not fails, fails


def test_l1ca_1ms_full(ctx, oracle):
    w = W.l1ca_1ms()
    say("case L1CA_1MS 32 PRN x 21 bins, N=4092")
    keys = put_replicas(ctx, w)
    gpu, surf, peaks = gpu_run(ctx, w, keys, surfaces=True, peaks=True)
    ref, rsurf, _ = oracle_run(oracle, w, surfaces=True)
    assert_parity(gpu, ref, "L1CA_1MS")
    scale = rsurf.max()
    assert np.max(np.abs(surf - rsurf)) < 2e-5 * scale
    for p, g in enumerate(gpu):
        if g["doppler_bin"] >= 0:
            np.testing.assert_array_equal(peaks[p], surf[p, g["doppler_bin"]])
    check_truth(w, ref)
    assert sum(r["detected"] for r in ref) >= 4
    for k in keys:
        ctx.replica_drop(k)


def test_l1ca_4ms_fold(ctx, oracle):
    w = W.l1ca_4ms()
    say("case L1CA_4MS fold 16368 -> 4092")
    keys = put_replicas(ctx, w)
    gpu, surf, _ = gpu_run(ctx, w, keys, surfaces=True)
    ref, rsurf, _ = oracle_run(oracle, w, surfaces=True)
    assert_parity(gpu, ref, "L1CA_4MS")
    assert np.max(np.abs(surf - rsurf)) < 2e-5 * rsurf.max()
    check_truth(w, ref)
    for k in keys:
        ctx.replica_drop(k)


def test_l1ca_4ms_noncoherent_fold(ctx, oracle):
    """GetOneMsPeak<reshape=true, coherent=false>: fold complex, then |.|"""
    w = W.l1ca_4ms()
    say("case L1CA_4MS coherent=false")
    keys = put_replicas(ctx, w)
    gpu, _, _ = gpu_run(ctx, w, keys, coherent=False)
    ref, _, _ = oracle_run(oracle, w, coherent=False)
    assert_parity(gpu, ref, "L1CA_4MS/nc")
    for k in keys:
        ctx.replica_drop(k)


def test_l1ca_wide_10ms(ctx, oracle):
    w = W.l1ca_wide()
    say("case L1CA_WIDE 32 PRN x 41 bins x 10 blocks, N=16368")
    keys = put_replicas(ctx, w)
    gpu, _, _ = gpu_run(ctx, w, keys)
    ref, _, _ = oracle_run(oracle, w)
    assert_parity(gpu, ref, "L1CA_WIDE")
    check_truth(w, ref)
    assert sum(r["detected"] for r in ref) >= 5
    for k in keys:
        ctx.replica_drop(k)


def test_boc_4ms_subset(ctx, oracle):
    """N = 81840 from the expanded codes' chips (the upsampled-code search: ten
    8184-point phase transforms per row); 8 planted PRNs + 4 absent ones, all
    97 bins."""
    w = W.boc_4ms()
    planted = sorted(s.prn for s in w.sats)
    absent = [p for p in range(36) if p not in planted][:4]
    prns = planted + absent
    say(f"case BOC_4MS PRNs {prns} x 97 bins x 5 blocks, N=81840")
    keys = [ctx.replica_put(w.codes[p], 1, w.n_fft) for p in prns]
    gpu, _, _ = gpu_run(ctx, w, keys)
    # the expanded code's 8184 half-chips held for 10 samples each
    assert ctx.last_search_form() == {"n_transform": 81840, "periods": 1, "samples_per_chip": 10}
    ref, _, _ = oracle_run(oracle, w, prns=prns)
    assert_parity(gpu, ref, "BOC_4MS")
    for i, p in enumerate(planted):
        s = next(s for s in w.sats if s.prn == p)
        r = ref[i]
        if r["snr"] > 3.3:
            idx = (s.code_phase * 10) % w.surface_len
            assert min(abs(r["index"] - idx), w.surface_len - abs(r["index"] - idx)) <= 1.5
    for k in keys:
        ctx.replica_drop(k)


def test_real8_kat(ctx, oracle):
    """The reference's real_8 acquisition KAT (test/integration_tests.cpp:253-263,
    :314, :395-398): real int8 @ 38.192 MHz, IF 9.55 MHz, FseConfig rate 8.192 MHz
    (-> 9.548 MHz), +-6 kHz / 200 Hz, 4 ms; >= 4 of 8 satellites within 400 Hz."""
    raw = np.fromfile(os.path.join(GOLDEN, "real_8_first4ms.bin"), np.int8)
    fs = 38.192e6
    fs_acq = U.adjust_sampling_rate(fs, 8.192e6)
    assert fs_acq == 9.548e6
    spm = int(fs_acq / 1e3)
    N = 4 * spm
    dop = U.doppler_bins(6000, 200)
    say("case real_8 KAT N=38192 fold 4, 32 PRN x 61 bins, R8 + mix + boxcar /4")
    codes = [U.gps_l1ca(s) for s in range(32)]
    keys = [ctx.replica_put(c, 4, N) for c in codes]
    gpu, _, _ = ctx.acquire(raw, U.FMT_R8, n_samples=raw.size, fs_in=fs, fs_acq=fs_acq, mix_frequency=9.55e6,
                            n_fft=N, n_blocks=1, surface_len=spm, doppler=dop, keys=keys)
    x = raw.astype(np.float32).astype(np.complex64)
    x = oracle.mix_inplace(x, fs, 9.55e6, 0.0)
    y = oracle.resample(x, int(fs_acq), int(fs))
    reps = np.stack([oracle.repeat_upsample(c, 4, N) for c in codes])
    ref, _, _ = oracle.acquire(y, reps, n_fft=N, n_blocks=1, surface_len=spm, fs_in=fs, fs_acq=fs_acq,
                               doppler=dop, precision=64)
    assert_parity(gpu, ref, "real_8")
    kat = {2: (0, 34213), 5: (6000, 28203), 8: (-1000, 4695), 14: (0, 36321), 17: (2000, 20725),
           20: (3000, 13404), 21: (0, 6289), 25: (5000, 26828)}
    match = sum(1 for sv, (d, c) in kat.items() if gpu[sv]["detected"] and abs(gpu[sv]["doppler"] - d) <= 400)
    say(f"real_8: {match}/8 KAT satellites matched")
    assert match >= 4
    for k in keys:
        ctx.replica_drop(k)


@pytest.mark.parametrize("n_fft", [1000, 2310, 4750, 20000, 33250, 32768, 65536])
def test_runtime_plan_lengths(ctx, oracle, n_fft):
    """lengths without a compile-time plan (runtime radix plan / four-step),
    and the power-of-two four-step lengths with compile-time factor plans"""
    rng = np.random.default_rng(n_fft)
    fs = n_fft * 1000.0
    code = np.where(rng.random(n_fft // 2) < 0.5, -1, 1).astype(np.int8)
    x = (rng.normal(0, 20, n_fft) + 1j * rng.normal(0, 20, n_fft)).astype(np.complex64)
    rep = oracle.repeat_upsample(code, 1, n_fft)
    x += (3.0 * np.roll(rep, 123) * np.exp(2j * np.pi * 1500 * np.arange(n_fft) / fs)).astype(np.complex64)
    dop = U.doppler_bins(3000, 500)
    say(f"case runtime plan N={n_fft}")
    key = ctx.replica_put(code, 1, n_fft)
    iq = np.ascontiguousarray(x).view(np.float32)
    gpu, surf, _ = ctx.acquire(iq, U.FMT_CF32, want_surfaces=True, fs_in=fs, fs_acq=fs, n_fft=n_fft, n_blocks=1,
                               surface_len=n_fft, doppler=dop, keys=[key])
    ref, rsurf, _ = oracle.acquire(x, rep[None], n_fft=n_fft, n_blocks=1, surface_len=n_fft, fs_in=fs, fs_acq=fs,
                                   doppler=dop, precision=64, want_surfaces=True)
    assert_parity(gpu, ref, f"N={n_fft}")
    assert np.max(np.abs(surf - rsurf)) < 2e-5 * rsurf.max()
    assert gpu[0]["index"] == 123 and gpu[0]["doppler"] == 1500.0
    ctx.replica_drop(key)


@pytest.mark.parametrize("n_fft,surface_len,coherent", [
    (4092, 4092, True), (8184, 8184, True), (8184, 4092, False), (12276, 12276, True), (12276, 4092, True),
    (16368, 16368, True), (16368, 4092, True), (16368, 4092, False)])
def test_corr_variants_match_oracle(ctx, oracle, n_fft, surface_len, coherent):
    """every correlation variant of a C/A length (the three-stage 1023*M
    kernel and the generic LDS kernels) against the oracle on one input.
    The replicas are stored as samples: replicas put from chips take the
    upsampled-code search (one code period) or the periodic fold (several),
    covered by test_upsampled_code_matches_oracle and
    test_periodic_fold_matches_oracle, which never run these variants."""
    rng = np.random.default_rng(n_fft + surface_len)
    fs = n_fft * 1000.0 * surface_len / n_fft
    P, dop = 3, U.doppler_bins(2000, 500)
    reps, keys = [], []
    for _ in range(P):
        code = np.where(rng.random(1023) < 0.5, -1, 1).astype(np.int8)
        reps.append(oracle.repeat_upsample(code, n_fft // surface_len, n_fft))
        keys.append(ctx.replica_put_samples(reps[-1]))
    x = (rng.normal(0, 20, n_fft) + 1j * rng.normal(0, 20, n_fft)).astype(np.complex64)
    x += (2.0 * np.roll(reps[1], 777) * np.exp(2j * np.pi * 1000 * np.arange(n_fft) / fs)).astype(np.complex64)
    iq = np.ascontiguousarray(x).view(np.float32)
    ref, rsurf, _ = oracle.acquire(x, np.stack(reps), n_fft=n_fft, n_blocks=1, surface_len=surface_len, fs_in=fs,
                                   fs_acq=fs, doppler=dop, coherent=coherent, precision=64, want_surfaces=True)
    variants = U.corr_variants(n_fft)
    assert len(variants) >= 2
    for t in variants:
        say(f"case corr variant N={n_fft} L={surface_len} coherent={coherent} threads={t}")
        ctx.set_threads(n_fft, t)
        gpu, surf, _ = ctx.acquire(iq, U.FMT_CF32, want_surfaces=True, fs_in=fs, fs_acq=fs, n_fft=n_fft,
                                   n_blocks=1, surface_len=surface_len, doppler=dop, keys=keys, coherent=coherent)
        assert ctx.last_search_form() == {"n_transform": n_fft, "periods": 1, "samples_per_chip": 0}
        assert_parity(gpu, ref, f"N={n_fft}/L={surface_len}/T={t}")
        assert np.max(np.abs(surf - rsurf)) < 2e-5 * rsurf.max()
        assert gpu[1]["index"] == 777 % surface_len and gpu[1]["doppler"] == 1000.0
    ctx.set_threads(n_fft, 0)
    with pytest.raises(U.UgsdrError):
        ctx.set_threads(n_fft, 333)
    for k in keys:
        ctx.replica_drop(k)


@pytest.mark.parametrize("n_fft,surface_len,n_blocks,coherent", [
    (16368, 4092, 2, True), (16368, 4092, 1, False), (8184, 4092, 3, True), (40920, 4092, 1, True),
    (81840, 4092, 1, False), (65472, 16368, 2, True)])
def test_periodic_fold_matches_oracle(ctx, oracle, n_fft, surface_len, n_blocks, coherent):
    """replicas put from chips with T = n_fft / surface_len code periods: the
    search runs at surface_len on the T periods' spectra combined coherently.
    Against the oracle's n_fft-point search folded to surface_len, on Doppler
    bins whose d * surface_len / fs fraction is nonzero (the combination's
    phase factors), and against the same search on sample replicas (no period:
    the n_fft-point transforms and the fold kernels)."""
    rng = np.random.default_rng(n_fft + surface_len + n_blocks)
    fs = surface_len * 1000.0
    dop = U.doppler_bins(2250, 125)
    codes, reps, x = planted_case(oracle, rng, n_fft, surface_len, 1023, fs, n_blocks=n_blocks, shift=1234,
                                  f_dop=875.0)
    T = n_fft // surface_len
    say(f"case periodic fold N={n_fft} L={surface_len} T={T} K={n_blocks} coherent={coherent}")
    keys = [ctx.replica_put(c, T, n_fft) for c in codes]
    skeys = [ctx.replica_put_samples(r) for r in reps]
    iq = np.ascontiguousarray(x).view(np.float32)
    kw = dict(want_surfaces=True, fs_in=fs, fs_acq=fs, n_fft=n_fft, n_blocks=n_blocks, surface_len=surface_len,
              doppler=dop, coherent=coherent)
    gpu, surf, _ = ctx.acquire(iq, U.FMT_CF32, keys=keys, **kw)
    assert ctx.last_search_form() == {"n_transform": surface_len, "periods": T,
                                      "samples_per_chip": surface_len // 1023}
    gs, ssurf, _ = ctx.acquire(iq, U.FMT_CF32, keys=skeys, **kw)
    assert ctx.last_search_form() == {"n_transform": n_fft, "periods": 1, "samples_per_chip": 0}
    ref, rsurf, _ = oracle.acquire(x, reps, n_fft=n_fft, n_blocks=n_blocks, surface_len=surface_len, fs_in=fs,
                                   fs_acq=fs, doppler=dop, coherent=coherent, precision=64, want_surfaces=True)
    assert_parity(gpu, ref, f"periodic N={n_fft}/L={surface_len}")
    assert_parity(gs, ref, f"samples N={n_fft}/L={surface_len}")
    assert np.max(np.abs(surf - rsurf)) < 2e-5 * rsurf.max()
    assert np.max(np.abs(ssurf - rsurf)) < 2e-5 * rsurf.max()
    assert gpu[1]["index"] == 1234 % surface_len and gpu[1]["doppler"] == 875.0
    for k in keys + skeys:
        ctx.replica_drop(k)


@pytest.mark.parametrize("m,n_blocks", [(4, 1), (4, 3), (8, 2), (12, 1), (16, 1), (16, 3)])
def test_upsampled_code_matches_oracle(ctx, oracle, m, n_blocks):
    """replicas of one 1023-chip code period held for m samples per chip: the
    row runs as m inverse 1023-point transforms of conj(C) Z_p, with Z formed
    once per (Doppler bin, block) from the class spectra and the boxcar's
    spectrum.  Against the oracle's N-point search and the same search on
    sample replicas, on bins of every class fraction (250 Hz steps on 1 kHz
    classes) and an odd Doppler span, surfaces included."""
    n_fft = 1023 * m
    fs = n_fft * 1000.0
    rng = np.random.default_rng(1000 + m * 10 + n_blocks)
    dop = U.doppler_bins(2750, 250)
    codes, reps, x = planted_case(oracle, rng, n_fft, n_fft, 1023, fs, n_blocks=n_blocks, shift=517 * m // 4,
                                  f_dop=750.0, n_prn=5)
    say(f"case upsampled code N={n_fft} U={m} K={n_blocks}")
    keys = [ctx.replica_put(c, 1, n_fft) for c in codes]
    skeys = [ctx.replica_put_samples(r) for r in reps]
    iq = np.ascontiguousarray(x).view(np.float32)
    kw = dict(want_surfaces=True, fs_in=fs, fs_acq=fs, n_fft=n_fft, n_blocks=n_blocks, surface_len=n_fft,
              doppler=dop)
    gpu, surf, _ = ctx.acquire(iq, U.FMT_CF32, keys=keys, **kw)
    assert ctx.last_search_form() == {"n_transform": n_fft, "periods": 1, "samples_per_chip": m}
    gs, ssurf, _ = ctx.acquire(iq, U.FMT_CF32, keys=skeys, **kw)
    assert ctx.last_search_form()["samples_per_chip"] == 0
    ref, rsurf, _ = oracle.acquire(x, reps, n_fft=n_fft, n_blocks=n_blocks, surface_len=n_fft, fs_in=fs, fs_acq=fs,
                                   doppler=dop, precision=64, want_surfaces=True)
    assert_parity(gpu, ref, f"upsampled N={n_fft}")
    assert_parity(gs, ref, f"samples N={n_fft}")
    assert np.max(np.abs(surf - rsurf)) < 2e-5 * rsurf.max()
    # 1 ms coherent: the planted 750 Hz may peak in a neighbouring 250 Hz bin
    assert gpu[1]["index"] == (517 * m // 4) % n_fft and abs(gpu[1]["doppler"] - 750.0) <= 250.0
    for k in keys + skeys:
        ctx.replica_drop(k)


@pytest.mark.parametrize("n_chips,m,n_blocks", [(4092, 2, 2), (4092, 5, 1), (8184, 2, 1), (8184, 10, 2),
                                                 (12276, 3, 1), (16368, 2, 1)])
def test_upsampled_long_code_matches_oracle(ctx, oracle, n_chips, m, n_blocks):
    """codes of 1023 M chips (M = 4, 8, 12, 16; 8184 is the BOC(1,1)-expanded
    4092-chip code) held for m samples per chip: the row runs as m inverse
    transforms of 1023 M points, one per phase, each over every block.
    Against the oracle's N-point search and the same search on sample
    replicas, surfaces included."""
    n_fft = n_chips * m
    fs = n_fft * 1000.0
    rng = np.random.default_rng(7000 + n_chips + m * 10 + n_blocks)
    dop = U.doppler_bins(2750, 250)
    shift = 517 * m + 3
    codes, reps, x = planted_case(oracle, rng, n_fft, n_fft, n_chips, fs, n_blocks=n_blocks, shift=shift,
                                  f_dop=750.0, n_prn=4)
    say(f"case upsampled long code N={n_fft} chips={n_chips} U={m} K={n_blocks}")
    keys = [ctx.replica_put(c, 1, n_fft) for c in codes]
    skeys = [ctx.replica_put_samples(r) for r in reps]
    iq = np.ascontiguousarray(x).view(np.float32)
    kw = dict(want_surfaces=True, fs_in=fs, fs_acq=fs, n_fft=n_fft, n_blocks=n_blocks, surface_len=n_fft,
              doppler=dop)
    gpu, surf, _ = ctx.acquire(iq, U.FMT_CF32, keys=keys, **kw)
    assert ctx.last_search_form() == {"n_transform": n_fft, "periods": 1, "samples_per_chip": m}
    gs, ssurf, _ = ctx.acquire(iq, U.FMT_CF32, keys=skeys, **kw)
    assert ctx.last_search_form()["samples_per_chip"] == 0
    ref, rsurf, _ = oracle.acquire(x, reps, n_fft=n_fft, n_blocks=n_blocks, surface_len=n_fft, fs_in=fs, fs_acq=fs,
                                   doppler=dop, precision=64, want_surfaces=True)
    assert_parity(gpu, ref, f"upsampled N={n_fft}")
    assert_parity(gs, ref, f"samples N={n_fft}")
    assert np.max(np.abs(surf - rsurf)) < 2e-5 * rsurf.max()
    assert gpu[1]["index"] == shift % n_fft and abs(gpu[1]["doppler"] - 750.0) <= 250.0
    for k in keys + skeys:
        ctx.replica_drop(k)


def planted_case(oracle, rng, n_fft, surface_len, n_chips, fs, n_blocks=2, shift=777, f_dop=1000.0, n_prn=3):
    """n_prn random +-1 codes of n_chips per surface_len samples, the second
    one planted at code phase `shift` and Doppler f_dop in complex AWGN"""
    repeats = n_fft // surface_len
    codes = [np.where(rng.random(n_chips) < 0.5, -1, 1).astype(np.int8) for _ in range(n_prn)]
    reps = np.stack([oracle.repeat_upsample(c, repeats, n_fft) for c in codes])
    n = n_blocks * n_fft
    t = np.arange(n)
    x = (rng.normal(0, 20, n) + 1j * rng.normal(0, 20, n)).astype(np.complex64)
    x += (2.0 * np.roll(np.tile(reps[1], n_blocks), shift) * np.exp(2j * np.pi * f_dop * t / fs)).astype(np.complex64)
    return codes, reps, x


# every transform length the workloads and recordings use beyond the C/A
# lengths above: the FP32 four-step plans N1 x 341 (N1 = 60, 120, 180, 240:
# L5-like 1-3 ms at 20.46 MHz and the BOC configuration), and the BbpDdc
# recording's N = 4200 (4.2 MHz, 1 ms) and 26600 (6.65 MHz, 4 ms) plans
@pytest.mark.parametrize("n_fft,surface_len,n_chips,fs", [
    (20460, 20460, 10230, 20.46e6), (40920, 40920, 10230, 20.46e6), (40920, 20460, 10230, 20.46e6),
    (61380, 61380, 10230, 20.46e6), (61380, 20460, 10230, 20.46e6), (81840, 81840, 8184, 20.46e6),
    (4200, 4200, 1023, 4.2e6), (26600, 26600, 1023, 6.65e6), (26600, 6650, 1023, 6.65e6),
    # T ms coherent L1 C/A requests beyond 16 ms-samples (runtime four-step plans)
    (24552, 4092, 1023, 4.092e6), (28644, 4092, 1023, 4.092e6), (36828, 4092, 1023, 4.092e6)])
def test_fourstep_and_recording_lengths_every_variant(ctx, oracle, n_fft, surface_len, n_chips, fs):
    rng = np.random.default_rng(n_fft * 7 + surface_len)
    dop = U.doppler_bins(2000, 500)
    codes, reps, x = planted_case(oracle, rng, n_fft, surface_len, n_chips, fs)
    repeats = n_fft // surface_len
    # replicas put as samples too: the 8184-chip code (BOC) would otherwise
    # run the upsampled-code search, not the transform of the whole row
    keys = [ctx.replica_put(c, repeats, n_fft) for c in codes]
    skeys = [ctx.replica_put_samples(r) for r in reps] if n_chips == 8184 else []
    iq = np.ascontiguousarray(x).view(np.float32)
    ref, rsurf, _ = oracle.acquire(x, reps, n_fft=n_fft, n_blocks=2, surface_len=surface_len, fs_in=fs, fs_acq=fs,
                                   doppler=dop, precision=64, want_surfaces=True)
    variants = U.corr_variants(n_fft)
    assert variants
    for t in variants:
        for kk in (keys, skeys):
            if not kk:
                continue
            say(f"case length N={n_fft} L={surface_len} K=2 threads={t} samples={kk is skeys}")
            ctx.set_threads(n_fft, t)
            gpu, surf, _ = ctx.acquire(iq, U.FMT_CF32, want_surfaces=True, fs_in=fs, fs_acq=fs, n_fft=n_fft,
                                       n_blocks=2, surface_len=surface_len, doppler=dop, keys=kk)
            assert_parity(gpu, ref, f"N={n_fft}/L={surface_len}/T={t}")
            assert np.max(np.abs(surf - rsurf)) < 2e-5 * rsurf.max()
            assert gpu[1]["index"] == 777 % surface_len and gpu[1]["doppler"] == 1000.0
    ctx.set_threads(n_fft, 0)
    for k in keys + skeys:
        ctx.replica_drop(k)


@pytest.mark.parametrize("upsampled", [True, False])
@pytest.mark.parametrize("n_prn,n_bins,n_blocks", [(3, 9, 3), (4, 70, 4)])
def test_corr1023_tail_rows_split_over_workgroups(ctx, oracle, n_prn, n_bins, n_blocks, upsampled):
    """N = 16368 with several blocks per row: whole rounds of rows run one row
    per workgroup, the rest are split block-wise over all workgroups and
    merged (3 x 9 rows: every row split; 4 x 70 rows: one round + a tail).
    upsampled: replicas put from chips (the upsampled-code search); else as
    samples (the three-stage transform of the whole row)."""
    n_fft = 16368
    rng = np.random.default_rng(n_prn * 100 + n_bins)
    fs = 16.368e6
    reps, keys = [], []
    for _ in range(n_prn):
        code = np.where(rng.random(1023) < 0.5, -1, 1).astype(np.int8)
        reps.append(oracle.repeat_upsample(code, 1, n_fft))
        keys.append(ctx.replica_put(code, 1, n_fft) if upsampled else ctx.replica_put_samples(reps[-1]))
    n = n_fft * n_blocks
    x = (rng.normal(0, 20, n) + 1j * rng.normal(0, 20, n)).astype(np.complex64)
    x += (1.0 * np.tile(np.roll(reps[1], 4321), n_blocks) *
          np.exp(2j * np.pi * 1000 * np.arange(n) / fs)).astype(np.complex64)
    iq = np.ascontiguousarray(x).view(np.float32)
    dop = np.arange(n_bins, dtype=np.float64) * 500.0 - 500.0 * (n_bins // 2)
    ref, rsurf, _ = oracle.acquire(x, np.stack(reps), n_fft=n_fft, n_blocks=n_blocks, surface_len=n_fft, fs_in=fs,
                                   fs_acq=fs, doppler=dop, precision=64, want_surfaces=True)
    t1023 = U.corr_variants(n_fft)[0]
    ctx.set_threads(n_fft, t1023)
    say(f"case corr1023 tail split P={n_prn} D={n_bins} K={n_blocks} upsampled={upsampled}")
    gpu, surf, _ = ctx.acquire(iq, U.FMT_CF32, want_surfaces=True, fs_in=fs, fs_acq=fs, n_fft=n_fft,
                               n_blocks=n_blocks, surface_len=n_fft, doppler=dop, keys=keys)
    assert ctx.last_search_form()["samples_per_chip"] == (16 if upsampled else 0)
    ctx.set_threads(n_fft, 0)
    assert_parity(gpu, ref, f"tail P={n_prn} D={n_bins} K={n_blocks}")
    assert np.max(np.abs(surf - rsurf)) < 2e-5 * rsurf.max()
    assert gpu[1]["index"] == 4321 and gpu[1]["doppler"] == 1000.0
    for k in keys:
        ctx.replica_drop(k)


def test_zero_signal_keeps_default_result(ctx):
    """all-zero input: no bin beats level 0, so the AcquisitionResult defaults stay
    (src/acquisition/acquisition_result.hpp:13-18, fse.hpp:265)"""
    say("case zero signal")
    key = ctx.replica_put(U.gps_l1ca(0), 1, 4092)
    iq = np.zeros(2 * 4092, np.int8)
    gpu, _, _ = ctx.acquire(iq, U.FMT_CI8, fs_in=4.092e6, fs_acq=4.092e6, n_fft=4092, n_blocks=1,
                            surface_len=4092, doppler=U.doppler_bins(1000, 500), keys=[key])
    r = gpu[0]
    assert r["level"] == 0 and r["sigma"] == 1.0 and r["doppler_bin"] == -1 and r["detected"] == 0
    ctx.replica_drop(key)


def test_errors(ctx):
    say("case error paths")
    key = ctx.replica_put(U.gps_l1ca(0), 1, 4092)
    iq = np.zeros(2 * 4092, np.int8)
    kw = dict(fs_in=4.092e6, fs_acq=4.092e6, n_fft=4092, n_blocks=1, doppler=[0.0])
    with pytest.raises(U.UgsdrError, match="SIZE"):   # reshape_and_sum.hpp:16-17
        ctx.acquire(iq, U.FMT_CI8, surface_len=1000, keys=[key], **kw)
    with pytest.raises(U.UgsdrError, match="STATE"):
        ctx.acquire(iq, U.FMT_CI8, surface_len=4092, keys=[key + 999], **kw)
    with pytest.raises(U.UgsdrError, match="SIZE"):
        ctx.acquire(iq[:1000], U.FMT_CI8, surface_len=4092, keys=[key], **kw)
    with pytest.raises(U.UgsdrError, match="ARG"):
        ctx.acquire(iq, U.FMT_CI8, surface_len=4092, keys=[key], **{**kw, "n_fft": 8184})
    ctx.replica_drop(key)
    with pytest.raises(U.UgsdrError, match="STATE"):
        ctx.replica_drop(key)


def test_replica_spectrum_matches_oracle(ctx, oracle):
    say("case replica spectrum")
    code = U.gps_l1ca(3)
    key = ctx.replica_put(code, 4, 16368)
    spec = ctx.replica_get(key, 16368)
    ref = np.conj(np.fft.fft(oracle.repeat_upsample(code, 4, 16368).astype(np.float64)))
    assert np.max(np.abs(spec - ref)) < 1e-5 * np.max(np.abs(ref))
    ctx.replica_drop(key)


def test_graph_replay_matches_eager(ctx):
    """repeated request shapes: eager, captured, then replayed from a hipGraph;
    every call (with other eager work between them) equals the graphs-off
    result bit for bit, for host IQ and for device IQ"""
    from conftest import DeviceArray

    w = W.l1ca_1ms()
    keys = put_replicas(ctx, w)
    rng = np.random.default_rng(7)
    iqs = [w.iq] + [np.ascontiguousarray(np.roll(w.iq, int(rng.integers(1, 4000)) * 2)) for _ in range(3)]
    kw = dict(fs_in=w.fs, fs_acq=w.fs, n_fft=w.n_fft, n_blocks=w.n_blocks, surface_len=w.surface_len,
              doppler=w.doppler, keys=keys)
    ctx.set_graphs(False)
    ref = [ctx.acquire(iq, w.fmt, want_peaks=True, **kw) for iq in iqs]
    ctx.set_graphs(True)
    r0 = ctx.graph_replays()
    say("case graph replay, host IQ")
    for rep in range(2):
        for i, iq in enumerate(iqs):
            res, _, peaks = ctx.acquire(iq, w.fmt, want_peaks=True, **kw)
            assert res == ref[i][0], (rep, i)
            np.testing.assert_array_equal(peaks, ref[i][2])
            ctx.dft(np.ones(64, np.complex64))   # unrelated eager work between replays
    assert ctx.graph_replays() >= r0 + 5
    say("case graph replay, device IQ")
    dev = DeviceArray(iqs[2])
    r1 = ctx.graph_replays()
    for _ in range(4):
        req, hold = ctx.make_request(None, w.fmt, n_samples=len(iqs[2]) // 2, iq_device_ptr=dev.ptr.value, **kw)
        res, _, _ = ctx.acquire_req(req)
        assert res == ref[2][0]
    assert ctx.graph_replays() >= r1 + 2
    dev.close()
    # a different shape in between drops the graph; results stay right
    res, _, _ = ctx.acquire(iqs[0], w.fmt, **{**kw, "doppler": w.doppler[:5]})
    res2, _, _ = ctx.acquire(iqs[1], w.fmt, want_peaks=True, **kw)
    assert res2 == ref[1][0]
    for k in keys:
        ctx.replica_drop(k)


def test_concurrent_callers_on_one_context(ctx):
    """the reference calls its operators from par_unseq worker threads
    (fse.hpp:294): 8 host threads share one context, each running 4 searches
    of its own (different PRN subsets, Doppler spans and inputs); every result
    equals the same search run serially"""
    from concurrent.futures import ThreadPoolExecutor
    w = W.l1ca_1ms()
    wide = W.l1ca_wide()
    say("case 8 threads x 4 searches on one context (L1CA_1MS and L1CA_WIDE shapes)")
    keys1 = put_replicas(ctx, w)
    keysw = put_replicas(ctx, wide)
    jobs = []
    for t in range(8):
        for i in range(4):
            k = (t * 4 + i) % 7
            if i % 2:
                jobs.append((wide, keysw[k:k + 9], U.doppler_bins(5000 - 250 * k, 250)))
            else:
                jobs.append((w, keys1[k:k + 12], U.doppler_bins(5000 - 500 * k, 500)))

    def run(job):
        wl, keys, dop = job
        return gpu_run(ctx, wl, keys, dop=dop)[0]

    serial = [run(j) for j in jobs]
    with ThreadPoolExecutor(8) as ex:
        parallel = list(ex.map(run, jobs))
    for a, b in zip(parallel, serial):
        assert a == b
    for k in keys1 + keysw:
        ctx.replica_drop(k)
