"""Galileo E1 code hypotheses on the GPU: one fused search runs every
satellite's four sign permutations (consecutive replicas, hypotheses=4 in
ugsdr_acq_request) and returns one result per satellite, the
FastSearchEngineBase::ProcessGalileo selection (src/acquisition/fse.hpp:
356-396).

Bar against the oracle's ProcessGalileo on the same int16 samples: the same
detections, code-phase index and Doppler bin, and level within 1e-3
relative (the north-star peak-metric tolerance).  The grouped results equal
the per-replica results reduced by the oracle's selection, and the request
server returns the same grouped results."""
import numpy as np
import pytest

import ugsdr_amd as U
from oracle import refcpu as R

from test_galileo import FS, SPM, e1_chips, e1_signal

pytestmark = pytest.mark.gpu

DOP = U.doppler_bins(5000, 250)


def say(msg):
    print(msg, flush=True)


def scene():
    rng = np.random.default_rng(77)
    chips = [e1_chips(100 + s) for s in range(10)]
    # six satellites in the signal (C/N0 about 44-50 dB-Hz at sigma 1), four absent
    sats = []
    for s in range(6):
        amp = np.sqrt(10 ** (rng.uniform(4.4, 5.0)) / FS)
        # Doppler on the search grid: a residual of up to half a bin turns the
        # carrier by up to pi over the 4 ms and can favour another permutation
        sats.append((chips[s], int(s % 4), int(rng.integers(0, 4 * SPM)), 250.0 * int(rng.integers(-18, 19)), amp,
                     float(rng.uniform(0, 6.28))))
    x = e1_signal(sats, seed=5, sigma=1.0)
    raw = np.empty(2 * x.size, np.int16)
    raw[0::2], raw[1::2] = np.round(300 * x.real), np.round(300 * x.imag)
    dec = (raw[0::2].astype(np.float32) + 1j * raw[1::2].astype(np.float32)).astype(np.complex64)
    return chips, sats, raw, dec


def put_hypotheses(ctx, chips):
    return [ctx.replica_put_samples(rep) for c in chips for rep in U.e1_sign_hypotheses(c, 4 * SPM)]


def request(ctx, raw, keys, hypotheses):
    return ctx.make_request(raw, U.FMT_CI16, fs_in=FS, fs_acq=FS, n_fft=4 * SPM, n_blocks=1, surface_len=4 * SPM,
                            doppler=DOP, keys=keys, reshape=False, hypotheses=hypotheses)


def test_hypotheses_match_oracle(ctx, oracle):
    chips, sats, raw, dec = scene()
    keys = put_hypotheses(ctx, chips)
    say("case Galileo hypotheses: 10 satellites x 4 permutations x 41 bins, 4 ms at 4.092 MHz vs oracle")
    try:
        req, hold = request(ctx, raw, keys, 4)
        got, _, _ = ctx.acquire_req(req)
        req1, hold1 = request(ctx, raw, keys, 0)
        per_replica, _, _ = ctx.acquire_req(req1)
    finally:
        for k in keys:
            ctx.replica_drop(k)
    ref = R.process_galileo(dec, chips, spm=SPM, fs_in=FS, fs_acq=FS, doppler=DOP)
    assert len(got) == 10
    for s, (a, b) in enumerate(zip(got, ref)):
        say(f"  sv {s}: detected {a['detected']} index {a['index']} bin {a['doppler_bin']} level {a['level']:.1f} "
            f"(oracle hypothesis {b['hypothesis']}, level {b['level']:.1f})")
        assert a["detected"] == b["detected"] and a["index"] == b["index"] and a["doppler_bin"] == b["doppler_bin"]
        assert a["level"] == pytest.approx(b["level"], rel=1e-3)
    assert [r["detected"] for r in got] == [1] * 6 + [0] * 4
    for s in range(6):
        assert ref[s]["hypothesis"] == sats[s][1] and got[s]["index"] == sats[s][2]
    # the library's grouping is the oracle's selection applied to its own per-replica results
    mine = R.reduce_hypotheses(per_replica, 4, 3.3)
    for a, b in zip(got, mine):
        assert all(a[k] == b[k] for k in ("index", "doppler_bin", "level", "detected"))


def test_hypotheses_through_server(ctx):
    chips, sats, raw, dec = scene()
    keys = put_hypotheses(ctx, chips)
    say("case Galileo hypotheses through the request server")
    try:
        req, hold = request(ctx, raw, keys, 4)
        solo, _, _ = ctx.acquire_req(req)
        with U.Server(ctx, n_streams=2, max_batch=4) as srv:
            t = srv.submit(req)
            res, _ = srv.poll(t, -1)
    finally:
        for k in keys:
            ctx.replica_drop(k)
    assert len(res) == 10
    for a, b in zip(res, solo):
        assert a["detected"] == b["detected"] and a["index"] == b["index"] and a["doppler_bin"] == b["doppler_bin"]
        assert a["level"] == pytest.approx(b["level"], rel=1e-5)


def test_hypotheses_refusals(ctx):
    chips, sats, raw, dec = scene()
    keys = put_hypotheses(ctx, chips[:2])
    try:
        req, hold = request(ctx, raw, keys[:7], 4)
        with pytest.raises(U.UgsdrError):
            ctx.acquire_req(req)
    finally:
        for k in keys:
            ctx.replica_drop(k)


def test_hypotheses_reduced_precision_match_fp32(ctx):
    """the grouping on the reduced (FP16 storage, matrix-core) search at the
    BOC configuration's length (4 ms at 20.46 MHz, N = 81840): the same
    satellites, code phases and Doppler bins as the FP32 search, peak
    metric within 0.1 dB"""
    from ugsdr_amd import workloads as W

    gw = W.galileo_e1(seed=0xE5, n_sv=6, n_present=3, fs=20.46e6, n_blocks=1, doppler_range=1000,
                      doppler_step=250, name="test")
    say("case Galileo hypotheses, FP16 vs FP32, 6 satellites x 4 permutations x 9 bins, N = 81840")
    out = {}
    for prec in (U.PREC_FP32, U.PREC_FP16):
        keys = [ctx.replica_put_samples(r, prec) for c in gw.codes for r in U.e1_sign_hypotheses(c, gw.n_fft)]
        try:
            res, _, _ = ctx.acquire(gw.iq, U.FMT_CI16, fs_in=gw.fs, fs_acq=gw.fs, n_fft=gw.n_fft, n_blocks=1,
                                    surface_len=gw.n_fft, doppler=gw.doppler, keys=keys, reshape=False,
                                    hypotheses=4, precision=prec)
        finally:
            for k in keys:
                ctx.replica_drop(k)
        out[prec] = res
    for a, b in zip(out[U.PREC_FP16], out[U.PREC_FP32]):
        assert a["detected"] == b["detected"] and a["index"] == b["index"] and a["doppler_bin"] == b["doppler_bin"]
        assert abs(20 * np.log10(a["snr"] / b["snr"])) < 0.1
    for s, h, tau, f in gw.planted:
        assert out[U.PREC_FP16][s]["detected"] and out[U.PREC_FP16][s]["index"] == tau


def test_hypotheses_as_chip_replicas_match_samples(ctx, oracle):
    """the four hypotheses put as signed chip sequences (e1_sign_chips, n_fft
    a multiple of the code length) run the upsampled-code search; per
    satellite the same selection as the sample replicas and the oracle"""
    chips, sats, raw, dec = scene()
    say("case Galileo hypotheses as chip replicas (upsampled-code search) vs sample replicas and the oracle")
    skeys = put_hypotheses(ctx, chips)
    ckeys = [ctx.replica_put(h, 1, 4 * SPM) for c in chips for h in U.e1_sign_chips(c)]
    try:
        req, hold = request(ctx, raw, ckeys, 4)
        got, _, _ = ctx.acquire_req(req)
        assert ctx.last_search_form()["samples_per_chip"] == 4 * SPM // len(chips[0])
        req_s, hold_s = request(ctx, raw, skeys, 4)
        ref_s, _, _ = ctx.acquire_req(req_s)
    finally:
        for k in skeys + ckeys:
            ctx.replica_drop(k)
    ref = R.process_galileo(dec, chips, spm=SPM, fs_in=FS, fs_acq=FS, doppler=DOP)
    for a, b, c in zip(got, ref_s, ref):
        assert a["detected"] == b["detected"] == c["detected"]
        assert a["index"] == b["index"] == c["index"] and a["doppler_bin"] == b["doppler_bin"] == c["doppler_bin"]
        assert a["level"] == pytest.approx(b["level"], rel=1e-5)
        assert a["level"] == pytest.approx(c["level"], rel=1e-3)
