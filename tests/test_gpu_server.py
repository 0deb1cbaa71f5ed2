"""Request server: continuous batching keeps each result the same as the same
request run alone through ugsdr_hip_acquire, requests admitted while others
run overlap on the device, and the memory budget is enforced.

Stated tolerance of "the same": detection, code-phase index, Doppler bin and
Doppler are equal; level, mean, sigma, metric and runner-up values agree to
1e-5 relative.  (A batch changes which rows the correlation splits across
workgroups, and so the FP32 summation order of a row's blocks.)"""
import numpy as np
import pytest

import ugsdr_amd as U
from ugsdr_amd import workloads as W

pytestmark = pytest.mark.gpu

EXACT = ("detected", "index", "doppler_bin", "doppler", "code_offset")
CLOSE = ("level", "mean", "stddev", "sigma", "snr", "runner_up_code", "runner_up_doppler")


def assert_same(res, ref):
    assert len(res) == len(ref)
    for a, b in zip(res, ref):
        for k in EXACT:
            assert a[k] == b[k], (k, a, b)
        for k in CLOSE:
            assert a[k] == pytest.approx(b[k], rel=1e-5, abs=1e-12), (k, a, b)


@pytest.mark.parametrize("boc_precision,coherent,graphs", [(U.PREC_FP32, False, False), (U.PREC_FP16, False, False),
                                                           (U.PREC_FP32, True, False), (U.PREC_FP32, True, True),
                                                           (U.PREC_FP16, False, True)])
def test_server_results_equal_solo(ctx, boc_precision, coherent, graphs):
    """coherent=True: L1 / L5 requests of T ms coherent (N = 4092 T or
    20460 T, Doppler step 500 / T Hz), so every coherent length of the stream
    shares rounds with the others.  graphs: the rounds' device steps captured
    and replayed as graphs instead of launched one by one."""
    print(f"case server isolation: mixed requests, BOC precision {boc_precision}, coherent T {coherent}, "
          f"round graphs {graphs}", flush=True)
    mix = W.ServeMix(ctx, seed=0xC5, pool=2, boc_precision=boc_precision, coherent=coherent)
    reqs = mix.stream(48 if coherent else 24, 2000.0)
    solo = []
    for r in reqs:
        req, hold = mix.request(r)
        res, _, _ = ctx.acquire_req(req)
        solo.append(res)
    with U.Server(ctx, n_streams=4, max_batch=8, round_graphs=graphs) as srv:
        tickets = []
        holds = []
        for r in reqs:
            req, hold = mix.request(r)
            holds.append((req, hold))
            tickets.append(srv.submit(req))
        lat = []
        for t, ref in zip(tickets, solo):
            got = srv.poll(t, -1)
            assert got is not None
            res, ms = got
            lat.append(ms)
            assert_same(res, ref)
        st = srv.stats()
    print(f"server stats {st}, latency ms p50 {np.median(lat):.3f} max {max(lat):.3f}", flush=True)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    if coherent:
        assert any(r.n_fft > 20460 and r.kind == "L1" for r in reqs)   # T_coh > 1 present
        assert any(r.n_fft > 20460 and r.kind == "L5" for r in reqs)
    assert st["max_batch_seen"] > 1   # requests shared correlation launches
    kinds = {r.kind for r in reqs}
    assert kinds == {"L1", "L5", "BOC"}
    if not graphs:
        assert st["graph_launches"] == 0 and st["graph_fallbacks"] == 0
    else:
        # a round runs as a graph once its kernel sequence has an idle executable
        assert st["graph_launches"] + st["graph_fallbacks"] == st["rounds"]


@pytest.mark.parametrize("kind", ["L1", "BOC"])
def test_server_graph_replay(ctx, kind):
    """one request shape submitted again and again on one stream: the first
    round runs eagerly while its capture is instantiated in the background,
    the later rounds update that executable and replay it, and every result
    equals the request run alone"""
    print(f"case server graph replay {kind}", flush=True)
    mix = W.ServeMix(ctx, seed=0xC8, pool=1)
    r = next(x for x in (mix.draw() for _ in range(100)) if x.kind == kind)
    req, hold = mix.request(r)
    ref, _, _ = ctx.acquire_req(req)
    with U.Server(ctx, n_streams=1, max_batch=1, round_graphs=True) as srv:
        for _ in range(12):
            res, _ = srv.poll(srv.submit(req), -1)
            assert_same(res, ref)
        st = srv.stats()
    print(f"server stats {st}", flush=True)
    assert st["rounds"] == 12 and st["graph_launches"] + st["graph_fallbacks"] == 12
    assert st["graph_launches"] >= 8 and st["graph_updates"] == st["graph_launches"]
    assert 1 <= st["graph_instantiations"] <= 4


def test_server_budget_and_errors(ctx):
    print("case server budget", flush=True)
    mix = W.ServeMix(ctx, seed=0xC6, pool=1)
    r = mix.draw()
    req, hold = mix.request(r)
    with U.Server(ctx, max_device_bytes=1024) as srv:
        with pytest.raises(U.UgsdrError, match="NOMEM"):
            srv.submit(req)
    with U.Server(ctx, max_inflight=2, max_batch=1) as srv:
        ts = [srv.submit(req) for _ in range(6)]
        outs = [srv.poll(t, -1)[0] for t in ts]
        for o in outs:
            assert_same(o, outs[0])
        with pytest.raises(U.UgsdrError, match="STATE"):
            srv.poll(ts[0], 0)


def test_server_pins_replicas(ctx):
    """a replica read by queued or running requests cannot be dropped; once
    their results are collected it can"""
    print("case server replica pin", flush=True)
    mix = W.ServeMix(ctx, seed=0xC7, pool=1)
    r = mix.draw()
    req, hold = mix.request(r)
    key = int(req.replica_keys[0])
    with U.Server(ctx, max_inflight=1, max_batch=1) as srv:
        ts = [srv.submit(req) for _ in range(8)]
        with pytest.raises(U.UgsdrError, match="STATE"):
            ctx.replica_drop(key)
        for t in ts:
            assert srv.poll(t, -1) is not None
    ctx.replica_drop(key)
    with pytest.raises(U.UgsdrError, match="STATE"):
        ctx.replica_drop(key)
