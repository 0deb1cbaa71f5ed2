"""Request server: continuous batching keeps each result identical to the same
request run alone through ugsdr_hip_acquire (bit-exact: the same kernels see
the same rows), requests admitted while others run overlap on the device, and
the memory budget is enforced."""
import numpy as np
import pytest

import ugsdr_amd as U
from ugsdr_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_server_results_equal_solo(ctx):
    print("case server isolation: 24 mixed requests", flush=True)
    mix = W.ServeMix(ctx, seed=0xC5, pool=2)
    reqs = mix.stream(24, 2000.0)
    solo = []
    for r in reqs:
        req, hold = mix.request(r)
        res, _, _ = ctx.acquire_req(req)
        solo.append(res)
    with U.Server(ctx, n_streams=4, max_batch=8) as srv:
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
            assert res == ref
        st = srv.stats()
    print(f"server stats {st}, latency ms p50 {np.median(lat):.3f} max {max(lat):.3f}", flush=True)
    assert st["completed"] == 24 and st["failed"] == 0
    assert st["max_batch_seen"] > 1   # requests shared correlation launches
    kinds = {r.kind for r in reqs}
    assert kinds == {"L1", "L5", "BOC"}


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
        assert all(o == outs[0] for o in outs)
        with pytest.raises(U.UgsdrError, match="STATE"):
            srv.poll(ts[0], 0)
