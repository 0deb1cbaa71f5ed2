#!/usr/bin/env python3
"""Benchmark of the fused acquisition hot path on one MI355X.

Workload (one step = one full search, IQ resident in HBM before timing):
  GPS L1 C/A wide search, plain FP32: complex int16 IQ at 16.368 MHz,
  1 ms coherent x 10 non-coherent blocks, Doppler +-5 kHz step 250 Hz
  (41 bins), 32 PRNs -> 21,474,816 cells (PRN x Doppler x code phase),
  8 seeded satellites at C/N0 38-48 dB-Hz with navigation-bit flips.

Prints one JSON line: whole-search throughput in cells/s, per-search p50/p99
latency, the correlation kernel's HBM roofline figure (algorithmic 8 bytes per
cell-block over the kernel's HIP-event duration) and the CPU baseline (the
oracle's FP32 restatement of the reference's per-PRN/per-Doppler loop, timed on
this host's CPU share on the same input, median of 5 after a warm-up, with the
CPU model and counts recorded; `cpu_baselines_other` holds the same for the
other acquisition configurations and the correlator bank).

`other_workloads` adds the other acquisition configurations (1 ms and 4 ms
L1 C/A, BOC 4 ms in FP32 and in the reduced FP16/BF16 form, Galileo E1 code
hypotheses), the correlator bank, closed-loop tracking, searches straight from
a packed BbpDdc recording, the DFE channel step and the anti-jamming front
end; `serving` and `serving_boc_fp16` are the request server under Poisson
arrivals (L1 / L5 requests of T ms coherent), `serving_noncoherent_1ms` the
same stream with T blocks of 1 ms summed non-coherently, `serving_round_graphs`
the first stream with the server's rounds run as captured hipGraphs.  `--leg bank|tracking|tracking_many|jse|galileo|galileo_c3` runs one of
those alone.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "acq cells/s (PRN×Doppler×phase) vs FP32 plain + p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
B_CB = 8  # algorithmic bytes per cell-block, FP32 plain


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--workload", default="L1CA_WIDE", choices=["L1CA_WIDE", "L1CA_1MS", "L1CA_4MS", "BOC_4MS"])
    ap.add_argument("--serve-requests", type=int, default=2000, help="serving leg: requests (0 skips)")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="serving leg: mean Poisson arrival rate")
    ap.add_argument("--no-extra", action="store_true", help="skip the other workloads")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16", "bf16"],
                    help="storage of the measured search (fp16/bf16: the reduced form)")
    ap.add_argument("--threads", type=int, default=0,
                    help="pin the correlation variant (workgroup size) instead of autotuning")
    ap.add_argument("--leg", default="", choices=["", "bank", "tracking", "tracking_many", "jse", "galileo", "galileo_c3"],
                    help="run only one of the other workloads and print its JSON (for profiling)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (rocprofv3 counter calibration)")
    return ap.parse_args()


def log(msg: str):
    """progress line on stderr (the JSON result goes to stdout)"""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def percentile(xs, q):
    xs = sorted(xs)
    k = (len(xs) - 1) * q
    f = int(k)
    c = min(f + 1, len(xs) - 1)
    return xs[f] + (xs[c] - xs[f]) * (k - f)


def traffic_from_profiles(workload: str):
    """HBM bytes per launch of the correlation kernel from committed rocprofv3
    --pmc summaries (FETCH_SIZE / WRITE_SIZE, copy-kernel calibrated)."""
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f).get("corr_bytes_per_launch")
    except Exception:
        return None


def host_cpu() -> dict:
    """the host's CPU model and counts: hardware_concurrency (os.cpu_count),
    the CPUs this process may run on, and the host's per-process share when
    it sets one (OMP_NUM_THREADS)"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return {"model": model, "hardware_concurrency": os.cpu_count(), "affinity": affinity,
            "omp_num_threads": int(omp) if omp.isdigit() else None}


def cpu_pool_threads(requested: int = 0) -> int:
    """threads for the CPU baseline pool: the CPUs this process may use,
    capped by the host's stated share"""
    if requested:
        return requested
    h = host_cpu()
    n = h["affinity"]
    if h["omp_num_threads"]:
        n = min(n, h["omp_num_threads"])
    return max(1, n)


def median_runs(fn, runs: int = 5):
    """one warm-up call, then `runs` timed calls: (median s, per-run s, last result)"""
    fn()
    times, out = [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t0)
    return statistics.median(times), times, out


def cpu_search_baseline(w, threads: int, budget_s: float = 1.5, runs: int = 5):
    """The oracle's FP32 search (the reference's structure: per PRN x Doppler x
    block: mix, forward DFT, conj-spectrum multiply, inverse DFT / N, |.|,
    fold, max, mean / sigma), PRN-parallel over `threads` host threads, on the
    workload's own input.  Bounded sample: the first PRNs and Doppler bins,
    sized so one run takes about budget_s; cells/s of the sample, median of
    `runs` after one warm-up."""
    import numpy as np

    from oracle import refcpu

    x = (w.iq[0::2].astype(np.float32) + 1j * w.iq[1::2].astype(np.float32)).astype(np.complex64)
    reps = w.replicas()
    n_prn, n_dop = len(reps), len(w.doppler)

    def search(n_p, n_d):
        return refcpu.acquire(x, reps[:n_p], n_fft=w.n_fft, n_blocks=w.n_blocks, surface_len=w.surface_len,
                              fs_in=w.fs, fs_acq=w.fs, doppler=w.doppler[:n_d], reshape=w.reshape, precision=32,
                              threads=threads)

    # probe: one PRN per thread over every bin (also warms the transform plans)
    n_p = min(n_prn, threads)
    t0 = time.perf_counter()
    search(n_p, n_dop)
    probe = time.perf_counter() - t0
    if probe > budget_s:
        n_d = max(1, min(n_dop, int(n_dop * budget_s / probe)))
    else:
        n_d = n_dop
        n_p = min(n_prn, n_p * max(1, int(budget_s / max(probe, 1e-4))))
    t, times, (res, _, _) = median_runs(lambda: search(n_p, n_d), runs)
    cells = n_p * n_d * w.surface_len
    whole = n_p == n_prn and n_d == n_dop
    return {
        "value": cells / t,
        "unit": "cells/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{'the whole search' if whole else 'sample'}: {n_p} of {n_prn} PRNs x {n_d} of {n_dop} "
                   f"bins x {w.n_blocks} blocks of N = {w.n_fft}, FP32, {threads} threads, median of {runs} "
                   f"after 1 warm-up ({t:.3f} s)"),
        "runs_s": times,
        "results": res if n_d == n_dop else None,   # per PRN comparable with the GPU search
    }


def cpu_bank_baseline(bw, threads: int, budget_s: float = 1.5, runs: int = 5):
    """The oracle's correlator bank (the reference's Correlator taps per channel
    with the FP64 carrier and code phases), channel-parallel over `threads`
    host threads: channel-samples/s over the first blocks of the workload's
    input, median of `runs` after one warm-up."""
    import numpy as np

    from oracle import refcpu

    def run(nb):
        x = (bw.iq[0:2 * nb * bw.spm:2].astype(np.float32) +
             1j * bw.iq[1:2 * nb * bw.spm:2].astype(np.float32)).astype(np.complex64)
        return refcpu.bank_nco(x, bw.spm, nb, bw.fs, bw.channels, bw.codes, bw.taps, threads=threads)

    nb = 4
    t0 = time.perf_counter()
    run(nb)
    probe = time.perf_counter() - t0
    nb = max(1, min(bw.n_blocks, int(nb * budget_s / max(probe, 1e-4))))
    t, times, _ = median_runs(lambda: run(nb), runs)
    cs = len(bw.channels) * nb * bw.spm
    return {
        "value": cs / t,
        "unit": "channel-samples/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{len(bw.channels)} channels x {len(bw.taps)} taps over {nb} of {bw.n_blocks} blocks of "
                   f"{bw.spm} samples, {threads} threads, median of {runs} after 1 warm-up ({t:.3f} s)"),
        "runs_s": times,
    }


def fft_calibration(lengths=(4092, 16368, 81840), reps: int = 20):
    """the oracle's FP32 transform time over numpy's pocketfft (complex64) per length"""
    import numpy as np

    from oracle import refcpu

    out = {}
    for n in lengths:
        rng = np.random.default_rng(n)
        z = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
        refcpu.fft(z, precision=32)
        np.fft.fft(z)
        t0 = time.perf_counter()
        for _ in range(reps):
            refcpu.fft(z, precision=32)
        t_or = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            np.fft.fft(z)
        t_np = (time.perf_counter() - t0) / reps
        out[str(n)] = {"oracle_us": 1e6 * t_or, "pocketfft_us": 1e6 * t_np, "ratio": t_or / t_np}
    return out


def time_search(ctx, w, steps: int, warmup: int, precision: int = 0):
    """seconds per search of workload w with its IQ resident in HBM"""
    import torch

    import ugsdr_amd as U

    keys = [ctx.replica_put(c, w.repeats, w.n_fft, precision) for c in w.codes]
    iq_dev = torch.from_numpy(w.iq).to("cuda:0")
    req, hold = ctx.make_request(None, w.fmt, fs_in=w.fs, fs_acq=w.fs, n_fft=w.n_fft, n_blocks=w.n_blocks,
                                 surface_len=w.surface_len, doppler=w.doppler, keys=keys, reshape=w.reshape,
                                 n_samples=w.n_samples, iq_device_ptr=iq_dev.data_ptr(), precision=precision)
    for _ in range(warmup):
        ctx.acquire_req(req, raw=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        res, _, _ = ctx.acquire_req(req, raw=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    res = U.results_to_dicts(res)
    ctx.set_profiling(True)
    ctx.acquire_req(req)
    stages = ctx.last_times()
    ctx.set_profiling(False)
    for k in keys:
        ctx.replica_drop(k)
    return dt, stages, res


def galileo_workload(ctx, c3: bool = False):
    """Galileo E1 code hypotheses: every satellite's four sign permutations in
    one fused search, one result per satellite (ProcessGalileo's selection).
    c3: the BOC acquisition configuration's shape (36 satellites, 4 ms x 5
    blocks at 20.46 MHz, +-6 kHz / 125 Hz) in the reduced FP16 form, its
    hypotheses put as signed chips (81840 = 10 x 8184: the upsampled-code
    search)"""
    import torch

    import ugsdr_amd as U
    from ugsdr_amd import workloads as W

    if c3:
        gw = W.galileo_e1(seed=0xE3, n_sv=36, fs=20.46e6, n_blocks=5, doppler_range=6000, doppler_step=125,
                          name="GALILEO_E1_HYP_C3_FP16")
        prec = U.PREC_FP16
    else:
        gw = W.galileo_e1()
        prec = U.PREC_FP32
    # signed chip replicas where the 4 ms code divides the transform (the
    # upsampled-code search), else the upsampled samples
    if all(gw.n_fft % len(c) == 0 and len(c) % 4 == 0 for c in gw.codes):
        gkeys = [ctx.replica_put(h, 1, gw.n_fft, prec) for c in gw.codes for h in U.e1_sign_chips(c)]
    else:
        gkeys = [ctx.replica_put_samples(r, prec) for c in gw.codes for r in U.e1_sign_hypotheses(c, gw.n_fft)]
    dev = torch.from_numpy(gw.iq).to("cuda:0")
    req, hold = ctx.make_request(None, U.FMT_CI16, fs_in=gw.fs, fs_acq=gw.fs, n_fft=gw.n_fft, n_blocks=gw.n_blocks,
                                 surface_len=gw.n_fft, doppler=gw.doppler, keys=gkeys, reshape=False,
                                 n_samples=gw.n_fft * gw.n_blocks, iq_device_ptr=dev.data_ptr(), hypotheses=4,
                                 precision=prec)
    for _ in range(3):
        ctx.acquire_req(req, raw=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        res, _, _ = ctx.acquire_req(req, raw=True)
    dt = (time.perf_counter() - t0) / 10
    ctx.set_profiling(True)
    ctx.acquire_req(req)
    stages = ctx.last_times()
    ctx.set_profiling(False)
    for k in gkeys:
        ctx.replica_drop(k)
    res = U.results_to_dicts(res)
    found = sorted(i for i, r in enumerate(res) if r["detected"])
    return {gw.name: {"ms_per_search": 1e3 * dt, "cells_per_s": gw.cells / dt, "cells": gw.cells,
                      "satellites": len(gw.codes), "hypotheses": 4, "n_fft": gw.n_fft, "blocks": gw.n_blocks,
                      "precision": {0: "fp32", 1: "fp16"}[prec], "stage_ms": stages, "detected_sv": found,
                      "planted_found_at_delay": sum(1 for s, h, tau, f in gw.planted
                                                    if res[s]["detected"] and res[s]["index"] == tau),
                      "planted": len(gw.planted)}}


def extra_workloads(ctx):
    """cells/s of the other acquisition workloads and the correlator bank rate"""
    import numpy as np
    import torch

    import ugsdr_amd as U
    from ugsdr_amd import workloads as W

    out = {}
    boc32 = None
    for name, steps in (("L1CA_1MS", 50), ("L1CA_4MS", 30), ("BOC_4MS", 3)):
        w = W.CONFIGS[name]()
        dt, stages, res = time_search(ctx, w, steps, 2)
        form = ctx.last_search_form()
        alg = 8 * w.cells * w.n_blocks
        out[name] = {"cells_per_s": w.cells / dt, "ms_per_search": 1e3 * dt, "cells": w.cells,
                     "cell_blocks": w.cells * w.n_blocks, "stage_ms": stages,
                     "detected": [i for i, r in enumerate(res) if r["detected"]]}
        if form["samples_per_chip"]:
            # the upsampled-code search's correlation kernel against 8 B per cell-block
            out[name]["roofline"] = {"bound": "hbm", "achieved": alg / (stages["corr"] * 1e-3) / 1e9,
                                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": alg / (stages["corr"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                     "bytes_per_cell_block": 8,
                                     "kernel": corr_kernel_name(form, w.surface_len, 0)}
        if name == "BOC_4MS":
            boc32 = (w, res)
    # the BOC search in the reduced form: 16-bit storage, matrix-core DFT stages
    w, res32 = boc32
    planted = sorted(s.prn for s in w.sats)
    for tag, prec in (("BOC_4MS_FP16", U.PREC_FP16), ("BOC_4MS_BF16", U.PREC_BF16)):
        dt, stages, res = time_search(ctx, w, 5, 2, prec)
        form = ctx.last_search_form()
        d_db = [20 * np.log10(res[p]["snr"] / res32[p]["snr"]) for p in planted]
        same = all(res[p]["index"] == res32[p]["index"] and res[p]["doppler_bin"] == res32[p]["doppler_bin"]
                   for p in planted)
        alg = 4 * w.cells * w.n_blocks
        out[tag] = {"cells_per_s": w.cells / dt, "ms_per_search": 1e3 * dt, "cells": w.cells,
                    "cell_blocks": w.cells * w.n_blocks, "stage_ms": stages,
                    "roofline": {"bound": "hbm", "achieved": alg / (stages["corr"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                                 "unit": "GB/s", "frac": alg / (stages["corr"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                 "bytes_per_cell_block": 4,
                                 "kernel": corr_kernel_name(form, w.surface_len, 0, "fp16" if prec == U.PREC_FP16
                                                            else "bf16")},
                    "pnr_delta_db_vs_fp32_planted": {"max_abs": float(np.max(np.abs(d_db))),
                                                     "values": [float(v) for v in d_db]},
                    "planted_index_and_bin_equal_fp32": same,
                    "speedup_vs_fp32": out["BOC_4MS"]["ms_per_search"] / (1e3 * dt)}
    out.update(galileo_workload(ctx))
    out.update(galileo_workload(ctx, c3=True))
    bw = W.bank_l5()
    iq_dev = torch.from_numpy(bw.iq).to("cuda:0")
    out["BANK_L5"] = bank_workload(ctx, bw, iq_dev)
    out["TRACK_L5_64CH"] = tracking_workload(ctx, bw, iq_dev)
    out.update(tracking_many_workload(ctx))
    return out


def bank_workload(ctx, bw, iq_dev):
    """The correlator bank in NCO mode: 64 channels x 5 taps over 1000 blocks of
    40920 samples resident in HBM; the kernel time is the median of 3 calls."""
    import numpy as np
    import torch

    import ugsdr_amd as U

    ctx.correlate_bank(None, U.FMT_CI16, spm=bw.spm, n_blocks=2, fs=bw.fs, channels=bw.channels, codes=bw.codes,
                       taps_chips=bw.taps, mode=U.BANK_NCO, iq_device_ptr=iq_dev.data_ptr())
    torch.cuda.synchronize()
    ctx.set_profiling(True)
    wall, kern = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        o, _ = ctx.correlate_bank(None, U.FMT_CI16, spm=bw.spm, n_blocks=bw.n_blocks, fs=bw.fs, channels=bw.channels,
                                  codes=bw.codes, taps_chips=bw.taps, mode=U.BANK_NCO, iq_device_ptr=iq_dev.data_ptr())
        wall.append(time.perf_counter() - t0)
        kern.append(ctx.last_times()["corr"] * 1e-3)
    ctx.set_profiling(False)
    k = statistics.median(kern)
    p = np.abs(o[:, :, 2])
    # kernel time from HIP events on the context's stream; the wall time adds the
    # channel-table upload and the copy of the 64 x 1000 x 5 outputs to the host
    return {"channel_samples_per_s": bw.channel_samples / k, "kernel_ms_per_call": 1e3 * k,
            "wall_ms_per_call": 1e3 * statistics.median(wall),
            "blocks": bw.n_blocks, "channels": len(bw.channels), "taps": len(bw.taps),
            "prompt_over_very_early_fraction": float(np.mean(p > np.abs(o[:, :, 0])))}


def tracking_workload(ctx, bw, iq_dev, cpu: bool = True):
    """closed-loop tracking (ugsdr_hip_track) on the correlator bank's signal:
    64 channels, each on its own seeded L5-like satellite (10230-chip codes,
    40.92 MHz complex int16, C/N0 40-50 dB-Hz, code Doppler included), started
    from an acquisition-size error (carrier +-60 Hz, code +-0.5 chip) and run
    through 1000 consecutive 1 ms epochs: mix, early / prompt / late split
    correlation, PLL/FLL and DLL update per epoch, all on the device.  Timed
    per call (1 s of signal, every channel) with HIP events around the loop
    kernel; the CPU leg is the oracle's loop (one channel, 1 core)."""
    import numpy as np

    import ugsdr_amd as U
    from oracle import refcpu as R

    rng = np.random.default_rng(0x7C4)
    n_ch, n_ep = len(bw.channels), bw.n_blocks
    spc = bw.fs / 10.23e6
    truth = [(s.doppler, (s.code_phase * spc) % bw.spm) for s in bw.sats]
    start = [(f + rng.uniform(-60, 60), tau + spc * rng.uniform(-0.5, 0.5)) for f, tau in truth]

    def fresh():
        return [U.track_state(cp, f, base_code_frequency=10.23e6) for f, cp in start]

    kw = dict(spm=bw.spm, n_blocks=n_ep, fs=bw.fs, codes=bw.codes, iq_device_ptr=iq_dev.data_ptr(), want_trace=True)
    ctx.track(None, U.FMT_CI16, states=fresh(), **kw)
    ctx.set_profiling(True)
    kern, wall = [], []
    for _ in range(3):
        states = fresh()
        t0 = time.perf_counter()
        epl, tr = ctx.track(None, U.FMT_CI16, states=states, **kw)
        wall.append(time.perf_counter() - t0)
        kern.append(ctx.last_times()["corr"] * 1e-3)
    ctx.set_profiling(False)
    k = statistics.median(kern)
    P = epl[-200:, :, 1]
    locked = [bool(abs(tr[-1, c, 1] - truth[c][0]) < 5.0 and np.mean(np.abs(P[:, c].imag)) < 0.3 * np.mean(np.abs(P[:, c].real)))
              for c in range(n_ch)]
    out = {"channels": n_ch, "epochs": n_ep, "taps": 3, "fs": bw.fs,
           "kernel_ms_per_call": 1e3 * k, "wall_ms_per_call": 1e3 * statistics.median(wall),
           "channel_epochs_per_s": n_ch * n_ep / k, "channel_samples_per_s": n_ch * n_ep * bw.spm / k,
           "us_per_epoch": 1e6 * k / n_ep, "realtime_factor": n_ep * 1e-3 / k,
           "locked_channels": int(sum(locked)),
           "note": "one call = every channel's loop over 1 s of signal resident in HBM; each epoch "
                   "depends on the previous epoch's loop update, so the epochs of a channel run in order"}
    if cpu:
        code3 = R.repeat_upsample(bw.codes[0], 3, 3 * bw.spm)
        x = bw.iq[:2 * bw.spm * 20].astype(np.float32).view(np.complex64)
        f, cp = start[0]
        st = R.track_state(cp, f, base_code_frequency=10.23e6)
        t0 = time.perf_counter()
        R.track(x, bw.spm, 20, bw.fs, code3, st)
        dt = time.perf_counter() - t0
        out["cpu_oracle"] = {"channel_epochs_per_s": 20 / dt, "cores": 1, "kind": "port",
                             "sample": "channel 0, 20 epochs"}
        out["speedup_vs_cpu_oracle_1core"] = out["channel_epochs_per_s"] / out["cpu_oracle"]["channel_epochs_per_s"]
    return out


def tracking_many_workload(ctx):
    """many-channel tracking at a low rate: 512 channels (two per CU) on a
    GPS L1 C/A signal of 8 seeded satellites (int16 at 4.092 MHz, C/N0
    44-48 dB-Hz), 64 channels per satellite started at different
    acquisition-size errors, 1000 epochs per call"""
    import numpy as np
    import torch

    import ugsdr_amd as U
    from ugsdr_amd import workloads as W

    rng = np.random.default_rng(0x7A1)
    fs, spm, n_ep, n_ch = 4.092e6, 4092, 1000, 512
    codes = W.l1ca_codes(32)
    sats = W._draw_sats(rng, 8, 32, lambda p: codes[p], 1.023e6, W.L1_HZ, 1023, 44, 48)
    iq = W.synthesize(sats, fs, spm * n_ep, 100.0, 0x7A1, 16)
    dev = torch.from_numpy(iq).to("cuda:0")
    chans = [(s, s.doppler + rng.uniform(-60, 60), (s.code_phase * 4.0) % spm + rng.uniform(-2, 2))
             for i in range(n_ch) for s in [sats[i % 8]]]

    def fresh():
        return [U.track_state(cp, f) for _, f, cp in chans]

    kw = dict(spm=spm, n_blocks=n_ep, fs=fs, codes=[codes[s.prn] for s, _, _ in chans], iq_device_ptr=dev.data_ptr())
    ctx.track(None, U.FMT_CI16, states=fresh(), **kw)
    ctx.set_profiling(True)
    kern = []
    for _ in range(3):
        epl, tr = ctx.track(None, U.FMT_CI16, states=fresh(), **kw)
        kern.append(ctx.last_times()["corr"] * 1e-3)
    ctx.set_profiling(False)
    k = statistics.median(kern)
    P = epl[-200:, :, 1]
    locked = sum(1 for c, (s, _, _) in enumerate(chans)
                 if abs(tr[-1, c, 1] - s.doppler) < 5.0
                 and np.mean(np.abs(P[:, c].imag)) < 0.3 * np.mean(np.abs(P[:, c].real)))
    return {"TRACK_L1_512CH": {"channels": n_ch, "epochs": n_ep, "fs": fs, "kernel_ms_per_call": 1e3 * k,
                               "channel_epochs_per_s": n_ch * n_ep / k, "us_per_epoch": 1e6 * k / n_ep,
                               "realtime_factor": n_ep * 1e-3 / k, "locked_channels": locked}}


def jse_workload(ctx):
    """the anti-jamming front end (ugsdr_hip_jse_process) on 100 ms at 4.2 MHz
    held in HBM: noise plus a CW jammer 21 dB above it (every column's
    strongest bin, so the reference's rule reports it narrowband and wideband
    and the STFT excision runs), and the same with a jammer-free input
    (detection only).  The CPU leg is the oracle on 10 ms."""
    import numpy as np
    import torch

    import ugsdr_amd as U
    from oracle import refcpu as R

    fs, n = 4.2e6, 420000
    rng = np.random.default_rng(0x15E)
    k = np.arange(n)
    noise = 0.5 * (rng.normal(size=n) + 1j * rng.normal(size=n))
    out = {}
    for tag, x in (("JSE_CW_100MS", noise + 6.0 * np.exp(2j * np.pi * 40 * k / 128)), ("JSE_CLEAN_100MS", noise)):
        x = x.astype(np.complex64)
        dev = torch.from_numpy(x).to("cuda:0")
        work = dev.clone()
        for _ in range(2):
            work.copy_(dev)
            torch.cuda.synchronize()
            ctx.jse_process(None, fs, device_ptr=work.data_ptr(), n=n)
        wall, stages = [], []
        for _ in range(10):
            work.copy_(dev)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, det = ctx.jse_process(None, fs, device_ptr=work.data_ptr(), n=n)
            wall.append(time.perf_counter() - t0)
        ctx.set_profiling(True)
        work.copy_(dev)
        torch.cuda.synchronize()
        ctx.jse_process(None, fs, device_ptr=work.data_ptr(), n=n)
        t = ctx.last_times()
        ctx.set_profiling(False)
        dt = statistics.median(wall)
        ncol = (n - 96) // 32
        out[tag] = {"ms_per_call": 1e3 * dt, "msamples_per_s": n / dt / 1e6, "realtime_factor": 0.1 / dt,
                    "columns": ncol, "detections": [(d[0], d[1]) for d in det],
                    "stage_ms": {"stft_detect": t["front"], "host_analysis": t["classes"], "suppression": t["corr"]},
                    "stft_detect_gb_per_s": (n * 8 + ncol * 128 * 8) / (t["front"] * 1e-3) / 1e9}
        m = 42000
        t0 = time.perf_counter()
        R.jse_process(x[:m], fs, R.WINDOW_BLACKMAN_HARRIS)
        cdt = time.perf_counter() - t0
        out[tag]["cpu_oracle"] = {"msamples_per_s": m / cdt / 1e6, "cores": 1, "kind": "port",
                                  "sample": "first 10 ms"}
    return out


def recorded_workloads(ctx):
    """acquisition straight from a packed BbpDdc recording: the first 4 ms of
    the reference's data/bbp_ddc_gps_L1.dat (tests/golden), and a 100 ms packed
    stream (those epochs repeated under consecutive ms counters) that sizes the
    front kernel's decode + resample rate.  2+2-bit samples at 33.25 MHz are
    decoded inside the front kernel and resampled to 4.2 MHz."""
    import numpy as np
    import torch

    import ugsdr_amd as U

    raw = np.fromfile(os.path.join(ROOT, "tests", "golden", "bbp_ddc_gps_L1_first4ms.bin"), np.uint8)
    fs_in, fs_acq = 33.25e6, 4.2e6
    spm_in, N = 33250, 4200
    eb = U.format_bytes(U.FMT_BBP_DDC, spm_in, fs_in)
    ep = raw.reshape(4, eb)
    stream = np.tile(ep, (25, 1))
    first = int.from_bytes(ep[0, :8].tobytes(), "little") & 0xFFFFFF
    for e in range(100):
        word = (int.from_bytes(stream[e, :8].tobytes(), "little") & ~0xFFFFFF) | ((first + e) & 0xFFFFFF)
        stream[e, :8] = np.frombuffer(word.to_bytes(8, "little"), np.uint8)
    dop = U.doppler_bins(5000, 250)
    keys = [ctx.replica_put(U.gps_l1ca(s), 1, N) for s in range(32)]
    out = {}
    for tag, n_ms, data in (("BBP_L1_RECORDING", 4, raw), ("BBP_STREAM_100MS", 100, stream.reshape(-1))):
        dev = torch.from_numpy(np.ascontiguousarray(data)).to("cuda:0")
        req, hold = ctx.make_request(None, U.FMT_BBP_DDC, fs_in=fs_in, fs_acq=fs_acq, n_fft=N, n_blocks=n_ms,
                                     surface_len=N, doppler=dop, keys=keys, n_samples=n_ms * spm_in,
                                     iq_device_ptr=dev.data_ptr())
        for _ in range(2):
            ctx.acquire_req(req, raw=True)
        torch.cuda.synchronize()
        steps = 10
        t0 = time.perf_counter()
        for _ in range(steps):
            res, _, _ = ctx.acquire_req(req, raw=True)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        ctx.set_profiling(True)
        ctx.acquire_req(req)
        stages = ctx.last_times()
        ctx.set_profiling(False)
        res = U.results_to_dicts(res)
        cells = 32 * len(dop) * N
        in_bytes = n_ms * eb
        front_s = stages["front"] * 1e-3
        out[tag] = {"ms_per_search": 1e3 * dt, "cells_per_s": cells / dt, "cells": cells, "blocks": n_ms,
                    "stage_ms": stages, "detected_sv": [i for i, r in enumerate(res) if r["detected"]],
                    "front": {"input_msamples_per_s": n_ms * spm_in / front_s / 1e6,
                              "hbm_bytes": in_bytes + n_ms * N * 8,
                              "gb_per_s": (in_bytes + n_ms * N * 8) / front_s / 1e9}}
    for k in keys:
        ctx.replica_drop(k)
    # the DFE channel step alone over the 100 ms packed stream, device in / out
    src = torch.from_numpy(np.ascontiguousarray(stream.reshape(-1))).to("cuda:0")
    dst = torch.zeros(100 * N, dtype=torch.complex64, device="cuda:0")
    with U.Dfe(ctx, U.FMT_BBP_DDC, fs_in, fs_acq, 2.5e3) as ch:
        for _ in range(3):
            ch.process(None, 100, in_device_ptr=src.data_ptr(), out_device_ptr=dst.data_ptr())
        steps = 20
        t0 = time.perf_counter()
        for _ in range(steps):
            ch.process(None, 100, in_device_ptr=src.data_ptr(), out_device_ptr=dst.data_ptr())
        dt = (time.perf_counter() - t0) / steps
    moved = 100 * eb + 100 * N * 8
    out["DFE_BBP_100MS"] = {"ms_per_call": 1e3 * dt, "input_msamples_per_s": 100 * spm_in / dt / 1e6,
                            "realtime_factor": 0.1 / dt, "hbm_bytes": moved, "gb_per_s": moved / dt / 1e9,
                            "note": "one call = 100 ms of 2+2-bit samples at 33.25 MHz decoded, mixed by 2.5 kHz "
                                    "and resampled to 4.2 MHz; wall time per call including its stream sync"}
    return out


def serving(ctx, n_requests: int, rate: float, boc_precision: int = 0, coherent: bool = True,
            round_graphs: bool = False):
    """Poisson arrivals of the mixed request stream into the continuous-batching
    server.  coherent: L1 / L5 requests of T ms coherent (N = 4092 T or
    20460 T, Doppler step 500 / T Hz); otherwise T blocks of 1 ms summed
    non-coherently (step 500 Hz).  round_graphs: the server's rounds run as
    captured and replayed hipGraphs."""
    import numpy as np

    import ugsdr_amd as U
    from ugsdr_amd import workloads as W

    mix = W.ServeMix(ctx, seed=0xC5, boc_precision=boc_precision, coherent=coherent)
    # warm-up: a shorter stream of the same mix at the same rate into the same
    # server (its replica cache, slot buffers, cost estimates and round graphs
    # then hold the steady state); not timed
    warm_reqs = mix.stream(max(64, n_requests // 4), rate)
    warm = [mix.request(r) for r in warm_reqs]
    reqs = mix.stream(n_requests, rate)
    built = [mix.request(r) for r in reqs]
    with U.Server(ctx, n_streams=4, max_batch=16, round_graphs=round_graphs) as srv:
        t0 = time.perf_counter()
        warm_t = []
        for (q, _), r in zip(warm, warm_reqs):
            now = time.perf_counter() - t0
            if r.arrival_s > now:
                time.sleep(r.arrival_s - now)
            warm_t.append(srv.submit(q))
        for t in warm_t:
            srv.poll(t, -1)
        st0 = srv.stats()
        t0 = time.perf_counter()
        tickets = []
        for (q, _), r in zip(built, reqs):
            now = time.perf_counter() - t0
            if r.arrival_s > now:
                time.sleep(r.arrival_s - now)
            tickets.append(srv.submit(q))
        submit_end = time.perf_counter() - t0
        lat = [srv.poll(t, -1)[1] for t in tickets]
        elapsed = time.perf_counter() - t0
        st = srv.stats()
    cells = sum(r.cells for r in reqs)
    # transform work of the searches as requested: rows x blocks x n_fft
    # points (a periodic-fold search runs shorter transforms for the same result)
    points = sum(len(r.keys) * len(r.doppler) * r.n_blocks * r.n_fft for r in reqs)
    kinds = {k: sum(1 for r in reqs if r.kind == k) for k in ("L1", "L5", "BOC")}
    lat_by = {k: [l for l, r in zip(lat, reqs) if r.kind == k] for k in kinds}
    return {
        "requests": n_requests, "offered_rate_hz": n_requests / submit_end, "mean_rate_hz": rate,
        "p50_ms": percentile(lat, 0.5), "p99_ms": percentile(lat, 0.99), "max_ms": max(lat),
        "mean_ms": statistics.mean(lat),
        "p99_ms_by_kind": {k: percentile(v, 0.99) for k, v in lat_by.items() if v},
        "cells_per_s": cells / elapsed, "mix": kinds, "rounds": st["rounds"] - st0["rounds"],
        "max_batch_seen": st["max_batch_seen"], "failed": st["failed"],
        "warmup_requests": len(warm_reqs),
        "graphs": {k: st[k] - st0[k]
                   for k in ("graph_launches", "graph_instantiations", "graph_updates", "graph_fallbacks")},
        "launch_host_us_per_round": (st["launch_host_us"] - st0["launch_host_us"]) /
                                    max(1, st["rounds"] - st0["rounds"]),
        "boc_precision": {0: "fp32", 1: "fp16", 2: "bf16"}[boc_precision],
        "l1_l5_integration": "T ms coherent" if coherent else "1 ms coherent x T non-coherent",
        "round_graphs": round_graphs,
        "transform_points_per_s": points / elapsed,
    }


def corr_kernel_name(form: dict, surface_len: int, threads: int, precision: str = "fp32") -> str:
    """the correlation kernel a search ran, from the form the library reports
    (transform length, periods folded, samples per chip of the upsampled-code
    search) and the tuned variant: the upsampled-code search runs the
    wave-autonomous kernel for 1023-chip codes and the phase transforms for
    longer ones (the reduced form: matrix-core phase transforms of N1 x 341
    points); otherwise the first variant of a 1023*M length (M = 4, 8, 12,
    16) is the three-stage kernel, and the reduced form the four-step
    matrix-core kernel"""
    import ugsdr_amd as U

    n = form["n_transform"]
    upc = form["samples_per_chip"]
    if precision != "fp32":
        t = threads or 1024
        if upc:
            lc = n // upc
            return (f"corr_upm<{precision},{lc // 341},{t}> (upsampled code: {upc} phase transforms of "
                    f"{lc} = {lc // 341} x 31 x 11 on the matrix cores)")
        return f"corr_mfma<{precision},{t}> (N={n})"
    if upc:
        lc = n // upc
        if lc == 1023:
            return f"corr1023w<{upc}> (upsampled code: {upc} phases of 1023 points, wave-autonomous)"
        return f"corr1023p<{lc // 1023}> (upsampled code: {upc} phase transforms of {lc} = 31 x 33 x {lc // 1023})"
    fold = f", {form['periods']} periods folded" if form["periods"] > 1 else ""
    if n % 1023 == 0 and n // 1023 in (4, 8, 12, 16) and threads == U.corr_variants(n)[0]:
        return f"corr1023<{n // 1023},{surface_len // 1023}> (31 x 33 x {n // 1023}, {threads} threads{fold})"
    if n <= 16384:
        return f"corr_small<{threads},...> (N={n}{fold})"
    return f"corr_large<{threads},...> (N={n}{fold})"


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the acquisition hot path is measured on a single MI355X "
              "(see the sharding notes in DESIGN.md); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import ugsdr_amd as U
    from ugsdr_amd import workloads as W

    if args.leg:
        ctx = U.Context(0)
        if args.leg == "bank":
            bw = W.bank_l5()
            out = {"BANK_L5": bank_workload(ctx, bw, torch.from_numpy(bw.iq).to("cuda:0"))}
        elif args.leg == "tracking":
            bw = W.bank_l5()
            out = {"TRACK_L5_64CH": tracking_workload(ctx, bw, torch.from_numpy(bw.iq).to("cuda:0"), cpu=False)}
        elif args.leg == "tracking_many":
            out = tracking_many_workload(ctx)
        elif args.leg == "jse":
            out = jse_workload(ctx)
        else:
            out = galileo_workload(ctx, c3=args.leg == "galileo_c3")
        print(json.dumps(out))
        return

    if args.copy_calibration_mib:
        # a copy kernel of known size: FETCH_SIZE / WRITE_SIZE per dispatch of it
        # should read the buffer size; the ratio scales the counters of the other kernels
        n = args.copy_calibration_mib << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_dispatch": n, "dispatches": 5}))
        return

    log(f"workload {args.workload}")
    w = W.CONFIGS[args.workload]()
    ctx = U.Context(0)
    prec = {"fp32": U.PREC_FP32, "fp16": U.PREC_FP16, "bf16": U.PREC_BF16}[args.precision]
    keys = [ctx.replica_put(c, w.repeats, w.n_fft, prec) for c in w.codes]
    iq_dev = torch.from_numpy(w.iq).to("cuda:0")
    torch.cuda.synchronize()
    variants = U.corr_variants(w.n_fft)
    if args.threads:
        ctx.set_threads(w.n_fft, args.threads)
        chosen, tune_ms = args.threads, []
    else:
        chosen, tune_ms = ctx.autotune(w.n_fft) if prec == U.PREC_FP32 else (0, [])
    log(f"autotune N={w.n_fft}: variants {variants} ms {tune_ms} -> {chosen}")
    req, hold = ctx.make_request(None, w.fmt, fs_in=w.fs, fs_acq=w.fs, n_fft=w.n_fft, n_blocks=w.n_blocks,
                                 surface_len=w.surface_len, doppler=w.doppler, keys=keys, reshape=w.reshape,
                                 n_samples=w.n_samples, iq_device_ptr=iq_dev.data_ptr(), precision=prec)
    for _ in range(args.warmup):
        res, _, _ = ctx.acquire_req(req, raw=True)
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    lat = []
    replays0 = ctx.graph_replays()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = time.perf_counter()
        res, _, _ = ctx.acquire_req(req, raw=True)
        lat.append(time.perf_counter() - s)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    replays = ctx.graph_replays() - replays0
    res = U.results_to_dicts(res)
    ms_per_step = 1e3 * elapsed / args.steps
    value = w.cells / (elapsed / args.steps)

    # kernel-level timing with HIP events on the context's stream
    ctx.set_profiling(True)
    stage = {"front": [], "classes": [], "corr": [], "final": []}
    for _ in range(min(args.steps, 20)):
        ctx.acquire_req(req)
        for k, v in ctx.last_times().items():
            stage[k].append(v)
    ctx.set_profiling(False)
    form = ctx.last_search_form()
    corr_ms = statistics.mean(stage["corr"])
    alg_bytes = (B_CB if prec == U.PREC_FP32 else 4) * w.cells * w.n_blocks
    achieved = alg_bytes / (corr_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles(args.workload)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {
            "workload": {"L1CA_WIDE": "GPS L1 C/A wide search: int16 IQ @ 16.368 MHz, 1 ms x 10 non-coherent, "
                                      "+-5 kHz / 250 Hz (41 bins), 32 PRNs",
                         "L1CA_1MS": "GPS L1 C/A: int8 IQ @ 4.092 MHz, 1 ms x 1, +-5 kHz / 500 Hz, 32 PRNs",
                         "L1CA_4MS": "GPS L1 C/A: int8 IQ @ 4.092 MHz, 4 ms coherent folded to 1 ms, 32 PRNs",
                         "BOC_4MS": "BOC(1,1) 4092 chips, int16 IQ @ 20.46 MHz, 4 ms x 5, +-6 kHz / 125 Hz, "
                                    "36 PRNs"}[args.workload],
            "prns": len(w.codes),
            "doppler_bins": len(w.doppler),
            "n_fft": w.n_fft,
            "noncoherent_blocks": w.n_blocks,
            "surface_len": w.surface_len,
            "cells_per_search": w.cells,
            "parallelism": "single GPU",
        },
        "latency_ms": {"p50": 1e3 * percentile(lat, 0.5), "p99": 1e3 * percentile(lat, 0.99)},
        "stage_ms": {k: statistics.mean(v) for k, v in stage.items()},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": corr_kernel_name(form, w.surface_len, chosen, args.precision),
            "algorithmic_bytes_per_launch": alg_bytes,
        },
        "search_form": form,
        "detected_prns": [i for i, r in enumerate(res) if r["detected"]],
        # thread variants of the transform of the whole row (sample replicas)
        "autotune": {"variants": variants, "ms_per_launch": tune_ms, "chosen": chosen},
        "graph_replays_in_timed_steps": replays,
    }
    log(f"headline {value:.4g} cells/s, {ms_per_step:.3f} ms/search, corr kernel {corr_ms:.3f} ms")
    if not args.no_extra:
        log("other workloads")
        try:
            out["other_workloads"] = extra_workloads(ctx)
            out["other_workloads"].update(recorded_workloads(ctx))
            out["other_workloads"].update(jse_workload(ctx))
        except Exception as e:  # the headline line must still print
            out["other_workloads"] = {"error": repr(e)}
    if args.serve_requests > 0:
        log(f"serving {args.serve_requests} requests at {args.serve_rate} req/s")
        try:
            out["serving"] = serving(ctx, args.serve_requests, args.serve_rate)
        except Exception as e:
            out["serving"] = {"error": repr(e)}
        log("serving again with the BOC requests in the reduced form (FP16)")
        try:
            out["serving_boc_fp16"] = serving(ctx, args.serve_requests, args.serve_rate, 1)
        except Exception as e:
            out["serving_boc_fp16"] = {"error": repr(e)}
        log("serving the stream with L1 / L5 as 1 ms x T non-coherent blocks")
        try:
            out["serving_noncoherent_1ms"] = serving(ctx, args.serve_requests, args.serve_rate, 0, False)
        except Exception as e:
            out["serving_noncoherent_1ms"] = {"error": repr(e)}
        log("serving the first stream again with its rounds as captured, replayed hipGraphs")
        try:
            out["serving_round_graphs"] = serving(ctx, args.serve_requests, args.serve_rate, 0, True, True)
        except Exception as e:
            out["serving_round_graphs"] = {"error": repr(e)}
    if not args.no_cpu:
        threads = cpu_pool_threads(args.cpu_threads)
        host = host_cpu()
        log(f"cpu baseline on {threads} threads ({host['model']}, hardware_concurrency {host['hardware_concurrency']})")
        cpu = cpu_search_baseline(w, threads)
        cpu_res = cpu.pop("results")
        cpu["host"] = host
        out["cpu_baseline"] = cpu
        out["speedup_vs_cpu"] = value / cpu["value"]
        if cpu_res is not None:
            agree = all(cpu_res[i]["detected"] == res[i]["detected"] and
                        (not res[i]["detected"] or (cpu_res[i]["index"] == res[i]["index"] and
                                                    cpu_res[i]["doppler_bin"] == res[i]["doppler_bin"]))
                        for i in range(len(cpu_res)))
            out["cpu_gpu_same_detections"] = agree
        out["cpu_fft_vs_pocketfft"] = fft_calibration()
        if not args.no_extra:
            # the other configurations' CPU figures beside their GPU ones
            ow = out.get("other_workloads", {})
            base = {}
            for name in ("L1CA_1MS", "L1CA_4MS", "BOC_4MS"):
                if name == args.workload:
                    continue
                log(f"cpu baseline {name}")
                c = cpu_search_baseline(W.CONFIGS[name](), threads)
                c.pop("results")
                g = ow.get(name, {}).get("cells_per_s")
                c["speedup_gpu_vs_cpu"] = g / c["value"] if g else None
                base[name] = c
            log("cpu baseline BANK_L5")
            c = cpu_bank_baseline(W.bank_l5(), threads)
            g = ow.get("BANK_L5", {}).get("channel_samples_per_s")
            c["speedup_gpu_vs_cpu"] = g / c["value"] if g else None
            base["BANK_L5"] = c
            out["cpu_baselines_other"] = base
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
