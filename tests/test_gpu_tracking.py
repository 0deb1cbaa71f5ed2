"""GPU-resident tracking loops (ugsdr_hip_track) through the C ABI.

* Parity with the oracle (refcpu_track, Tracker::TrackSingleSatellite with
  Pll / Dll): eight channels on one signal holding eight satellites, every
  epoch's early / prompt / late within 2e-4 of the prompt's scale and the
  loop trace (carrier frequency, code phase, carrier phase) within 1e-3 Hz,
  1e-3 samples and 1e-5 cycles.  The correlations are FP32 sums in another
  order on each side; the loop state is FP64 on both.
* The device loops lock from acquisition-size errors like the oracle's
  (tests/test_tracking.py: carrier pulled in, prompt in I, early and late at
  3/4 of the prompt, code phase in [delay, delay + 1)).
* The hand-off: a fused acquisition of the same signal seeds the states
  (code_phase = code_offset, carrier = Doppler bin; TrackingParameters'
  constructor, tracking_parameters.hpp:118-124) and every detected
  satellite locks.
"""
import numpy as np
import pytest

import ugsdr_amd as U
from oracle import refcpu as R

from test_tracking import FS, SPM, code3_table, l1_signal

pytestmark = pytest.mark.gpu

SATS = [(0, 1500.3, 1234.0), (3, -2210.7, 77.0), (5, 3120.2, 4000.0), (8, -450.9, 2048.0),
        (11, 890.1, 15.0), (14, -3999.4, 3071.0), (19, 2501.6, 512.0), (27, -1275.2, 3600.0)]


def say(msg):
    print(msg, flush=True)


def multi_signal(n_ms, seed=9):
    x = np.zeros(n_ms * SPM, np.complex128)
    for i, (sv, f, tau) in enumerate(SATS):
        x += l1_signal(sv, n_ms, f, tau, phase0=0.4 * i, amp=1.0, sigma=0.0)
    rng = np.random.default_rng(seed)
    x += 1.5 * (rng.normal(size=x.size) + 1j * rng.normal(size=x.size))
    return x.astype(np.complex64)


def test_track_matches_oracle(ctx, oracle):
    n_ms = 300
    x = multi_signal(n_ms)
    errs = [(30.0, 0.7), (-45.0, -1.1), (12.0, 0.2), (-20.0, 1.4), (50.0, -0.5), (-8.0, 0.9), (35.0, 1.1),
            (-25.0, -0.3)]
    states = [U.track_state(tau + dc, f + df) for (sv, f, tau), (df, dc) in zip(SATS, errs)]
    codes = [U.gps_l1ca(sv) for sv, _, _ in SATS]
    say(f"case track 8 channels x {n_ms} epochs @ 4.092 MHz vs oracle")
    iq = x.view(np.float32)
    epl, tr = ctx.track(iq, U.FMT_CF32, spm=SPM, n_blocks=n_ms, fs=FS, states=states, codes=codes)
    for c, ((sv, f, tau), (df, dc)) in enumerate(zip(SATS, errs)):
        st = R.track_state(tau + dc, f + df)
        repl, rtr = R.track(x, SPM, n_ms, FS, code3_table(sv), st)
        scale = np.max(np.abs(repl[:, 1]))
        e_epl = np.max(np.abs(epl[:, c, :] - repl)) / scale
        e_f = np.max(np.abs(tr[:, c, 1] - rtr[:, 1]))
        e_cp = np.max(np.abs(tr[:, c, 2] - rtr[:, 2]))
        e_ph = np.max(np.abs(tr[:, c, 0] - rtr[:, 0]))
        say(f"  sv {sv}: epl {e_epl:.2e}, carrier {e_f:.2e} Hz, code {e_cp:.2e} samples, phase {e_ph:.2e} cycles")
        assert e_epl < 2e-4 and e_f < 1e-3 and e_cp < 1e-3 and e_ph < 1e-5
        assert states[c].epochs == n_ms and abs(states[c].carrier_frequency - st.carrier_frequency) < 1e-3


def test_track_locks_like_oracle(ctx):
    n_ms = 1000
    sv, f0, tau = 4, 1500.3, 1234.0
    x = l1_signal(sv, n_ms, f0, tau)
    starts = [(40.0, 0.6), (-60.0, -1.2), (25.0, 1.6)]
    states = [U.track_state(tau + dc, f0 + df) for df, dc in starts]
    say("case track lock from acquisition-size errors, 3 channels x 1000 epochs")
    epl, tr = ctx.track(x.view(np.float32), U.FMT_CF32, spm=SPM, n_blocks=n_ms, fs=FS, states=states,
                        codes=[U.gps_l1ca(sv)] * 3)
    for c, st in enumerate(states):
        P = epl[-200:, c, 1]
        assert abs(tr[-1, c, 1] - f0) < 3.0
        assert np.mean(np.abs(P.imag)) < 0.15 * np.mean(np.abs(P.real))
        assert abs(np.mean(np.abs(P)) / SPM - 1) < 0.01
        for tap in (0, 2):
            assert abs(np.mean(np.abs(epl[-200:, c, tap])) / np.mean(np.abs(P)) - 0.75) < 0.02
        assert tau - 0.25 <= st.code_phase <= tau + 1.25


def test_acquisition_hands_off_to_tracking(ctx):
    n_ms = 1000
    x = multi_signal(n_ms, seed=21)
    keys = [ctx.replica_put(U.gps_l1ca(s), 1, SPM) for s in range(32)]
    say("case acquisition (4 ms, 32 PRN, 125 Hz bins) -> tracking hand-off -> 1000 epochs")
    try:
        res, _, _ = ctx.acquire(x[:4 * SPM].view(np.float32), U.FMT_CF32, n_samples=4 * SPM, fs_in=FS, fs_acq=FS,
                                n_fft=SPM, n_blocks=4, surface_len=SPM, doppler=U.doppler_bins(5000, 125),
                                keys=keys)
    finally:
        for k in keys:
            ctx.replica_drop(k)
    found = [p for p, r in enumerate(res) if r["detected"]]
    assert sorted(found) == sorted(sv for sv, _, _ in SATS)
    states = [U.track_state(res[p]["code_offset"], res[p]["doppler"]) for p in found]
    epl, tr = ctx.track(x.view(np.float32), U.FMT_CF32, spm=SPM, n_blocks=n_ms, fs=FS, states=states,
                        codes=[U.gps_l1ca(p) for p in found])
    truth = {sv: (f, tau) for sv, f, tau in SATS}
    for c, p in enumerate(found):
        f0, tau = truth[p]
        P = epl[-200:, c, 1]
        say(f"  sv {p}: acq {res[p]['doppler']:.0f} Hz / {res[p]['code_offset']:.0f} -> "
            f"{tr[-1, c, 1]:.1f} Hz / {tr[-1, c, 2]:.2f}")
        assert abs(tr[-1, c, 1] - f0) < 5.0
        assert np.mean(np.abs(P.imag)) < 0.2 * np.mean(np.abs(P.real))
        assert tau - 0.5 <= tr[-1, c, 2] <= tau + 1.5


@pytest.mark.parametrize("fmt,fs", [(U.FMT_CI16, 4.093e6), (U.FMT_CI8, 4.094e6), (U.FMT_R8, 4.095e6)])
def test_track_formats_and_ragged_epochs(ctx, oracle, fmt, fs):
    """integer sample formats (the compiled CI8 / CI16 loaders and the
    run-time decoder behind R8) at rates whose epochs are not whole groups of
    four samples (4093, 4094, 4095 per ms): the same bar against the oracle
    fed the decoded samples"""
    spm, n_ms = int(fs / 1e3), 200
    sv, f0, tau = 6, -1730.4, 2900.0
    x = l1_signal(sv, n_ms, f0, tau, fs=fs)
    if fmt == U.FMT_CI16:
        raw = np.empty(2 * x.size, np.int16)
        raw[0::2], raw[1::2] = np.round(100 * x.real), np.round(100 * x.imag)
        dec = (raw[0::2] + 1j * raw[1::2]).astype(np.complex64)
    elif fmt == U.FMT_CI8:
        raw = np.empty(2 * x.size, np.int8)
        raw[0::2], raw[1::2] = np.clip(np.round(30 * x.real), -127, 127), np.clip(np.round(30 * x.imag), -127, 127)
        dec = (raw[0::2].astype(np.float32) + 1j * raw[1::2].astype(np.float32)).astype(np.complex64)
    else:
        raw = np.clip(np.round(30 * x.real), -127, 127).astype(np.int8)
        dec = raw.astype(np.float32).astype(np.complex64)
    starts = [(30.0, 0.4), (-45.0, -0.8), (10.0, 1.3)]
    states = [U.track_state(tau + dc, f0 + df) for df, dc in starts]
    say(f"case track fmt {fmt}, {spm} samples per epoch, 3 channels x {n_ms} epochs vs oracle")
    epl, tr = ctx.track(raw, fmt, spm=spm, n_blocks=n_ms, fs=fs, states=states, codes=[U.gps_l1ca(sv)] * 3)
    for c, (df, dc) in enumerate(starts):
        st = R.track_state(tau + dc, f0 + df)
        repl, rtr = R.track(dec, spm, n_ms, fs, code3_table(sv, spm), st)
        scale = np.max(np.abs(repl[:, 1]))
        e_epl = np.max(np.abs(epl[:, c, :] - repl)) / scale
        e_f = np.max(np.abs(tr[:, c, 1] - rtr[:, 1]))
        e_cp = np.max(np.abs(tr[:, c, 2] - rtr[:, 2]))
        say(f"  channel {c}: epl {e_epl:.2e}, carrier {e_f:.2e} Hz, code {e_cp:.2e} samples")
        assert e_epl < 2e-4 and e_f < 1e-3 and e_cp < 1e-3


def test_track_four_ms_code_matches_oracle(ctx, oracle):
    """a 4 ms code period (Galileo E1-like BOC(1,1), 8184 half-chips: the
    epoch's code phase steps back by a ms of samples each epoch within the
    period, tracker.hpp:165) and a wider early/late spacing of 0.5 chip"""
    from test_galileo import e1_chips

    spm, n_ms, fs = SPM, 240, FS
    rng = np.random.default_rng(44)
    sats = [(e1_chips(201), 700.4, 5000.0), (e1_chips(202), -1900.8, 12345.0)]
    k = np.arange(n_ms * spm)
    x = 1.2 * (rng.normal(size=k.size) + 1j * rng.normal(size=k.size))
    for chips, f, tau in sats:
        idx = np.floor((k - tau) * 2.046e6 / fs).astype(np.int64) % 8184
        x = x + chips[idx].astype(np.float64) * np.exp(2j * np.pi * f * k / fs + 0.7j)
    x = x.astype(np.complex64)
    states = [U.track_state(tau + 0.6, f + 25.0, code_period_ms=4) for _, f, tau in sats]
    say(f"case track 4 ms BOC codes, spacing 0.5, 2 channels x {n_ms} epochs vs oracle")
    epl, tr = ctx.track(x.view(np.float32), U.FMT_CF32, spm=spm, n_blocks=n_ms, fs=fs, states=states,
                        codes=[c for c, _, _ in sats], spacing=0.5)
    for c, (chips, f, tau) in enumerate(sats):
        st = R.track_state(tau + 0.6, f + 25.0, code_period_ms=4)
        repl, rtr = R.track(x, spm, n_ms, fs, R.repeat_upsample(chips, 3, 3 * 4 * spm), st, spacing=0.5)
        scale = np.max(np.abs(repl[:, 1]))
        e_epl = np.max(np.abs(epl[:, c, :] - repl)) / scale
        e_f = np.max(np.abs(tr[:, c, 1] - rtr[:, 1]))
        e_cp = np.max(np.abs(tr[:, c, 2] - rtr[:, 2]))
        say(f"  channel {c}: epl {e_epl:.2e}, carrier {e_f:.2e} Hz, code {e_cp:.2e} samples, "
            f"final {tr[-1, c, 1]:.1f} Hz")
        assert e_epl < 2e-4 and e_f < 1e-3 and e_cp < 1e-3
        assert states[c].epochs == n_ms


@pytest.mark.parametrize("n_ch", [300, 600])
def test_track_more_channels_than_cus(ctx, oracle, n_ch):
    """more channels than the GPU's CUs (the two- and four-channels-per-CU
    workgroup shapes) over 120 epochs; every 37th channel against the oracle"""
    n_ms = 120
    x = multi_signal(n_ms, seed=31)
    rng = np.random.default_rng(5)
    chans = []
    for i in range(n_ch):
        sv, f, tau = SATS[i % len(SATS)]
        chans.append((sv, f + rng.uniform(-40, 40), tau + rng.uniform(-0.8, 0.8)))
    states = [U.track_state(tau, f) for _, f, tau in chans]
    say(f"case track {n_ch} channels x {n_ms} epochs (more channels than CUs)")
    epl, tr = ctx.track(x.view(np.float32), U.FMT_CF32, spm=SPM, n_blocks=n_ms, fs=FS, states=states,
                        codes=[U.gps_l1ca(sv) for sv, _, _ in chans])
    for c in range(0, n_ch, 37):
        sv, f, tau = chans[c]
        st = R.track_state(tau, f)
        repl, rtr = R.track(x, SPM, n_ms, FS, code3_table(sv), st)
        scale = np.max(np.abs(repl[:, 1]))
        assert np.max(np.abs(epl[:, c, :] - repl)) / scale < 2e-4
        assert np.max(np.abs(tr[:, c, 1] - rtr[:, 1])) < 1e-3
        assert np.max(np.abs(tr[:, c, 2] - rtr[:, 2])) < 1e-3


def test_dfe_output_tracked_on_device(ctx):
    """the DFE channel's output left in device memory feeds the loops
    directly (iq_on_device): the same result as tracking its host copy"""
    from conftest import DeviceArray
    from test_formats import BBP_FS, BBP_L1

    raw = np.fromfile(BBP_L1, np.uint8)
    fs, n_ms = 4.2e6, 4
    spm = int(fs / 1e3)
    buf = DeviceArray(np.zeros(n_ms * spm, np.complex64))
    say("case DFE (BbpDdc recording) -> device buffer -> tracking loops")
    try:
        with U.Dfe(ctx, U.FMT_BBP_DDC, BBP_FS, fs, 0.0) as ch:
            ch.process(raw, n_ms, out_device_ptr=buf.ptr.value)
        host = buf.get()
        starts = [(1375.0, 1000.0), (-625.0, 3000.0)]
        s_dev = [U.track_state(cp, f) for f, cp in starts]
        s_host = [U.track_state(cp, f) for f, cp in starts]
        codes = [U.gps_l1ca(22), U.gps_l1ca(8)]
        epl_d, tr_d = ctx.track(None, U.FMT_CF32, spm=spm, n_blocks=n_ms, fs=fs, states=s_dev, codes=codes,
                                iq_device_ptr=buf.ptr.value)
        epl_h, tr_h = ctx.track(host.view(np.float32), U.FMT_CF32, spm=spm, n_blocks=n_ms, fs=fs, states=s_host,
                                codes=codes)
    finally:
        buf.close()
    assert np.array_equal(epl_d, epl_h) and np.array_equal(tr_d, tr_h)
