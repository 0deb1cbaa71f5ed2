"""DFE channel step on the GPU, through the C ABI (ugsdr_hip_dfe_*).

* Each call against the oracle's chain of Channel::GetSeveralEpochs
  (src/dfe/dfe.hpp:192-203): decode, Mixer::Translate member overload
  (mixer.hpp:53-60, the phase moves on by 2 pi frac(f/1000) once per call,
  whatever the call's length), Resampler::Transform; output within 1e-5 of the
  largest sample (FP32 data, FP64 tone phase on both sides) and the running
  phase equal to the oracle's bit for bit.
* 1 ms calls are one continuous mix: four 1 ms calls equal one 4 ms call.
* Chained into the acquisition on the device: the DFE output stays in HBM
  and feeds ugsdr_hip_acquire (iq_on_device), which finds the recording's
  strong satellites (PRNs 9 and 23).
* Refusals: rates that are not whole kHz, a broken BbpDdc ms counter.
"""
import numpy as np
import pytest

import ugsdr_amd as U

from conftest import DeviceArray
from test_formats import BBP_FS, BBP_L1

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_dfe_calls_match_oracle(ctx, oracle):
    fs_in, fs_out, f, ph0 = 16.368e6, 4.092e6, -1.2345678e6, 0.7
    spm = int(fs_in / 1e3)
    rng = np.random.default_rng(5)
    n_ms = 6
    t = np.arange(n_ms * spm) / fs_in
    sig = 900 * np.exp(2j * np.pi * 1.2345678e6 * t) + rng.normal(0, 300, t.size) + 1j * rng.normal(0, 300, t.size)
    raw = np.empty(2 * t.size, np.int16)
    raw[0::2], raw[1::2] = np.round(sig.real), np.round(sig.imag)
    say("case DFE CI16 16.368 -> 4.092 MHz, f = -1.2345678 MHz, 3 calls of 2 ms")
    C_phase = ph0
    with U.Dfe(ctx, U.FMT_CI16, fs_in, fs_out, f, ph0) as ch:
        for call in range(3):
            chunk = raw[2 * call * 2 * spm: 2 * (call + 1) * 2 * spm]
            got = ch.process(chunk, 2)
            x = (chunk[0::2].astype(np.float32) + 1j * chunk[1::2].astype(np.float32)).astype(np.complex64)
            x, C_phase = oracle.mixer_member_translate(x, fs_in, f, C_phase)
            ref = oracle.resample(x, int(fs_out), int(fs_in))
            assert got.shape == ref.shape
            err = np.max(np.abs(got - ref))
            say(f"  call {call}: max |err| {err:.3g} of {np.max(np.abs(ref)):.3g}")
            assert err <= 1e-5 * np.max(np.abs(ref))
            assert ch.phase == C_phase
        # the tone at +1.2345678 MHz is moved to DC: the output is mostly its mean
        assert abs(np.mean(got)) > 0.5 * 900


def test_dfe_one_ms_calls_are_continuous(ctx):
    raw = np.fromfile(BBP_L1, np.uint8)
    eb = U.format_bytes(U.FMT_BBP_DDC, 33250, BBP_FS)
    f = 1.2e6 + 250.5                    # not a whole number of kHz: every call moves the phase
    say("case DFE BbpDdc 33.25 -> 4.2 MHz, 4 x 1 ms vs 1 x 4 ms")
    with U.Dfe(ctx, U.FMT_BBP_DDC, BBP_FS, 4.2e6, f) as a, U.Dfe(ctx, U.FMT_BBP_DDC, BBP_FS, 4.2e6, f) as b:
        parts = [a.process(raw[e * eb:(e + 1) * eb], 1) for e in range(4)]
        whole = b.process(raw, 4)
    got = np.concatenate(parts)
    assert np.max(np.abs(got - whole)) <= 1e-5 * np.max(np.abs(whole))


def test_dfe_into_acquisition_on_device(ctx):
    raw = np.fromfile(BBP_L1, np.uint8)
    fs_out = 4.2e6
    N = int(fs_out / 1e3)
    say("case DFE output in HBM -> ugsdr_hip_acquire (iq_on_device), 32 PRN x 41 bins x 4 ms")
    buf = DeviceArray(np.zeros(4 * N, np.complex64))
    keys = [ctx.replica_put(U.gps_l1ca(s), 1, N) for s in range(32)]
    try:
        with U.Dfe(ctx, U.FMT_BBP_DDC, BBP_FS, fs_out, 0.0) as ch:
            assert ch.process(raw, 4, out_device_ptr=buf.ptr.value) is None
        res, _, _ = ctx.acquire(None, U.FMT_CF32, fs_in=fs_out, fs_acq=fs_out, n_fft=N, n_blocks=4, surface_len=N,
                                doppler=U.doppler_bins(5000, 250), keys=keys, n_samples=4 * N,
                                iq_device_ptr=buf.ptr.value)
        direct, _, _ = ctx.acquire(raw, U.FMT_BBP_DDC, n_samples=4 * 33250, fs_in=BBP_FS, fs_acq=fs_out, n_fft=N,
                                   n_blocks=4, surface_len=N, doppler=U.doppler_bins(5000, 250), keys=keys)
    finally:
        for k in keys:
            ctx.replica_drop(k)
        buf.close()
    det = {p for p, r in enumerate(res) if r["detected"]}
    say(f"  detected sv {sorted(det)}")
    assert {8, 22} <= det
    # the same front end runs inside the fused acquisition: identical results,
    # except that code_offset counts samples of each request's own input rate
    # (ratio fs_in / fs_acq times the index, src/acquisition/fse.hpp:262)
    for a, b in zip(res, direct):
        assert {k: v for k, v in a.items() if k != "code_offset"} == {k: v for k, v in b.items() if k != "code_offset"}
        assert b["code_offset"] == a["code_offset"] * (BBP_FS / fs_out)


def test_dfe_refusals(ctx):
    with pytest.raises(U.UgsdrError):
        U.Dfe(ctx, U.FMT_CI16, 16.3685e6, 4.092e6, 0.0)
    raw = np.fromfile(BBP_L1, np.uint8).copy()
    eb = U.format_bytes(U.FMT_BBP_DDC, 33250, BBP_FS)
    raw[3 * eb] ^= 4                      # epoch 3's counter jumps
    with U.Dfe(ctx, U.FMT_BBP_DDC, BBP_FS, 4.2e6, 0.0) as ch:
        with pytest.raises(U.UgsdrError) as e:
            ch.process(raw, 4)
        assert e.value.code == -1
