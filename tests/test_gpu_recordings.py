"""The reference's iq_16_plus_16 and GPS L5 recordings through the fused
acquisition on the GPU, against the oracle (parity rules of tests/parity.py).

* iq_16_plus_16 (4 MHz int16, 1575.42 MHz): the first sample decodes to
  (-150, 16) on the device (test/integration_tests.cpp:105-121); the GPS L1
  C/A search of the reference's acquisition test (:401-404, +-15 kHz / 200 Hz,
  4 ms at the DefaultFseConfig rule's 4 MHz) finds the oracle's satellites.
* bbp_ddc_gps_L5 (66.5 MHz 2+2-bit, 1176.45 MHz): AcquireL5 with
  coherent=false (src/acquisition/fse.hpp:302-325): decode and the /4 boxcar
  to 16.625 MHz in the front kernel, N = 66500 (runtime four-step plan), the
  complex 4 -> 1 ms fold before |.|, for L5I and L5Q codes.
"""
import os

import numpy as np
import pytest

import ugsdr_amd as U

from parity import compare
from test_recordings import BBP_L5, IQ16, IQ16_SVS, L5_FS, L5_SVS

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_iq16_recording(ctx, oracle):
    say("case iq_16_plus_16 recording: first sample + 32 PRN x 151 bins, N=16000 fold 4000")
    raw = np.fromfile(IQ16, np.int16)
    first = ctx.unpack(raw, U.FMT_CI16, 4, 4e6)
    assert first[0] == complex(-150, 16)
    fs, spm, n = 4e6, 4000, 16000
    keys = [ctx.replica_put(U.gps_l1ca(s), 4, n) for s in range(32)]
    dop = np.arange(-15000, 15001, 200.0)
    gpu, _, _ = ctx.acquire(raw, U.FMT_CI16, fs_in=fs, fs_acq=fs, n_fft=n, n_blocks=1, surface_len=spm,
                            doppler=dop, keys=keys)
    x = (raw[0::2].astype(np.float32) + 1j * raw[1::2].astype(np.float32)).astype(np.complex64)
    reps = np.stack([oracle.repeat_upsample(U.gps_l1ca(s), 4, n) for s in range(32)])
    ref, _, _ = oracle.acquire(x, reps, n_fft=n, n_blocks=1, surface_len=spm, fs_in=fs, fs_acq=fs, doppler=dop,
                               precision=64)
    fails, notes = compare(gpu, ref)
    say(f"iq16: notes {notes}")
    assert not fails, fails
    assert [i for i, r in enumerate(gpu) if r["detected"]] == IQ16_SVS
    for k in keys:
        ctx.replica_drop(k)


@pytest.mark.parametrize("signal", ["l5i", "l5q"])
def test_l5_recording(ctx, oracle, signal):
    say(f"case bbp_ddc_gps_L5 recording {signal}: 32 PRN x 51 bins, 66.5 -> 16.625 MHz, N=66500, complex fold")
    raw = np.fromfile(BBP_L5, np.uint8)
    fa = U.adjust_sampling_rate(L5_FS, 20.46e6)
    spm_in, spm, n = 66500, 16625, 66500
    gen = getattr(U, f"gps_{signal}")
    keys = [ctx.replica_put(gen(s), 4, n) for s in range(32)]
    dop = np.arange(-5000, 5001, 200.0)
    gpu, _, _ = ctx.acquire(raw, U.FMT_BBP_DDC, n_samples=4 * spm_in, fs_in=L5_FS, fs_acq=fa, n_fft=n,
                            n_blocks=1, surface_len=spm, doppler=dop, keys=keys, coherent=False)
    x = oracle.unpack_bbp(raw, 4 * spm_in, L5_FS)
    y = oracle.resample(x, int(fa), int(L5_FS))
    reps = np.stack([oracle.repeat_upsample(gen(s), 4, n) for s in range(32)])
    ref, _, _ = oracle.acquire(y, reps, n_fft=n, n_blocks=1, surface_len=spm, fs_in=L5_FS, fs_acq=fa, doppler=dop,
                               coherent=False, precision=64)
    fails, notes = compare(gpu, ref)
    say(f"L5 {signal}: notes {notes}")
    assert not fails, fails
    assert [i for i, r in enumerate(gpu) if r["detected"]] == L5_SVS
    for k in keys:
        ctx.replica_drop(k)
