"""The oracle's tracking loop (refcpu_track: Tracker::TrackSingleSatellite with
TrackingParameters::Pll / Dll, src/tracking/tracker.hpp:156-177,
tracking_parameters.hpp:495-528), pinned by what a working loop must do on a
synthetic GPS L1 C/A signal of known Doppler and delay, not by its formulas:

* started off by an acquisition-size error (tens of Hz, about a chip), the
  PLL/FLL pulls the carrier to the true Doppler and locks the phase: the
  prompt's energy ends up in I;
* the prompt reaches the full-epoch correlation, and early / late the
  triangle's value a quarter chip off the peak (3/4 of the prompt);
* the DLL settles the code phase into [delay, delay + 1): the replica of
  CorrelateSplit is shifted by whole samples (ceil of the phase,
  tracking_parameters.hpp:480-483), so early minus late is zero across the
  sample that holds the true delay;
* a loop whose frequency update is reversed (planted in the test, on the
  oracle's state between epochs) does not lock.
"""
import numpy as np
import pytest

from oracle import refcpu as R

FS, SPM = 4.092e6, 4092


def l1_signal(sv, n_ms, doppler, delay, phase0=0.3, amp=1.0, sigma=1.0, seed=1, fs=FS):
    """x[k] = amp c((k - delay) 1.023e6 / fs) exp(j (2 pi doppler k / fs + phase0)) + noise"""
    n = n_ms * int(fs / 1e3)
    code = R.gps_l1ca(sv).astype(np.float64)
    k = np.arange(n)
    chip = np.floor((k - delay) * 1.023e6 / fs).astype(np.int64) % 1023
    rng = np.random.default_rng(seed)
    x = amp * code[chip] * np.exp(1j * (2 * np.pi * doppler * k / fs + phase0))
    x = x + sigma * (rng.normal(size=n) + 1j * rng.normal(size=n))
    return x.astype(np.complex64)


def code3_table(sv, spm=SPM):
    return R.repeat_upsample(R.gps_l1ca(sv), 3, 3 * spm)


@pytest.mark.parametrize("df,dcp", [(40.0, 0.6), (-60.0, -1.2), (25.0, 1.6)])
def test_oracle_loop_locks(oracle, df, dcp):
    sv, f0, tau = 4, 1500.3, 1234.0
    x = l1_signal(sv, 1000, f0, tau)
    st = R.track_state(tau + dcp, f0 + df)
    epl, tr = R.track(x, SPM, 1000, FS, code3_table(sv), st)
    P = epl[-200:, 1]
    assert abs(tr[-1, 1] - f0) < 3.0                                    # carrier pulled in
    assert np.mean(np.abs(P.imag)) < 0.15 * np.mean(np.abs(P.real))      # phase lock
    assert abs(np.mean(np.abs(P)) / SPM - 1) < 0.01                     # full correlation
    for tap in (0, 2):
        assert abs(np.mean(np.abs(epl[-200:, tap])) / np.mean(np.abs(P)) - 0.75) < 0.02
    assert tau - 0.25 <= st.code_phase <= tau + 1.25
    assert st.epochs == 1000


def test_oracle_state_bookkeeping(oracle):
    """UpdatePhase and the PLL increments add up: after n epochs the carrier
    phase is the sum of 2 pi f / 1000 over the epochs plus pi/2 times the
    sum of the phase discriminators (the trace records -phase / 2 pi)"""
    sv = 7
    x = l1_signal(sv, 50, -2200.0, 300.0, seed=3)
    st = R.track_state(300.4, -2210.0, carrier_phase=0.25)
    epl, tr = R.track(x, SPM, 50, FS, code3_table(sv), st)
    freqs = np.concatenate([[-2210.0], tr[:-1, 1]])
    # discriminator of each prompt: atan(Q / I) / 2 pi (I, Q in float)
    q = (epl[:, 1].imag / epl[:, 1].real).astype(np.float32)
    e = np.arctan(q.astype(np.float64)) / (2 * np.pi)
    phase = 0.25 + np.cumsum(2 * np.pi * freqs / 1000.0 + e * np.pi / 2)
    assert np.allclose(-tr[:, 0] * 2 * np.pi, phase, rtol=0, atol=1e-9)


def test_flipped_pll_sign_does_not_lock(oracle):
    """the same loop with its frequency update reversed after every epoch
    (done here on the oracle's state): it walks away from the carrier"""
    sv, f0, tau = 4, 1500.3, 1234.0
    x = l1_signal(sv, 300, f0, tau)
    code3 = code3_table(sv)
    st = R.track_state(tau, f0 + 20.0)
    for b in range(300):
        saved_f = st.carrier_frequency
        R.track(x[b * SPM:(b + 1) * SPM], SPM, 1, FS, code3, st)
        st.carrier_frequency = saved_f - (st.carrier_frequency - saved_f)
    assert abs(st.carrier_frequency - f0) > 50.0
