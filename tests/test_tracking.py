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
  oracle's state between epochs) does not lock;
* the loop gains (tracking_parameters.hpp:72-87) against the textbook PLL /
  FLL / DLL designs at the reference's design values: every epoch's update is
  predicted from the epoch's discriminators, and the PLL's closed-loop noise
  bandwidth from a phase-step response equals the linear model's.
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


# --------------------------------------------------------------- loop gains --
# Textbook loop designs (Kaplan & Hegarty, Understanding GPS, loop filter
# table: second-order PLL with omega_n = Bn / 0.53 and a2 = 1.414 (zeta 0.707);
# first-order FLL aid omega_f = Bf / 0.25; DLL omega_n = 8 zeta Bn / (4 zeta^2 + 1),
# tau1 = k / omega_n^2, tau2 = 2 zeta / omega_n) at the reference's design
# values: Bn 25 Hz, Bf 250 Hz, T 1 ms; DLL Bn 0.7 Hz, zeta 1, k 0.1.
T_LOOP = 1e-3
W_PLL = 25.0 / 0.53
K1_PLL = T_LOOP * W_PLL ** 2 + 1.414 * W_PLL
K2_PLL = 1.414 * W_PLL
K3_FLL = T_LOOP * 250.0 / 0.25
W_DLL = 8 * 1.0 * 0.7 / (4 * 1.0 ** 2 + 1)
TAU1, TAU2 = 0.1 / W_DLL ** 2, 2 * 1.0 / W_DLL
K1_DLL, K2_DLL = TAU2 / TAU1, T_LOOP / TAU1


def discriminators(epl):
    """the loops' inputs from the oracle's own E / P / L output, in the
    reference's float arithmetic: PLL atan(Q/I)/2pi, FLL atan2(cross, dot)/pi
    of consecutive prompts, DLL (|E| - |L|) / (|E| + |L|)"""
    P = epl[:, 1]
    I, Q = P.real.astype(np.float32), P.imag.astype(np.float32)
    e = np.arctan((Q / I).astype(np.float64)) / (2 * np.pi)
    Ip = np.concatenate([[np.float32(0)], I[:-1]])
    Qp = np.concatenate([[np.float32(0)], Q[:-1]])
    cross = (I * Qp - Ip * Q).astype(np.float32)
    dot = np.abs(I * Ip + Qp * Q).astype(np.float32)
    fe = np.arctan2(cross, dot).astype(np.float32).astype(np.float64) / np.pi
    Ea, La = np.abs(epl[:, 0]).astype(np.float32), np.abs(epl[:, 2]).astype(np.float32)
    d = ((Ea - La) / (Ea + La)).astype(np.float32).astype(np.float64)
    return e, fe, d


def test_loop_filter_gains_pinned_against_textbook_designs(oracle):
    """Over a noisy run, every epoch's update of the oracle's loops is
    predicted from its own discriminators by the textbook designs above: the
    carrier frequency step is K1 e_n - K2 e_(n-1) - K3 fe_n, the carrier phase
    gains (pi/2) e_n beyond 2 pi f T, and the code NCO step is
    (K1d (d_n - d_(n-1)) + K2d d_n) base/1.023e6.  A mistyped constant, a
    swapped pair or a dropped term leaves residuals far above the float
    rounding allowed here (discriminators of tens of millicycles move the
    carrier by tens of Hz)."""
    sv, f0, tau = 11, -1750.0, 2222.0
    x = l1_signal(sv, 400, f0, tau, amp=1.0, sigma=4.0, seed=21)
    st = R.track_state(tau + 0.4, f0 + 15.0)
    f_start, nco_start = st.carrier_frequency, st.code_nco
    epl, tr = R.track(x, SPM, 400, FS, code3_table(sv), st)
    e, fe, d = discriminators(epl)
    assert np.std(e) > 0.01 and np.std(d) > 0.01          # the loops see varied inputs
    f = np.concatenate([[f_start], tr[:, 1]])
    e_prev = np.concatenate([[0.0], e[:-1]])
    pred = K1_PLL * e - K2_PLL * e_prev - K3_FLL * fe
    assert np.max(np.abs(np.diff(f) - pred)) < 1e-4          # Hz
    # phase: psi_n = psi_(n-1) + 2 pi f_(n-1) T + (pi/2) e_n
    psi = np.concatenate([[0.0], -tr[:, 0] * 2 * np.pi])
    assert np.max(np.abs(np.diff(psi) - 2 * np.pi * f[:-1] * T_LOOP - np.pi / 2 * e)) < 1e-9
    # DLL: code_nco = base - code_frequency (base 1.023 MHz)
    nco = np.concatenate([[nco_start], 1.023e6 - tr[:, 3]])
    d_prev = np.concatenate([[0.0], d[:-1]])
    pred_d = K1_DLL * (d - d_prev) + K2_DLL * d
    assert np.max(np.abs(np.diff(nco) - pred_d)) < 1e-5      # chips/s


def closed_loop_bandwidth(h):
    """one-sided noise bandwidth of a unit-DC-gain discrete closed loop from
    its impulse response: sum h^2 / (2 T)"""
    return float(np.sum(h ** 2) / (2 * T_LOOP) / np.sum(h) ** 2)


def test_pll_closed_loop_noise_bandwidth(oracle):
    """A clean, locked signal takes a 0.005-cycle carrier phase step; the
    oracle loop's phase estimate is the step response.  Its noise bandwidth
    equals that of the linear loop built from the textbook gains with the
    reference's quarter-cycle phase feed and FLL aid (89 Hz, not the PI
    filter's own 25 Hz design: the PI part alone closes at 27.3 Hz in discrete
    time at T = 1 ms, and the phase feed (pi/2) e dominates)."""
    sv, f0, tau, ph0, n_ms, n0, step = 4, 1500.0, 1234.0, 0.3, 400, 50, 0.005
    n = n_ms * SPM
    code = R.gps_l1ca(sv).astype(np.float64)
    k = np.arange(n)
    chip = np.floor((k - tau) * 1.023e6 / FS).astype(np.int64) % 1023
    x = (code[chip] * np.exp(1j * (2 * np.pi * f0 * k / FS + ph0 + 2 * np.pi * step * (k >= n0 * SPM))))
    st = R.track_state(tau, f0, carrier_phase=ph0)
    epl, tr = R.track(x.astype(np.complex64), SPM, n_ms, FS, code3_table(sv), st)
    est = (-tr[:, 0] * 2 * np.pi - ph0 - 2 * np.pi * f0 * (np.arange(n_ms) + 1) / 1000.0) / (2 * np.pi)
    s = est[n0 - 1:] / step
    assert abs(s[0]) < 1e-3 and abs(s[1] - 0.25) < 1e-3 and abs(s[-1] - 1) < 2e-3
    b_oracle = closed_loop_bandwidth(np.diff(s))

    def model(feed, k3):
        th, fr, ep, out = 0.0, 0.0, 0.0, []
        for _ in range(4000):
            e = 1.0 - th - fr * T_LOOP / 2     # mean phase error over the epoch
            th += fr * T_LOOP
            fr += K1_PLL * e - K2_PLL * ep + k3 * 2 * (e - ep)
            th += feed * e
            ep = e
            out.append(th)
        return closed_loop_bandwidth(np.diff(np.concatenate([[0.0], out])))

    b_pi, b_full = model(0.0, 0.0), model(0.25, K3_FLL)
    print(f"PLL noise bandwidth: oracle {b_oracle:.2f} Hz, linear model {b_full:.2f} Hz, PI part alone {b_pi:.2f} Hz")
    assert abs(b_pi - 25.0) < 0.1 * 25.0
    assert abs(b_oracle - b_full) < 0.02 * b_full
