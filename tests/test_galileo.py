"""Galileo E1 code hypotheses (FastSearchEngineBase::ProcessGalileo,
src/acquisition/fse.hpp:356-396): the oracle's selection, pinned by hand
tables and by planted sign patterns, not by its formulas.

* std::max_element over a satellite's hypotheses: the first of the largest
  level wins; the winner alone is thresholded (hand-made result tables).
* A noise-free 4 ms BOC(1,1) signal whose milliseconds carry one of the four
  sign permutations: the oracle's search picks that permutation, at the
  planted code phase, with the full coherent level 4 spm A (closed form);
  every other permutation scores lower.
* The package's replica builder (e1_sign_hypotheses) equals the oracle's
  upsample-then-sign construction.
"""
import numpy as np
import pytest

import ugsdr_amd as U
from oracle import refcpu as R

FS = 4.092e6
SPM = 4092
PERMS = [[1, 1, 1, 1], [1, 1, 1, -1], [1, 1, -1, -1], [1, -1, -1, -1]]


def e1_chips(seed):
    """a seeded 4092-chip primary code, BOC(1,1)-expanded to 8184 half-chips"""
    rng = np.random.default_rng(seed)
    prim = (2 * rng.integers(0, 2, 4092) - 1).astype(np.int8)
    return R.boc_expand(prim)


def e1_signal(sats, seed=0, sigma=0.0):
    """x[k] = sum A rep_h[(k - tau) mod 4 spm] exp(j (2 pi f k / fs + phi)) + noise,
    rep_h = the floor-index upsampled code with permutation h on its ms"""
    n = 4 * SPM
    k = np.arange(n)
    x = np.zeros(n, np.complex128)
    for chips, h, tau, f, amp, phi in sats:
        rep = R.repeat_upsample(chips, 1, n) * np.repeat(np.asarray(PERMS[h], np.float32), SPM)
        x += amp * np.roll(rep, tau) * np.exp(1j * (2 * np.pi * f * k / FS + phi))
    rng = np.random.default_rng(seed)
    x += sigma * (rng.normal(size=n) + 1j * rng.normal(size=n))
    return x.astype(np.complex64)


def _res(level, snr):
    return {"level": level, "snr": snr, "index": 0, "doppler_bin": 0, "detected": 0}


def test_reduce_hypotheses_hand_tables(oracle):
    runs = [_res(3, 4.0), _res(5, 2.0), _res(5, 9.0), _res(1, 9.0),      # tie: the first 5 wins, snr 2
            _res(2, 1.0), _res(2, 1.0), _res(2, 1.0), _res(2, 1.0),      # all equal: hypothesis 0
            _res(0, 0.0), _res(1, 0.5), _res(7, 3.4), _res(6, 9.0)]      # the largest level, snr 3.4
    out = R.reduce_hypotheses(runs, 4, 3.3)
    assert [r["hypothesis"] for r in out] == [1, 0, 2]
    assert [r["detected"] for r in out] == [0, 0, 1]


@pytest.mark.parametrize("h", [0, 1, 2, 3])
def test_oracle_picks_planted_permutation(oracle, h):
    chips = e1_chips(11)
    tau, amp = 5123, 0.75
    x = e1_signal([(chips, h, tau, 1500.0, amp, 0.4)])
    out = R.process_galileo(x, [chips], spm=SPM, fs_in=FS, fs_acq=FS, doppler=U.doppler_bins(3000, 250))
    r = out[0]
    assert r["hypothesis"] == h and r["index"] == tau and r["doppler"] == 1500.0
    assert abs(r["level"] - 4 * SPM * amp) < 1e-4 * 4 * SPM * amp
    # every other permutation on its own scores lower
    res, _, _ = R.acquire(x, np.stack([R.repeat_upsample(chips, 1, 4 * SPM) * np.repeat(np.asarray(p, np.float32), SPM)
                                       for p in PERMS]),
                          n_fft=4 * SPM, n_blocks=1, surface_len=4 * SPM, fs_in=FS, fs_acq=FS,
                          doppler=U.doppler_bins(3000, 250), reshape=False)
    for k, rk in enumerate(res):
        if k != h:
            assert rk["level"] < 0.8 * r["level"]


def test_replica_builder_matches_oracle_construction(oracle):
    chips = e1_chips(3)
    for n in (4 * SPM, 4 * 4200, 4 * 6650):
        got = U.e1_sign_hypotheses(chips, n)
        base = R.repeat_upsample(chips, 1, n)
        for k, p in enumerate(PERMS):
            assert np.array_equal(got[k], base * np.repeat(np.asarray(p, np.float32), n // 4))
