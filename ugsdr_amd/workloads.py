"""Synthetic workloads of the acquisition / correlator benchmarks.

Every input is synthesised deterministically (libugsdr_siggen.so: counter-based
xoshiro256** + FP64 Box-Muller), with a truth record per satellite:

  L1CA_1MS   GPS L1 C/A, int8 IQ @ 4.092 MHz, 1 ms x 1, +-5 kHz / 500 Hz (21 bins),
             32 PRNs, 6 satellites, C/N0 U[42,48] dB-Hz, sigma 8 LSB
  L1CA_4MS   same signal model, 4 ms coherent x 1 folded to 1 ms (the reference's
             own ms_to_process = 4 search)
  L1CA_WIDE  GPS L1 C/A, int16 IQ @ 16.368 MHz, 1 ms x 10 non-coherent,
             +-5 kHz / 250 Hz (41 bins), 32 PRNs, 8 satellites C/N0 U[38,48],
             20 ms navigation bits, sigma 512 LSB
  BOC_4MS    BOC(1,1) seeded 4092-chip codes, int16 IQ @ 20.46 MHz, 4 ms x 5,
             +-6 kHz / 125 Hz (97 bins), 36 PRNs, 8 satellites
  BANK_L5    64 channels x 5 taps (+-0.5, +-0.25, 0 chip), seeded 10230-chip codes,
             int16 IQ @ 40.92 MHz, 1000 x 1 ms blocks
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import FMT_CI8, FMT_CI16, HERE, doppler_bins

GEN_LIB_PATH = os.path.join(HERE, "libugsdr_siggen.so")
L1_HZ = 1575.42e6
L5_HZ = 1176.45e6


class SigSat(C.Structure):
    _fields_ = [
        ("code", C.POINTER(C.c_int8)), ("code_len", C.c_int64), ("chip_rate", C.c_double),
        ("carrier_hz", C.c_double), ("code_phase", C.c_double), ("doppler", C.c_double),
        ("carrier_phase", C.c_double), ("cn0_dbhz", C.c_double), ("nav_period_s", C.c_double),
        ("nav_offset_s", C.c_double), ("nav_seed", C.c_uint64),
    ]


_gen = None


def gen():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_LIB_PATH):
            raise RuntimeError(f"{GEN_LIB_PATH} not built: run python -m ugsdr_amd.build")
        g = C.CDLL(GEN_LIB_PATH)
        g.ugsdr_siggen_generate.argtypes = [C.POINTER(SigSat), C.c_int, C.c_double, C.c_int64, C.c_double,
                                            C.c_uint64, C.c_int, C.c_void_p, C.c_int]
        g.ugsdr_siggen_generate.restype = C.c_int
        g.ugsdr_siggen_random_code.argtypes = [C.c_uint64, C.c_int64, C.c_void_p]
        _gen = g
    return _gen


def random_code(seed: int, length: int) -> np.ndarray:
    out = np.zeros(length, np.int8)
    gen().ugsdr_siggen_random_code(seed, length, out.ctypes.data)
    return out


@dataclass
class Sat:
    prn: int                 # index into the searched code list
    code: np.ndarray         # +-1 chips as transmitted (BOC expanded if BOC)
    chip_rate: float
    carrier_hz: float
    code_phase: float        # chips of delay
    doppler: float
    carrier_phase: float     # cycles
    cn0: float
    nav_period_s: float = 0.0
    nav_offset_s: float = 0.0
    nav_seed: int = 0


def synthesize(sats: list[Sat], fs: float, n: int, sigma: float, seed: int, fmt_bits: int,
               threads: int | None = None) -> np.ndarray:
    """Interleaved I/Q array (int8 / int16 / float32) of n complex samples."""
    arr_t = (SigSat * max(1, len(sats)))()
    keep = []
    for i, s in enumerate(sats):
        code = np.ascontiguousarray(s.code, np.int8)
        keep.append(code)
        arr_t[i] = SigSat(code.ctypes.data_as(C.POINTER(C.c_int8)), len(code), s.chip_rate, s.carrier_hz,
                          s.code_phase, s.doppler, s.carrier_phase, s.cn0, s.nav_period_s, s.nav_offset_s,
                          s.nav_seed)
    dtype = {8: np.int8, 16: np.int16, 32: np.float32}[fmt_bits]
    out = np.zeros(2 * n, dtype)
    rc = gen().ugsdr_siggen_generate(arr_t, len(sats), fs, n, sigma, seed, fmt_bits, out.ctypes.data,
                                     threads or min(os.cpu_count() or 1, 16))
    if rc:
        raise RuntimeError("signal synthesis failed")
    del keep
    return out


@dataclass
class AcqWorkload:
    name: str
    fs: float
    fmt: int
    iq: np.ndarray                  # interleaved I/Q
    n_fft: int
    n_blocks: int
    surface_len: int
    reshape: bool
    doppler: np.ndarray
    codes: list                     # per searched PRN: +-1 chips (replica source)
    repeats: int                    # RepeatCodeNTimes factor for the replica
    sats: list = field(default_factory=list)
    chip_rate: float = 1.023e6      # of `codes`
    seed: int = 0

    @property
    def n_samples(self) -> int:
        return self.iq.size // 2

    @property
    def cells(self) -> int:
        return len(self.codes) * len(self.doppler) * self.surface_len

    def truth(self):
        """(prn, expected surface index, Doppler) of each planted satellite."""
        spc = self.fs / self.chip_rate
        out = []
        for s in self.sats:
            idx = (s.code_phase * spc) % self.surface_len
            out.append((s.prn, idx, s.doppler, s.cn0))
        return out

    def replicas(self) -> np.ndarray:
        """(n_prn, n_fft) real replica samples (RepeatCodeNTimes + floor upsampler)."""
        reps = []
        for c in self.codes:
            L = len(c) * self.repeats
            i = np.arange(self.n_fft, dtype=np.int64)
            reps.append(c[((i * L) // self.n_fft) % L % len(c)].astype(np.float32))
        return np.stack(reps)


def _draw_sats(rng, n_sats, n_prn, code_of, chip_rate, carrier, code_len, cn0_lo, cn0_hi, dop=4000.0,
               nav=False, code_ms=1):
    prns = rng.choice(n_prn, n_sats, replace=False)
    sats = []
    for p in prns:
        tau = rng.uniform(0, code_len)
        s = Sat(int(p), code_of(int(p)), chip_rate, carrier, tau, rng.uniform(-dop, dop), rng.uniform(0, 1),
                rng.uniform(cn0_lo, cn0_hi))
        if nav:
            s.nav_period_s = 0.02
            s.nav_offset_s = int(rng.integers(0, 20)) * 1e-3 * code_ms + tau / chip_rate
            s.nav_seed = int(rng.integers(1, 2**62))
        sats.append(s)
    return sats


def l1ca_codes(n_prn: int = 32):
    from . import gps_l1ca
    return [gps_l1ca(i) for i in range(n_prn)]


def l1ca_1ms(seed: int = 0xC1, n_prn: int = 32) -> AcqWorkload:
    rng = np.random.default_rng(seed)
    codes = l1ca_codes(32)
    fs = 4.092e6
    sats = _draw_sats(rng, 6, 32, lambda p: codes[p], 1.023e6, L1_HZ, 1023, 42, 48)
    iq = synthesize(sats, fs, 4092, 8.0, seed, 8)
    return AcqWorkload("L1CA_1MS", fs, FMT_CI8, iq, 4092, 1, 4092, True, doppler_bins(5000, 500),
                       codes[:n_prn], 1, sats, seed=seed)


def l1ca_4ms(seed: int = 0xC0, n_prn: int = 32) -> AcqWorkload:
    rng = np.random.default_rng(seed)
    codes = l1ca_codes(32)
    fs = 4.092e6
    sats = _draw_sats(rng, 6, 32, lambda p: codes[p], 1.023e6, L1_HZ, 1023, 42, 48)
    iq = synthesize(sats, fs, 16368, 8.0, seed, 8)
    return AcqWorkload("L1CA_4MS", fs, FMT_CI8, iq, 16368, 1, 4092, True, doppler_bins(5000, 500),
                       codes[:n_prn], 4, sats, seed=seed)


def l1ca_wide(seed: int = 0xC2, n_prn: int = 32, n_blocks: int = 10, doppler_range: float = 5000,
              doppler_step: float = 250) -> AcqWorkload:
    rng = np.random.default_rng(seed)
    codes = l1ca_codes(32)
    fs = 16.368e6
    sats = _draw_sats(rng, 8, 32, lambda p: codes[p], 1.023e6, L1_HZ, 1023, 38, 48, nav=True)
    iq = synthesize(sats, fs, 16368 * n_blocks, 512.0, seed, 16)
    return AcqWorkload("L1CA_WIDE", fs, FMT_CI16, iq, 16368, n_blocks, 16368, True,
                       doppler_bins(doppler_range, doppler_step), codes[:n_prn], 1, sats, seed=seed)


def boc_codes(seed: int = 0xC3, n_prn: int = 36):
    return [random_code(seed * 1000 + p, 4092) for p in range(n_prn)]


def boc_4ms(seed: int = 0xC3, n_prn: int = 36, n_blocks: int = 5, doppler_range: float = 6000,
            doppler_step: float = 125) -> AcqWorkload:
    from . import boc
    rng = np.random.default_rng(seed)
    prim = boc_codes(seed, 36)
    codes = [boc(c) for c in prim]       # 8184 half-chips per 4 ms
    fs = 20.46e6
    sats = _draw_sats(rng, 8, 36, lambda p: codes[p], 2.046e6, L1_HZ, 8184, 40, 48, dop=5000)
    iq = synthesize(sats, fs, 81840 * n_blocks, 512.0, seed, 16)
    return AcqWorkload("BOC_4MS", fs, FMT_CI16, iq, 81840, n_blocks, 81840, False,
                       doppler_bins(doppler_range, doppler_step), codes[:n_prn], 1, sats,
                       chip_rate=2.046e6, seed=seed)


@dataclass
class BankWorkload:
    name: str
    fs: float
    spm: int
    n_blocks: int
    iq: np.ndarray
    codes: np.ndarray          # (n_ch, code_len)
    channels: list             # (f, theta_cycles, code_rate, code_phase_chips)
    taps: np.ndarray
    sats: list

    @property
    def channel_samples(self) -> int:
        return self.spm * self.n_blocks * len(self.channels)


def bank_l5(seed: int = 0xC4, n_ch: int = 64, n_blocks: int = 1000, fs: float = 40.92e6) -> BankWorkload:
    rng = np.random.default_rng(seed)
    codes = np.stack([random_code(seed * 7919 + c, 10230) for c in range(n_ch)])
    spm = int(round(fs / 1000))
    sats, chans = [], []
    for c in range(n_ch):
        f = rng.uniform(-5000, 5000)
        tau = rng.uniform(0, 10230)
        th = rng.uniform(0, 1)
        s = Sat(c, codes[c], 10.23e6, L5_HZ, tau, f, th, rng.uniform(40, 50))
        sats.append(s)
        rate = 10.23e6 * (1 + f / L5_HZ)
        # code phase at t = 0 in the NCO convention: chip(t) = floor(tau_nco + rate t)
        chans.append((f, th, rate, (-tau) % 10230))
    iq = synthesize(sats, fs, spm * n_blocks, 512.0, seed, 16)
    taps = np.array([-0.5, -0.25, 0.0, 0.25, 0.5])
    return BankWorkload("BANK_L5", fs, spm, n_blocks, iq, codes, chans, taps, sats)


@dataclass
class GalileoWorkload:
    name: str
    fs: float
    n_fft: int
    iq: np.ndarray             # interleaved int16 I/Q, 4 ms
    codes: list                # per satellite: BOC(1,1)-expanded 4092-chip code (8184)
    planted: list              # (satellite, permutation, delay in samples, Doppler)
    doppler: np.ndarray
    n_blocks: int = 1          # 4 ms blocks summed after |.|

    @property
    def cells(self) -> int:    # satellites x permutations x Doppler x code phase
        return len(self.codes) * 4 * len(self.doppler) * self.n_fft


def galileo_e1(seed: int = 0xE1, n_sv: int = 50, n_present: int = 6, fs: float = 8.192e6, n_blocks: int = 1,
               doppler_range: float = 5000, doppler_step: float = 250,
               name: str = "GALILEO_E1_HYP") -> GalileoWorkload:
    """E1-like code-hypothesis search: n_sv seeded 4092-chip codes (BOC(1,1)),
    n_present of them in n_blocks x 4 ms of int16 IQ at fs, each with one of
    the four per-ms sign permutations (the same in every 4 ms block), C/N0
    U[44, 50] dB-Hz, Doppler on the search grid, sigma 300 LSB"""
    from . import boc, e1_sign_hypotheses
    rng = np.random.default_rng(seed)
    n = int(round(4 * fs / 1e3))
    codes = [boc(random_code(seed * 7919 + s, 4092)) for s in range(n_sv)]
    sigma = 300.0
    k = np.arange(n * n_blocks)
    x = sigma * (rng.normal(size=k.size) + 1j * rng.normal(size=k.size))
    planted = []
    span = int(doppler_range // doppler_step)
    for s in rng.choice(n_sv, n_present, replace=False):
        h, tau = int(rng.integers(0, 4)), int(rng.integers(0, n))
        f = doppler_step * int(rng.integers(-span, span + 1))
        amp = sigma * np.sqrt(2 * 10 ** (rng.uniform(44, 50) / 10) / fs)
        rep = np.tile(np.roll(e1_sign_hypotheses(codes[s], n)[h], tau), n_blocks)
        x += amp * rep * np.exp(1j * (2 * np.pi * f * k / fs + rng.uniform(0, 2 * np.pi)))
        planted.append((int(s), h, tau, f))
    iq = np.empty(2 * k.size, np.int16)
    iq[0::2] = np.clip(np.round(x.real), -32767, 32767)
    iq[1::2] = np.clip(np.round(x.imag), -32767, 32767)
    return GalileoWorkload(name, fs, n, iq, codes, sorted(planted), doppler_bins(doppler_range, doppler_step),
                           n_blocks)


CONFIGS = {
    "L1CA_1MS": l1ca_1ms,
    "L1CA_4MS": l1ca_4ms,
    "L1CA_WIDE": l1ca_wide,
    "BOC_4MS": boc_4ms,
}


# ------------------------------------------------------------------ serving
@dataclass
class ServeRequest:
    kind: str
    iq: np.ndarray
    fmt: int
    fs: float
    n_fft: int
    n_blocks: int
    surface_len: int
    reshape: bool
    doppler: np.ndarray
    keys: list
    arrival_s: float = 0.0
    precision: int = 0       # PREC_FP32; BOC requests may run the reduced form

    @property
    def cells(self) -> int:
        return len(self.keys) * len(self.doppler) * self.surface_len


class ServeMix:
    """Mixed acquisition request stream with Poisson arrivals:
    50 % L1 C/A-like (int8 @ 4.092 MHz, T = 1-10 ms),
    30 % L5-like (seeded 10230-chip codes, int16 @ 20.46 MHz, T = 1-4 ms),
    20 % BOC(1,1) (seeded 4092-chip codes, int16 @ 20.46 MHz, 4 ms coherent);
    each with 1-16 PRNs and a Doppler span of +-2..+-6 kHz, Doppler step
    1/(2 T_coh).  coherent=True takes T as the coherent length: one block of
    N = 4092 T (L1) or 20460 T (L5) samples folded to the 1 ms code period
    (GetOneMsPeak<reshape, coherent>), Doppler step 500/T Hz; coherent=False
    takes T as 1 ms coherent blocks summed non-coherently (K = T, step 500 Hz),
    about 1/T of the work.  Signals come from a small pool of pre-synthesised
    records (4 planted satellites each, C/N0 U[44,48]); replica spectra are
    put in the context's cache once per (code, length).  boc_precision
    selects the storage of the BOC requests (FP32 plain, or FP16/BF16: the
    reduced form with matrix-core DFT stages)."""

    def __init__(self, ctx, seed: int = 0xC5, pool: int = 4, boc_precision: int = 0, coherent: bool = False):
        from . import gps_l1ca, boc
        self.ctx = ctx
        self.rng = np.random.default_rng(seed)
        self.l1_codes = [gps_l1ca(i) for i in range(32)]
        self.l5_codes = [random_code(seed * 31 + i, 10230) for i in range(32)]
        self.boc_codes = [boc(random_code(seed * 37 + i, 4092)) for i in range(36)]
        self.keys = {"L1": [ctx.replica_put(c, 1, 4092) for c in self.l1_codes],
                     "L5": [ctx.replica_put(c, 1, 20460) for c in self.l5_codes],
                     "BOC": [ctx.replica_put(c, 1, 81840, boc_precision) for c in self.boc_codes]}
        self.coherent = coherent
        if coherent:   # replicas of T code periods at N = T x spm
            for T in range(2, 11):
                self.keys[("L1", T)] = [ctx.replica_put(c, T, 4092 * T) for c in self.l1_codes]
            for T in range(2, 5):
                self.keys[("L5", T)] = [ctx.replica_put(c, T, 20460 * T) for c in self.l5_codes]
        self.boc_precision = boc_precision
        self.pool = {"L1": [], "L5": [], "BOC": []}
        for i in range(pool):
            r = np.random.default_rng(seed * 101 + i)
            sats = _draw_sats(r, 4, 32, lambda p: self.l1_codes[p], 1.023e6, L1_HZ, 1023, 44, 48)
            self.pool["L1"].append(synthesize(sats, 4.092e6, 4092 * 10, 8.0, seed + i, 8))
            sats = _draw_sats(r, 4, 32, lambda p: self.l5_codes[p], 10.23e6, L5_HZ, 10230, 44, 48)
            self.pool["L5"].append(synthesize(sats, 20.46e6, 20460 * 4, 512.0, seed + 100 + i, 16))
            sats = _draw_sats(r, 4, 36, lambda p: self.boc_codes[p], 2.046e6, L1_HZ, 8184, 44, 48)
            self.pool["BOC"].append(synthesize(sats, 20.46e6, 81840, 512.0, seed + 200 + i, 16))

    def draw(self) -> ServeRequest:
        u = self.rng.random()
        n_prn = int(self.rng.integers(1, 17))
        span = float(self.rng.integers(2, 7)) * 1000.0
        sig = self.pool_pick
        if u < 0.5:
            T = int(self.rng.integers(1, 11))
            prns = self.rng.choice(32, n_prn, replace=False)
            if self.coherent:
                keys = self.keys["L1"] if T == 1 else self.keys[("L1", T)]
                return ServeRequest("L1", sig("L1")[: 2 * 4092 * T], FMT_CI8, 4.092e6, 4092 * T, 1, 4092, True,
                                    doppler_bins(span, 500.0 / T), [keys[p] for p in prns])
            return ServeRequest("L1", sig("L1")[: 2 * 4092 * T], FMT_CI8, 4.092e6, 4092, T, 4092, True,
                                doppler_bins(span, 500.0), [self.keys["L1"][p] for p in prns])
        if u < 0.8:
            T = int(self.rng.integers(1, 5))
            prns = self.rng.choice(32, n_prn, replace=False)
            if self.coherent:
                keys = self.keys["L5"] if T == 1 else self.keys[("L5", T)]
                return ServeRequest("L5", sig("L5")[: 2 * 20460 * T], FMT_CI16, 20.46e6, 20460 * T, 1, 20460, True,
                                    doppler_bins(span, 500.0 / T), [keys[p] for p in prns])
            return ServeRequest("L5", sig("L5")[: 2 * 20460 * T], FMT_CI16, 20.46e6, 20460, T, 20460, True,
                                doppler_bins(span, 500.0), [self.keys["L5"][p] for p in prns])
        prns = self.rng.choice(36, n_prn, replace=False)
        return ServeRequest("BOC", sig("BOC"), FMT_CI16, 20.46e6, 81840, 1, 81840, False,
                            doppler_bins(span, 125.0), [self.keys["BOC"][p] for p in prns],
                            precision=self.boc_precision)

    def pool_pick(self, kind):
        return self.pool[kind][int(self.rng.integers(0, len(self.pool[kind])))]

    def stream(self, n: int, rate_hz: float):
        t = 0.0
        out = []
        for _ in range(n):
            t += float(self.rng.exponential(1.0 / rate_hz))
            r = self.draw()
            r.arrival_s = t
            out.append(r)
        return out

    def request(self, r: ServeRequest):
        """C request struct + the arrays it points to (keep both alive)"""
        return self.ctx.make_request(r.iq, r.fmt, fs_in=r.fs, fs_acq=r.fs, n_fft=r.n_fft, n_blocks=r.n_blocks,
                                     surface_len=r.surface_len, doppler=r.doppler, keys=r.keys, reshape=r.reshape,
                                     precision=r.precision)
