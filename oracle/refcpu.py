"""ctypes bindings for the oracle library (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, never by the product package.  Each
wrapper mirrors one function of refcpu.h, which cites the reference file:line
it restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class AcqParams(C.Structure):
    _fields_ = [
        ("n_fft", C.c_int64), ("n_blocks", C.c_int64), ("surface_len", C.c_int64),
        ("reshape", C.c_int32), ("coherent", C.c_int32),
        ("fs_in", C.c_double), ("fs_acq", C.c_double), ("intermediate_frequency", C.c_double),
        ("threshold", C.c_double), ("n_doppler", C.c_int64), ("doppler", C.POINTER(C.c_double)),
        ("n_prn", C.c_int64), ("replicas", C.POINTER(C.c_float)),
    ]


class AcqResult(C.Structure):
    _fields_ = [
        ("level", C.c_double), ("sigma", C.c_double), ("mean", C.c_double), ("stddev", C.c_double),
        ("doppler", C.c_double), ("code_offset", C.c_double), ("snr", C.c_double),
        ("runner_up_code", C.c_double), ("runner_up_doppler", C.c_double),
        ("index", C.c_int64), ("doppler_bin", C.c_int32), ("detected", C.c_int32),
    ]


class Channel(C.Structure):
    _fields_ = [("carrier_frequency", C.c_double), ("carrier_phase", C.c_double),
                ("code_rate", C.c_double), ("code_phase", C.c_double)]


P = np.ctypeslib.ndpointer
F32 = P(dtype=np.float32, flags="C_CONTIGUOUS")
F64 = P(dtype=np.float64, flags="C_CONTIGUOUS")
I8 = P(dtype=np.int8, flags="C_CONTIGUOUS")
U8 = P(dtype=np.uint8, flags="C_CONTIGUOUS")


class TrackState(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("code_phase", "code_frequency", "base_code_frequency", "carrier_phase",
                                          "carrier_frequency", "carrier_phase_error", "code_nco", "code_error")] + [
        ("previous_prompt", C.c_float * 2), ("epochs", C.c_int64), ("code_period_ms", C.c_int32),
        ("reserved", C.c_int32)]


def _declare(L):
    for name in ("refcpu_gps_l1ca", "refcpu_gps_l5i", "refcpu_gps_l5q"):
        getattr(L, name).argtypes = [C.c_int, I8]
        getattr(L, name).restype = C.c_int
    L.refcpu_glonass_of.argtypes = [I8]
    L.refcpu_boc_expand.argtypes = [I8, C.c_size_t, I8]
    L.refcpu_bbp_epoch_bytes.argtypes = [C.c_double]
    L.refcpu_bbp_epoch_bytes.restype = C.c_int64
    L.refcpu_bbp_verify_headers.argtypes = [U8, C.c_int64, C.c_int64]
    L.refcpu_bbp_verify_headers.restype = C.c_int
    L.refcpu_unpack_bbp.argtypes = [U8, C.c_int64, C.c_int64, C.c_int64, F32]
    L.refcpu_unpack_nt1065.argtypes = [U8, C.c_int64, C.c_int, F32]
    L.refcpu_track.argtypes = [F32, C.c_int64, C.c_int64, C.c_double, F32, C.c_double, C.POINTER(TrackState), F32,
                               C.c_void_p]
    L.refcpu_repeat_upsample.argtypes = [I8, C.c_size_t, C.c_size_t, F32, C.c_size_t]
    L.refcpu_upsample_f.argtypes = [F32, C.c_size_t, F32, C.c_size_t]
    L.refcpu_upsample_hold_c.argtypes = [F32, C.c_size_t, F32, C.c_size_t]
    L.refcpu_accumulate_c.argtypes = [F32, C.c_size_t, C.c_size_t, F32]
    L.refcpu_accumulate_c.restype = C.c_size_t
    L.refcpu_resample_len.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t]
    L.refcpu_resample_len.restype = C.c_size_t
    L.refcpu_resample_c.argtypes = [F32, C.c_size_t, C.c_size_t, C.c_size_t, F32]
    L.refcpu_resample_c.restype = C.c_size_t
    L.refcpu_adjust_sampling_rate.argtypes = [C.c_double, C.c_double]
    L.refcpu_adjust_sampling_rate.restype = C.c_double
    L.refcpu_mix_c.argtypes = [F32, C.c_size_t, C.c_double, C.c_double, C.c_double]
    L.refcpu_mix_inplace_c.argtypes = [F32, C.c_size_t, C.c_double, C.c_double, C.c_double]
    L.refcpu_mixer_member_translate_c.argtypes = [F32, C.c_size_t, C.c_double, C.c_double,
                                                  C.POINTER(C.c_double)]
    L.refcpu_table_mixer_table.argtypes = [F32]
    L.refcpu_table_mix_c.argtypes = [F32, C.c_size_t, C.c_double, C.c_double, C.c_double]
    L.refcpu_table_mixer_member_c.argtypes = [F32, C.c_size_t, C.c_double, C.c_double, C.POINTER(C.c_double)]
    L.refcpu_fft_d.argtypes = [F64, C.c_size_t, C.c_int]
    L.refcpu_fft_f.argtypes = [F32, C.c_size_t, C.c_int]
    L.refcpu_mf_prepare_f.argtypes = [F32, C.c_size_t, F32]
    L.refcpu_mf_process_f.argtypes = [F32, F32, C.c_size_t]
    L.refcpu_abs_c.argtypes = [F32, C.c_size_t, F32]
    L.refcpu_reshape_sum_f.argtypes = [F32, C.c_size_t, C.c_size_t, F32]
    L.refcpu_reshape_sum_f.restype = C.c_int
    L.refcpu_max_index_f.argtypes = [F32, C.c_size_t, C.POINTER(C.c_float), C.POINTER(C.c_int64)]
    L.refcpu_mean_std_f.argtypes = [F32, C.c_size_t, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.refcpu_acquire.argtypes = [C.POINTER(AcqParams), F32, C.c_int, C.c_int,
                                 C.POINTER(AcqResult), C.c_void_p, C.c_void_p]
    L.refcpu_acquire.restype = C.c_int
    L.refcpu_correlate.argtypes = [F32, F32, C.c_size_t, F64]
    L.refcpu_correlate_split.argtypes = [F32, C.c_size_t, F32, C.c_double, C.c_double, F64]
    L.refcpu_epl.argtypes = [F32, C.c_size_t, C.c_double, C.c_double, C.c_double, F32, C.c_double,
                             C.c_double, C.c_double, F64, C.c_int, F64]
    L.refcpu_bank_nco.argtypes = [F32, C.c_int64, C.c_int64, C.c_double, C.POINTER(Channel), C.c_int,
                                  I8, C.c_int64, F64, C.c_int, C.c_int, F64]


# ------------------------------------------------------------------ helpers --
def c2f(x: np.ndarray) -> np.ndarray:
    """complex array -> interleaved float32"""
    return np.ascontiguousarray(np.asarray(x, np.complex64)).view(np.float32).copy()


def f2c(x: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(x, np.float32).view(np.complex64).copy()


def gps_l1ca(sv: int) -> np.ndarray:
    out = np.zeros(1023, np.int8)
    if lib().refcpu_gps_l1ca(sv, out):
        raise RuntimeError("SV number for the GPS C/A code exceed available values")
    return out


def gps_l5i(sv: int) -> np.ndarray:
    out = np.zeros(10230, np.int8)
    if lib().refcpu_gps_l5i(sv, out):
        raise RuntimeError("SV number for the GPS L5I code exceed available values")
    return out


def gps_l5q(sv: int) -> np.ndarray:
    out = np.zeros(10230, np.int8)
    if lib().refcpu_gps_l5q(sv, out):
        raise RuntimeError("SV number for the GPS L5Q code exceed available values")
    return out


def glonass_of() -> np.ndarray:
    out = np.zeros(511, np.int8)
    lib().refcpu_glonass_of(out)
    return out


def boc_expand(chips: np.ndarray) -> np.ndarray:
    chips = np.ascontiguousarray(chips, np.int8)
    out = np.zeros(2 * len(chips), np.int8)
    lib().refcpu_boc_expand(chips, len(chips), out)
    return out


def bbp_epoch_bytes(fs: float) -> int:
    return int(lib().refcpu_bbp_epoch_bytes(fs))


def bbp_verify_headers(data: np.ndarray, n_epochs: int, epoch_bytes: int) -> bool:
    data = np.ascontiguousarray(data, np.uint8)
    assert data.size >= n_epochs * epoch_bytes
    return lib().refcpu_bbp_verify_headers(data, n_epochs, epoch_bytes) == 0


def unpack_bbp(data: np.ndarray, n: int, fs: float) -> np.ndarray:
    """GetPartialSignal for BbpDdc: n complex samples from epoch 0"""
    data = np.ascontiguousarray(data, np.uint8)
    spm, eb = int(round(fs / 1e3)), bbp_epoch_bytes(fs)
    assert data.size >= -(-n // spm) * eb
    out = np.zeros(2 * n, np.float32)
    lib().refcpu_unpack_bbp(data, n, spm, eb, out)
    return f2c(out)


def unpack_nt1065(data: np.ndarray, n: int, offset_bits: int) -> np.ndarray:
    data = np.ascontiguousarray(data, np.uint8)
    assert data.size >= n
    out = np.zeros(2 * n, np.float32)
    lib().refcpu_unpack_nt1065(data, n, offset_bits, out)
    return f2c(out)


def track_state(code_phase: float, carrier_frequency: float, base_code_frequency: float = 1.023e6,
                code_period_ms: int = 1, carrier_phase: float = 0.0) -> TrackState:
    """TrackingParameters right after the acquisition hand-off"""
    st = TrackState()
    st.code_phase, st.carrier_frequency, st.carrier_phase = code_phase, carrier_frequency, carrier_phase
    st.code_frequency = st.base_code_frequency = base_code_frequency
    st.code_period_ms = code_period_ms
    return st


def track(x: np.ndarray, spm: int, n_blocks: int, fs: float, code3: np.ndarray, state: TrackState,
          spacing: float = 0.25):
    """Tracker::Track for one channel; state is advanced in place.  Returns
    (epl [n_blocks, 3] complex64, trace [n_blocks, 4])"""
    xb = c2f(np.ascontiguousarray(x[:spm * n_blocks], np.complex64))
    code3 = np.ascontiguousarray(code3, np.float32)
    assert code3.size == 3 * spm * state.code_period_ms
    epl = np.zeros(6 * n_blocks, np.float32)
    trace = np.zeros((n_blocks, 4), np.float64)
    lib().refcpu_track(xb, spm, n_blocks, fs, code3, spacing, C.byref(state), epl, trace.ctypes.data)
    return f2c(epl).reshape(n_blocks, 3), trace


def repeat_upsample(code: np.ndarray, repeats: int, samples: int) -> np.ndarray:
    code = np.ascontiguousarray(code, np.int8)
    out = np.zeros(samples, np.float32)
    lib().refcpu_repeat_upsample(code, len(code), repeats, out, samples)
    return out


def upsample(src: np.ndarray, samples: int) -> np.ndarray:
    src = np.ascontiguousarray(src, np.float32)
    out = np.zeros(samples, np.float32)
    lib().refcpu_upsample_f(src, len(src), out, samples)
    return out


def upsample_hold(src: np.ndarray, samples: int) -> np.ndarray:
    s = c2f(src)
    out = np.zeros(2 * samples, np.float32)
    lib().refcpu_upsample_hold_c(s, len(src), out, samples)
    return f2c(out)


def accumulate(src: np.ndarray, ratio: int) -> np.ndarray:
    s = c2f(src)
    out = np.zeros(2 * (len(src) // ratio + 1), np.float32)
    n = lib().refcpu_accumulate_c(s, len(src), ratio, out)
    return f2c(out[: 2 * n])


def resample(src: np.ndarray, new_fs: int, old_fs: int) -> np.ndarray:
    s = c2f(src)
    n_out = lib().refcpu_resample_len(len(src), new_fs, old_fs)
    out = np.zeros(2 * n_out, np.float32)
    n = lib().refcpu_resample_c(s, len(src), new_fs, old_fs, out)
    return f2c(out[: 2 * n])


def adjust_sampling_rate(fs: float, target: float) -> float:
    return lib().refcpu_adjust_sampling_rate(fs, target)


def mix(x: np.ndarray, fs: float, f: float, phase: float = 0.0) -> np.ndarray:
    b = c2f(x)
    lib().refcpu_mix_c(b, len(x), fs, f, phase)
    return f2c(b)


def mix_inplace(x: np.ndarray, fs: float, f: float, phase: float = 0.0) -> np.ndarray:
    b = c2f(x)
    lib().refcpu_mix_inplace_c(b, len(x), fs, f, phase)
    return f2c(b)


def mixer_member_translate(x: np.ndarray, fs: float, f: float, phase: float):
    b = c2f(x)
    ph = C.c_double(phase)
    lib().refcpu_mixer_member_translate_c(b, len(x), fs, f, C.byref(ph))
    return f2c(b), ph.value


def table_mixer_table() -> np.ndarray:
    t = np.zeros(80, np.float32)
    lib().refcpu_table_mixer_table(t)
    return t


def table_mix(x: np.ndarray, fs: float, f: float, phase: float = 0.0) -> np.ndarray:
    b = c2f(x)
    lib().refcpu_table_mix_c(b, len(x), fs, f, phase)
    return f2c(b)


def table_mixer_member(x: np.ndarray, fs: float, f: float, phase: float):
    b = c2f(x)
    ph = C.c_double(phase)
    lib().refcpu_table_mixer_member_c(b, len(x), fs, f, C.byref(ph))
    return f2c(b), ph.value


def fft(x: np.ndarray, inverse: bool = False, precision: int = 64) -> np.ndarray:
    if precision == 64:
        b = np.ascontiguousarray(np.asarray(x, np.complex128)).view(np.float64).copy()
        lib().refcpu_fft_d(b, len(x), int(inverse))
        return b.view(np.complex128)
    b = c2f(x)
    lib().refcpu_fft_f(b, len(x), int(inverse))
    return f2c(b)


def mf_prepare(code: np.ndarray) -> np.ndarray:
    code = np.ascontiguousarray(code, np.float32)
    out = np.zeros(2 * len(code), np.float32)
    lib().refcpu_mf_prepare_f(code, len(code), out)
    return f2c(out)


def mf_process(x: np.ndarray, spec: np.ndarray) -> np.ndarray:
    b = c2f(x)
    lib().refcpu_mf_process_f(b, c2f(spec), len(x))
    return f2c(b)


def abs_(x: np.ndarray) -> np.ndarray:
    out = np.zeros(len(x), np.float32)
    lib().refcpu_abs_c(c2f(x), len(x), out)
    return out


def reshape_sum(x: np.ndarray, block: int) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(block, np.float32)
    if lib().refcpu_reshape_sum_f(x, len(x), block, out):
        raise RuntimeError("Signal size is not divided by block size")
    return out


def max_index(x: np.ndarray):
    x = np.ascontiguousarray(x, np.float32)
    v, i = C.c_float(), C.c_int64()
    lib().refcpu_max_index_f(x, len(x), C.byref(v), C.byref(i))
    return v.value, i.value


def mean_std(x: np.ndarray):
    x = np.ascontiguousarray(x, np.float32)
    m, s = C.c_double(), C.c_double()
    lib().refcpu_mean_std_f(x, len(x), C.byref(m), C.byref(s))
    return m.value, s.value


RESULT_FIELDS = [f for f, _ in AcqResult._fields_]


def acquire(signal: np.ndarray, replicas: np.ndarray, *, n_fft: int, n_blocks: int, surface_len: int,
            fs_in: float, fs_acq: float, doppler, reshape: bool = True, coherent: bool = True,
            intermediate_frequency: float = 0.0, threshold: float = 3.3, precision: int = 64,
            threads: int | None = None, want_surfaces: bool = False, want_peaks: bool = False):
    """FastSearchEngineBase::ProcessBpsk generalised to K blocks (refcpu.h).

    signal: complex, n_blocks*n_fft samples at fs_acq.  replicas: (n_prn, n_fft)
    real.  Returns (list of dict per PRN, surfaces or None, peaks or None)."""
    sig = c2f(np.asarray(signal)[: n_blocks * n_fft])
    reps = np.ascontiguousarray(replicas, np.float32)
    n_prn = reps.shape[0]
    dop = np.ascontiguousarray(doppler, np.float64)
    prm = AcqParams(n_fft, n_blocks, surface_len, int(reshape), int(coherent), fs_in, fs_acq,
                    intermediate_frequency, threshold, len(dop),
                    dop.ctypes.data_as(C.POINTER(C.c_double)), n_prn,
                    reps.ctypes.data_as(C.POINTER(C.c_float)))
    out = (AcqResult * n_prn)()
    surf = np.zeros((n_prn, len(dop), surface_len), np.float32) if want_surfaces else None
    peaks = np.zeros((n_prn, surface_len), np.float32) if want_peaks else None
    threads = threads or min(os.cpu_count() or 1, 16)
    rc = lib().refcpu_acquire(C.byref(prm), sig, precision, threads, out,
                              surf.ctypes.data if surf is not None else None,
                              peaks.ctypes.data if peaks is not None else None)
    if rc:
        raise RuntimeError(f"refcpu_acquire failed ({rc})")
    res = [{f: getattr(out[i], f) for f in RESULT_FIELDS} for i in range(n_prn)]
    return res, surf, peaks


def reduce_hypotheses(results: list, h: int, threshold: float = 3.3) -> list:
    """FastSearchEngineBase::ProcessGalileo's choice (src/acquisition/fse.hpp:
    382-388): per run of h results, std::max_element by level (the first of
    the largest), pushed when its SNR clears the threshold"""
    out = []
    for g in range(len(results) // h):
        run = results[g * h:(g + 1) * h]
        best = 0
        for k in range(1, h):
            if run[best]["level"] < run[k]["level"]:
                best = k
        r = dict(run[best])
        r["detected"] = int(r["snr"] > threshold)
        r["hypothesis"] = best
        out.append(r)
    return out


def process_galileo(signal: np.ndarray, chips_list, *, spm: int, fs_in: float, fs_acq: float, doppler,
                    intermediate_frequency: float = 0.0, threshold: float = 3.3, threads: int | None = None):
    """FastSearchEngineBase::ProcessGalileo (src/acquisition/fse.hpp:356-396)
    on a signal already at fs_acq: per satellite the 4 ms code is upsampled to
    4 spm samples (Upsampler::Transform, floor index), each of the four sign
    permutations (:374-379) multiplies its milliseconds, ProcessBpsk<false>
    (no fold, no threshold) runs every hypothesis, and the largest level wins"""
    perms = [[1, 1, 1, 1], [1, 1, 1, -1], [1, 1, -1, -1], [1, -1, -1, -1]]
    n = 4 * spm
    reps = []
    for chips in chips_list:
        base = repeat_upsample(np.asarray(chips, np.int8), 1, n)
        for p in perms:
            reps.append(base * np.repeat(np.asarray(p, np.float32), spm))
    res, _, _ = acquire(signal, np.stack(reps), n_fft=n, n_blocks=1, surface_len=n, fs_in=fs_in, fs_acq=fs_acq,
                        doppler=doppler, reshape=False, intermediate_frequency=intermediate_frequency,
                        threshold=threshold, threads=threads)
    return reduce_hypotheses(res, 4, threshold)


def correlate(signal: np.ndarray, code: np.ndarray) -> complex:
    if len(signal) != len(code):
        raise RuntimeError("Size mismatch")
    out = np.zeros(2)
    lib().refcpu_correlate(c2f(signal), np.ascontiguousarray(code, np.float32), len(code), out)
    return complex(out[0], out[1])


def correlate_split(signal: np.ndarray, spm: int, code3: np.ndarray, code_period_ms: float,
                    code_phase: float) -> complex:
    out = np.zeros(2)
    lib().refcpu_correlate_split(c2f(signal), spm, np.ascontiguousarray(code3, np.float32),
                                 code_period_ms, code_phase, out)
    return complex(out[0], out[1])


def epl(x: np.ndarray, fs: float, carrier_frequency: float, carrier_phase: float, code3: np.ndarray,
        code_period_ms: float, code_phase: float, samples_per_chip: float, taps) -> np.ndarray:
    b = c2f(x)
    taps = np.ascontiguousarray(taps, np.float64)
    out = np.zeros(2 * len(taps))
    lib().refcpu_epl(b, len(x), fs, carrier_frequency, carrier_phase,
                     np.ascontiguousarray(code3, np.float32), code_period_ms, code_phase,
                     samples_per_chip, taps, len(taps), out)
    return out.view(np.complex128)


def bank_nco(x: np.ndarray, spm: int, n_blocks: int, fs: float, channels, codes: np.ndarray, taps,
             threads: int | None = None) -> np.ndarray:
    """channels: list of (carrier_frequency, carrier_phase_cycles, code_rate, code_phase_chips)."""
    n_ch = len(channels)
    ch = (Channel * n_ch)(*[Channel(*c) for c in channels])
    codes = np.ascontiguousarray(codes, np.int8)
    taps = np.ascontiguousarray(taps, np.float64)
    out = np.zeros(2 * n_blocks * n_ch * len(taps))
    xb = np.ascontiguousarray(x, np.float32) if x.dtype == np.float32 else c2f(x)
    lib().refcpu_bank_nco(xb, spm, n_blocks, fs, ch, n_ch, codes, codes.shape[1], taps, len(taps),
                          threads or min(os.cpu_count() or 1, 16), out)
    return out.view(np.complex128).reshape(n_blocks, n_ch, len(taps))


# ======================================================= anti-jamming (STFT)
# Restatement of ShortTimeFourierTransform (src/math/stft.hpp, ipp_stft.hpp),
# JammingDetector (src/antijamming/jamming_detection.hpp) and
# JammingSuppressionEngine (src/antijamming/jse.hpp).  Samples and STFT cells
# are complex64 like the reference's std::complex<float>; the DFTs run in
# FP64 (numpy pocketfft) and are rounded to complex64.

STFT_WINDOW, STFT_OVERLAP = 128, 96            # stft.hpp:18 defaults
WINDOW_BLACKMAN_HARRIS, WINDOW_HANN = 0, 1


def stft_window(kind: int, n: int = STFT_WINDOW) -> np.ndarray:
    """IppStft's 4-term Blackman-Harris over n - 1 (ipp_stft.hpp:26-35,
    accumulated into a float term by term) or SequentialStft's periodic Hann
    (stft.hpp:31-37)"""
    w = np.zeros(n, np.float32)
    if kind == WINDOW_BLACKMAN_HARRIS:
        a = (0.3635819, 0.4891775, 0.1365995, 0.0106411)
        for i in range(n):
            acc = np.float32(0.0)          # float += double: summed in double, rounded per term
            for j in range(4):
                acc = np.float32(float(acc) + a[j] * (-1.0) ** j * float(np.cos(2.0 * j * np.pi * i / (n - 1))))
            w[i] = acc
    else:
        for i in range(n):
            w[i] = np.float32(0.5 * (1 - np.cos(2.0 * np.pi * i / n)))
    return w


def stft(x: np.ndarray, kind: int) -> np.ndarray:
    """SplitIntoColumns (stft.hpp:40-58): (n - 96) / 32 columns of 128
    samples at stride 32, windowed, forward DFT without scaling"""
    x = np.asarray(x, np.complex64)
    stride = STFT_WINDOW - STFT_OVERLAP
    ncol = (len(x) - STFT_OVERLAP) // stride
    w = stft_window(kind)
    idx = np.arange(ncol)[:, None] * stride + np.arange(STFT_WINDOW)[None, :]
    cols = (x[idx] * w[None, :]).astype(np.complex64)
    return np.fft.fft(cols.astype(np.complex128), axis=1).astype(np.complex64)


def istft(cols: np.ndarray) -> np.ndarray:
    """OverlapAndAdd (stft.hpp:60-72): the inverse DFT (1/128) of each column
    added at stride 32, in column order; 128 + 32 (ncol - 1) samples"""
    cols = np.asarray(cols, np.complex64)
    stride = STFT_WINDOW - STFT_OVERLAP
    out = np.zeros(STFT_WINDOW + (len(cols) - 1) * stride, np.complex64)
    t = np.fft.ifft(cols.astype(np.complex128), axis=1).astype(np.complex64)
    for i in range(len(cols)):
        out[i * stride:i * stride + STFT_WINDOW] += t[i]
    return out


def null_region_detector(index: int, null: int, data: np.ndarray) -> None:
    """JammingDetector::NullRegion (jamming_detection.hpp:105-113), size_t
    arithmetic: start and finish taken modulo the size"""
    n = len(data)
    start = (index - null + n) % n
    finish = (index + null + n) % n
    if finish < start:
        finish += n
    for i in range(start, finish + 1):
        data[i % n] = 0


def null_region_jse(index: int, null: int, data: np.ndarray) -> None:
    """JammingSuppressionEngine::NullRegion (jse.hpp:50-59), signed
    arithmetic: start = index - null + size, finish = index + null"""
    n = len(data)
    start = index - null + n
    finish = index + null
    if finish < start:
        finish += n
    for i in range(start, finish + 1):
        data[i % n] = 0


def jse_intermediate(cols: np.ndarray, fs: float, max_cnt: int = 8, null: int = 3, js_threshold: float = 6.0):
    """IntermediateDetection (jamming_detection.hpp:115-137): per column the
    eight strongest bins, each nulled +-3 after it is taken, J/N = value /
    sigma of the nulled magnitudes, candidates under 6 dropped (order kept).
    Returns per column a list of (bin, frequency, j/n)."""
    out = []
    for col in cols:
        mag = np.abs(col).astype(np.float32)
        cand = []
        for _ in range(max_cnt):
            idx = int(np.argmax(mag))           # first index of the maximum
            cand.append((idx, idx * fs / len(mag), float(mag[idx])))
            null_region_detector(idx, null, mag)
        _, sd = mean_std(mag)
        cand = [(b, f, v / sd) for b, f, v in cand]
        out.append([c for c in cand if not (c[2] < js_threshold)])
    return out


NARROWBAND, WIDEBAND = 1, 2


def jse_analyze(intermediate, probability_threshold: float = 0.9):
    """AnalyzeStftData (jamming_detection.hpp:139-190), including its integer
    mean count, its frequency-ordered map and its wideband rule (cnt >=
    size_t(map size * 0.9), so a one-entry map also reports wideband).
    Returns a list of (type, bin, frequency, j/n)."""
    dst = []
    mean_jammers = sum(len(el) for el in intermediate) // len(intermediate)
    for i in range(mean_jammers):
        occ = {}                              # frequency -> [count, bin, frequency, j/n sum]
        quantity = 0.0
        for el in intermediate:
            if len(el) <= i:
                continue
            b, f, jn = el[i]
            e = occ.setdefault(f, [0.0, 0, 0.0, 0.0])
            e[0] += 1
            e[1], e[2] = b, f
            e[3] += jn
            quantity += 1
        for e in occ.values():
            e[0] /= quantity
            e[3] /= quantity
        keys = sorted(occ)
        cnt_wideband = 0
        wb = occ[keys[0]]
        for k in keys:
            e = occ[k]
            if e[0] >= probability_threshold:
                dst.append((NARROWBAND, e[1], e[2], e[3]))
                break
            if e[0] >= probability_threshold / len(occ):
                cnt_wideband += 1
                wb = e
        if cnt_wideband >= int(len(occ) * probability_threshold):
            dst.append((WIDEBAND, wb[1], wb[2], wb[3]))
    return dst


def jse_process(x: np.ndarray, fs: float, kind: int, null_window: int = 128, abs_threshold: float = 20.0):
    """JammingSuppressionEngine::Process (jse.hpp:93-103): detect; no
    detections leave the samples; only narrowband ones notch +-128 bins of
    the whole-signal DFT around each (FilterInFreqencyDomain, :61-69);
    otherwise STFT cells with |X| >= 20 are zeroed and the columns overlap-
    added back, resized to the input length (FilterWideband, :71-87).
    Returns (samples, detections)."""
    x = np.asarray(x, np.complex64).copy()
    cols = stft(x, kind)
    det = jse_analyze(jse_intermediate(cols, fs))
    if not det:
        return x, det
    if all(d[0] != WIDEBAND for d in det):
        X = np.fft.fft(x.astype(np.complex128)).astype(np.complex64)
        for _, _, f, _ in det:
            null_region_jse(int(f / fs * len(X)), null_window, X)
        return np.fft.ifft(X.astype(np.complex128)).astype(np.complex64), det
    cols = cols.copy()
    cols[np.abs(cols) >= abs_threshold] = 0
    y = istft(cols)
    out = np.zeros(len(x), np.complex64)
    m = min(len(x), len(y))
    out[:m] = y[:m]
    return out, det
