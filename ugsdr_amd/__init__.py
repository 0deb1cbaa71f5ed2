"""ugsdr_amd: MI355X-native acquisition hot path of the ugsdr GNSS receiver.

Python bindings over the C ABI of libugsdr_hip.so (include/ugsdr_hip.h).  The
native library is required: importing works without it, but every GPU entry
point raises if it cannot be loaded — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(HERE, "libugsdr_hip.so")

FMT_CI8, FMT_CI16, FMT_CF32, FMT_R8 = 1, 2, 3, 4
FMT_BBP_DDC, FMT_NT1065_1, FMT_NT1065_2, FMT_NT1065_3, FMT_NT1065_4 = 5, 6, 7, 8, 9
PACKED_FORMATS = (FMT_BBP_DDC, FMT_NT1065_1, FMT_NT1065_2, FMT_NT1065_3, FMT_NT1065_4)
PREC_FP32, PREC_FP16, PREC_BF16 = 0, 1, 2
STFT_BLACKMAN_HARRIS, STFT_HANN = 0, 1
JAMMER_NARROWBAND, JAMMER_WIDEBAND = 1, 2
BANK_REF_SPLIT, BANK_NCO = 0, 1

ERRORS = {-1: "ARG", -2: "HIP", -3: "SIZE", -4: "UNSUPPORTED", -5: "NOMEM", -6: "STATE"}


class UgsdrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class AcqRequest(C.Structure):
    _fields_ = [
        ("iq", C.c_void_p), ("fmt", C.c_int32), ("iq_on_device", C.c_int32), ("n_samples", C.c_int64),
        ("fs_in", C.c_double), ("fs_acq", C.c_double), ("mix_frequency", C.c_double), ("mix_phase", C.c_double),
        ("intermediate_frequency", C.c_double), ("n_fft", C.c_int64), ("n_blocks", C.c_int64),
        ("surface_len", C.c_int64), ("reshape", C.c_int32), ("coherent", C.c_int32),
        ("doppler", C.POINTER(C.c_double)), ("n_doppler", C.c_int64),
        ("replica_keys", C.POINTER(C.c_uint64)), ("n_prn", C.c_int64), ("threshold", C.c_double),
        ("precision", C.c_int32), ("hypotheses", C.c_int32),
    ]


class AcqResult(C.Structure):
    _fields_ = [
        ("level", C.c_double), ("sigma", C.c_double), ("mean", C.c_double), ("stddev", C.c_double),
        ("doppler", C.c_double), ("code_offset", C.c_double), ("snr", C.c_double),
        ("runner_up_code", C.c_double), ("runner_up_doppler", C.c_double),
        ("index", C.c_int64), ("doppler_bin", C.c_int32), ("detected", C.c_int32),
    ]


class ChannelState(C.Structure):
    _fields_ = [("carrier_frequency", C.c_double), ("carrier_phase", C.c_double),
                ("code_rate", C.c_double), ("code_phase", C.c_double)]


class BankRequest(C.Structure):
    _fields_ = [
        ("iq", C.c_void_p), ("fmt", C.c_int32), ("iq_on_device", C.c_int32), ("spm", C.c_int64),
        ("n_blocks", C.c_int64), ("fs", C.c_double), ("mode", C.c_int32), ("n_channels", C.c_int32),
        ("n_taps", C.c_int32), ("reserved", C.c_int32), ("taps_chips", C.POINTER(C.c_double)),
        ("codes", C.POINTER(C.c_int8)), ("code_len", C.c_int64), ("samples_per_chip", C.c_double),
        ("code_period_ms", C.c_double),
    ]


class Dfe:
    """One DFE channel (Channel, src/dfe/dfe.hpp:110-203): recorded samples in
    any FMT_* at fs_in, mixed by mixer_frequency with a running phase that
    continues across calls (Mixer::Translate member overload,
    src/mixer/mixer.hpp:53-60), resampled to fs_out."""

    def __init__(self, ctx, fmt: int, fs_in: float, fs_out: float, mixer_frequency: float, phase: float = 0.0):
        self.ctx, self.fmt, self.fs_in, self.fs_out = ctx, fmt, fs_in, fs_out
        p = C.c_void_p()
        _check(lib().ugsdr_hip_dfe_create(ctx._p, fmt, fs_in, fs_out, mixer_frequency, phase, C.byref(p)))
        self._p = p

    def process(self, raw, n_epochs: int, in_device_ptr: int | None = None, out_device_ptr: int | None = None):
        """Channel::GetSeveralEpochs: n_epochs ms -> complex64 at fs_out (None
        when written to out_device_ptr)"""
        n_out = n_epochs * int(round(self.fs_out / 1e3))
        if in_device_ptr is None:
            raw = np.ascontiguousarray(raw)
            need = format_bytes(self.fmt, n_epochs * int(round(self.fs_in / 1e3)), self.fs_in)
            if raw.nbytes < need:
                raise ValueError(f"{n_epochs} ms need {need} bytes, got {raw.nbytes}")
            src, src_dev = raw.ctypes.data, 0
        else:
            src, src_dev = in_device_ptr, 1
        if out_device_ptr is None:
            out = np.zeros(n_out, np.complex64)
            _check(lib().ugsdr_hip_dfe_process(self._p, src, src_dev, n_epochs, _ptr(out), 0))
            return out
        _check(lib().ugsdr_hip_dfe_process(self._p, src, src_dev, n_epochs, out_device_ptr, 1))
        return None

    @property
    def phase(self) -> float:
        v = C.c_double()
        _check(lib().ugsdr_hip_dfe_phase(self._p, C.byref(v)))
        return v.value

    def close(self):
        if self._p:
            lib().ugsdr_hip_dfe_destroy(self._p)
            self._p = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class TrackState(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("code_phase", "code_frequency", "base_code_frequency", "carrier_phase",
                                          "carrier_frequency", "carrier_phase_error", "code_nco", "code_error")] + [
        ("previous_prompt", C.c_float * 2), ("epochs", C.c_int64), ("code_period_ms", C.c_int32),
        ("reserved", C.c_int32)]


class Jammer(C.Structure):
    _fields_ = [("type", C.c_int32), ("frequency_bin", C.c_int32), ("frequency", C.c_double),
                ("jammer_to_noise", C.c_double)]


class TrackRequest(C.Structure):
    _fields_ = [
        ("iq", C.c_void_p), ("fmt", C.c_int32), ("iq_on_device", C.c_int32), ("spm", C.c_int64),
        ("n_blocks", C.c_int64), ("fs", C.c_double), ("n_channels", C.c_int32), ("reserved", C.c_int32),
        ("codes", C.POINTER(C.c_int8)), ("code_len", C.c_int64), ("spacing_chips", C.c_double),
    ]


def track_state(code_phase: float, carrier_frequency: float, base_code_frequency: float = 1.023e6,
                code_period_ms: int = 1, carrier_phase: float = 0.0) -> TrackState:
    """TrackingParameters right after the acquisition hand-off"""
    st = TrackState()
    st.code_phase, st.carrier_frequency, st.carrier_phase = code_phase, carrier_frequency, carrier_phase
    st.code_frequency = st.base_code_frequency = base_code_frequency
    st.code_period_ms = code_period_ms
    return st


class ServerConfig(C.Structure):
    _fields_ = [("max_device_bytes", C.c_int64), ("max_inflight", C.c_int32), ("n_streams", C.c_int32),
                ("max_batch", C.c_int32), ("round_budget_us", C.c_int32), ("round_graphs", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("submitted", "completed", "failed", "inflight", "inflight_bytes",
                                          "rounds", "batches", "max_batch_seen", "graph_launches",
                                          "graph_instantiations", "graph_updates", "graph_fallbacks",
                                          "launch_host_us")]


PENDING = 1
RESULT_FIELDS = [f for f, _ in AcqResult._fields_]

# every exported symbol with its ctypes signature (checked against include/ugsdr_hip.h by the tests)
VP, I64, I32, D, F = C.c_void_p, C.c_int64, C.c_int32, C.c_double, C.c_float
SIGNATURES = {
    "ugsdr_hip_last_error": ([], C.c_char_p),
    "ugsdr_hip_version": ([], C.c_char_p),
    "ugsdr_hip_ctx_create": ([C.c_int, C.POINTER(VP)], C.c_int),
    "ugsdr_hip_ctx_destroy": ([VP], C.c_int),
    "ugsdr_hip_ctx_stream": ([VP, C.POINTER(VP)], C.c_int),
    "ugsdr_hip_ctx_device_bytes": ([VP], I64),
    "ugsdr_hip_ctx_set_profiling": ([VP, C.c_int], C.c_int),
    "ugsdr_hip_ctx_last_times": ([VP, VP], C.c_int),
    "ugsdr_hip_autotune": ([VP, I64, C.POINTER(I32), VP, I32], C.c_int),
    "ugsdr_hip_set_autotune": ([VP, C.c_int], C.c_int),
    "ugsdr_hip_corr_variants": ([I64, VP, I32], C.c_int),
    "ugsdr_hip_set_graphs": ([VP, C.c_int], C.c_int),
    "ugsdr_hip_graph_replays": ([VP], I64),
    "ugsdr_hip_last_search_form": ([VP, C.POINTER(I64), C.POINTER(I32), C.POINTER(I32)], C.c_int),
    "ugsdr_hip_set_threads": ([VP, I64, I32], C.c_int),
    "ugsdr_format_bytes": ([C.c_int, I64, D], I64),
    "ugsdr_hip_track": ([VP, VP, VP, VP, VP], C.c_int),
    "ugsdr_hip_stft": ([VP, VP, I64, C.c_int, VP], C.c_int),
    "ugsdr_hip_istft": ([VP, VP, I64, VP], C.c_int),
    "ugsdr_hip_jse_process": ([VP, VP, I64, D, C.c_int, C.c_int, VP, I32, C.POINTER(I32)], C.c_int),
    "ugsdr_hip_dfe_create": ([VP, C.c_int, D, D, D, D, C.POINTER(VP)], C.c_int),
    "ugsdr_hip_dfe_process": ([VP, VP, C.c_int, I64, VP, C.c_int], C.c_int),
    "ugsdr_hip_dfe_phase": ([VP, C.POINTER(D)], C.c_int),
    "ugsdr_hip_dfe_destroy": ([VP], C.c_int),
    "ugsdr_hip_unpack": ([VP, VP, C.c_int, I64, D, VP], C.c_int),
    "ugsdr_code_gps_l1ca": ([C.c_int, VP], C.c_int),
    "ugsdr_code_gps_l5i": ([C.c_int, VP], C.c_int),
    "ugsdr_code_gps_l5q": ([C.c_int, VP], C.c_int),
    "ugsdr_code_glonass_of": ([VP], C.c_int),
    "ugsdr_code_boc": ([VP, I64, VP], C.c_int),
    "ugsdr_adjust_sampling_rate": ([D, D], D),
    "ugsdr_hip_replica_put": ([VP, VP, I64, I64, I64, C.c_int, C.POINTER(C.c_uint64)], C.c_int),
    "ugsdr_hip_replica_put_samples": ([VP, VP, I64, C.c_int, C.POINTER(C.c_uint64)], C.c_int),
    "ugsdr_hip_replica_drop": ([VP, C.c_uint64], C.c_int),
    "ugsdr_hip_replica_get": ([VP, C.c_uint64, VP, I64], C.c_int),
    "ugsdr_hip_acquire": ([VP, C.POINTER(AcqRequest), C.POINTER(AcqResult), VP, VP], C.c_int),
    "ugsdr_hip_acquire_noalloc": ([VP, C.POINTER(AcqRequest), C.POINTER(AcqResult), VP, VP], C.c_int),
    "ugsdr_hip_mix": ([VP, VP, VP, I64, D, D, D], C.c_int),
    "ugsdr_hip_mix_inplace": ([VP, VP, I64, D, D, D], C.c_int),
    "ugsdr_hip_table_mix": ([VP, VP, I64, D, D, D, C.c_int], C.c_int),
    "ugsdr_hip_dft": ([VP, VP, VP, I64, C.c_int], C.c_int),
    "ugsdr_hip_dft_reduced": ([VP, C.c_int, C.c_int, VP, VP, I64, I64], C.c_int),
    "ugsdr_hip_mf_prepare": ([VP, VP, VP, I64], C.c_int),
    "ugsdr_hip_mf_filter": ([VP, VP, VP, VP, I64], C.c_int),
    "ugsdr_hip_mf_filter_code": ([VP, VP, VP, VP, I64], C.c_int),
    "ugsdr_hip_abs": ([VP, VP, VP, I64], C.c_int),
    "ugsdr_hip_reshape_sum": ([VP, VP, VP, I64, I64, C.c_int], C.c_int),
    "ugsdr_hip_max_index": ([VP, VP, I64, C.POINTER(F), C.POINTER(I64)], C.c_int),
    "ugsdr_hip_mean_std": ([VP, VP, I64, C.POINTER(F), C.POINTER(F)], C.c_int),
    "ugsdr_resample_len": ([I64, I64, I64], I64),
    "ugsdr_hip_resample": ([VP, VP, I64, I64, I64, VP], C.c_int),
    "ugsdr_hip_upsample": ([VP, VP, I64, VP, I64], C.c_int),
    "ugsdr_hip_correlate": ([VP, VP, I64, VP, I64, VP], C.c_int),
    "ugsdr_hip_correlate_bank": ([VP, C.POINTER(BankRequest), C.POINTER(ChannelState), VP], C.c_int),
    "ugsdr_hip_server_create": ([VP, C.POINTER(ServerConfig), C.POINTER(VP)], C.c_int),
    "ugsdr_hip_server_submit": ([VP, C.POINTER(AcqRequest), C.POINTER(C.c_uint64)], C.c_int),
    "ugsdr_hip_server_poll": ([VP, C.c_uint64, C.c_int, C.POINTER(AcqResult), C.POINTER(C.c_double)], C.c_int),
    "ugsdr_hip_server_stats": ([VP, C.POINTER(ServerStats)], C.c_int),
    "ugsdr_hip_server_destroy": ([VP], C.c_int),
}

_lib = None


def lib():
    """Loads libugsdr_hip.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(HIP_LIB_PATH):
            raise RuntimeError(f"{HIP_LIB_PATH} not built: run python -m ugsdr_amd.build")
        L = C.CDLL(HIP_LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise UgsdrError(rc, lib().ugsdr_hip_last_error().decode())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ------------------------------------------------------------------ host codes
def gps_l1ca(sv: int) -> np.ndarray:
    out = np.zeros(1023, np.int8)
    _check(lib().ugsdr_code_gps_l1ca(sv, _ptr(out)))
    return out


def gps_l5i(sv: int) -> np.ndarray:
    out = np.zeros(10230, np.int8)
    _check(lib().ugsdr_code_gps_l5i(sv, _ptr(out)))
    return out


def gps_l5q(sv: int) -> np.ndarray:
    out = np.zeros(10230, np.int8)
    _check(lib().ugsdr_code_gps_l5q(sv, _ptr(out)))
    return out


def glonass_of() -> np.ndarray:
    out = np.zeros(511, np.int8)
    _check(lib().ugsdr_code_glonass_of(_ptr(out)))
    return out


def boc(chips: np.ndarray) -> np.ndarray:
    chips = np.ascontiguousarray(chips, np.int8)
    out = np.zeros(2 * len(chips), np.int8)
    _check(lib().ugsdr_code_boc(_ptr(chips), len(chips), _ptr(out)))
    return out


# The per-ms sign permutations FastSearchEngineBase::ProcessGalileo tries on
# the 4 ms E1 code (src/acquisition/fse.hpp:374-379)
E1_SIGN_PERMUTATIONS = np.array([[1, 1, 1, 1], [1, 1, 1, -1], [1, 1, -1, -1], [1, -1, -1, -1]], np.int8)


def e1_sign_hypotheses(chips: np.ndarray, n_fft: int) -> np.ndarray:
    """one satellite's four replicas (fse.hpp:371-385): the code upsampled to
    n_fft samples by floor index (Upsampler::Transform), then millisecond i
    (n_fft / 4 samples) multiplied by sign i of each permutation.  Acquire them
    as consecutive replicas with hypotheses=4."""
    if n_fft % 4:
        raise ValueError("the 4 ms code needs n_fft divisible by 4")
    chips = np.asarray(chips, np.float32)
    i = np.arange(n_fft, dtype=np.int64)
    rep = chips[(i * len(chips)) // n_fft]
    signs = np.repeat(E1_SIGN_PERMUTATIONS.astype(np.float32), n_fft // 4, axis=1)
    return np.ascontiguousarray(rep[None, :] * signs)


def e1_sign_chips(chips: np.ndarray) -> np.ndarray:
    """the same four hypotheses as chip sequences (a code of a multiple of 4
    chips: millisecond i is chips [i L/4, (i + 1) L/4)), for replica_put with
    repeats=1 when n_fft is a multiple of the code length: upsampled by
    floor index they equal e1_sign_hypotheses sample for sample, and they take
    the upsampled-code search"""
    chips = np.asarray(chips, np.int8)
    if len(chips) % 4:
        raise ValueError("the 4 ms code needs a multiple of 4 chips")
    signs = np.repeat(E1_SIGN_PERMUTATIONS, len(chips) // 4, axis=1)
    return np.ascontiguousarray(chips[None, :] * signs)


def adjust_sampling_rate(fs: float, target: float) -> float:
    return lib().ugsdr_adjust_sampling_rate(fs, target)


def resample_len(n: int, new_fs: int, old_fs: int) -> int:
    return lib().ugsdr_resample_len(n, new_fs, old_fs)


def format_bytes(fmt: int, n_samples: int, fs: float) -> int:
    """bytes holding n samples of a format (BBP_DDC: whole epochs with headers)"""
    b = lib().ugsdr_format_bytes(fmt, n_samples, fs)
    if b < 0:
        _check(b)
    return int(b)


def results_to_dicts(out) -> list:
    """ugsdr_acq_result array -> one dict per PRN"""
    return [{f: getattr(r, f) for f in RESULT_FIELDS} for r in out]


def corr_variants(n_fft: int) -> list:
    """workgroup sizes naming the correlation variants built for a length"""
    out = np.zeros(8, np.int32)
    n = lib().ugsdr_hip_corr_variants(n_fft, _ptr(out), 8)
    if n < 0:
        _check(n)
    return [int(v) for v in out[:min(n, 8)]]


def doppler_bins(doppler_range: float, doppler_step: float, center: float = 0.0) -> np.ndarray:
    """Doppler loop of FastSearchEngineBase::ProcessBpsk (src/acquisition/fse.hpp:250-252):
    f = -range; f <= range; f += step (accumulated in FP64)."""
    out = []
    f = -doppler_range
    while f <= doppler_range:
        out.append(f + center)
        f += doppler_step
    return np.array(out, np.float64)


def _c64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, np.complex64))


# ------------------------------------------------------------------ context
class Context:
    """One HIP device context (stream + device-resident caches)."""

    def __init__(self, device: int = 0):
        self._p = C.c_void_p()
        _check(lib().ugsdr_hip_ctx_create(device, C.byref(self._p)))
        self.device = device

    def close(self):
        if self._p:
            lib().ugsdr_hip_ctx_destroy(self._p)
            self._p = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._p

    def stream(self) -> int:
        s = C.c_void_p()
        _check(lib().ugsdr_hip_ctx_stream(self._p, C.byref(s)))
        return s.value or 0

    def device_bytes(self) -> int:
        return lib().ugsdr_hip_ctx_device_bytes(self._p)

    def autotune(self, n_fft: int):
        """(chosen workgroup size, per-candidate ms) for a transform length"""
        t = C.c_int32()
        ms = np.zeros(4, np.float32)
        _check(lib().ugsdr_hip_autotune(self._p, n_fft, C.byref(t), _ptr(ms), 4))
        return t.value, [float(v) for v in ms if v > 0]

    def set_autotune(self, enable: bool):
        _check(lib().ugsdr_hip_set_autotune(self._p, int(enable)))

    def set_graphs(self, enable: bool):
        """hipGraph replay of repeated request shapes (on by default)"""
        _check(lib().ugsdr_hip_set_graphs(self._p, int(enable)))

    def graph_replays(self) -> int:
        return int(lib().ugsdr_hip_graph_replays(self._p))

    def last_search_form(self) -> dict:
        """form of the last acquire: per-row transform points, code periods
        combined by the periodic fold, the upsampled-code factor (0: off)"""
        n, t, u = C.c_int64(), C.c_int32(), C.c_int32()
        _check(lib().ugsdr_hip_last_search_form(self._p, C.byref(n), C.byref(t), C.byref(u)))
        return {"n_transform": n.value, "periods": t.value, "samples_per_chip": u.value}

    def set_threads(self, n_fft: int, threads: int):
        """pin the correlation variant for a length (0 unpins)"""
        _check(lib().ugsdr_hip_set_threads(self._p, n_fft, threads))

    def set_profiling(self, enable: bool = True):
        _check(lib().ugsdr_hip_ctx_set_profiling(self._p, int(enable)))

    def last_times(self) -> dict:
        """HIP-event stage times (ms) of the last acquire: front, classes, corr, final"""
        ms = np.zeros(4, np.float32)
        _check(lib().ugsdr_hip_ctx_last_times(self._p, _ptr(ms)))
        return dict(zip(("front", "classes", "corr", "final"), map(float, ms)))

    # ---- replica cache
    def replica_put(self, chips: np.ndarray, repeats: int, n_fft: int, precision: int = PREC_FP32) -> int:
        chips = np.ascontiguousarray(chips, np.int8)
        key = C.c_uint64()
        _check(lib().ugsdr_hip_replica_put(self._p, _ptr(chips), len(chips), repeats, n_fft, precision,
                                           C.byref(key)))
        return key.value

    def replica_put_samples(self, samples: np.ndarray, precision: int = PREC_FP32) -> int:
        samples = np.ascontiguousarray(samples, np.float32)
        key = C.c_uint64()
        _check(lib().ugsdr_hip_replica_put_samples(self._p, _ptr(samples), len(samples), precision,
                                                   C.byref(key)))
        return key.value

    def replica_drop(self, key: int):
        _check(lib().ugsdr_hip_replica_drop(self._p, key))

    def replica_get(self, key: int, n_fft: int) -> np.ndarray:
        out = np.zeros(n_fft, np.complex64)
        _check(lib().ugsdr_hip_replica_get(self._p, key, _ptr(out), n_fft))
        return out

    # ---- fused acquisition
    def make_request(self, iq, fmt: int, *, fs_in: float, fs_acq: float, n_fft: int, n_blocks: int,
                     surface_len: int, doppler, keys, reshape: bool = True, coherent: bool = True,
                     mix_frequency: float = 0.0, mix_phase: float = 0.0, intermediate_frequency: float = 0.0,
                     threshold: float = 3.3, precision: int = PREC_FP32, n_samples: int | None = None,
                     iq_device_ptr: int | None = None, hypotheses: int = 0):
        dop = np.ascontiguousarray(doppler, np.float64)
        keys = np.ascontiguousarray(keys, np.uint64)
        req = AcqRequest()
        if iq_device_ptr is not None:
            req.iq = iq_device_ptr
            req.iq_on_device = 1
            req.n_samples = int(n_samples)
            hold = None
        else:
            arr = np.ascontiguousarray(iq)
            req.iq = arr.ctypes.data
            req.iq_on_device = 0
            if fmt in PACKED_FORMATS and n_samples is None:
                raise ValueError("packed recordings need n_samples")
            per = {FMT_CI8: 2, FMT_CI16: 2, FMT_CF32: 2, FMT_R8: 1}.get(fmt, 1)
            req.n_samples = int(n_samples if n_samples is not None else arr.size // per
                                if arr.dtype != np.complex64 else arr.size)
            hold = arr
        req.fmt = fmt
        req.fs_in, req.fs_acq = fs_in, fs_acq
        req.mix_frequency, req.mix_phase = mix_frequency, mix_phase
        req.intermediate_frequency = intermediate_frequency
        req.n_fft, req.n_blocks, req.surface_len = n_fft, n_blocks, surface_len
        req.reshape, req.coherent = int(reshape), int(coherent)
        req.doppler = dop.ctypes.data_as(C.POINTER(C.c_double))
        req.n_doppler = len(dop)
        req.replica_keys = keys.ctypes.data_as(C.POINTER(C.c_uint64))
        req.n_prn = len(keys)
        req.threshold = threshold
        req.precision = precision
        req.hypotheses = hypotheses
        return req, (hold, dop, keys)

    def acquire_req(self, req, want_surfaces: bool = False, want_peaks: bool = False, raw: bool = False,
                    noalloc: bool = False):
        """results as dicts per PRN, or per satellite with hypotheses > 1 (raw:
        the ugsdr_acq_result ctypes array, no conversion); noalloc runs the
        search through ugsdr_hip_acquire_noalloc"""
        P, D, L = req.n_prn, req.n_doppler, req.surface_len
        out = (AcqResult * (P // max(1, req.hypotheses)))()
        surf = np.zeros((P, D, L), np.float32) if want_surfaces else None
        peaks = np.zeros((P, L), np.float32) if want_peaks else None
        fn = lib().ugsdr_hip_acquire_noalloc if noalloc else lib().ugsdr_hip_acquire
        _check(fn(self._p, C.byref(req), out,
                                       _ptr(surf) if surf is not None else None,
                                       _ptr(peaks) if peaks is not None else None))
        res = out if raw else results_to_dicts(out)
        return res, surf, peaks

    def acquire(self, iq, fmt: int, want_surfaces: bool = False, want_peaks: bool = False, **kw):
        req, hold = self.make_request(iq, fmt, **kw)
        return self.acquire_req(req, want_surfaces, want_peaks)

    def acquire_noalloc(self, iq, fmt: int, want_surfaces: bool = False, want_peaks: bool = False, **kw):
        """the same search through the entry point that allocates no device
        memory (its fixed workspace; FP32 only)"""
        req, hold = self.make_request(iq, fmt, **kw)
        return self.acquire_req(req, want_surfaces, want_peaks, noalloc=True)

    # ---- tracking loops
    def track(self, iq, fmt: int, *, spm: int, n_blocks: int, fs: float, states, codes, spacing: float = 0.25,
              iq_device_ptr: int | None = None, want_trace: bool = True):
        """Tracker::Track on the device for every channel; states (a TrackState
        array or list) are advanced in place.  Returns (epl [n_blocks, n_ch, 3]
        complex64, trace [n_blocks, n_ch, 4] or None)"""
        n_ch = len(states)
        arr = (TrackState * n_ch)(*states) if not isinstance(states, C.Array) else states
        codes = np.ascontiguousarray(np.stack([np.asarray(c, np.int8) for c in codes]))
        req = TrackRequest()
        if iq_device_ptr is not None:
            req.iq, req.iq_on_device = iq_device_ptr, 1
            hold = None
        else:
            hold = np.ascontiguousarray(iq)
            req.iq, req.iq_on_device = hold.ctypes.data, 0
        req.fmt, req.spm, req.n_blocks, req.fs = fmt, spm, n_blocks, fs
        req.n_channels = n_ch
        req.codes = codes.ctypes.data_as(C.POINTER(C.c_int8))
        req.code_len = codes.shape[1]
        req.spacing_chips = spacing
        epl = np.zeros((n_blocks, n_ch, 3), np.complex64)
        trace = np.zeros((n_blocks, n_ch, 4), np.float64) if want_trace else None
        _check(lib().ugsdr_hip_track(self._p, C.byref(req), arr, _ptr(epl),
                                     trace.ctypes.data if trace is not None else None))
        if not isinstance(states, C.Array):
            for i in range(n_ch):
                states[i] = arr[i]
        return epl, trace

    # ---- anti-jamming (STFT)
    def stft(self, x, window: int = STFT_BLACKMAN_HARRIS) -> np.ndarray:
        """ShortTimeFourierTransform::Transform: (n - 96) // 32 columns of 128"""
        x = _c64(x)
        ncol = (len(x) - 96) // 32
        out = np.zeros((max(ncol, 0), 128), np.complex64)
        _check(lib().ugsdr_hip_stft(self._p, _ptr(x), len(x), window, _ptr(out)))
        return out

    def istft(self, cols) -> np.ndarray:
        """OverlapAndAdd of the columns: 128 + 32 (n_cols - 1) samples"""
        cols = np.ascontiguousarray(cols, np.complex64)
        out = np.zeros(128 + 32 * (len(cols) - 1), np.complex64)
        _check(lib().ugsdr_hip_istft(self._p, _ptr(cols), len(cols), _ptr(out)))
        return out

    def jse_process(self, x, fs: float, window: int = STFT_BLACKMAN_HARRIS, device_ptr: int | None = None,
                    n: int | None = None):
        """JammingSuppressionEngine::Process.  Host samples: returns (cleaned
        copy, detections); device_ptr (n complex64 in device memory): cleaned
        in place, returns (None, detections).  Detections are tuples (type,
        bin, frequency, j/n)."""
        det = (Jammer * 16)()
        cnt = C.c_int32()
        if device_ptr is not None:
            _check(lib().ugsdr_hip_jse_process(self._p, device_ptr, int(n), fs, window, 1, det, 16, C.byref(cnt)))
            y = None
        else:
            y = _c64(x).copy()
            _check(lib().ugsdr_hip_jse_process(self._p, _ptr(y), len(y), fs, window, 0, det, 16, C.byref(cnt)))
        return y, [(d.type, d.frequency_bin, d.frequency, d.jammer_to_noise) for d in det[:min(cnt.value, 16)]]

    # ---- recorded formats
    def unpack(self, raw, fmt: int, n: int, fs: float) -> np.ndarray:
        """SignalParametersBase::GetPartialSignal from the start of a recording:
        n samples of fmt (bytes in raw) decoded on the device to complex64"""
        raw = np.ascontiguousarray(raw)
        need = format_bytes(fmt, n, fs)
        if raw.nbytes < need:
            raise ValueError(f"{n} samples of format {fmt} need {need} bytes, got {raw.nbytes}")
        out = np.zeros(n, np.complex64)
        _check(lib().ugsdr_hip_unpack(self._p, raw.ctypes.data, fmt, n, fs, _ptr(out)))
        return out

    # ---- per-operator drop-ins
    def mix(self, x, fs: float, f: float, phase: float = 0.0) -> np.ndarray:
        x = _c64(x)
        out = np.empty_like(x)
        _check(lib().ugsdr_hip_mix(self._p, _ptr(x), _ptr(out), len(x), fs, f, phase))
        return out

    def mix_inplace(self, x, fs: float, f: float, phase: float = 0.0) -> np.ndarray:
        x = _c64(x).copy()
        _check(lib().ugsdr_hip_mix_inplace(self._p, _ptr(x), len(x), fs, f, phase))
        return x

    def table_mix(self, x, fs: float, f: float, phase: float = 0.0) -> np.ndarray:
        """TableMixer through Mixer::Translate's in-place overload
        (src/mixer/table_mixer.hpp:78-101, mixer.hpp:33-41)."""
        x = _c64(x).copy()
        _check(lib().ugsdr_hip_table_mix(self._p, _ptr(x), len(x), fs, f, phase, 0))
        return x

    def table_mix_device(self, ptr: int, n: int, fs: float, f: float, phase: float = 0.0) -> None:
        """the same on n complex float samples in device memory, in place"""
        _check(lib().ugsdr_hip_table_mix(self._p, ptr, int(n), fs, f, phase, 1))

    def table_mixer_member(self, x, fs: float, f: float, phase: float):
        """Mixer<TableMixer>'s member Translate (mixer.hpp:53-60): mixes with
        the stored phase, returns (samples, phase + 2 pi frac(f/1000))."""
        y = self.table_mix(x, fs, f, phase)
        m = math.fmod(f / 1000.0, 1.0)
        if m < 0:
            m += 1
        return y, phase + 2 * math.pi * m

    def dft(self, x, inverse: bool = False) -> np.ndarray:
        x = _c64(x)
        out = np.empty_like(x)
        _check(lib().ugsdr_hip_dft(self._p, _ptr(x), _ptr(out), len(x), int(inverse)))
        return out

    def dft_reduced(self, x, precision: int, inverse: bool = False) -> np.ndarray:
        """batched [batch, n] transform on the matrix cores with 16-bit operands"""
        x = np.ascontiguousarray(x, np.complex64)
        if x.ndim == 1:
            x = x[None]
        out = np.empty_like(x)
        _check(lib().ugsdr_hip_dft_reduced(self._p, precision, int(inverse), _ptr(x), _ptr(out), x.shape[1],
                                           x.shape[0]))
        return out

    def mf_prepare(self, code) -> np.ndarray:
        code = np.ascontiguousarray(code, np.float32)
        out = np.zeros(len(code), np.complex64)
        _check(lib().ugsdr_hip_mf_prepare(self._p, _ptr(code), _ptr(out), len(code)))
        return out

    def mf_filter(self, x, spec) -> np.ndarray:
        x, spec = _c64(x), _c64(spec)
        out = np.empty_like(x)
        _check(lib().ugsdr_hip_mf_filter(self._p, _ptr(x), _ptr(spec), _ptr(out), len(x)))
        return out

    def mf_filter_code(self, x, code) -> np.ndarray:
        x = _c64(x)
        code = np.ascontiguousarray(code, np.float32)
        out = np.empty_like(x)
        _check(lib().ugsdr_hip_mf_filter_code(self._p, _ptr(x), _ptr(code), _ptr(out), len(x)))
        return out

    def abs(self, x) -> np.ndarray:
        x = _c64(x)
        out = np.zeros(len(x), np.float32)
        _check(lib().ugsdr_hip_abs(self._p, _ptr(x), _ptr(out), len(x)))
        return out

    def reshape_sum(self, x, block: int) -> np.ndarray:
        if np.iscomplexobj(x):
            x = _c64(x)
            out = np.zeros(block, np.complex64)
            _check(lib().ugsdr_hip_reshape_sum(self._p, _ptr(x), _ptr(out), len(x), block, 1))
        else:
            x = np.ascontiguousarray(x, np.float32)
            out = np.zeros(block, np.float32)
            _check(lib().ugsdr_hip_reshape_sum(self._p, _ptr(x), _ptr(out), len(x), block, 0))
        return out

    def max_index(self, x):
        x = np.ascontiguousarray(x, np.float32)
        v, i = C.c_float(), C.c_int64()
        _check(lib().ugsdr_hip_max_index(self._p, _ptr(x), len(x), C.byref(v), C.byref(i)))
        return v.value, i.value

    def mean_std(self, x):
        x = np.ascontiguousarray(x, np.float32)
        m, s = C.c_float(), C.c_float()
        _check(lib().ugsdr_hip_mean_std(self._p, _ptr(x), len(x), C.byref(m), C.byref(s)))
        return m.value, s.value

    def resample(self, x, new_fs: int, old_fs: int) -> np.ndarray:
        x = _c64(x)
        out = np.zeros(resample_len(len(x), new_fs, old_fs), np.complex64)
        _check(lib().ugsdr_hip_resample(self._p, _ptr(x), len(x), new_fs, old_fs, _ptr(out)))
        return out

    def upsample(self, x, samples: int) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros(samples, np.float32)
        _check(lib().ugsdr_hip_upsample(self._p, _ptr(x), len(x), _ptr(out), samples))
        return out

    def correlate(self, signal, code) -> complex:
        s = _c64(signal)
        c = np.ascontiguousarray(code, np.float32)
        out = np.zeros(1, np.complex64)
        _check(lib().ugsdr_hip_correlate(self._p, _ptr(s), len(s), _ptr(c), len(c), _ptr(out)))
        return complex(out[0])

    # ---- correlator bank
    def correlate_bank(self, iq, fmt: int, *, spm: int, n_blocks: int, fs: float, channels, codes, taps_chips,
                       mode: int = BANK_NCO, samples_per_chip: float = 0.0, code_period_ms: float = 1.0,
                       iq_device_ptr: int | None = None):
        """channels: list of (carrier_frequency, carrier_phase, code_rate, code_phase); returns
        (out[n_blocks, n_ch, n_taps] complex64, advanced channel states)."""
        codes = np.ascontiguousarray(codes, np.int8)
        taps = np.ascontiguousarray(taps_chips, np.float64)
        n_ch = len(channels)
        ch = (ChannelState * n_ch)(*[ChannelState(*c) for c in channels])
        req = BankRequest()
        hold = None
        if iq_device_ptr is not None:
            req.iq, req.iq_on_device = iq_device_ptr, 1
        else:
            hold = np.ascontiguousarray(iq)
            req.iq, req.iq_on_device = hold.ctypes.data, 0
        req.fmt, req.spm, req.n_blocks, req.fs, req.mode = fmt, spm, n_blocks, fs, mode
        req.n_channels, req.n_taps = n_ch, len(taps)
        req.taps_chips = taps.ctypes.data_as(C.POINTER(C.c_double))
        req.codes = codes.ctypes.data_as(C.POINTER(C.c_int8))
        req.code_len = codes.shape[1]
        req.samples_per_chip, req.code_period_ms = samples_per_chip, code_period_ms
        out = np.zeros((n_blocks, n_ch, len(taps)), np.complex64)
        _check(lib().ugsdr_hip_correlate_bank(self._p, C.byref(req), ch, _ptr(out)))
        del hold
        states = [(c.carrier_frequency, c.carrier_phase, c.code_rate, c.code_phase) for c in ch]
        return out, states


# ------------------------------------------------------------------ server
class Server:
    """Continuous-batching acquisition server on a context (ugsdr_hip_server_*)."""

    def __init__(self, ctx: Context, max_device_bytes: int = 0, max_inflight: int = 0, n_streams: int = 0,
                 max_batch: int = 0, round_budget_us: int = 0, round_graphs: bool = False):
        self._p = C.c_void_p()
        cfg = ServerConfig(max_device_bytes, max_inflight, n_streams, max_batch, round_budget_us, int(round_graphs))
        _check(lib().ugsdr_hip_server_create(ctx.handle, C.byref(cfg), C.byref(self._p)))
        self._n_prn = {}

    def submit(self, req: AcqRequest) -> int:
        t = C.c_uint64()
        _check(lib().ugsdr_hip_server_submit(self._p, C.byref(req), C.byref(t)))
        self._n_prn[t.value] = req.n_prn // max(1, req.hypotheses)
        return t.value

    def poll(self, ticket: int, wait_ms: int = -1):
        """(results, latency_ms) when finished, None while pending"""
        n = self._n_prn.get(ticket, 0)
        out = (AcqResult * max(n, 1))()
        lat = C.c_double()
        rc = lib().ugsdr_hip_server_poll(self._p, ticket, wait_ms, out, C.byref(lat))
        if rc == PENDING:
            return None
        self._n_prn.pop(ticket, None)
        _check(rc)
        return [{f: getattr(out[i], f) for f in RESULT_FIELDS} for i in range(n)], lat.value

    def stats(self) -> dict:
        st = ServerStats()
        _check(lib().ugsdr_hip_server_stats(self._p, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self._p:
            lib().ugsdr_hip_server_destroy(self._p)
            self._p = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
