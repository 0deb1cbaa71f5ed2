"""Recorded sample formats on the CPU: the oracle's decoders and the epoch
geometry, pinned independently of the oracle's own code.

* BbpDdc and Nt1065 byte layouts: every byte value decoded through ctypes
  bit-field structures (the platform ABI lays them out, as it does the
  reference's structs, src/helpers/BbpPackedSpan.hpp:11-19 and
  src/helpers/NtlabPackedSpan.hpp:10-21);
* the reference's Nt1065 first-sample KAT (test/integration_tests.cpp:144-176:
  channels 1-4 read 1, 1, 1, 3);
* the epoch length rule (src/signal_parameters.hpp:74-80) against the sizes of
  the reference's BbpDdc recordings (data/bbp_ddc_gps_L1.dat and _L2: 100 epochs
  of 16640 bytes; _L5: 100 of 33264) and the C-ABI's ugsdr_format_bytes;
* the header check (signal_parameters.hpp:36-45) on the recording's first 4
  epochs (tests/golden/bbp_ddc_gps_L1_first4ms.bin) and on a damaged copy;
* the oracle's acquisition of those 4 ms against a separate numpy search (mask
  decoding, numpy FFT) at the recording's own rate.
"""
import ctypes as C
import os

import numpy as np
import pytest

import ugsdr_amd as U

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BBP_L1 = os.path.join(GOLDEN, "bbp_ddc_gps_L1_first4ms.bin")
BBP_FS = 33.25e6


class _Bbp(C.Structure):
    _fields_ = [("im_0", C.c_int8, 2), ("re_0", C.c_int8, 2), ("im_1", C.c_int8, 2), ("re_1", C.c_int8, 2)]


def _nt_struct(offset):
    fields = ([("skip", C.c_uint8, offset)] if offset else []) + [("sign", C.c_uint8, 1), ("magnitude", C.c_uint8, 1)]
    return type(f"_Nt{offset}", (C.Structure,), {"_fields_": fields})


def test_bbp_bitfield_layout(oracle):
    spm = 512
    data = np.zeros(oracle.bbp_epoch_bytes(spm * 1e3), np.uint8)   # header, 256 bytes, padding
    data[8:8 + 256] = np.arange(256, dtype=np.uint8)
    got = oracle.unpack_bbp(data, spm, spm * 1e3)
    for b in range(256):
        s = _Bbp.from_buffer_copy(bytes([b]))
        assert got[2 * b] == complex(s.re_0 | 1, s.im_0 | 1), b
        assert got[2 * b + 1] == complex(s.re_1 | 1, s.im_1 | 1), b
    # every sample is one of the four unit-magnitude points
    assert set(np.unique(got)) == {1 + 1j, 1 - 1j, -1 + 1j, -1 - 1j}


@pytest.mark.parametrize("channel,offset", [(1, 6), (2, 4), (3, 2), (4, 0)])
def test_nt1065_bitfield_layout(oracle, channel, offset):
    data = np.arange(256, dtype=np.uint8)
    got = oracle.unpack_nt1065(data, 256, offset)
    S = _nt_struct(offset)
    for b in range(256):
        s = S.from_buffer_copy(bytes([b]))
        assert got[b] == (2 * s.sign - 1) * (1 + 2 * s.magnitude), (channel, b)
    assert np.all(got.imag == 0)


def test_nt1065_reference_kat(oracle):
    """the four channels of the reference recording's first sample read 1, 1, 1, 3;
    exactly one byte value decodes to that, and the oracle agrees on it"""
    offsets = (6, 4, 2, 0)
    data = np.arange(256, dtype=np.uint8)
    vals = np.stack([oracle.unpack_nt1065(data, 256, o).real for o in offsets], axis=1)
    hits = [b for b in range(256) if tuple(vals[b]) == (1, 1, 1, 3)]
    assert hits == [0x57]


@pytest.mark.parametrize("fs,epoch", [(33.25e6, 16640), (66.5e6, 33264)])
def test_bbp_epoch_bytes_match_recordings(oracle, fs, epoch):
    # data/bbp_ddc_gps_L1.dat, _L2.dat: 1664000 bytes; _L5.dat: 3326400; 100 epochs each
    assert oracle.bbp_epoch_bytes(fs) == epoch
    assert U.format_bytes(U.FMT_BBP_DDC, int(fs / 1e3), fs) == epoch


def test_format_bytes_abi_matches_oracle(oracle):
    rng = np.random.default_rng(7)
    for _ in range(200):
        khz = int(rng.integers(2, 120000))
        fs = khz * 1e3
        n = int(rng.integers(1, 5)) * khz - int(rng.integers(0, khz))
        assert U.format_bytes(U.FMT_BBP_DDC, n, fs) == -(-n // khz) * oracle.bbp_epoch_bytes(fs), fs
    assert U.format_bytes(U.FMT_CI16, 10, 4e6) == 40
    assert U.format_bytes(U.FMT_NT1065_3, 10, 79.5e6) == 10
    with pytest.raises(U.UgsdrError):
        U.format_bytes(U.FMT_BBP_DDC, 10, 33.2505e6)   # not a whole number of kHz
    with pytest.raises(U.UgsdrError):
        U.format_bytes(42, 10, 1e6)


def test_bbp_recording_headers(oracle):
    raw = np.fromfile(BBP_L1, np.uint8)
    eb = oracle.bbp_epoch_bytes(BBP_FS)
    assert raw.size == 4 * eb
    assert oracle.bbp_verify_headers(raw, 4, eb)
    bad = raw.copy()
    bad[2 * eb] ^= 1                   # epoch 2's counter off by one
    assert not oracle.bbp_verify_headers(bad, 4, eb)


def _numpy_decode(raw, spm, eb, n_ep):
    out = []
    for e in range(n_ep):
        d = raw[e * eb + 8: e * eb + 8 + spm // 2]
        s = np.empty(spm, np.complex128)
        s[0::2] = np.where(d & 0x08, -1, 1) + 1j * np.where(d & 0x02, -1, 1)
        s[1::2] = np.where(d & 0x80, -1, 1) + 1j * np.where(d & 0x20, -1, 1)
        out.append(s)
    return np.concatenate(out)


def bbp_recording_search(oracle, precision=64):
    """oracle chain on the 4 ms recording: decode, resample 33.25 -> 4.2 MHz
    (Resampler::Transform: lcm hold / boxcar), acquire GPS L1 C/A 1 ms x 4"""
    raw = np.fromfile(BBP_L1, np.uint8)
    spm_in = int(BBP_FS / 1e3)
    x = oracle.unpack_bbp(raw, 4 * spm_in, BBP_FS)
    fs_acq = 4.2e6
    y = oracle.resample(x, int(fs_acq), int(BBP_FS))
    N = int(fs_acq / 1e3)
    codes = [oracle.gps_l1ca(s) for s in range(32)]
    reps = np.stack([oracle.repeat_upsample(c, 1, N) for c in codes])
    dop = U.doppler_bins(5000, 250)
    res, _, _ = oracle.acquire(y, reps, n_fft=N, n_blocks=4, surface_len=N, fs_in=BBP_FS, fs_acq=fs_acq,
                               doppler=dop, precision=precision)
    return res, dop


def test_bbp_recording_acquisition_oracle(oracle):
    """PRNs 9 and 23 (sv 8, 22) stand out in the recording; the oracle's code
    offsets and Doppler agree with a plain numpy search at 33.25 MHz"""
    res, dop = bbp_recording_search(oracle)
    det = {p for p, r in enumerate(res) if r["detected"]}
    assert {8, 22} <= det
    raw = np.fromfile(BBP_L1, np.uint8)
    spm = int(BBP_FS / 1e3)
    x = _numpy_decode(raw, spm, oracle.bbp_epoch_bytes(BBP_FS), 4)
    t = np.arange(4 * spm) / BBP_FS
    for sv in (8, 22):
        c = oracle.gps_l1ca(sv).astype(np.float64)
        rep = c[(np.arange(spm) * 1.023e6 / BBP_FS).astype(np.int64) % 1023]
        R = np.conj(np.fft.fft(rep))
        fd = res[sv]["doppler"]
        y = x * np.exp(-2j * np.pi * fd * t)
        acc = sum(np.abs(np.fft.ifft(np.fft.fft(y[k * spm:(k + 1) * spm]) * R)) for k in range(4))
        peak = int(np.argmax(acc))
        # one output sample at 4.2 MHz spans 7.9 input samples
        d = (res[sv]["code_offset"] - peak + spm / 2) % spm - spm / 2
        assert abs(d) <= 12, (sv, res[sv]["code_offset"], peak)
