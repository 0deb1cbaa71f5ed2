"""Parity rules of the acquisition path (GPU result vs the FP64 oracle).

* same detection flag unless the oracle metric is marginal (|snr - thr| <= 1e-3 thr)
* bit-exact surface index unless the oracle's best row has a near-tie
  (runner-up within NEAR_TIE relative of the max)
* bit-exact Doppler bin unless another bin's level is within NEAR_TIE
* level and snr within 1e-3 relative
"""
import numpy as np

NEAR_TIE = 1e-5
METRIC_RTOL = 1e-3


def compare(gpu, ref, threshold=3.3, rtol=METRIC_RTOL):
    """Returns (failures, notes)."""
    fails, notes = [], []
    for p, (g, r) in enumerate(zip(gpu, ref)):
        lvl = r["level"]
        if lvl <= 0:
            if g["level"] != 0 or g["doppler_bin"] != -1:
                fails.append(f"prn{p}: oracle level 0, gpu {g['level']}")
            continue
        code_tie = (lvl - r["runner_up_code"]) / lvl < NEAR_TIE
        dop_tie = (lvl - r["runner_up_doppler"]) / lvl < NEAR_TIE
        marginal = abs(r["snr"] - threshold) <= 1e-3 * threshold
        if dop_tie:
            notes.append(f"prn{p}: Doppler near-tie")
        elif g["doppler_bin"] != r["doppler_bin"]:
            fails.append(f"prn{p}: doppler bin {g['doppler_bin']} != {r['doppler_bin']}")
        if code_tie:
            notes.append(f"prn{p}: code near-tie")
        elif not dop_tie and g["index"] != r["index"]:
            fails.append(f"prn{p}: index {g['index']} != {r['index']}")
        if marginal:
            notes.append(f"prn{p}: marginal metric {r['snr']:.6f}")
        elif g["detected"] != r["detected"]:
            fails.append(f"prn{p}: detected {g['detected']} != {r['detected']}")
        for k in ("level", "snr", "sigma"):
            if abs(g[k] - r[k]) > rtol * abs(r[k]):
                fails.append(f"prn{p}: {k} {g[k]} vs {r[k]}")
        if not dop_tie and g["doppler"] != r["doppler"]:
            fails.append(f"prn{p}: doppler {g['doppler']} != {r['doppler']}")
        if not dop_tie and not code_tie and g["code_offset"] != r["code_offset"]:
            fails.append(f"prn{p}: code_offset {g['code_offset']} != {r['code_offset']}")
    return fails, notes


def iq_to_complex(iq: np.ndarray) -> np.ndarray:
    return (iq[0::2].astype(np.float32) + 1j * iq[1::2].astype(np.float32)).astype(np.complex64)
