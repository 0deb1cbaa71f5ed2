"""Writes two more slices of the reference's recordings (data only; run in a
checkout that has the reference mounted):

* bbp_ddc_gps_L5_first4ms.bin: the first 4 epochs (4 x 33264 bytes) of the
  BbpDdc recording data/bbp_ddc_gps_L5.dat.  33264-byte epochs bound its rate
  to (66.480, 66.512] MHz (epoch length 16 * ceil((spm + 16) / 32) bytes,
  src/signal_parameters.hpp:74-80); 66.5 MHz, twice the L1 recording's
  33.25 MHz, is the whole-kHz rate used, at the L5 centre 1176.45 MHz.
* iq_16_plus_16_first4ms.bin: the first 4 ms (4 x 4000 interleaved int16 I/Q
  pairs, 64000 bytes) of data/iq_16_plus_16.bin, recorded at 4 MHz centred on
  1575.42 MHz (test/integration_tests.cpp:75, :105-121); its first sample is
  (-150, 16)."""
import os
import sys

REF = os.environ.get("UGSDR_REFERENCE", "/root/reference") + "/data/"
HERE = os.path.dirname(os.path.abspath(__file__))
SLICES = [("bbp_ddc_gps_L5.dat", "bbp_ddc_gps_L5_first4ms.bin", 4 * 33264),
          ("iq_16_plus_16.bin", "iq_16_plus_16_first4ms.bin", 4 * 4000 * 4)]

if __name__ == "__main__":
    for src, dst, n in SLICES:
        with open(REF + src, "rb") as f:
            data = f.read(n)
        if len(data) != n:
            sys.exit(f"short read of {src}")
        with open(os.path.join(HERE, dst), "wb") as f:
            f.write(data)
        print(dst, len(data))
