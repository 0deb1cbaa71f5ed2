"""Writes tests/golden/bbp_ddc_gps_L1_first4ms.bin: the first 4 epochs
(4 x 16640 bytes) of the reference's BbpDdc recording data/bbp_ddc_gps_L1.dat.
The recording is 100 epochs of 16640 bytes: an 8-byte header whose low 24
bits count milliseconds, then 2+2-bit complex samples (src/signal_parameters.hpp:
74-80, src/helpers/BbpPackedSpan.hpp:9-86).  16640-byte epochs fit a rate in
(33.232, 33.264] MHz; 33.25 MHz is the whole-kHz rate at which GPS L1 C/A
acquisition on it gives the largest peaks (PRNs 9 and 23).  Data only; run in
a checkout that has the reference mounted."""
import os
import sys

SRC = os.environ.get("UGSDR_REFERENCE", "/root/reference") + "/data/bbp_ddc_gps_L1.dat"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bbp_ddc_gps_L1_first4ms.bin")

if __name__ == "__main__":
    with open(SRC, "rb") as f:
        data = f.read(4 * 16640)
    if len(data) != 4 * 16640:
        sys.exit("short read")
    with open(DST, "wb") as f:
        f.write(data)
    print(DST, len(data))
