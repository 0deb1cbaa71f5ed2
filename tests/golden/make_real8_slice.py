"""Writes tests/golden/real_8_first4ms.bin: the first 4 ms (4 x 38192 bytes)
of the reference's recording data/real_8.bin (real int8 @ 38.192 MHz, IF 9.55
MHz), the input of its real_8 acquisition KAT (test/integration_tests.cpp:253-263,
:395-398).  Data only; run in a checkout that has the reference mounted."""
import os
import sys

SRC = os.environ.get("UGSDR_REFERENCE", "/root/reference") + "/data/real_8.bin"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "real_8_first4ms.bin")

if __name__ == "__main__":
    with open(SRC, "rb") as f:
        data = f.read(4 * 38192)
    if len(data) != 4 * 38192:
        sys.exit("short read")
    with open(DST, "wb") as f:
        f.write(data)
    print(DST, len(data))
