import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def ctx():
    import ugsdr_amd

    c = ugsdr_amd.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def oracle():
    from oracle import refcpu

    refcpu.lib()
    return refcpu


class DeviceArray:
    """Device memory from the HIP runtime the library itself uses (for tests
    of the device-pointer entry points)."""

    def __init__(self, host):
        import ctypes as C

        import numpy as np

        import ugsdr_amd

        ugsdr_amd.lib()   # the library's HIP runtime is loaded first
        self._hip = C.CDLL("libamdhip64.so.7")
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert self._hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.host.nbytes)) == 0
        assert self._hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), C.c_size_t(self.host.nbytes),
                                   1) == 0   # hipMemcpyHostToDevice
        assert self._hip.hipDeviceSynchronize() == 0

    def get(self):
        """a host copy of the device contents"""
        import ctypes as C

        import numpy as np

        out = np.empty_like(self.host)
        assert self._hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(out.nbytes), 2) == 0
        return out

    def close(self):
        if self.ptr:
            self._hip.hipFree(self.ptr)
            self.ptr = None
