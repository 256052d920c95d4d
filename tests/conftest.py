import ctypes as C
import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, GOLDEN)

FIXTURES = sorted(os.path.splitext(os.path.basename(p))[0]
                  for p in glob.glob(os.path.join(GOLDEN, "*.npz")))
LIST_KEYS = (("t", "inpt", "jnpt"), ("s", "inps", "jnps"),
             ("te", "inpte", "jnpte"), ("se", "inpse", "jnpse"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcm3d_nb.so)")


def load_fixture(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


@pytest.fixture(params=FIXTURES)
def fixture(request):
    return request.param, load_fixture(request.param)


def rel_rms(a, b, floor=1.0):
    """RMS(a - b) / max(RMS(b), floor).  The floor (1 K/A for forces) keeps
    the measure meaningful where forces vanish by symmetry (a perfect
    lattice), where both sides are pure rounding noise."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(float(np.sqrt(np.mean(b * b))), floor)
    return float(np.sqrt(np.mean((a - b) ** 2)) / den)


H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip_runtime():
    """The HIP runtime the library itself runs on (device buffers for the
    GPU tests; no second runtime in the process)."""
    from cm3d_amd import nb
    nb.lib()
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


class DevBuf:
    """A device copy of a float64 host array."""

    def __init__(self, h, host):
        self.h, self.n = h, host.nbytes
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), self.n) == 0
        assert h.hipMemcpy(self.p, host.ctypes.data, self.n, H2D) == 0

    def get(self):
        out = np.empty(self.n // 8)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, self.n, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)
