"""CPU check of the per-cell logarithm used by the free-energy kernels
(ampe_amd/csrc/mathfn.h): the header is compiled with the host compiler (the
device build runs the same IEEE operations: frexp, correctly rounded
division, fused multiply-adds) and compared with the C library's log, which
is accurate to within an ulp, over the ranges the kernels see (phase
compositions in (1e-12, 1), temperatures, and the full double range)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HDR = os.path.join(HERE, "..", "ampe_amd", "csrc", "mathfn.h")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("mathfn")
    src = d / "m.cpp"
    src.write_text(f'#include "{os.path.abspath(HDR)}"\n'
                   'extern "C" void logv(const double* x, double* y, long n)\n'
                   '{ for (long i = 0; i < n; i++) y[i] = ampe::log_pos(x[i]); }\n')
    so = d / "libm.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-o", str(so),
                    str(src)], check=True)
    L = C.CDLL(str(so))
    L.logv.argtypes = [C.c_void_p, C.c_void_p, C.c_long]
    return L


def _log(L, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    L.logv(x.ctypes.data, y.ctypes.data, x.size)
    return y


@pytest.mark.parametrize("lo,hi", [(1e-12, 1e-6), (1e-6, 0.5), (0.5, 1.5), (0.999, 1.001),
                                   (800.0, 2000.0), (1e-300, 1e300)])
def test_log_within_one_ulp(lib, lo, hi):
    rng = np.random.default_rng(5)
    x = np.exp(rng.uniform(np.log(lo), np.log(hi), 400_000))
    y, ref = _log(lib, x), np.log(x)
    ulp = np.abs(y.view(np.int64) - ref.view(np.int64))
    assert ulp.max() <= 1
    assert np.mean(ulp == 0) > 0.9


def test_log_special_values(lib):
    x = np.array([1.0, 2.0, 0.5, np.e, 1e-310, 5e-324, 0.0, np.inf, -1.0, np.nan])
    y = _log(lib, x)
    assert y[0] == 0.0
    assert y[1] == np.log(2.0) and y[2] == -np.log(2.0)
    assert abs(y[3] - 1.0) <= 2.3e-16
    for i in (4, 5):  # subnormal arguments
        assert abs(y[i] - np.log(x[i])) <= 1e-15 * abs(np.log(x[i]))
    assert y[6] == -np.inf and y[7] == np.inf
    assert np.isnan(y[8]) and np.isnan(y[9])
