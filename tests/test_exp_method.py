"""The bilateral filter's FP64 exponential (Tang's table-driven method,
kinectfusion_amd/csrc/kf_exp.hpp), compiled from the same header for the host
with g++, against Python's math.exp: relative error below 4e-16 over the
filter's argument range, and the float it rounds to equal to the float
math.exp rounds to."""
import ctypes as C
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = r"""
#include "kf_exp.hpp"
static const double kTable[64] = KF_EXP2_J64_TABLE;
extern "C" void exp_batch(const double* x, double* y, long n) {
  for (long i = 0; i < n; ++i) y[i] = kf::exp_nonpositive(x[i], kTable);
}
"""


@pytest.fixture(scope="module")
def exp_batch(tmp_path_factory):
    d = tmp_path_factory.mktemp("exp")
    (d / "exp.cpp").write_text(SRC)
    lib = d / "libexp.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC",
                    f"-I{ROOT / 'kinectfusion_amd' / 'csrc'}", str(d / "exp.cpp"), "-o", str(lib)], check=True)
    so = C.CDLL(str(lib))
    so.exp_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_long]

    def run(x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        so.exp_batch(x.ctypes.data, y.ctypes.data, x.size)
        return y

    return run


def test_relative_error_and_float_rounding(exp_batch):
    rng = np.random.default_rng(3)
    # the filter's exponents: spatial (-d^2 / 2 sigma_s^2) plus range terms, down to underflow
    x = np.concatenate([-rng.exponential(2.0, 200000), -rng.uniform(0.0, 110.0, 200000),
                        -rng.uniform(0.0, 1e-6, 20000)])
    y = exp_batch(x)
    ref = np.array([math.exp(v) for v in x])
    rel = np.abs(y - ref) / ref
    assert rel.max() < 4e-16
    assert np.array_equal(y.astype(np.float32), ref.astype(np.float32))


def test_special_arguments(exp_batch):
    ln2_64 = math.log(2.0) / 64
    x = np.array([0.0, -0.0, -1e-300, -ln2_64 / 2, -ln2_64, -math.log(2.0), -1.0, -87.3, -103.9, -109.99])
    y = exp_batch(x)
    ref = np.array([math.exp(v) for v in x])
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == 1.0
    assert y[5] == 0.5  # exp(-ln2) = 2^-1 to the last bit
    assert np.all(np.abs(y - ref) <= 4e-16 * ref)
    # below -110 the float result is 0
    assert np.all(exp_batch(np.array([-110.5, -1e3, -np.inf])) == 0.0)
    assert np.float32(math.exp(-110.0)) == 0.0
