"""The FP32 erfc used by the pair kernels (csrc/common.h: kErfcP, kErfcQ).

erfc(x) = exp(-x^2) * t * Q(t), t = 1 / (1 + p x).  The coefficients are read
from the header that the kernels compile, evaluated here with FP32 Horner
steps as the kernels do, and checked against scipy's erfc / erfcx.
"""
import os
import re

import numpy as np
from scipy.special import erfc, erfcx

HDR = os.path.join(os.path.dirname(__file__), "..", "mpidopenmmplugin_amd", "csrc", "common.h")


def _coefficients():
    text = open(HDR).read()
    p = float(re.search(r"kErfcP\s*=\s*([-0-9.eE]+)f", text).group(1))
    body = re.search(r"kErfcQ\[9\]\s*=\s*\{([^}]*)\}", text).group(1)
    q = [float(v.strip().rstrip("f")) for v in body.split(",") if v.strip()]
    assert len(q) == 9
    return p, q


def _erfcx_fp32(x):
    p, q = _coefficients()
    x = np.asarray(x, dtype=np.float32)
    t = np.float32(1.0) / (np.float32(1.0) + np.float32(p) * x)
    acc = np.full_like(x, np.float32(q[8]))
    for c in reversed(q[:8]):
        acc = acc * t + np.float32(c)
    return (acc * t).astype(np.float64)


def test_fitted_erfcx_relative_error():
    x = np.linspace(0.0, 9.0, 200001).astype(np.float32)
    rel = np.abs(_erfcx_fp32(x) / erfcx(x.astype(np.float64)) - 1.0)
    assert rel.max() < 4e-7


def test_fitted_erfc_with_fp32_exponential():
    # the Ewald range of the workloads: alpha * r <= alpha * cutoff ~ 2.6, with margin
    x = np.linspace(0.0, 5.0, 100001).astype(np.float32)
    e = np.exp(-(x * x)).astype(np.float32)
    approx = e.astype(np.float64) * _erfcx_fp32(x)
    rel = np.abs(approx / erfc(x.astype(np.float64)) - 1.0)
    assert rel.max() < 3e-6  # exp(-x^2) in FP32 carries ~x^2 ulp; the fit adds < 4e-7


def test_fit_reproduces_closed_values():
    # erfc(0) = 1, and the large-x asymptote erfcx(x) ~ 1 / (x sqrt(pi))
    assert abs(_erfcx_fp32(np.array([0.0]))[0] - 1.0) < 2e-7
    x = 9.0
    assert abs(_erfcx_fp32(np.array([x]))[0] * x * np.sqrt(np.pi) - (1 - 1 / (2 * x * x))) < 5e-3
