"""Pin of the slow-force factor used by the pair kernel (no GPU).

The kernel evaluates fSlow/qq = beta^3 H(beta^2 r^2) - 1/rc^3 with the
rational P6/Q6 of namd-3.0_amd/csrc/slowforce.h.  Here the coefficients are
read from that header and evaluated in FP32 exactly as the kernel's Horner
scheme does, then compared with H computed in FP64 from its closed form
(erf and the Taylor series at small u), and the resulting force factor with
NAMD's formula (1 - erfc(a) - (2/sqrt(pi)) a e^{-a^2})/r^3 - 1/rc^3.
"""
import math
import os
import re

import numpy as np
from scipy.special import erf, erfc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _coeffs():
    txt = open(os.path.join(ROOT, "namd-3.0_amd", "csrc", "slowforce.h")).read()
    P = [float(v) for _, v in sorted((int(k), v) for k, v in re.findall(r"HP(\d) = ([-0-9.e+]+)f", txt))]
    Q = [1.0] + [float(v) for _, v in sorted((int(k), v) for k, v in re.findall(r"HQ(\d) = ([-0-9.e+]+)f", txt))]
    return np.array(P, np.float32), np.array(Q, np.float32)


def _H_exact(u):
    u = np.asarray(u, np.float64)
    out = np.empty_like(u)
    small = u < 0.5
    us = u[small]
    s = np.zeros_like(us)
    for n in range(30):
        s += (-1) ** n * us ** n / (math.factorial(n) * (2 * n + 3))
    out[small] = 4 / math.sqrt(math.pi) * s
    a = np.sqrt(u[~small])
    out[~small] = (erf(a) - 2 / math.sqrt(math.pi) * a * np.exp(-a * a)) / a ** 3
    return out


def _H_f32(u32, P, Q):
    p = P[-1]
    for c in P[-2::-1]:
        p = p * u32 + c
    q = Q[-1]
    for c in Q[-2::-1]:
        q = q * u32 + c
    return (p / q).astype(np.float64)


def test_rational_matches_closed_form():
    P, Q = _coeffs()
    assert P.size == 7 and Q.size == 7
    u = np.linspace(0, 16, 200001).astype(np.float32)
    rel = np.abs(_H_f32(u, P, Q) / _H_exact(u.astype(np.float64)) - 1)
    assert rel.max() < 5e-7
    assert np.sqrt((rel ** 2).mean()) < 1.5e-7
    assert abs(_H_exact(np.array([0.0]))[0] - 4 / (3 * math.sqrt(math.pi))) < 1e-15


def test_force_factor_matches_namd_formula():
    P, Q = _coeffs()
    beta, rc = 0.257952, 12.0
    r = np.linspace(0.5, rc, 50000)
    a = beta * r
    namd = (1 - erfc(a) - 2 / math.sqrt(math.pi) * a * np.exp(-a * a)) / r ** 3 - 1 / rc ** 3
    u = (beta * beta * r * r).astype(np.float32)
    ours = beta ** 3 * _H_f32(u, P, Q) - 1 / rc ** 3
    # absolute error relative to the Coulomb scale 1/r^3 of the same pair
    assert np.max(np.abs(ours - namd) * r ** 3) < 1e-6
