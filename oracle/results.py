"""ORACLE (test infrastructure only): S-parameter conversions, restated in numpy.

Follows:
* ``results.cpp:419-453``  Result::get_k           k = diag(sqrt(Zo))
* ``results.cpp:455-511``  Result::S2Z             Z = k (I + S)(I - S)^-1 k
* ``results.cpp:513-560``  Result::Z2S             S = (Z k^-1 + k)^-1 (Z k^-1 - k)
* ``results.cpp:562-578``  Result::renormalize     S2Z with the port impedances, Z2S with Zref at every port

Pinned in tests/test_results.py by closed forms that do not use these formulas: the impedance matrix of a lossless
line section, the one-port reflection coefficient, a matched load seen from another reference impedance.
"""
from __future__ import annotations

import numpy as np


def _k(Zo, n):
    return np.diag(np.sqrt(np.broadcast_to(np.asarray(Zo, np.complex128), (n,))))


def s2z(S, Zo):
    n = S.shape[0]
    k = _k(Zo, n)
    eye = np.eye(n)
    return k @ (eye + S) @ np.linalg.inv(eye - S) @ k


def z2s(Z, Zo):
    n = Z.shape[0]
    k = _k(Zo, n)
    zk = Z @ np.linalg.inv(k)
    return np.linalg.inv(zk + k) @ (zk - k)


def renormalize(S, Zo, Zref):
    return z2s(s2z(S, Zo), Zref)
