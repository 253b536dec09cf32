"""Interpolative decompositions on the GPU (Common/ID.m, Common/LRID_rand.m)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int64)


def ID(A, par):
    """[T, I] = ID(A, par): par < 1 accuracy, par >= 1 fixed rank (Common/ID.m:1).
    Returns (T, I) with T k x (n-k) and I a 0-based column permutation (skeleton first)."""
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    perm = np.empty(n, dtype=np.int64)
    T = np.empty(n * n)
    k = C.c_int64()
    check(lib().hssc_id(A.ctypes.data_as(_DP), m, n, m, float(par), perm.ctypes.data_as(_IP),
                        T.ctypes.data_as(_DP), C.byref(k)))
    kk = k.value
    return T[:kk * (n - kk)].reshape((n - kk, kk)).T.copy(), perm


def LRID_rand(A, acc, k0, Cinc, seed=0):
    """K = LRID_rand(...) with dense matvecs (Common/LRID_rand.m:1): A ~ U @ V.T."""
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    r = min(m, n)
    U = np.empty((r, m)).T
    V = np.empty((r, n)).T
    Ub = np.zeros(m * r)
    Vb = np.zeros(n * r)
    J = np.empty(n, dtype=np.int64)
    k = C.c_int64()
    check(lib().hssc_lrid_rand(A.ctypes.data_as(_DP), m, n, m, float(acc), int(k0), int(Cinc),
                               int(seed), C.byref(k), J.ctypes.data_as(_IP),
                               Ub.ctypes.data_as(_DP), Vb.ctypes.data_as(_DP)))
    kk = k.value
    U = Ub[:m * kk].reshape((kk, m)).T.copy()
    V = Vb[:n * kk].reshape((kk, n)).T.copy()
    return U, V, J
