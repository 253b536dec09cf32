"""Python mirror of the reference's HSS1D MATLAB entry points, backed by libhssc.so.

    K_HSS = HSS1D_compress_nsym('green', curve, nbox_max, acc)   # HSS1D_compress_nsym.m
    HSS1D_invert_nsym(K_HSS)                                      # HSS1D_invert_nsym.m
    x = HSS1D_applyinv_nsym(K_HSS, f)                             # HSS1D_applyinv_nsym.m
    u = HSS1D_apply_nsym(K_HSS, q)                                # HSS1D_apply_nsym.m

``curve`` is a Curve (points, outward normals, quadrature weights, curvature); the
operator is the second-kind Laplace double layer -1/2 I + D (see include/hssc.h).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import PREC, check, lib

_DP = C.POINTER(C.c_double)


@dataclass
class Curve:
    X: np.ndarray      # N x 2
    nrm: np.ndarray    # N x 2 outward unit normals
    w: np.ndarray      # N quadrature weights
    kappa: np.ndarray  # N curvatures

    @property
    def N(self):
        return self.X.shape[0]


def star_contour(N, a=0.3, m=5) -> Curve:
    """r(t) = 1 + a cos(m t) sampled at t_j = 2 pi j / N (trapezoidal rule)."""
    t = 2 * np.pi * np.arange(N) / N
    r = 1 + a * np.cos(m * t)
    rp = -a * m * np.sin(m * t)
    rpp = -a * m * m * np.cos(m * t)
    xp = rp * np.cos(t) - r * np.sin(t)
    yp = rp * np.sin(t) + r * np.cos(t)
    xpp = rpp * np.cos(t) - 2 * rp * np.sin(t) - r * np.cos(t)
    ypp = rpp * np.sin(t) + 2 * rp * np.cos(t) - r * np.sin(t)
    s = np.hypot(xp, yp)
    return Curve(np.stack([r * np.cos(t), r * np.sin(t)], 1), np.stack([yp / s, -xp / s], 1),
                 s * 2 * np.pi / N, (xp * ypp - yp * xpp) / s ** 3)


class HSS1D:
    def __init__(self, handle, N):
        self.handle = handle
        self.N = N

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            lib().hssc_hss1d_destroy(h)
            self.handle = None

    @property
    def nboxes(self):
        nb, d = C.c_int32(), C.c_int32()
        check(lib().hssc_hss1d_info(self.handle, C.byref(nb), C.byref(d)))
        return nb.value

    @property
    def depth(self):
        nb, d = C.c_int32(), C.c_int32()
        check(lib().hssc_hss1d_info(self.handle, C.byref(nb), C.byref(d)))
        return d.value

    def box(self, b):
        lev, n, k = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().hssc_hss1d_box(self.handle, int(b), C.byref(lev), C.byref(n), C.byref(k),
                                   None, None))
        up = np.empty(k.value, dtype=np.int64)
        dn = np.empty(k.value, dtype=np.int64)
        check(lib().hssc_hss1d_box(self.handle, int(b), None, None, None,
                                   up.ctypes.data_as(C.POINTER(C.c_int64)),
                                   dn.ctypes.data_as(C.POINTER(C.c_int64))))
        return dict(level=lev.value, nskel=n.value, kskel=k.value, isk_up=up, isk_dn=dn)

    def features(self, b):
        """Geometric features of box b fed to the rank network (hssc_hss1d_box_features)."""
        f = np.empty(8, dtype=np.float32)
        check(lib().hssc_hss1d_box_features(self.handle, int(b),
                                            f.ctypes.data_as(C.POINTER(C.c_float))))
        return f

    def id_stats(self):
        """(blocks, fallbacks to the plain ID, sketch rows, rows factored by the plain ID)."""
        v = [C.c_int64() for _ in range(4)]
        check(lib().hssc_hss1d_id_stats(self.handle, *[C.byref(x) for x in v]))
        return dict(zip(("blocks", "fallbacks", "sketch_rows", "full_rows"), (x.value for x in v)))


RANK_MODES = {"id": 0, "table": 1, "mlp": 2}


def HSS1D_compress_nsym(type_, curve: Curve, nbox_max, acc, nproxy=64, rank="id", mlp=None,
                        oversample=8) -> HSS1D:
    """HSS1D/HSS1D_compress_nsym.m:1 with type 'green'.

    ``rank`` picks how the IDs are sized: 'id' runs the plain ID on every block; 'table'
    sketches each block with the closed-form rank guess of HSS2D/predict_rank.m; 'mlp' with
    the prediction of ``mlp`` (a RankMLP).  Sketches that fail fall back to the plain ID."""
    if type_ != "green":
        raise ValueError("only the 'green' (proxy) compression is implemented")
    X = np.asfortranarray(curve.X, dtype=np.float64)
    Nm = np.asfortranarray(curve.nrm, dtype=np.float64)
    w = np.ascontiguousarray(curve.w, dtype=np.float64)
    k = np.ascontiguousarray(curve.kappa, dtype=np.float64)
    h = C.c_void_p()
    check(lib().hssc_hss1d_compress_ranked(
        X.ctypes.data_as(_DP), Nm.ctypes.data_as(_DP), w.ctypes.data_as(_DP),
        k.ctypes.data_as(_DP), X.shape[0], int(nbox_max), float(acc), int(nproxy),
        RANK_MODES[rank], mlp.handle if mlp is not None else None, int(oversample), C.byref(h)))
    return HSS1D(h, X.shape[0])


def HSS1D_compress_batched(curves, nbox_max, acc, nproxy=64, rank="id", mlp=None,
                           oversample=8) -> HSS1D:
    """Many curves in one operator (one tree per curve, shared launches per level); vectors
    are the curves' vectors concatenated in order (see split)."""
    Ns = np.array([c.N for c in curves], dtype=np.int64)
    X = np.asfortranarray(np.concatenate([c.X for c in curves]), dtype=np.float64)
    Nm = np.asfortranarray(np.concatenate([c.nrm for c in curves]), dtype=np.float64)
    w = np.ascontiguousarray(np.concatenate([c.w for c in curves]), dtype=np.float64)
    k = np.ascontiguousarray(np.concatenate([c.kappa for c in curves]), dtype=np.float64)
    h = C.c_void_p()
    check(lib().hssc_hss1d_compress_batched(
        len(curves), Ns.ctypes.data_as(C.POINTER(C.c_int64)), X.ctypes.data_as(_DP),
        Nm.ctypes.data_as(_DP), w.ctypes.data_as(_DP), k.ctypes.data_as(_DP), int(nbox_max),
        float(acc), int(nproxy), RANK_MODES[rank], mlp.handle if mlp is not None else None,
        int(oversample), C.byref(h)))
    H = HSS1D(h, int(Ns.sum()))
    H.sizes = Ns
    return H


def split(K_HSS: HSS1D, v):
    """The per-curve pieces of a concatenated vector of a batched operator."""
    sizes = getattr(K_HSS, "sizes", np.array([K_HSS.N]))
    return np.split(np.asarray(v), np.cumsum(sizes)[:-1])


def HSS1D_invert_nsym(K_HSS: HSS1D, acc=None, precision="fp64") -> HSS1D:
    """HSS1D/HSS1D_invert_nsym.m:1 (the inverse factors are stored in the same handle).

    ``precision="fp32"`` builds FP32 factors beside the FP64 operator for HSS1D_solve."""
    check(lib().hssc_hss1d_invert_prec(K_HSS.handle, PREC[precision]))
    return K_HSS


def HSS1D_solve(K_HSS: HSS1D, f, tol=1e-10, maxit=10):
    """Iterative refinement (Common/Iterative_Refinement.m:20-35): FP32 factors (FP64 if
    only those exist) against the FP64 compressed operator.  Returns (x, solves, relres)."""
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    if f.shape[0] != K_HSS.N:
        raise ValueError(f"right-hand side must have N = {K_HSS.N} entries, got {f.shape[0]}")
    x = np.empty_like(f)
    it, rel = C.c_int32(), C.c_double()
    check(lib().hssc_hss1d_solve(K_HSS.handle, f.ctypes.data_as(_DP), x.ctypes.data_as(_DP),
                                 float(tol), int(maxit), C.byref(it), C.byref(rel)))
    return x, it.value, rel.value


def _vec(fn, K_HSS, v):
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
    if v.shape[0] != K_HSS.N:
        raise ValueError(f"vector must have N = {K_HSS.N} entries, got {v.shape[0]}")
    out = np.empty_like(v)
    check(fn(K_HSS.handle, v.ctypes.data_as(_DP), out.ctypes.data_as(_DP), 1))
    return out


def HSS1D_applyinv_nsym(K_HSS: HSS1D, f):
    """HSS1D/HSS1D_applyinv_nsym.m:1 (one right-hand side)."""
    return _vec(lib().hssc_hss1d_applyinv, K_HSS, f)


def HSS1D_apply_nsym(K_HSS: HSS1D, q):
    """HSS1D/HSS1D_apply_nsym.m:1 (one vector)."""
    return _vec(lib().hssc_hss1d_apply, K_HSS, q)
