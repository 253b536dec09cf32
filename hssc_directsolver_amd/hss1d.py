"""Python mirror of the reference's HSS1D MATLAB entry points, backed by libhssc.so.

    K_HSS = HSS1D_compress_nsym('green', X, nbox_max, acc, params)  # HSS1D_compress_nsym.m
    K_HSS = HSS1D_compress_nsym('brute', K, nbox_max, acc)
    K_HSS = HSS1D_compress_nsym('green', curve, nbox_max, acc)
    HSS1D_invert_nsym(K_HSS)                                         # HSS1D_invert_nsym.m
    x = HSS1D_applyinv_nsym(K_HSS, f)                                # HSS1D_applyinv_nsym.m
    u = HSS1D_apply_nsym(K_HSS, q)                                   # HSS1D_apply_nsym.m

'green' with points X (N x 2) and a params object compresses Kernel_Eval(X, X, params), as
the reference does; 'brute' compresses a given dense matrix; 'green' with a Curve (points,
outward normals, quadrature weights, curvature) compresses the second-kind Laplace double
layer -1/2 I + D (see include/hssc.h).  Vectors are N or N x m (m right-hand sides).
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


def HSS1D_compress_nsym(type_, curve, nbox_max, acc, params=None, nproxy=None, rank="id",
                        mlp=None, oversample=8) -> HSS1D:
    """HSS1D/HSS1D_compress_nsym.m:1 (K_HSS = HSS1D_compress_nsym(type,X,TREE,acc,params)).

    type 'green': ``curve`` is an N x 2 point array with ``params`` (a Params; the kernel is
    Kernel_Eval(X, X, params), proxy circles of ``nproxy`` points, 50 as the reference), or
    a Curve (the double layer, 64 proxy points by default).  type 'brute': ``curve`` is the
    dense N x N matrix.  For curves, ``rank`` picks how the IDs are sized: 'id' runs the
    plain ID on every block; 'table' sketches each block with the closed-form rank guess of
    HSS2D/predict_rank.m; 'mlp' with the prediction of ``mlp`` (a RankMLP).  Sketches that
    fail fall back to the plain ID."""
    h = C.c_void_p()
    if type_ == "brute":
        K = np.asfortranarray(curve, dtype=np.float64)
        if K.ndim != 2 or K.shape[0] != K.shape[1]:
            raise ValueError("'brute' needs a square matrix")
        check(lib().hssc_hss1d_compress_dense(K.ctypes.data_as(_DP), K.shape[0], K.shape[0],
                                              int(nbox_max), float(acc), C.byref(h)))
        return HSS1D(h, K.shape[0])
    if type_ != "green":
        raise ValueError(f"unknown compression type {type_!r} ('green' or 'brute')")
    if not isinstance(curve, Curve):
        if params is None:
            raise ValueError("'green' on points needs the kernel's params")
        X = np.asfortranarray(curve, dtype=np.float64)
        if X.ndim != 2 or X.shape[1] != 2:
            raise ValueError("points must be N x 2")
        raw = params.raw if hasattr(params, "raw") else params
        check(lib().hssc_hss1d_compress_kernel(C.byref(raw), X.ctypes.data_as(_DP), X.shape[0],
                                               int(nbox_max), float(acc),
                                               int(50 if nproxy is None else nproxy), C.byref(h)))
        return HSS1D(h, X.shape[0])
    nproxy = 64 if nproxy is None else nproxy
    X = np.asfortranarray(curve.X, dtype=np.float64)
    Nm = np.asfortranarray(curve.nrm, dtype=np.float64)
    w = np.ascontiguousarray(curve.w, dtype=np.float64)
    k = np.ascontiguousarray(curve.kappa, dtype=np.float64)
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
    v = np.asarray(v, dtype=np.float64)
    if v.ndim not in (1, 2) or v.shape[0] != K_HSS.N:
        raise ValueError(f"vectors must be N = {K_HSS.N} or N x m, got shape {v.shape}")
    m = 1 if v.ndim == 1 else v.shape[1]
    vf = np.asfortranarray(v)
    out = np.empty_like(vf, order="F")
    check(fn(K_HSS.handle, vf.ctypes.data_as(_DP), out.ctypes.data_as(_DP), m))
    return out


def HSS1D_applyinv_nsym(K_HSS: HSS1D, f):
    """HSS1D/HSS1D_applyinv_nsym.m:1 (N or N x m right-hand sides)."""
    return _vec(lib().hssc_hss1d_applyinv, K_HSS, f)


def HSS1D_apply_nsym(K_HSS: HSS1D, q):
    """HSS1D/HSS1D_apply_nsym.m:1 (N or N x m vectors)."""
    return _vec(lib().hssc_hss1d_apply, K_HSS, q)
