"""CPU oracle for the HSS2D compress / invert / apply path (test infrastructure only).

This module is a plain numpy/scipy restatement of the reference MATLAB package
HSSC_DirectSolver (Corona, Martinsson & Zorin) in its dense-block ("HSS-D") mode.
It exists to check the HIP implementation; only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it.  The product library never
calls into it.

Every function cites the reference file:line it follows (paths relative to the
reference repository root).  Indices are 0-based here; the reference is 1-based.
Matrices are numpy arrays (row-major storage, but only their mathematical values matter).

What pins this oracle (see tests/test_oracle_*.py):
  * geometry (A1, A4-A6): per-level (n, k) table derived from the reference's geometry
    rules, J is a permutation, skeleton of a box is a subset of its I_src, heap numbering;
  * kernel (A2): closed form log / Hankel, diagonal special cases, symmetry;
  * interpolation (A8): least-squares residual <= acc on the proxy
    (HSS2D/Test/Test_HSS2D_skeletons.m:52-70);
  * invert / applyinv / apply (A13-A17): dense solve A\\b and dense A@x at small N
    (HSS2D/Test/Test_HSS2D_InverseCompression.m:300-311, Test_HSS2D_Matrix_Apply.m:227-235),
    and the published error magnitudes (INSTALL.txt:57-63).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla

# ----------------------------------------------------------------------------------------
# Parameters and grid
# ----------------------------------------------------------------------------------------


def tree_parameters(pot, kh, sym, Ng, np_, acc, lay, n_cut, TI, corr0=None):
    """Restates Common/HSS_tree_parameters.m:108-175.

    Grid: np_ x np_ panels of Ng x Ng cell-centred nodes on [-1,1]^2, panel-major order,
    y fastest inside a panel and across panels (meshgrid + column-major reshape,
    :144-160).  ``corr0`` replaces the Duan-Rokhlin side-car (:167-175) for kh>0.
    """
    hp = 2.0 / np_
    hh = 2.0 / (Ng * np_)
    nlevels = int(round(2 * (math.log(np_) / math.log(2) - 1) + 3))
    p = dict(
        Ng=Ng, np=np_, bisec_rule="uniform", depth=nlevels - 1, max_particles=Ng * Ng,
        skel_rule="bdry", skel_extra_pts=0, proxy_rule="bdry", lambda_=2, layers=lay,
        h=hh, hp=hp, flag_pot=pot, kh=kh, acc=acc, n_cut=n_cut, INTERform="lowrank",
        transinv=TI, sym=sym, k_cut=2500, dim=2, m=1, proxy=0,
    )
    if kh > 0:
        if corr0 is None:
            raise ValueError("kh>0 needs the Duan-Rokhlin diagonal correction corr0")
        p["order"] = 4
        p["dr_weights"] = complex(corr0)
    p["X_source"] = source_points(Ng, np_)
    return p


def source_points(Ng, np_):
    """X_source of Common/HSS_tree_parameters.m:144-160 (0-based order).

    Point g = p*Ng^2 + (jx*Ng + iy) with panel p = py + px*np_;
    x = -1 + px*hp + (jx+0.5)*hh, y = -1 + py*hp + (iy+0.5)*hh.
    The arithmetic order here is the one the C++ tree builder uses, so both produce
    bit-identical coordinates.
    """
    hp = 2.0 / np_
    hh = 2.0 / (Ng * np_)
    N = np_ * np_ * Ng * Ng
    X = np.empty((N, 2))
    g = np.arange(N)
    p = g // (Ng * Ng)
    loc = g % (Ng * Ng)
    px, py = p // np_, p % np_
    jx, iy = loc // Ng, loc % Ng
    X[:, 0] = (-1.0 + px * hp) + (jx + 0.5) * hh
    X[:, 1] = (-1.0 + py * hp) + (iy + 0.5) * hh
    return X


# ----------------------------------------------------------------------------------------
# Kernel blocks
# ----------------------------------------------------------------------------------------


def b_func(x, y, x0, y0, alpha):
    """Common/Kernel_Eval.m:99-108: b = 1 + alpha*exp(-|x-x0|^2)."""
    if alpha == 0:
        return np.ones_like(x)
    d1 = x - x0
    d2 = y - y0
    return 1.0 + alpha * np.exp(-d1 * d1 - d2 * d2)


def bump_function(x, y, h):
    """Common/Bump_function.m:53-62 (option 4, tanh plateau)."""
    c = 4 * h
    D = 10.0
    f1 = 0.5 * np.tanh(D * (-np.abs(x) + 1 - c)) + 0.5
    f2 = 0.5 * np.tanh(D * (-np.abs(y) + 1 - c)) + 0.5
    return f1 * f2


def gaussian_contrast(x, y, amp, width):
    """The Gaussian contrast of the Lippmann-Schwinger configuration (no reference
    counterpart: Kernel_Eval.m:133,143 hard-wire Bump_function option 4):
    q = amp * exp(-|x|^2 / (2 width^2))."""
    return amp * np.exp(-(x * x + y * y) / (2.0 * width * width))


def point_weights(X, params, role):
    """Per-point factor b(x) (role 'row') or c(y) (role 'col').

    Restates Common/Kernel_Eval.m:111-159 (LOCAL_get_bc).  The reference evaluates b on the
    row points X1 and c on the column points X2; both depend only on their own point, so
    they are per-point vectors.  ``params["proxy"]`` (default 0) > 0 replaces b by ones
    (:117-121, :131-139), < 0 replaces c by ones (:141-154).  ``params["contrast"] =
    ("gaussian", amp, width)`` replaces the tanh plateau of the symmetric kh > 0 case by
    sqrt(gaussian_contrast) (the configuration's contrast; the reference has no such option).
    """
    x, y = X[:, 0], X[:, 1]
    kh, h = params["kh"], params["h"]
    proxy = params.get("proxy", 0)
    if params["sym"] == 0 and kh > 0:
        if role == "row" and proxy <= 0:
            return bump_function(x, y, 0)  # Bump_function(...,opts) with h=0 (:118)
        return np.ones_like(x)
    if params["transinv"] == 0:
        if (role == "row" and proxy > 0) or (role == "col" and proxy < 0):
            return np.ones_like(x)
        if kh > 0:
            con = params.get("contrast")
            if con is not None and con[0] == "gaussian":
                return np.sqrt(gaussian_contrast(x, y, con[1], con[2]))
            return np.sqrt(bump_function(x, y, h))
        if role == "row" or params["sym"]:
            return b_func(x, y, 0.3, 0.6, 0.25)
        return b_func(x, y, 0.2, -0.6, 0.25)
    return np.ones_like(x)


def kernel_eval(X1, X2, params):
    """Common/Kernel_Eval.m:29-96 for SL_L_2D and SL_H_2D (2D plane).

    SL_L_2D: A = (den==0) + h^2 * 0.5 * b .* log(den + (den==0)) .* c    (:60-61)
    SL_H_2D: A = w*(den==0) + h^2 k^2 (i/4) b .* (H0(k sqrt(den+(den==0))) - H0(k)(den==0)) .* c
             with w = 1 + k^2 (i/4) corr0                                  (:44-49)

    SL_L_2D options without a reference counterpart (the configuration's literal kernel):
    ``kernel_scale`` s multiplies the kernel log r (1: the reference's h^2 log r; 1/(2 pi):
    (1/2 pi) log r) and ``length_scale`` L makes the grid's physical spacing L h and
    distances L r, so the off-diagonal entries become s (L h)^2 log(L r); L = 1/2 maps the
    reference's [-1,1]^2 onto [0,1]^2.
    """
    from scipy.special import hankel1

    h = params["h"]
    b = point_weights(X1, params, "row")
    c = point_weights(X2, params, "col")
    dx = X1[:, 0][:, None] - X2[:, 0][None, :]
    dy = X1[:, 1][:, None] - X2[:, 1][None, :]
    den = dx * dx + dy * dy
    z = den == 0
    pot = params["flag_pot"]
    if pot == "SL_L_2D":
        s = params.get("kernel_scale", 1.0)
        L = params.get("length_scale", 1.0)
        if s == 1.0 and L == 1.0:
            return z + (h * h) * 0.5 * b[:, None] * np.log(den + z) * c[None, :]
        logr = 0.5 * np.log(den + z) + np.log(L) * (~z)
        return z + s * (L * h) ** 2 * b[:, None] * logr * c[None, :]
    if pot == "SL_H_2D":
        kh = params["kh"]
        C = hankel1(0, kh)
        w = 1 + kh * kh * 0.25j * params["dr_weights"]
        H = hankel1(0, kh * np.sqrt(den + z)) - C * z
        return w * z + (h * h) * kh * kh * 0.25j * b[:, None] * H * c[None, :]
    raise NotImplementedError(pot)


# ----------------------------------------------------------------------------------------
# Tree, skeletons, proxies
# ----------------------------------------------------------------------------------------


def level_radii(lev):
    """HSS2D/HSS2D_build_tree_fsym.m:322-326: (colrad, rowrad) = (x, y) half-widths."""
    scl = lev // 2
    rowrad = 2.0 ** (-scl)
    colrad = rowrad / (2 ** (lev % 2))
    return colrad, rowrad


def box_centers(depth):
    """Heap-numbered centres, HSS2D/HSS2D_build_tree_fsym.m:1289-1344 (bisect_level).

    Box b (1-based) at level l has children 2b, 2b+1.  Even l splits in x (2b left),
    odd l splits in y (2b on top).  Returns cent[1..nboxes] (row 0 unused).
    """
    nboxes = 2 ** (depth + 1) - 1
    cent = np.zeros((nboxes + 1, 2))
    for lev in range(depth):
        s = 0.5 * 0.5 ** (lev // 2)
        for b in range(2 ** lev, 2 ** (lev + 1)):
            if lev % 2 == 0:
                cent[2 * b] = cent[b] + (-s, 0.0)
                cent[2 * b + 1] = cent[b] + (s, 0.0)
            else:
                cent[2 * b] = cent[b] + (0.0, s)
                cent[2 * b + 1] = cent[b] + (0.0, -s)
    return cent


def leaf_slabs(params, cent):
    """Build_I_source_unif, HSS2D/HSS2D_build_tree_fsym.m:1203-1230: leaf b owns the
    contiguous slab of the panel that lies in it."""
    depth, Ng, np_ = params["depth"], params["Ng"], params["np"]
    hp = params["hp"]
    nmx = Ng * Ng
    out = {}
    for b in range(2 ** depth, 2 ** (depth + 1)):
        px = int(math.floor((cent[b, 0] + 1.0) / hp))
        py = int(math.floor((cent[b, 1] + 1.0) / hp))
        p = py + px * np_
        out[b] = np.arange(p * nmx, (p + 1) * nmx)
    return out


def orient_skeleton(Xs, cB, th0):
    """LOCAL_orient_skeleton, HSS2D/HSS2D_build_tree_fsym.m:866-874: stable sort by
    radius, then stable sort by angle measured from th0."""
    dx = Xs[:, 0] - cB[0]
    dy = Xs[:, 1] - cB[1]
    th = np.arctan2(dy, dx)
    rho = np.hypot(dx, dy)
    th = np.where(th < th0, th + 2 * np.pi, th)
    J1 = np.argsort(rho, kind="stable")
    J2 = np.argsort(th[J1], kind="stable")
    return J1[J2]


def orient_residual(Xr, nrm, tan):
    """LOCAL_orient_residual, HSS2D/HSS2D_build_tree_fsym.m:877-883."""
    J1 = np.argsort(Xr[:, nrm], kind="stable")
    J2 = np.argsort(-Xr[J1, tan], kind="stable")
    return J1[J2]


def level_orientation(lev):
    """(norm_int, tan_int, th0) of HSS2D/HSS2D_build_tree_fsym.m:840-844 (0-based axes)."""
    if lev % 2 == 0:
        return 0, 1, np.pi / 2
    return 1, 0, 0.0


def preset_skeleton(Xsrc, cB, lev, params, pad_to=0):
    """'bdry' rule of LOCAL_Preset_Skeleton_Info, HSS2D/HSS2D_build_tree_fsym.m:804-864.

    Returns (J, k): J = [skeleton (polar order) ; residual (interface order)] as local
    0-based indices into Xsrc.  ``pad_to`` > count moves the residual points nearest the
    box boundary (distance min(colrad - |x - cx|, rowrad - |y - cy|), ties by index) into
    the skeleton until it has pad_to points (the library's uniform-level padding above
    ID-selected levels; the reference keeps ragged counts).
    """
    colrad, rowrad = level_radii(lev)
    width = params["layers"] * params["h"]
    bdry = (np.abs(Xsrc[:, 0] - cB[0]) > (colrad - width)) | (
        np.abs(Xsrc[:, 1] - cB[1]) > (rowrad - width))
    idx = np.arange(Xsrc.shape[0])
    I_sk, I_rs = idx[bdry], idx[~bdry]
    if pad_to > I_sk.size and I_rs.size:
        d = np.minimum(colrad - np.abs(Xsrc[I_rs, 0] - cB[0]),
                       rowrad - np.abs(Xsrc[I_rs, 1] - cB[1]))
        o = np.argsort(d, kind="stable")[:pad_to - I_sk.size]
        I_sk = np.concatenate([I_sk, I_rs[o]])
        I_rs = np.delete(I_rs, o)
    nrm, tan, th0 = level_orientation(lev)
    I_rs = I_rs[orient_residual(Xsrc[I_rs], nrm, tan)]
    I_sk = I_sk[orient_skeleton(Xsrc[I_sk], cB, th0)]
    return np.concatenate([I_sk, I_rs]), I_sk.size


def bdry_count(Xsrc, cB, lev, params):
    colrad, rowrad = level_radii(lev)
    width = params["layers"] * params["h"]
    return int(np.sum((np.abs(Xsrc[:, 0] - cB[0]) > (colrad - width)) | (
        np.abs(Xsrc[:, 1] - cB[1]) > (rowrad - width))))


def non_preset_skeleton(Xp, Xsrc, cB, lev, params, rng, k=None):
    """LOCAL_Non_Preset_Skeleton_Info, HSS2D/HSS2D_build_tree_fsym.m:686-729.

    ID of A21 = K(X_ext, X_src) with the adaptive rank of Common/ID.m (``k`` fixes the
    rank instead, the library's padding to the level's largest rank), then the skeleton in
    polar order around cB and the residual along the interface (:705-724), T.U =
    Ttmp(Js, Jr) and T.V = I (:726-727; stored dense here).  Returns (J, k, T, rank): rank
    is the adaptive rank of the box.
    """
    A21 = kernel_eval(Xp, Xsrc, params)
    Ttmp, I = id_acc(A21, params["acc"], rng)
    rank = Ttmp.shape[0]
    if k is None:
        k = rank
    elif k != rank:
        Ttmp, I = id_fixed(A21, k)
    nrm, tan, th0 = level_orientation(lev)
    I_sk, I_rs = I[:k], I[k:]
    Js = orient_skeleton(Xsrc[I_sk], cB, th0)
    Jr = orient_residual(Xsrc[I_rs], nrm, tan)
    J = np.concatenate([I_sk[Js], I_rs[Jr]])
    return J, k, Ttmp[np.ix_(Js, Jr)], rank


def _colon(a, d, b):
    """MATLAB a:d:b for spans that are integer multiples of d (up to rounding)."""
    n = int(math.floor((b - a) / d + 1e-10)) + 1
    if n <= 0:
        return np.zeros(0)
    return a + d * np.arange(n)


def box_proxy(cB, colrad, rowrad, layers, h):
    """HSS2D/build_box_proxy.m:52-85: four strips of exterior points at spacing h,
    ``layers`` deep, split into X_proxy (corners trimmed, Remark 3.4 of the paper) and the
    trimmed X_extra.  Returns [X_proxy; X_extra], the proxy LOCAL_Build_Proxy uses."""
    cx, cy = cB
    width = (layers - 0.5) * h
    # strip 1 (left): meshgrid then transpose -> x fastest
    x1 = _colon(cx - colrad - width, h, cx - colrad - h / 2)
    y1 = _colon(cy - rowrad - width, h, cy + rowrad + width)
    X1 = np.stack(np.meshgrid(x1, y1, indexing="xy"), -1).reshape(-1, 2)
    # strip 2 (top): meshgrid column-major -> y fastest
    x2 = _colon(cx - colrad + h / 2, h, cx + colrad - h / 2)
    y2 = _colon(cy + rowrad + width, -h, cy + rowrad + h / 2)
    X2 = np.stack(np.meshgrid(x2, y2, indexing="ij"), -1).reshape(-1, 2)
    # strip 3 (right): transposed -> x fastest
    x3 = _colon(cx + colrad + h / 2, h, cx + colrad + width)
    y3 = _colon(cy + rowrad + width, -h, cy - rowrad - width)
    X3 = np.stack(np.meshgrid(x3, y3, indexing="xy"), -1).reshape(-1, 2)
    # strip 4 (bottom): column-major -> y fastest
    x4 = _colon(cx + colrad - h / 2, -h, cx - colrad + h / 2)
    y4 = _colon(cy - rowrad - h / 2, -h, cy - rowrad - width)
    X4 = np.stack(np.meshgrid(x4, y4, indexing="ij"), -1).reshape(-1, 2)

    r1 = layers * layers - 0.5
    r2 = layers * layers + 0.5
    ew = colrad + width - h / 2
    eh = rowrad + width - h / 2
    d = X1 - (cx, cy)
    I1 = ((np.abs(d[:, 0]) > ew) & ((d[:, 1] > rowrad + width - r1 * h) |
                                     (d[:, 1] < -(rowrad + width - r2 * h)))) | (np.abs(d[:, 1]) > eh)
    d = X2 - (cx, cy)
    I2 = (np.abs(d[:, 1]) > eh) & ((d[:, 0] > colrad + width - r1 * h) |
                                   (d[:, 0] < -(colrad + width - r2 * h)))
    d = X3 - (cx, cy)
    I3 = ((np.abs(d[:, 0]) > ew) & ((d[:, 1] > rowrad + width - r2 * h) |
                                     (d[:, 1] < -(rowrad + width - r1 * h)))) | (np.abs(d[:, 1]) > eh)
    d = X4 - (cx, cy)
    I4 = (np.abs(d[:, 1]) > eh) & ((d[:, 0] > colrad + width - r2 * h) |
                                   (d[:, 0] < -(colrad + width - r1 * h)))
    # LOCAL_Build_Proxy (HSS2D_build_tree_fsym.m:962-965, proxy_rule 'bdry') keeps the
    # trimmed corner points too, appended after X_proxy: X_ext = [X_proxy; X_extra].
    return np.concatenate([X1[~I1], X2[~I2], X3[~I3], X4[~I4], X1[I1], X2[I2], X3[I3], X4[I4]])


def default_rand(n):
    """MATLAB's rand(n, 1) right after rng('default'): MT19937 seeded 5489 with 53-bit doubles,
    the stream numpy's RandomState(5489) produces."""
    return np.random.RandomState(5489).random_sample(n)


def random_sampler(X, cB, colrad, rowrad, lam, lay, h):
    """LOCAL_random_sampler, HSS2D_build_tree_fsym.m:979-1006."""
    D = np.abs(np.minimum(colrad - np.abs(X[:, 0] - cB[0]), rowrad - np.abs(X[:, 1] - cB[1])))
    width = lay * h
    rad = D.max()
    sz = -math.floor(math.log2(width / rad)) + 1
    prob = 2.0 ** (-lam * np.arange(sz - 1, -1, -1, dtype=float))
    B = D < width
    if sz > 0:
        Ld = np.minimum(-np.ceil(np.log2(D / rad)), sz - 2) + 1
        Ld[B] = sz
        P = prob[Ld.astype(int) - 1]
        B = default_rand(P.size) < P
    return X[B]


def build_proxy(cB, colrad, rowrad, lev, n_skel, lay_prox, params):
    """LOCAL_Build_Proxy, HSS2D_build_tree_fsym.m:886-977, for the 'bdry' and 'rand' rules
    (:893-921): 'rand' thins the ring until it holds k0 points, each pass on the previous
    pass's points with lambda x 0.9."""
    Xe = box_proxy(cB, colrad, rowrad, lay_prox, params["h"])
    if params["proxy_rule"] != "rand":
        return Xe
    k0 = math.ceil((2 / 3 if lev % 2 == 0 else 3 / 4) * n_skel)
    k0 = min(k0, Xe.shape[0])
    lam, k, c = params["lambda_"], 0, 0
    while k0 > k and c < 10:
        Xe = random_sampler(Xe, cB, colrad, rowrad, lam, params["layers"], params["h"])
        k = Xe.shape[0]
        if k0 > k:
            lam = 0.9 * lam
        c += 1
    return Xe


def proxy_layers(lev, params):
    """lay_prox of HSS2D/HSS2D_build_tree_fsym.m:339-347.

    ``params['proxy_layers'] > 0`` replaces the rule for lev > 0 (the reference's
    2*round(rad/h) for non-Laplace kernels grows with the box: 256 layers at level 1 of a
    512^2 grid)."""
    colrad, rowrad = level_radii(lev)
    rad = min(rowrad, colrad)
    if lev > 0 or params["m"] < 1:
        if params.get("proxy_layers", 0) > 0 and lev > 0:
            return params["proxy_layers"]
        if params["flag_pot"] == "SL_L_2D":
            return params["layers"] + 1
        return 2 * int(round(rad / params["h"]))
    return params["layers"]


# ----------------------------------------------------------------------------------------
# Interpolative decompositions (Common/ID.m, ID_rand.m, LRID_rand.m, QR_findrank.m)
# ----------------------------------------------------------------------------------------


def ls_solve(A, B):
    """MATLAB A\\B for a tall A: column-pivoted Householder QR basic solution (xGEQP3)."""
    Q, R, P = sla.qr(A, mode="economic", pivoting=True)
    Y = sla.solve_triangular(R, Q.T.conj() @ B if np.iscomplexobj(Q) else Q.T @ B)
    X = np.empty_like(Y)
    X[P] = Y
    return X


def _tsolve(R11, R12):
    return sla.solve_triangular(R11, R12)


def id_acc(A, acc, rng):
    """skel_col_qr of Common/ID.m:40-86 (rank by |R_ii| > acc |R_11|, randomized check).

    Returns (T, I): T is k x (n-k), I a column permutation (skeleton first)."""
    Q, R, I = sla.qr(A, mode="economic", pivoting=True)
    Rdg = np.abs(np.diag(R))
    if Rdg.size == 0:
        return np.zeros((0, A.shape[1])), I
    Rmx = Rdg[0] * acc
    k = int(np.sum(Rdg > Rmx))
    T = _tsolve(R[:k, :k], R[:k, k:])
    l = min(A.shape)
    while k < l and _rand_error(A, T, I, k, rng) > _rand_norm(A, I, k, rng) * acc:
        Rmx *= 0.75
        k = int(np.sum(Rdg > Rmx))
        T = _tsolve(R[:k, :k], R[:k, k:])
    return T, I


def id_fixed(A, k):
    """Fixed-rank ID as the library computes it: T = R11^-1 R12 of the column-pivoted QR
    (back substitution).  The reference's skel_col_fixed (Common/ID.m:106-113) applies the
    SVD pseudo-inverse of R11 with singular values above 1e-12 instead; the two agree while
    R11 is well conditioned and otherwise differ only along directions at rounding level
    (the HSS1D oracle, oracle/hss1d.py id_fixed, follows the reference's form)."""
    Q, R, I = sla.qr(A, mode="economic", pivoting=True)
    return _tsolve(R[:k, :k], R[:k, k:]), I


def _rand_norm(A, I, k, rng):
    """Common/ID.m:68-74."""
    x = rng.random(A.shape[1] - k)
    x /= np.linalg.norm(x)
    Ar = A[:, I[k:]]
    y = Ar.conj().T @ (Ar @ x)
    return math.sqrt(np.linalg.norm(y))


def _rand_error(A, T, I, k, rng):
    """Common/ID.m:77-86."""
    x = rng.random(A.shape[1] - k)
    x /= np.linalg.norm(x)
    As, Ar = A[:, I[:k]], A[:, I[k:]]
    y = As @ (T @ x) - Ar @ x
    y = T.conj().T @ (As.conj().T @ y) - Ar.conj().T @ y
    return math.sqrt(np.linalg.norm(y))


def qr_findrank(m, n, matvecAt, k0, C, rng, eps):
    """Common/QR_findrank.m:45-92: adaptive Gaussian range finder on A^T.

    ``eps`` is the routine's tolerance argument (the caller passes acc,
    interpolation_operator.m:119 -> Rand_AT_sample.m:26): |R_ii| > eps counts toward the rank.
    Returns (krank, Y) with Y = A^T G (n x krank)."""
    k0 = min(k0, min(m, n))
    G = rng.standard_normal((m, k0))
    Y = matvecAt(G)
    Q, R, _ = sla.qr(Y, mode="economic", pivoting=True)
    D = np.abs(np.diag(R))
    residual = D[-1] if D.size else 0.0
    krank = int(np.sum(D > eps))
    while residual > eps and krank < min(m, n):
        num = min(C, min(m, n) - krank)
        G = rng.standard_normal((m, num))
        Ynew = matvecAt(G)
        if Y.shape[1] < krank + num:
            Y = np.concatenate([Y[:, :krank], np.zeros((Y.shape[0], krank + num - krank),
                                                       dtype=Y.dtype)], axis=1)
        Y[:, krank:krank + num] = Ynew
        M = Ynew - Q @ (Q.conj().T @ Ynew)
        Qc, R, _ = sla.qr(M, mode="economic", pivoting=True)
        D = np.abs(np.diag(R))
        residual = D[-1] if D.size else 0.0
        krank += int(np.sum(D > eps))
        Qc = Qc - Q @ (Q.conj().T @ Qc)
        Q = np.concatenate([Q, Qc], axis=1)
    return krank, Y[:, :krank]


def skel_colk(Z, k):
    """Common/ID_rand.m:98-106: fixed-rank ID of Z via CPQR + SVD pseudo-inverse of R11."""
    Q, R, I = sla.qr(Z, mode="economic", pivoting=True)
    if k == 0:
        return np.zeros((0, Z.shape[1]), dtype=Z.dtype), I
    U, D, Vh = np.linalg.svd(R[:k, :k])
    q = int(np.sum(D > 1e-12))
    T = Vh[:q].conj().T @ ((U[:, :q].conj().T @ R[:k, k:]) / D[:q, None])
    return T, I


def id_rand_update(T0, J0, k0, Xp, Xsrc, params, rng):
    """LOCAL_ID_rand_update, HSS2D_build_tree_fsym.m:1069-1124, dense T (T.V = I): a
    randomized ID (ID_rand 'norm' with q0 = 8 layers ceil(log2 k0) samples and increments
    C = q0 / 2, absolute tolerance acc, :1093-1097) of the interpolation residual
    Res = K(Xp, X_rs) - K(Xp, X_sk) T0 adds its k1 skeleton columns to the skeleton set:
    J = [J0_sk; J0_rs(J1)], T = [T0(:, J1_rs) - T0(:, J1_sk) T1; T1] (:1100-1106).
    Returns (T, J, k)."""
    J0 = np.asarray(J0)
    J0_sk, J0_rs = J0[:k0], J0[k0:]
    A_rs = kernel_eval(Xp, Xsrc[J0_rs], params)
    A_sk = kernel_eval(Xp, Xsrc[J0_sk], params)
    Res = A_rs - A_sk @ T0
    m, n = Res.shape
    q0 = 8 * params["layers"] * math.ceil(math.log2(k0))
    krank, Y = qr_findrank(m, n, lambda G: Res.T @ G, q0, q0 // 2, rng, params["acc"])
    T1, J1 = skel_colk(Y.T, krank)
    k1 = T1.shape[0]
    J1_sk, J1_rs = J1[:k1], J1[k1:]
    Tf = np.concatenate([T0[:, J1_rs] - T0[:, J1_sk] @ T1, T1])
    return Tf, np.concatenate([J0_sk, J0_rs[J1]]), k0 + k1


def lrid_rand(K21, k0, C, rng, tol):
    """Common/LRID_rand.m:10-40 with ID_rand.m:26-49 and Rand_AT_sample.m:24-41.

    K21 ~ U V^T with U = K21(:, J(1:k)) and V(J,:) = [I T]^T.  Returns (U, V)."""
    m, n = K21.shape
    krank, Y = qr_findrank(m, n, lambda G: K21.T @ G, k0, C, rng, tol)
    Z = Y.T
    T, J = skel_colk(Z, krank)
    k = T.shape[0]
    U = K21[:, J[:k]]
    V = np.zeros((n, k))
    V[J, :] = np.concatenate([np.eye(k), T], axis=1).T
    return U, V


def interpolation_operator(params, Xp, X_sk, X_rs, lowrank, q, lev, rng):
    """HSS2D/interpolation_operator.m:50-127 ('box_2_proxy').

    Solves K(Xp, X_sk) T = K(Xp, X_rs) in least squares.  lowrank=False returns
    (T, None); lowrank=True compresses K21 by randomized ID first (:80-127) and returns
    (T.U, T.V) with T = T.U @ T.V.T."""
    K11 = kernel_eval(Xp, X_sk, params)
    K21 = kernel_eval(Xp, X_rs, params)
    if not lowrank:
        return ls_solve(K11, K21), None
    m, n = K21.shape
    if q[lev] > 0:
        k0, C = q[lev] + 1, 10
    elif np.sum(q > 0) < 5:
        k0 = min(int(math.ceil(9 * math.log2(min(m, n)) - 20)), min(m, n))
        C = 5
    else:
        k0 = min(max(int(math.ceil(2 * q[lev + 2] - q[lev + 4] + 2)), q[lev + 2]), min(m, n))
        C = int(math.ceil(abs(q[lev + 2] - q[lev + 4]) / 2)) + 1
    k0 = max(k0, 1)
    U, V = lrid_rand(K21, k0, C, rng, params["acc"])
    return ls_solve(K11, U), V


# ----------------------------------------------------------------------------------------
# Tree build (HSS2D/HSS2D_build_tree_fsym.m)
# ----------------------------------------------------------------------------------------


@dataclass
class Tree:
    params: dict
    depth: int
    cent: np.ndarray
    I_src: dict = field(default_factory=dict)   # box -> global indices (0-based)
    I_sk: dict = field(default_factory=dict)    # box -> global skeleton indices
    J: dict = field(default_factory=dict)       # key -> local permutation (skel first)
    k: dict = field(default_factory=dict)       # key -> skeleton size
    T: dict = field(default_factory=dict)       # key -> (U, V) with T = U if V is None else U V^T
    Xp: dict = field(default_factory=dict)      # key -> proxy points
    box_rank: dict = field(default_factory=dict)  # key -> own skeleton count (before padding)

    def key(self, b):
        """TI trees store (J, k, T) per level, NTI trees per box (:308-312)."""
        if self.params["transinv"]:
            return ("lev", level_of(b))
        return ("box", b)

    def Tdense(self, b):
        U, V = self.T[self.key(b)]
        return U if V is None else U @ V.T


def level_of(b):
    return int(b).bit_length() - 1


def levhss(tree, n_cut, trinv):
    """HSS2D/levHSS.m:33-65: walk down from the root until the smallest skeleton of a level
    is at most n_cut (never past depth - 1); return that level rounded down to an even one,
    or -1 if the root's level already qualifies."""
    lev = 0
    if trinv == 0:
        n_min = max(tree.k[tree.key(b)] for b in (1, 2, 3))
    else:
        n_min = tree.k[tree.key(1)]
    while n_min > n_cut and lev < tree.depth - 1:
        lev += 1
        n_min = min(tree.k[tree.key(b)] for b in range(2 ** lev, 2 ** (lev + 1)))
    if lev > 0:
        return lev - 1 if lev % 2 == 1 else lev
    return -1


def build_index_sets(params):
    """The tree, skeleton index sets and proxies of build_tree, without T."""
    return build_tree(params, with_T=False)


def residual_pivots(A, k0, params):
    """The library's extra-point choice for a preset box (no reference counterpart; it
    replaces the randomized ID of id_rand_update, whose Gaussian stream is MATLAB's): the
    column-pivoted QR of the interpolation residual A_rs - A_sk T = (I - P_sk) A_rs, its
    points counted while |R_jj| > acc / sqrt(q0) (a Gaussian sketch of q0 samples, as the
    reference draws, scales the residual's singular values by about sqrt(q0)).
    Returns (pivot order of the residual columns, count)."""
    Q, _ = np.linalg.qr(A[:, :k0])
    Res = A[:, k0:] - Q @ (Q.conj().T @ A[:, k0:])
    _, R, P = sla.qr(Res, mode="economic", pivoting=True)
    q0 = 8 * params["layers"] * math.ceil(math.log2(max(k0, 2)))
    return P, int(np.sum(np.abs(np.diag(R)) > params["acc"] / math.sqrt(q0)))


def extra_points(T0, J0, k0, Xp, Xsrc, cB, lev, b, params, rng, TI):
    """The skel_extra_pts update of a preset box (HSS2D_build_tree_fsym.m:545-564 NTI,
    :582-597 TI): id_rand_update, then the skeleton re-oriented in polar order from th0 (TI
    and even boxes) or th1 = th0 - pi (odd boxes), the residual along the interface (TI) or
    left in the ID's order (NTI).  Returns dense (T, J, k)."""
    T, J, k = id_rand_update(T0, J0, k0, Xp, Xsrc, params, rng)
    nrm, tan, th0 = level_orientation(lev)
    th = th0 if (TI or b % 2 == 0) else th0 - np.pi
    Jsk, Jrs = J[:k], J[k:]
    Js = orient_skeleton(Xsrc[Jsk], cB, th)
    Jr = orient_residual(Xsrc[Jrs], nrm, tan) if TI else np.arange(Jrs.size)
    return T[np.ix_(Js, Jr)], np.concatenate([Jsk[Js], Jrs[Jr]]), k


def build_tree(params, interform=None, seed=0, with_T=True, timings=None, preset_all=False,
               stop_level=0, pad=False):
    """HSS2D_build_tree_fsym (uniform tree, 'bdry' or 'rand' proxies; 'bdry' or ID skeletons).

    Restates :81-287 (tree), :292-623 (upward skeleton pass) and
    LOCAL_Interpolation_Operators :1010-1064.  ``interform`` defaults to params
    INTERform ('lowrank' in the reference; 'dense' gives exact LS T at every level).
    For kh > 0 the skeletons of the two bottom levels come from an ID of K(X_ext, X_src)
    (:374-395 selects, LOCAL_Non_Preset_Skeleton_Info :686-729), with ragged per-box ranks
    as in the reference; ``preset_all=True`` keeps the 'bdry' rule there instead.
    ``pad=True`` restates the library's uniform levels: every box of an ID level keeps the
    level's largest adaptive rank (fixed-rank ID), and preset levels above pad their 'bdry'
    sets to the level's largest count.  ``stop_level`` > 0 ends the upward pass there
    (levels below the root only; used to time a bounded sample of a large tree).
    """
    if params["skel_rule"] != "bdry" or params["proxy_rule"] not in ("bdry", "rand"):
        raise NotImplementedError("only the 'bdry' skeleton rule and the 'bdry' / 'rand' "
                                  "proxy rules are restated")
    interform = interform or params["INTERform"]
    rng = np.random.default_rng(seed)
    depth = params["depth"]
    X = params["X_source"]
    cent = box_centers(depth)
    tree = Tree(params=params, depth=depth, cent=cent)
    tree.I_src.update(leaf_slabs(params, cent))
    q = np.zeros(depth + 5, dtype=int)  # ranks of low-rank T per level (:299)
    TI = params["transinv"]
    ragged_below = False
    import time as _time
    for lev in range(depth, stop_level - 1, -1):
        _t0 = _time.perf_counter()
        colrad, rowrad = level_radii(lev)
        lay_prox = proxy_layers(lev, params)
        id_level = params["kh"] > 0 and not preset_all and lev >= depth - 1
        boxes = range(2 ** lev, 2 ** (lev + 1))
        for b in boxes:
            if lev < depth:
                tree.I_src[b] = np.concatenate([tree.I_sk[2 * b], tree.I_sk[2 * b + 1]])
        keyboxes = [2 ** lev] if TI else list(boxes)
        pad_to = None

        if pad and id_level:
            ranks = [id_acc(kernel_eval(build_proxy(cent[b], colrad, rowrad, lev,
                                                    tree.I_src[b].size, lay_prox, params),
                                        X[tree.I_src[b]], params), params["acc"],
                            np.random.default_rng([seed, b]))[0].shape[0]
                     for b in keyboxes]
            pad_to = max(ranks)
        elif pad and ragged_below and not TI:
            pad_to = max(bdry_count(X[tree.I_src[b]], cent[b], lev, params) for b in keyboxes)
        extra_order = {}
        if pad and params.get("skel_extra_pts", 0) and not id_level and lev > 0:
            # the library's uniform extra points: every key keeps the level's largest count,
            # the next pivots of its own residual order (residual_pivots)
            counts = {}
            for b in keyboxes:
                Xsrc = X[tree.I_src[b]]
                J0, k0 = preset_skeleton(Xsrc, cent[b], lev, params, pad_to or 0)
                Xp = build_proxy(cent[b], colrad, rowrad, lev, Xsrc.shape[0], lay_prox, params)
                if Xsrc.shape[0] > k0:
                    P, k1 = residual_pivots(kernel_eval(Xp, Xsrc[J0], params), k0, params)
                else:
                    P, k1 = np.zeros(0, int), 0
                extra_order[b] = (J0, k0, P)
                counts[b] = k0 + k1
            pad_to = max(counts.values())
        for b in boxes:
            Xsrc = X[tree.I_src[b]]
            key = tree.key(b)
            if not TI or b == 2 ** lev:
                cB = cent[b]
                n = Xsrc.shape[0]
                Xp = build_proxy(cB, colrad, rowrad, lev, n, lay_prox, params)
                if id_level:
                    J, k, Tid, rank = non_preset_skeleton(Xp, Xsrc, cB, lev, params,
                                                          np.random.default_rng([seed, b]), pad_to)
                    tree.box_rank[key] = rank
                    U, V = (Tid, None) if with_T else (None, None)
                elif b in extra_order:
                    J0, k0, P = extra_order[b]
                    nrm, tan, th0 = level_orientation(lev)
                    order = np.concatenate([J0[:k0], J0[k0:][P]])
                    I_sk, I_rs = order[:pad_to], order[pad_to:]
                    Js = orient_skeleton(Xsrc[I_sk], cB, th0)
                    Jr = orient_residual(Xsrc[I_rs], nrm, tan)
                    J, k = np.concatenate([I_sk[Js], I_rs[Jr]]), pad_to
                    tree.box_rank[key] = counts[b]
                    U, V = (interpolation_operator(params, Xp, Xsrc[J[:k]], Xsrc[J[k:]], False,
                                                   q, lev, rng) if with_T and n > k
                            else (np.zeros((k, 0)), None) if with_T else (None, None))
                else:
                    J, k = preset_skeleton(Xsrc, cB, lev, params, pad_to or 0)
                    tree.box_rank[key] = bdry_count(Xsrc, cB, lev, params)
                    if not with_T:
                        U, V = None, None
                    elif n > k:
                        extra = params.get("skel_extra_pts", 0) and lev > 0 and not pad
                        lowrank = interform == "lowrank" and lev < depth - 1 and not extra
                        # k >= k_cut selects interpolation_operator_HSS in the reference:
                        # the same T to tolerance, restated by the dense least squares
                        # (the HSS1D least squares is out of scope)
                        U, V = interpolation_operator(params, Xp, Xsrc[J[:k]], Xsrc[J[k:]],
                                                      lowrank, q, lev, rng)
                        if extra:
                            U, J, k = extra_points(U, J, k, Xp, Xsrc, cB, lev, b, params, rng,
                                                   TI)
                    else:
                        U, V = np.zeros((k, 0)), None
                tree.J[key], tree.k[key], tree.T[key], tree.Xp[key] = J, k, (U, V), Xp
                if with_T and V is not None and lev > 0:
                    q[lev] = max(q[lev], U.shape[1]) if not TI else U.shape[1]
            J, k = tree.J[key], tree.k[key]
            tree.I_sk[b] = tree.I_src[b][J[:k]]
        if (id_level or extra_order) and not TI:
            ragged_below = True
        if timings is not None:
            timings[lev] = timings.get(lev, 0.0) + _time.perf_counter() - _t0
    return tree


# ----------------------------------------------------------------------------------------
# Inversion (HSS2D/HSS2D_invert_fsym.m, build_Finv_fsym.m, build_E_fsym.m)
# ----------------------------------------------------------------------------------------


@dataclass
class Inverse:
    FI: dict = field(default_factory=dict)  # key -> n x n (I_src order)
    E: dict = field(default_factory=dict)   # key -> k x k (skeleton order)


def R_matrix(J, k, T):
    """R(:,J) = [I T] of HSS2D/HSS2D_build_E_fsym.m:46."""
    n = J.size
    R = np.zeros((k, n), dtype=T.dtype if T.size else float)
    R[:, J] = np.concatenate([np.eye(k), T], axis=1)
    return R


def invert(tree, timings=None, stop_level=0):
    """HSS2D/HSS2D_invert_fsym.m:73-195 with Fopts='dense'.

    Leaf: FI = inv(K(I_src, I_src)) (:102-111).  Non-leaf: FI = inv([E1 K12; K12^T E2])
    (build_Finv_fsym.m:93-101).  lev>0: E = inv(R FI R^T) (build_E_fsym.m:43-75; the
    LR=1 branch is the same product evaluated through T.U, T.V), symmetrised as the
    reference's LOCAL_Invert_sym does: E = (Ai + Ai.')/2 (build_E_fsym.m:216,244,255-259)."""
    params = tree.params
    X = params["X_source"]
    import time as _time
    out = Inverse()
    TI = params["transinv"]
    for lev in range(tree.depth, stop_level - 1, -1):
        _t0 = _time.perf_counter()
        boxes = [2 ** lev] if TI else range(2 ** lev, 2 ** (lev + 1))
        for b in boxes:
            key = tree.key(b)
            if lev == tree.depth:
                Xs = X[tree.I_src[b]]
                F = kernel_eval(Xs, Xs, params)
            else:
                c1, c2 = 2 * b, 2 * b + 1
                E1, E2 = out.E[tree.key(c1)], out.E[tree.key(c2)]
                K12 = kernel_eval(X[tree.I_sk[c1]], X[tree.I_sk[c2]], params)
                F = np.block([[E1, K12], [K12.T, E2]])
            FI = np.linalg.inv(F)
            out.FI[key] = FI
            if lev > 0:
                R = R_matrix(tree.J[key], tree.k[key], tree.Tdense(b))
                Ai = np.linalg.inv(R @ FI @ R.T)
                out.E[key] = 0.5 * (Ai + Ai.T)
        if timings is not None:
            timings[lev] = timings.get(lev, 0.0) + _time.perf_counter() - _t0
    return out


def applyinv(tree, inv, Xrhs):
    """HSS2D/HSS2D_applyinv_TI.m:63-225 and HSS2D_applyinv_fsym.m (identical per box).

    Up (lev = depth..1): psi = FI q; u = psi(J_s) + T psi(J_r); phi_up = E u.
    Down (lev = 0..depth): lam = q - L phi_up; eta = L E phi_dn; phi = FI (eta + lam)."""
    depth = tree.depth
    x = Xrhs.reshape(Xrhs.shape[0], -1)
    Y = np.zeros(x.shape, dtype=np.result_type(x, inv.FI[tree.key(1)]))
    q, phi_up, phi_dn = {}, {}, {}
    for lev in range(depth, 0, -1):
        for b in range(2 ** lev, 2 ** (lev + 1)):
            key = tree.key(b)
            q[b] = x[tree.I_src[b]] if lev == depth else np.concatenate(
                [phi_up[2 * b], phi_up[2 * b + 1]])
            J, k = tree.J[key], tree.k[key]
            psi = inv.FI[key] @ q[b]
            u = psi[J[:k]] + tree.Tdense(b) @ psi[J[k:]]
            phi_up[b] = inv.E[key] @ u
    for lev in range(0, depth + 1):
        for b in range(2 ** lev, 2 ** (lev + 1)):
            key = tree.key(b)
            if lev == 0:
                qb = np.concatenate([phi_up[2], phi_up[3]]) if depth > 0 else x[tree.I_src[1]]
                v = qb
            else:
                J, k = tree.J[key], tree.k[key]
                T = tree.Tdense(b)
                Lphi = np.zeros_like(q[b], dtype=Y.dtype)
                Lphi[J] = np.concatenate([phi_up[b], T.T @ phi_up[b]])
                lam = q[b] - Lphi
                t = inv.E[key] @ phi_dn[b]
                eta = np.zeros_like(lam)
                eta[J] = np.concatenate([t, T.T @ t])
                v = eta + lam
            phi = inv.FI[key] @ v
            if lev < depth:
                kc = tree.I_sk[2 * b].size
                phi_dn[2 * b], phi_dn[2 * b + 1] = phi[:kc], phi[kc:]
            else:
                Y[tree.I_src[b]] = phi
    return Y.reshape(Xrhs.shape)


# ----------------------------------------------------------------------------------------
# Forward operator (HSS2D/HSS2D_compress_fsym.m, HSS2D_apply_TI.m, HSS2D_apply_fsym.m)
# ----------------------------------------------------------------------------------------


def compress(tree):
    """HSS2D/HSS2D_compress_fsym.m:73-138 (dense Bself at leaves, B12 at siblings)."""
    params = tree.params
    X = params["X_source"]
    TI = params["transinv"]
    Bself, B12 = {}, {}
    for lev in range(0, tree.depth + 1):
        boxes = [2 ** lev] if TI else range(2 ** lev, 2 ** (lev + 1))
        for b in boxes:
            key = tree.key(b)
            if lev == tree.depth:
                Xs = X[tree.I_src[b]]
                Bself[key] = kernel_eval(Xs, Xs, params)
            else:
                B12[key] = kernel_eval(X[tree.I_sk[2 * b]], X[tree.I_sk[2 * b + 1]], params)
    return Bself, B12


def apply(tree, comp, Xvec):
    """HSS2D/HSS2D_apply_TI.m:74-189: up phi_up = R q; down phi = B q + L phi_dn."""
    Bself, B12 = comp
    depth = tree.depth
    x = Xvec.reshape(Xvec.shape[0], -1)
    dt = np.result_type(x, B12[tree.key(1)] if depth > 0 else Bself[tree.key(1)])
    Y = np.zeros(x.shape, dtype=dt)
    q, phi_up, phi_dn = {}, {}, {}
    for lev in range(depth, -1, -1):
        for b in range(2 ** lev, 2 ** (lev + 1)):
            key = tree.key(b)
            q[b] = x[tree.I_src[b]] if lev == depth else np.concatenate(
                [phi_up[2 * b], phi_up[2 * b + 1]])
            J, k = tree.J[key], tree.k[key]
            phi_up[b] = q[b][J[:k]] + tree.Tdense(b) @ q[b][J[k:]]
    for lev in range(0, depth + 1):
        for b in range(2 ** lev, 2 ** (lev + 1)):
            key = tree.key(b)
            if lev == depth:
                lam = Bself[key] @ q[b]
            else:
                k1 = tree.I_sk[2 * b].size
                B = B12[key]
                lam = np.concatenate([B @ q[b][k1:], B.T @ q[b][:k1]])
            if lev > 0:
                J, k = tree.J[key], tree.k[key]
                eta = np.zeros_like(lam)
                eta[J] = np.concatenate([phi_dn[b], tree.Tdense(b).T @ phi_dn[b]])
                phi = eta + lam
            else:
                phi = lam
            if lev < depth:
                k1 = tree.I_sk[2 * b].size
                phi_dn[2 * b], phi_dn[2 * b + 1] = phi[:k1], phi[k1:]
            else:
                Y[tree.I_src[b]] = phi
    return Y.reshape(Xvec.shape)


def level_shapes(tree):
    """(n, k, M) per level, for the per-level shape table and flop counts."""
    out = {}
    for lev in range(tree.depth + 1):
        b = 2 ** lev
        key = tree.key(b)
        out[lev] = (tree.I_src[b].size, tree.k[key], tree.Xp[key].shape[0])
    return out


def dense_matrix(params):
    X = params["X_source"]
    return kernel_eval(X, X, params)


def random_rhs(N, seed=0, m=1):
    """Unit-norm Gaussian RHS (Test_HSS2D_InverseCompression.m:145-146)."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((N, m))
    return v / np.sqrt(np.sum(v * v, axis=0))
