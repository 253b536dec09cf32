"""CPU oracle for the HSS1D compress / invert / apply path (test infrastructure only).

numpy/scipy restatement of HSSC_DirectSolver's HSS1D solver (nsym):
  HSS1D/HSS1D_compress_nsym.m   tree, neighbour lists, proxy circles, ID skeletons, blocks
                                ('green' compression, :84-154) and the 'brute' compression
                                of a given dense matrix (:158-181, :210-214)
  HSS1D/HSS1D_invert_nsym.m     telescoping inverse (Finv, E, Linv, Rinv, Dinv, FTinv)
  HSS1D/HSS1D_applyinv_nsym.m   inverse apply
  HSS1D/HSS1D_apply_nsym.m      forward apply
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU baseline may use it.

Kernels of the 'green' compression (``compress``):
  * PointKernel: the reference's own, Common/Kernel_Eval.m on point sets (proxy rows and
    columns by the same Kernel_Eval, :152-156), e.g. SL_L_2D with params.proxy = 1 on the
    skeleton points of a 2D box as HSS1D/Test/Test_HSS1D_compress.m does;
  * a Curve: the second-kind Laplace double layer on a closed contour (the HSS1D workload of
    the benchmark configurations; no counterpart in Kernel_Eval, which takes positions
    only), A = -1/2 I + D,  D_ij = (1/2pi) (x_i - x_j).n_j / |x_i - x_j|^2 w_j  (i != j),
    D_jj = -kappa_j w_j / (4 pi); proxy sources on the circles use the log kernel
    (1/2pi) log|x - z| (equivalent charges), proxy targets the double-layer kernel.
Pinned by: dense A @ x and dense A \\ b (Test_HSS1D_compress.m:79-109 semantics), the
identity D 1 = -1/2 (Gauss' lemma: the double layer of the constant density is -1/2
inside), the ID reconstruction bound, and 'green' against 'brute' on the same matrix.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla

from . import hss2d as _h2

RADIUS_REL = 1.75  # HSS1D/HSS1D_compress_nsym.m:51


# ----------------------------------------------------------------------------------------
# Geometry and kernels
# ----------------------------------------------------------------------------------------


@dataclass
class Curve:
    X: np.ndarray       # N x 2 points
    nrm: np.ndarray     # N x 2 outward unit normals
    w: np.ndarray       # N quadrature weights
    kappa: np.ndarray   # N signed curvatures


def star_contour(N, a=0.3, m=5):
    """r(t) = 1 + a cos(m t), trapezoidal nodes t_j = 2 pi j / N."""
    t = 2 * np.pi * np.arange(N) / N
    r = 1 + a * np.cos(m * t)
    rp = -a * m * np.sin(m * t)
    rpp = -a * m * m * np.cos(m * t)
    x, y = r * np.cos(t), r * np.sin(t)
    xp = rp * np.cos(t) - r * np.sin(t)
    yp = rp * np.sin(t) + r * np.cos(t)
    xpp = rpp * np.cos(t) - 2 * rp * np.sin(t) - r * np.cos(t)
    ypp = rpp * np.sin(t) + 2 * rp * np.cos(t) - r * np.sin(t)
    s = np.hypot(xp, yp)
    nrm = np.stack([yp / s, -xp / s], axis=1)
    kappa = (xp * ypp - yp * xpp) / s ** 3
    return Curve(np.stack([x, y], axis=1), nrm, s * 2 * np.pi / N, kappa)


def dl_block(C: Curve, I, J):
    """A[I, J] of the second-kind double-layer operator (targets I, sources J)."""
    X = C.X
    dx = X[I, 0][:, None] - X[J, 0][None, :]
    dy = X[I, 1][:, None] - X[J, 1][None, :]
    r2 = dx * dx + dy * dy
    same = np.asarray(I)[:, None] == np.asarray(J)[None, :]
    num = dx * C.nrm[J, 0][None, :] + dy * C.nrm[J, 1][None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        K = np.where(same, 0.0, num / np.where(same, 1.0, r2)) / (2 * np.pi) * C.w[J][None, :]
    diag = -C.kappa[J] * C.w[J] / (4 * np.pi)
    K = K + same * (diag[None, :] - 0.5)
    return K


def dl_from_points(C: Curve, Z, J):
    """Double-layer kernel at off-curve targets Z (proxy rows) from curve sources J."""
    dx = Z[:, 0][:, None] - C.X[J, 0][None, :]
    dy = Z[:, 1][:, None] - C.X[J, 1][None, :]
    r2 = dx * dx + dy * dy
    num = dx * C.nrm[J, 0][None, :] + dy * C.nrm[J, 1][None, :]
    return num / r2 / (2 * np.pi) * C.w[J][None, :]


def log_to_points(C: Curve, I, Z):
    """Log kernel from off-curve sources Z (proxy columns) at curve targets I."""
    dx = C.X[I, 0][:, None] - Z[:, 0][None, :]
    dy = C.X[I, 1][:, None] - Z[:, 1][None, :]
    return np.log(dx * dx + dy * dy) / (4 * np.pi)


def dense_matrix(C):
    I = np.arange(C.X.shape[0])
    if isinstance(C, PointKernel):
        return C.block(I, I)
    return dl_block(C, I, I)


class DoubleLayer:
    """The double-layer operator of a Curve as the three kinds of blocks the 'green'
    compression evaluates: curve x curve, proxy targets x curve, curve x proxy sources."""

    def __init__(self, C: Curve):
        self.C, self.X = C, C.X

    def block(self, I, J):
        return dl_block(self.C, I, J)

    def rows_from(self, Z, J):
        return dl_from_points(self.C, Z, J)

    def cols_to(self, I, Z):
        return log_to_points(self.C, I, Z)


class PointKernel:
    """Kernel_Eval(X1, X2, params) (Common/Kernel_Eval.m:29-159) on the points X: every block,
    proxy rows and proxy columns included, is one Kernel_Eval call (HSS1D_compress_nsym.m:
    152-156, 222-249)."""

    def __init__(self, X, params):
        self.X, self.params = np.asarray(X, dtype=np.float64), params

    def block(self, I, J):
        return _h2.kernel_eval(self.X[I], self.X[J], self.params)

    def rows_from(self, Z, J):
        return _h2.kernel_eval(Z, self.X[J], self.params)

    def cols_to(self, I, Z):
        return _h2.kernel_eval(self.X[I], Z, self.params)


# ----------------------------------------------------------------------------------------
# Tree, neighbours, proxy circles (HSS1D/HSS1D_compress_nsym.m:277-412)
# ----------------------------------------------------------------------------------------


@dataclass
class Box:
    level: int
    parent: int
    c1: int = -1
    c2: int = -1
    start: int = 0       # 0-based first index of the box's points
    count: int = 0


def get_tree(ntot, nbox_max):
    """LOCAL_get_tree :277-334: halve (first half = ceil) until <= nbox_max, BFS numbering.
    Boxes are 1-based (index 0 unused)."""
    boxes = [None, Box(level=0, parent=0, start=0, count=ntot)]
    first, last, created, level = 1, 0, 1, 0
    while created > 0:
        first, last, created = last + 1, last + created, 0
        for b in range(first, last + 1):
            nb = boxes[b].count
            if nb > nbox_max:
                nh = (nb + 1) // 2
                s1, s2 = last + created + 1, last + created + 2
                boxes[b].c1, boxes[b].c2 = s1, s2
                boxes.append(Box(level + 1, b, start=boxes[b].start, count=nh))
                boxes.append(Box(level + 1, b, start=boxes[b].start + nh, count=nb - nh))
                created += 2
        level += 1
    return boxes


def circum_circle(P):
    """LOCAL_get_circum_circle :339-359."""
    xxc = 0.5 * (P[0] + P[-1])
    R = math.sqrt(np.max((P[:, 0] - xxc[0]) ** 2 + (P[:, 1] - xxc[1]) ** 2))
    if 1.2 * R * R > (P[0, 0] - P[-1, 0]) ** 2 + (P[0, 1] - P[-1, 1]) ** 2:
        xxc = P.mean(axis=0)
        R = math.sqrt(np.max((P[:, 0] - xxc[0]) ** 2 + (P[:, 1] - xxc[1]) ** 2))
    return xxc, R


def circle(xxc, R, n):
    """Proxy points of LOCAL_construct_circle :364-371 (rows 1 and 4)."""
    t = np.linspace(0, 2 * np.pi * (1 - 1 / n), n)
    return np.stack([xxc[0] + R * np.cos(t), xxc[1] + R * np.sin(t)], axis=1)


def neighbour_lists(X, boxes):
    """LOCAL_construct_potential_neighborlist :376-412."""
    nb = len(boxes) - 1
    nei = [None, [1]] + [[] for _ in range(nb - 1)]
    for b in range(2, nb + 1):
        B = boxes[b]
        xxc, R = circum_circle(X[B.start:B.start + B.count])
        pot = []
        for j in nei[B.parent]:
            if boxes[j].c1 > 0:
                pot.append(boxes[j].c1)
            if boxes[j].c2 > 0:
                pot.append(boxes[j].c2)
        for j in pot:
            Bj = boxes[j]
            P = X[Bj.start:Bj.start + Bj.count]
            if np.min((P[:, 0] - xxc[0]) ** 2 + (P[:, 1] - xxc[1]) ** 2) < RADIUS_REL ** 2 * R * R:
                nei[b].append(j)
    for b in range(2, nb + 1):
        nei[b] = sorted(j for j in nei[b] if j != b)
    return nei


# ----------------------------------------------------------------------------------------
# ID (Common/ID.m skel_col_qr and skel_col_fixed)
# ----------------------------------------------------------------------------------------


def id_fixed(A, k):
    """skel_col_fixed, Common/ID.m:106-113 (SVD pseudo-inverse of R11, cutoff 1e-12)."""
    Q, R, I = sla.qr(A, mode="economic", pivoting=True)
    U, D, Vh = np.linalg.svd(R[:k, :k])
    q = int(np.sum(D > 1e-12))
    T = Vh[:q].T @ ((U[:, :q].T @ R[:k, k:]) / D[:q, None])
    return T, I


# ----------------------------------------------------------------------------------------
# Compression, inversion, apply
# ----------------------------------------------------------------------------------------


@dataclass
class HSS1D:
    K: object            # DoubleLayer / PointKernel ('green') or the dense matrix ('brute')
    boxes: list
    nproxy: int
    acc: float
    isk_up: dict = field(default_factory=dict)
    isk_dn: dict = field(default_factory=dict)
    T_up: dict = field(default_factory=dict)
    J_up: dict = field(default_factory=dict)
    T_dn: dict = field(default_factory=dict)
    J_dn: dict = field(default_factory=dict)
    nskel: dict = field(default_factory=dict)
    kskel: dict = field(default_factory=dict)
    M_self: dict = field(default_factory=dict)
    M_sib: dict = field(default_factory=dict)

    def leaf(self, b):
        return self.boxes[b].c1 <= 0 and self.boxes[b].c2 <= 0

    def ind(self, b):
        B = self.boxes[b]
        return np.arange(B.start, B.start + B.count)


def compress(C, nbox_max, acc, nproxy=64, seed=0):
    """HSS1D_compress_nsym('green', X, nbox_max, acc, params), :55-273.  ``C`` is a Curve
    (the double layer) or a PointKernel (Kernel_Eval; the reference's nproxy is 50, :52)."""
    rng = np.random.default_rng(seed)
    kern = C if isinstance(C, PointKernel) else DoubleLayer(C)
    X = kern.X
    boxes = get_tree(X.shape[0], nbox_max)
    H = HSS1D(kern, boxes, nproxy, acc)
    nei = neighbour_lists(X, boxes)
    nb = len(boxes) - 1
    for b in range(nb, 1, -1):
        B = boxes[b]
        if H.leaf(b):
            out = H.ind(b)
            inn = out
        else:
            out = np.concatenate([H.isk_up[B.c1], H.isk_up[B.c2]])
            inn = np.concatenate([H.isk_dn[B.c1], H.isk_dn[B.c2]])
        xxc, R = circum_circle(X[np.sort(np.concatenate([out, inn]))])
        Zp = circle(xxc, RADIUS_REL * R, nproxy)
        out_maybe, in_maybe = [], []
        for j in nei[b]:
            Bj = boxes[j]
            if H.leaf(j):
                out_maybe.append(H.ind(j))
                in_maybe.append(H.ind(j))
            else:
                out_maybe += [H.isk_up[Bj.c1], H.isk_up[Bj.c2]]
                in_maybe += [H.isk_dn[Bj.c1], H.isk_dn[Bj.c2]]
        out_maybe = np.concatenate(out_maybe) if out_maybe else np.zeros(0, int)
        in_maybe = np.concatenate(in_maybe) if in_maybe else np.zeros(0, int)
        rr = (RADIUS_REL * R) ** 2
        out_inside = out_maybe[((X[out_maybe] - xxc) ** 2).sum(1) < rr]
        in_inside = in_maybe[((X[in_maybe] - xxc) ** 2).sum(1) < rr]
        A_out = np.concatenate([kern.block(in_inside, out), kern.rows_from(Zp, out)])
        A_in = np.concatenate([kern.block(inn, out_inside), kern.cols_to(inn, Zp)], axis=1)
        _skeletons(H, b, out, inn, A_out, A_in, rng)
    for b in range(nb, 0, -1):
        B = boxes[b]
        if H.leaf(b):
            H.M_self[b] = kern.block(H.ind(b), H.ind(b))
        else:
            H.M_sib[B.c1] = kern.block(H.isk_dn[B.c1], H.isk_up[B.c2])
            H.M_sib[B.c2] = kern.block(H.isk_dn[B.c2], H.isk_up[B.c1])
    return H


def _skeletons(H, b, out, inn, A_out, A_in, rng):
    """:185-208: IDs of A_out and A_in^T, ranks equalised by a fixed-rank ID of the side
    with the smaller rank, skeleton sets and interpolation matrices recorded."""
    acc = H.acc
    Tout, Jout = _h2.id_acc(A_out, acc, rng)
    Tin, Jin = _h2.id_acc(A_in.T, acc, rng)
    k = max(Tout.shape[0], Tin.shape[0])
    if k != Tout.shape[0]:
        Tout, Jout = id_fixed(A_out, k)
    elif k != Tin.shape[0]:
        Tin, Jin = id_fixed(A_in.T, k)
    H.isk_up[b], H.T_up[b], H.J_up[b] = out[Jout[:k]], Tout, Jout
    H.isk_dn[b], H.T_dn[b], H.J_dn[b] = inn[Jin[:k]], Tin, Jin
    H.nskel[b], H.kskel[b] = out.size, k


def compress_brute(K, nbox_max, acc, seed=0):
    """HSS1D_compress_nsym('brute', K, nbox_max, acc), :158-181, :210-214, :220-271.

    The off-diagonal row of a leaf is every index outside it; that of a parent box is the
    union of the skeletons of the level below (I_src{lev+1}, accumulated in the order the
    boxes are compressed, nboxes down to 2) without the box's own indices."""
    rng = np.random.default_rng(seed)
    K = np.asarray(K)
    N = K.shape[0]
    boxes = get_tree(N, nbox_max)
    H = HSS1D(K, boxes, 0, acc)
    nb = len(boxes) - 1
    I_src = {}
    for b in range(nb, 1, -1):
        B = boxes[b]
        lev = B.level
        lo, hi = B.start, B.start + B.count
        if H.leaf(b):
            out = H.ind(b)
            inn = out
            off_out = np.concatenate([np.arange(0, lo), np.arange(hi, N)])
            off_in = off_out
        else:
            out = np.concatenate([H.isk_up[B.c1], H.isk_up[B.c2]])
            inn = np.concatenate([H.isk_dn[B.c1], H.isk_dn[B.c2]])
            s1, s2 = I_src[lev + 1]
            off_out = s1[(s1 < lo) | (s1 >= hi)]
            off_in = s2[(s2 < lo) | (s2 >= hi)]
        A_out = K[np.ix_(off_in, out)]
        A_in = K[np.ix_(inn, off_out)]
        _skeletons(H, b, out, inn, A_out, A_in, rng)
        if lev > 0:
            s1, s2 = I_src.get(lev, (np.zeros(0, int), np.zeros(0, int)))
            I_src[lev] = (np.concatenate([s1, H.isk_up[b]]), np.concatenate([s2, H.isk_dn[b]]))
    for b in range(nb, 0, -1):
        B = boxes[b]
        if H.leaf(b):
            H.M_self[b] = K[np.ix_(H.ind(b), H.ind(b))]
        else:
            H.M_sib[B.c1] = K[np.ix_(H.isk_dn[B.c1], H.isk_up[B.c2])]
            H.M_sib[B.c2] = K[np.ix_(H.isk_dn[B.c2], H.isk_up[B.c1])]
    return H


def _RL(H: HSS1D, b):
    n, k = H.nskel[b], H.kskel[b]
    R = np.zeros((k, n))
    R[:, H.J_up[b][:k]] = np.eye(k)
    R[:, H.J_up[b][k:]] = H.T_up[b]
    L = np.zeros((n, k))
    L[H.J_dn[b][:k], :] = np.eye(k)
    L[H.J_dn[b][k:], :] = H.T_dn[b].T
    return R, L


def invert(H: HSS1D):
    """HSS1D_invert_nsym.m:44-91."""
    inv = {"Linv": {}, "Rinv": {}, "Dinv": {}}
    E = {}
    nb = len(H.boxes) - 1
    for b in range(nb, 1, -1):
        B = H.boxes[b]
        F = H.M_self[b] if H.leaf(b) else np.block(
            [[E[B.c1], H.M_sib[B.c1]], [H.M_sib[B.c2], E[B.c2]]])
        R, L = _RL(H, b)
        Finv = np.linalg.inv(F)
        E[b] = np.linalg.inv(R @ Finv @ L)
        inv["Linv"][b] = Finv @ L @ E[b]
        inv["Rinv"][b] = E[b] @ R @ Finv
        inv["Dinv"][b] = Finv - inv["Linv"][b] @ (R @ Finv)
    FT = np.block([[E[2], H.M_sib[2]], [H.M_sib[3], E[3]]])
    inv["FTinv"] = np.linalg.inv(FT)
    return inv


def applyinv(H: HSS1D, inv, f):
    """HSS1D_applyinv_nsym.m:43-86."""
    f = f.reshape(f.shape[0], -1)
    nb = len(H.boxes) - 1
    up, dn = {}, {}
    for b in range(nb, 1, -1):
        B = H.boxes[b]
        src = f[H.ind(b)] if H.leaf(b) else np.concatenate([up[B.c1], up[B.c2]])
        up[b] = inv["Rinv"][b] @ src
    xt = inv["FTinv"] @ np.concatenate([up[2], up[3]])
    k2 = inv["Linv"][2].shape[1]
    dn[2], dn[3] = xt[:k2], xt[k2:]
    x = np.zeros_like(f)
    for b in range(2, nb + 1):
        B = H.boxes[b]
        if H.leaf(b):
            x[H.ind(b)] = inv["Linv"][b] @ dn[b] + inv["Dinv"][b] @ f[H.ind(b)]
        else:
            xl = inv["Linv"][b] @ dn[b] + inv["Dinv"][b] @ np.concatenate([up[B.c1], up[B.c2]])
            k1 = H.kskel[B.c1]
            dn[B.c1], dn[B.c2] = xl[:k1], xl[k1:]
    return x


def apply(H: HSS1D, q):
    """HSS1D_apply_nsym.m:38-130."""
    q = q.reshape(q.shape[0], -1)
    nb = len(H.boxes) - 1
    up, dn = {}, {}

    def proj(b, v):
        k = H.kskel[b]
        J = H.J_up[b]
        return v[J[:k]] + H.T_up[b] @ v[J[k:]]

    def evl(b, u):
        k = H.kskel[b]
        J = H.J_dn[b]
        out = np.zeros((J.size, u.shape[1]))
        out[J[:k]] = u
        out[J[k:]] = H.T_dn[b].T @ u
        return out

    for b in range(nb, 1, -1):
        B = H.boxes[b]
        src = q[H.ind(b)] if H.leaf(b) else np.concatenate([up[B.c1], up[B.c2]])
        up[b] = proj(b, src)
    dn[2] = H.M_sib[2] @ up[3]
    dn[3] = H.M_sib[3] @ up[2]
    for b in range(2, nb + 1):
        B = H.boxes[b]
        if not H.leaf(b):
            ul = evl(b, dn[b])
            n1 = H.isk_dn[B.c1].size
            dn[B.c1] = ul[:n1] + H.M_sib[B.c1] @ up[B.c2]
            dn[B.c2] = ul[n1:] + H.M_sib[B.c2] @ up[B.c1]
    u = np.zeros_like(q)
    for b in range(2, nb + 1):
        if H.leaf(b):
            u[H.ind(b)] = evl(b, dn[b]) + H.M_self[b] @ q[H.ind(b)]
    return u
