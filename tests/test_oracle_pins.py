"""Pins of the oracle's definitional functions by properties that a plausible mistake breaks:
the grid ordering against a literal MATLAB meshgrid / column-major reshape, the tanh plateau
and b_func by their shapes, the non-symmetric weights, the configuration-literal kernel
options by their closed forms, and the Gaussian-contrast finding (no GPU needed)."""
import math

import numpy as np
import pytest

from oracle import hss2d as o
from oracle.dr_weights import corr0


# ---- A1: X_source ordering (Common/HSS_tree_parameters.m:144-160) ------------------------

def matlab_grid(Ng, np_):
    """HSS_tree_parameters.m:144-160 statement by statement, with numpy's meshgrid (MATLAB's
    meshgrid(z) = meshgrid(z, z), 'xy' indexing) and MATLAB's column-major reshape/(:)."""
    hp = 2.0 / np_
    hh = 2.0 / (Ng * np_)
    z = np.linspace(0.5 * hh, Ng * hh - 0.5 * hh, Ng)
    ZZ1, ZZ2 = np.meshgrid(z, z)
    J = Ng * Ng
    zzloc = np.vstack([ZZ1.reshape(1, J, order="F"), ZZ2.reshape(1, J, order="F")])
    corners = np.arange(np_) * hp - 1.0  # -1:hp:1-hp
    X1, Y1 = np.meshgrid(corners, corners)
    Xmin = np.column_stack([X1.ravel(order="F"), Y1.ravel(order="F")])
    X = np.empty((np_ * np_ * J, 2))
    for i in range(np_ * np_):
        X[J * i:J * (i + 1)] = Xmin[i] + zzloc.T
    return X


def test_source_points_tiny_grid_by_hand():
    # Ng = 2, np = 2 (h = 1/2), worked out by hand from meshgrid + (:): panels in the order
    # (-1,-1), (-1,0), (0,-1), (0,0) of their lower-left corners, and inside a panel the
    # points (x, y) with y running fastest
    want = np.array([
        [-0.75, -0.75], [-0.75, -0.25], [-0.25, -0.75], [-0.25, -0.25],
        [-0.75, 0.25], [-0.75, 0.75], [-0.25, 0.25], [-0.25, 0.75],
        [0.25, -0.75], [0.25, -0.25], [0.75, -0.75], [0.75, -0.25],
        [0.25, 0.25], [0.25, 0.75], [0.75, 0.25], [0.75, 0.75]])
    assert np.array_equal(o.source_points(2, 2), want)
    assert np.array_equal(matlab_grid(2, 2), want)


@pytest.mark.parametrize("Ng,np_", [(2, 2), (3, 4), (8, 4), (5, 8)])
def test_source_points_vs_literal_meshgrid(Ng, np_):
    X = o.source_points(Ng, np_)
    assert np.max(np.abs(X - matlab_grid(Ng, np_))) < 1e-15
    # and the library (host code, no device needed) agrees bit for bit with the oracle
    import hssc_directsolver_amd as H
    P = H.HSS_tree_parameters("SL_L_2D", 0, 1, Ng, np_, 1e-10, 2, 400, 1)
    assert np.array_equal(P.X_source, X)


# ---- A2: the tanh plateau (Common/Bump_function.m:53-62) ---------------------------------

def test_bump_function_shape():
    h = 1.0 / 64
    D, c = 10.0, 4 * h  # delta = 1/10, c = 4h
    f = o.bump_function
    y = np.array([0.0, 0.3, -0.7])
    shoulder = 1 - c
    # product of one-variable factors: f(x,y) f(0,0) = f(x,0) f(0,y)
    x = np.array([0.2, -0.9, 0.97])
    assert np.allclose(f(x, y, h) * f(0.0, 0.0, h), f(x, 0.0, h) * f(0.0, y, h), rtol=1e-14)
    # even in each coordinate
    assert np.array_equal(f(-x, y, h), f(x, y, h)) and np.array_equal(f(x, -y, h), f(x, y, h))
    # half height at |x| = 1 - 4h: the x factor there is exactly 1/2 (tanh 0)
    fx0 = f(0.0, y, h)  # = f1(0) f2(y), with f1(0) = 1 - O(exp(-2 D (1 - c)))
    one_minus = 0.5 * math.tanh(D * (1 - c)) + 0.5
    assert np.allclose(f(shoulder, y, h), 0.5 * fx0 / one_minus, rtol=1e-14)
    assert np.allclose(f(-shoulder, y, h), 0.5 * fx0 / one_minus, rtol=1e-14)
    # slope of the x factor at the shoulder: d/dx [tanh(D (1 - c - x)) / 2] = -D / 2
    d = 1e-6
    g = lambda t: f(t, 0.0, h) / f(0.0, 0.0, h) * one_minus
    slope = (g(shoulder + d) - g(shoulder - d)) / (2 * d)
    assert abs(slope - (-D / 2)) < 1e-6
    # plateau inside; at the edge |x| = 1 the factor is the logistic 1 / (1 + e^(2 D c))
    assert abs(f(0.0, 0.0, h) - 1.0) < 1e-7
    assert abs(g(1.0) - 1 / (1 + math.exp(2 * D * c))) < 1e-14
    # h = 0 (the non-symmetric option, Kernel_Eval.m:118 with Bump_function's nargin == 3):
    # the shoulder moves to |x| = 1
    assert abs(f(1.0, 0.0, 0.0) / f(0.0, 0.0, 0.0) - 0.5 / (0.5 * math.tanh(D) + 0.5)) < 1e-14


# ---- A2: b_func (Common/Kernel_Eval.m:99-108) and the per-point weights (:111-159) ---------

def test_b_func_radial_bump():
    x0, y0 = 0.3, 0.6
    assert o.b_func(np.array([x0]), np.array([y0]), x0, y0, 0.25)[0] == 1.25
    r = np.array([0.1, 0.5, 1.3])
    for th in (0.0, 1.0, 2.5, 4.0):
        v = o.b_func(x0 + r * math.cos(th), y0 + r * math.sin(th), x0, y0, 0.25)
        assert np.allclose(v, 1 + 0.25 * np.exp(-r * r), rtol=1e-15)
    assert np.array_equal(o.b_func(r, r, x0, y0, 0), np.ones(3))
    # on the reference grid the weight of the symmetric NTI Laplace kernel peaks at (0.3, 0.6)
    p = o.tree_parameters("SL_L_2D", 0, 1, 10, 4, 1e-10, 2, 400, 0)
    X = p["X_source"]
    b = o.point_weights(X, p, "row")
    i = np.argmax(b)
    assert np.hypot(X[i, 0] - 0.3, X[i, 1] - 0.6) <= p["h"] / math.sqrt(2) + 1e-15
    assert np.array_equal(o.point_weights(X, p, "col"), b)  # sym: c = b


def test_nonsymmetric_weights():
    X = np.array([[0.3, 0.6], [0.2, -0.6], [0.0, 0.0], [0.95, -0.2]])
    # sym = 0 Laplace NTI: b centred at (0.3, 0.6), c at (0.2, -0.6) (Kernel_Eval.m:146-150)
    p = o.tree_parameters("SL_L_2D", 0, 0, 8, 4, 1e-10, 2, 400, 0)
    b, c = o.point_weights(X, p, "row"), o.point_weights(X, p, "col")
    assert b[0] == 1.25 and c[1] == 1.25 and b[1] < 1.25 and c[0] < 1.25
    # sym = 0 Helmholtz: b = Bump_function with h = 0, c = 1 (:115-128)
    ph = o.tree_parameters("SL_H_2D", 20.0, 0, 8, 4, 1e-10, 3, 800, 0, corr0=corr0(1 / 16, 20.0))
    bh, ch = o.point_weights(X, ph, "row"), o.point_weights(X, ph, "col")
    assert np.array_equal(ch, np.ones(4))
    assert np.allclose(bh, o.bump_function(X[:, 0], X[:, 1], 0.0), rtol=0)
    # TI: ones
    pt = o.tree_parameters("SL_L_2D", 0, 1, 8, 4, 1e-10, 2, 400, 1)
    assert np.array_equal(o.point_weights(X, pt, "row"), np.ones(4))


# ---- configuration-literal kernel options (no reference counterpart) ----------------------

def test_kernel_scale_and_length_scale_closed_form():
    # the configuration's I + h^2 (1/2 pi) log|x - y| on [0,1]^2 with h = 1/256 is the
    # reference grid (h = 2/256 on [-1,1]^2) with L = 1/2 and s = 1/(2 pi)
    p = o.tree_parameters("SL_L_2D", 0, 1, 8, 32, 1e-10, 2, 400, 1)
    p["kernel_scale"] = 1 / (2 * math.pi)
    p["length_scale"] = 0.5
    X = p["X_source"][::997][:40]
    A = o.kernel_eval(X, X, p)
    Y = (X + 1.0) / 2.0  # the same points on [0,1]^2
    r = np.hypot(Y[:, 0][:, None] - Y[:, 0], Y[:, 1][:, None] - Y[:, 1])
    np.fill_diagonal(r, 1.0)
    want = (1 / 256) ** 2 * np.log(r) / (2 * math.pi)
    np.fill_diagonal(want, 1.0)
    assert np.max(np.abs(A - want)) < 1e-18
    # s = L = 1 is the reference's kernel
    p["kernel_scale"] = p["length_scale"] = 1.0
    q = o.tree_parameters("SL_L_2D", 0, 1, 8, 32, 1e-10, 2, 400, 1)
    assert np.array_equal(o.kernel_eval(X, X, p), o.kernel_eval(X, X, q))


def test_gaussian_contrast_weights():
    ph = o.tree_parameters("SL_H_2D", 20.0, 1, 8, 4, 1e-10, 3, 800, 0, corr0=corr0(1 / 16, 20.0))
    ph["contrast"] = ("gaussian", 0.5, 0.15)
    X = np.array([[0.0, 0.0], [0.15, 0.0], [0.0, -0.3]])
    b = o.point_weights(X, ph, "row")
    assert np.allclose(b ** 2, [0.5, 0.5 * math.exp(-0.5), 0.5 * math.exp(-2.0)], rtol=1e-15)


def test_gaussian_contrast_defeats_proxy_compression():
    # The configuration's Gaussian contrast (amplitude 0.5, width 0.15) instead of the
    # reference's tanh plateau: the reference's algorithm ('bdry' skeletons and proxies at
    # every level, restated by the oracle) misses A\b by orders of magnitude although A is
    # well conditioned, while the plateau solves to 1e-13 on the same grid.  (With ID
    # skeletons at the two bottom levels the Gaussian case reaches only ~6e-3.)
    def solve(contrast):
        h = 2 / 32
        p = o.tree_parameters("SL_H_2D", 20.0, 1, 8, 4, 1e-10, 3, 800, 0, corr0=corr0(h, 20.0))
        p["proxy_layers"] = 4
        if contrast:
            p["contrast"] = ("gaussian", 0.5, 0.15)
        tree = o.build_tree(p, interform="dense", preset_all=True)
        A = o.dense_matrix(p)
        rng = np.random.default_rng(3)
        b = rng.standard_normal(A.shape[0]) + 1j * rng.standard_normal(A.shape[0])
        y = o.applyinv(tree, o.invert(tree), b).ravel()
        x = np.linalg.solve(A, b)
        return np.linalg.norm(y - x) / np.linalg.norm(x), np.linalg.cond(A)

    err_g, cond_g = solve(True)
    err_p, _ = solve(False)
    assert cond_g < 10
    assert err_g > 1.0
    assert err_p < 1e-12


# ---- A12: levHSS (HSS2D/levHSS.m:33-65) ---------------------------------------------------

def test_levhss_known_answers():
    import hssc_directsolver_amd as H
    # 512^2 Laplace TI: level skeletons 4080, 3056, 2032, 1520, 1008, 752, 496, 368, ... (the
    # 'bdry' rule); the first level with k <= 400 is 7, odd, so levHSS returns 6
    P = H.HSS_tree_parameters("SL_L_2D", 0, 1, 8, 64, 1e-10, 2, 400, 1)
    T = H.HSS2D_build_index_sets(P)
    assert [T.level_info(l)["k"] for l in range(8)] == [4080, 3056, 2032, 1520, 1008, 752, 496,
                                                       368]
    assert H.levHSS(T, 400) == 6
    # a threshold above the root's skeleton: -1 (the dense mode of HSS2D_invert_fsym.m:61-65)
    assert H.levHSS(T, 5000) == -1
    # never past depth - 1 (= 11 here, odd -> 10)
    assert H.levHSS(T, 1) == 10


@pytest.mark.parametrize("Ng,np_,TI,n_cut", [(8, 8, 1, 100), (8, 8, 0, 100), (8, 16, 1, 150),
                                             (8, 16, 0, 300), (8, 4, 1, 10), (6, 8, 0, 60)])
def test_levhss_library_vs_oracle(Ng, np_, TI, n_cut):
    import hssc_directsolver_amd as H
    p = o.tree_parameters("SL_L_2D", 0, 1, Ng, np_, 1e-10, 2, n_cut, TI)
    P = H.HSS_tree_parameters("SL_L_2D", 0, 1, Ng, np_, 1e-10, 2, n_cut, TI)
    assert H.levHSS(H.HSS2D_build_index_sets(P), n_cut) == o.levhss(o.build_index_sets(p), n_cut,
                                                                     TI)
