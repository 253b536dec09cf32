"""The oracle (CPU restatement of the reference) pinned against the reference's published
outputs, closed forms and dense linear algebra (no GPU needed)."""
import numpy as np
import pytest
import scipy.special as sp

from oracle import hss2d as o
from oracle.dr_weights import corr0


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_kernel_closed_forms():
    # SL_L_2D: (den==0) + h^2 * 0.5 * b log(den) c  (Kernel_Eval.m:60-61, b_func :99-108)
    p = o.tree_parameters("SL_L_2D", 0, 1, 8, 4, 1e-10, 2, 400, 0)
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, (30, 2))
    A = o.kernel_eval(X, X, p)
    b = 1 + 0.25 * np.exp(-(X[:, 0] - 0.3) ** 2 - (X[:, 1] - 0.6) ** 2)
    r = np.hypot(X[:, 0][:, None] - X[:, 0], X[:, 1][:, None] - X[:, 1])
    np.fill_diagonal(r, 1.0)
    want = p["h"] ** 2 * b[:, None] * np.log(r) * b[None, :]
    np.fill_diagonal(want, 1.0)
    assert np.max(np.abs(A - want)) < 1e-15
    # SL_H_2D: H0 from AMOS (hankel1) against J0 + i Y0 (cephes); diagonal w = 1 + k^2 i/4 corr0
    h = 2 / 32
    c0 = corr0(h, 20.0)
    ph = o.tree_parameters("SL_H_2D", 20.0, 1, 8, 4, 1e-10, 3, 800, 0, corr0=c0)
    A = o.kernel_eval(X, X, ph)
    bb = np.sqrt(o.bump_function(X[:, 0], X[:, 1], h))
    z = 20.0 * r
    want = h * h * 400 * 0.25j * bb[:, None] * (sp.j0(z) + 1j * sp.y0(z)) * bb[None, :]
    np.fill_diagonal(want, 1 + 400 * 0.25j * c0)
    assert np.max(np.abs(A - want)) < 1e-14 * np.max(np.abs(want))
    assert np.array_equal(A, A.T) or np.max(np.abs(A - A.T)) < 1e-16


def test_install_sample_dense_inverse_application():
    # INSTALL.txt:57-63 (Run_HSS2D_tests: Ng = 16, np = 8, TI Laplace): the dense-block
    # inverse applied to a random unit RHS leaves a residual against the compressed operator
    # of 7.28e-15 ("Error of Dense Inverse Application")
    p = o.tree_parameters("SL_L_2D", 0, 1, 16, 8, 1e-10, 2, 400, 1)
    tree = o.build_tree(p, interform="dense")
    inv = o.invert(tree)
    comp = o.compress(tree)
    b = o.random_rhs(p["X_source"].shape[0], seed=0)
    y = o.applyinv(tree, inv, b)
    assert np.linalg.norm(b - o.apply(tree, comp, y)) < 5e-14


@pytest.mark.parametrize("TI", [1, 0])
def test_hss2d_laplace_vs_dense(TI):
    # Test_HSS2D_InverseCompression.m:300-311 (Y vs A\b) and Test_HSS2D_Matrix_Apply.m:227-235
    p = o.tree_parameters("SL_L_2D", 0, 1, 8, 8, 1e-10, 2, 400, TI)
    tree = o.build_tree(p, interform="dense")
    inv = o.invert(tree)
    A = o.dense_matrix(p)
    b = o.random_rhs(A.shape[0], seed=1)[:, 0]
    y = o.applyinv(tree, inv, b).ravel()
    assert rel(y, np.linalg.solve(A, b)) < 1e-9
    x = np.random.default_rng(2).standard_normal(A.shape[0])
    assert rel(o.apply(tree, o.compress(tree), x).ravel(), A @ x) < 1e-10


def test_hss2d_helmholtz_vs_dense():
    h = 2 / 32
    p = o.tree_parameters("SL_H_2D", 20.0, 1, 8, 4, 1e-10, 3, 800, 0, corr0=corr0(h, 20.0))
    p["proxy_layers"] = 4
    tree = o.build_tree(p, interform="dense", preset_all=True)
    inv = o.invert(tree)
    A = o.dense_matrix(p)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(A.shape[0]) + 1j * rng.standard_normal(A.shape[0])
    y = o.applyinv(tree, inv, b).ravel()
    assert rel(y, np.linalg.solve(A, b)) < 1e-10
    # a plausible mistake -- conjugating instead of transposing the complex symmetric
    # sibling block -- is caught by the same check
    K = o.kernel_eval(p["X_source"][:5], p["X_source"][5:9], p)
    assert not np.allclose(K.T, K.conj().T)


def test_lowrank_interpolation_matches_dense():
    # the reference's default INTERform 'lowrank' (interpolation_operator.m:80-127) gives
    # the same compressed operator to tolerance as the dense least squares
    p = o.tree_parameters("SL_L_2D", 0, 1, 8, 8, 1e-10, 2, 400, 1)
    A = o.dense_matrix(p)
    b = o.random_rhs(A.shape[0], seed=4)[:, 0]
    yd = o.applyinv(t := o.build_tree(p, interform="dense"), o.invert(t), b).ravel()
    yl = o.applyinv(t2 := o.build_tree(p, interform="lowrank"), o.invert(t2), b).ravel()
    assert rel(yl, yd) < 1e-8
