"""GPU runs at the configurations' own sizes and options, checked the way the reference's tests
check them at that size, plus the kernel options without a reference counterpart against the
oracle's closed forms.

  * N = 65 536 Laplace (Ng 8, np 32), TI and NTI: the residual of the dense-block inverse
    application against the compressed operator, and A_HSS applied to unit vectors against
    exact kernel columns (Test_HSS2D_Matrix_Apply.m:239-247: for N >= 1e4 the reference
    compares the columns Kernel_Eval(X, X(Jp,:)) instead of a dense A).
  * The published sample of Run_HSS2D_tests.m:7 (Ng 16, np 8, TI Laplace, INSTALL.txt:57-63):
    "Error of Dense Inverse Application" 7.28e-15.
  * A leaf shape whose least-squares panels take the merged sub-panel split-K products with
    a one-panel skeleton (n = 81 > 64, k = 32, M = 208).
"""
import math

import numpy as np
import pytest

import hssc_directsolver_amd as H
from oracle import hss2d as o
from oracle.dr_weights import corr0


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def colnorm_err(Y, Yt):
    # the reference's measure sqrt(max(sum((Y_true - Y).^2))) (largest column 2-norm)
    return float(np.sqrt(np.max(np.sum(np.abs(Yt - Y) ** 2, axis=0))))


@pytest.mark.gpu
@pytest.mark.parametrize("TI", [1, 0])
def test_laplace_256_full_size(TI):
    print(f"laplace 256^2 TI={TI}", flush=True)
    P = H.HSS_tree_parameters("SL_L_2D", 0, 1, 8, 32, 1e-10, 2, 400, TI)
    p = o.tree_parameters("SL_L_2D", 0, 1, 8, 32, 1e-10, 2, 400, TI)
    N = P.N
    assert N == 65536
    tree = H.HSS2D_build_tree_fsym(P)
    inv = H.HSS2D_invert_fsym(tree)
    op = H.HSS2D_compress_fsym(tree)
    b = o.random_rhs(N, seed=0)[:, 0]
    y = H.HSS2D_applyinv_TI(tree, inv, b)
    assert np.linalg.norm(b - H.HSS2D_apply_TI(tree, op, y)) < 1e-12
    # exact columns: A_HSS e_j against K(X, x_j)
    cols = np.random.default_rng(5).choice(N, 6, replace=False)
    E = np.zeros((N, 6))
    E[cols, np.arange(6)] = 1.0
    Y = H.HSS2D_apply_TI(tree, op, E)
    X = p["X_source"]
    Yt = o.kernel_eval(X, X[cols], p)
    assert colnorm_err(Y, Yt) < 1e-9
    # the solution of the compressed system solves the exact one to the compression error
    # (the exact residual through the exact columns of a second right-hand side, A y = sum of
    # K(X, x_j) y_j is too costly here; instead A_HSS ~ A is checked above column by column)


@pytest.mark.gpu
def test_install_sample_ng16_np8():
    print("install sample Ng=16 np=8", flush=True)
    P = H.HSS_tree_parameters("SL_L_2D", 0, 1, 16, 8, 1e-10, 2, 400, 1)
    p = o.tree_parameters("SL_L_2D", 0, 1, 16, 8, 1e-10, 2, 400, 1)
    tree = H.HSS2D_build_tree_fsym(P)
    inv = H.HSS2D_invert_fsym(tree)
    op = H.HSS2D_compress_fsym(tree)
    b = o.random_rhs(P.N, seed=0)[:, 0]
    y = H.HSS2D_applyinv_TI(tree, inv, b)
    # INSTALL.txt:59: mean Error of Dense Inverse Application 7.284727e-15
    assert np.linalg.norm(b - H.HSS2D_apply_TI(tree, op, y)) < 5e-14
    otree = o.build_tree(p, interform="dense")
    yo = o.applyinv(otree, o.invert(otree), b).ravel()
    assert rel(y, yo) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("TI", [1, 0])
def test_merged_split_k_shape(TI):
    # leaf n = 81, k = 32 (one 32-column panel), M = 208: the sub-panel product of the merged
    # trailing block needs ceil(208/128) * 32 * 49 partial sums per box
    print(f"merged split-K shape TI={TI}", flush=True)
    kw = dict(proxy_layers=4)
    P = H.HSS_tree_parameters("SL_L_2D", 0, 1, 9, 4, 1e-10, 1, 400, TI, **kw)
    p = o.tree_parameters("SL_L_2D", 0, 1, 9, 4, 1e-10, 1, 400, TI)
    p["proxy_layers"] = 4
    tree = H.HSS2D_build_tree_fsym(P)
    leaf = tree.level_info(tree.depth)
    assert (leaf["n"], leaf["k"], leaf["M"]) == (81, 32, 208)
    inv = H.HSS2D_invert_fsym(tree)
    A = o.dense_matrix(p)
    b = o.random_rhs(A.shape[0], seed=7)[:, 0]
    y = H.HSS2D_applyinv_TI(tree, inv, b)
    otree = o.build_tree(p, interform="dense")
    yo = o.applyinv(otree, o.invert(otree), b).ravel()
    x = np.linalg.solve(A, b)
    assert rel(y, yo) < 1e-9
    assert rel(y, x) <= 1.5 * rel(yo, x) + 1e-10


# ---- kernel options --------------------------------------------------------------------

@pytest.mark.gpu
def test_kernel_eval_nonsymmetric_and_options():
    rng = np.random.default_rng(11)
    X1 = rng.uniform(-1, 1, (33, 2))
    X2 = np.concatenate([X1[:4], rng.uniform(-1, 1, (21, 2))])
    h = 1 / 16
    cases = []
    # sym = 0 (Kernel_Eval.m:115-128 and :146-150): b and c differ
    cases.append((H.HSS_tree_parameters("SL_L_2D", 0, 0, 8, 4, 1e-10, 2, 400, 0),
                  o.tree_parameters("SL_L_2D", 0, 0, 8, 4, 1e-10, 2, 400, 0)))
    c0 = corr0(h, 20.0)
    cases.append((H.HSS_tree_parameters("SL_H_2D", 20.0, 0, 8, 4, 1e-10, 3, 800, 0, corr0=c0),
                  o.tree_parameters("SL_H_2D", 20.0, 0, 8, 4, 1e-10, 3, 800, 0, corr0=c0)))
    # Gaussian contrast
    pg = o.tree_parameters("SL_H_2D", 20.0, 1, 8, 4, 1e-10, 3, 800, 0, corr0=c0)
    pg["contrast"] = ("gaussian", 0.5, 0.15)
    cases.append((H.HSS_tree_parameters("SL_H_2D", 20.0, 1, 8, 4, 1e-10, 3, 800, 0, corr0=c0,
                                        contrast=("gaussian", 0.5, 0.15)), pg))
    # configuration-literal Laplace: (1/2 pi) log on [0,1]^2
    pl = o.tree_parameters("SL_L_2D", 0, 1, 8, 4, 1e-10, 2, 400, 1)
    pl["kernel_scale"], pl["length_scale"] = 1 / (2 * math.pi), 0.5
    cases.append((H.HSS_tree_parameters("SL_L_2D", 0, 1, 8, 4, 1e-10, 2, 400, 1,
                                        kernel_scale=1 / (2 * math.pi), length_scale=0.5), pl))
    for P, p in cases:
        A = H.Kernel_Eval(X1, X2, P)
        R = o.kernel_eval(X1, X2, p)
        assert np.max(np.abs(A - R)) <= 1e-13 * np.max(np.abs(R)), p["flag_pot"]


@pytest.mark.gpu
def test_config_literal_laplace_solve():
    # I + h^2 (1/2 pi) log|x - y| on [0,1]^2, h = 1/64 here (the configuration's operator at
    # a size the dense solve checks), TI and NTI trees of the reference
    for TI in (1, 0):
        P = H.HSS_tree_parameters("SL_L_2D", 0, 1, 8, 8, 1e-10, 2, 400, TI,
                                  kernel_scale=1 / (2 * math.pi), length_scale=0.5)
        p = o.tree_parameters("SL_L_2D", 0, 1, 8, 8, 1e-10, 2, 400, TI)
        p["kernel_scale"], p["length_scale"] = 1 / (2 * math.pi), 0.5
        tree = H.HSS2D_build_tree_fsym(P)
        inv = H.HSS2D_invert_fsym(tree)
        A = o.dense_matrix(p)
        b = o.random_rhs(A.shape[0], seed=2)[:, 0]
        y = H.HSS2D_applyinv_TI(tree, inv, b)
        assert rel(y, np.linalg.solve(A, b)) < 1e-9, TI
