"""Least squares with tall proxies (more than 1024 rows per panel), the longest CholeskyQR3
panels (row-parallel, with Householder reconstruction) of the interpolation operators.  Pinned
as every other level:
the least-squares residual of T on the proxy (HSS2D/Test/Test_HSS2D_skeletons.m:52-70) against
the oracle's LAPACK least squares (interpolation_operator.m:57-66), and the solution against the
oracle's.  TI trees at 128^2 (one box per level, so the oracle stays quick): Laplace (level 1:
M = 1188) and Helmholtz k = 20 with 'bdry' skeletons (levels 1-2: M = 1600, 1088).
"""
import numpy as np
import pytest

import hssc_directsolver_amd as H
from oracle import hss2d as o
from oracle.dr_weights import corr0

CASES = ["laplace", "helmholtz"]
_cache = {}


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def case(kind):
    if kind not in _cache:
        print(f"[tall-panels] building {kind} 128^2 TI", flush=True)
        if kind == "laplace":
            P = H.HSS_tree_parameters("SL_L_2D", 0, 1, 8, 16, 1e-10, 2, 400, 1)
            p = o.tree_parameters("SL_L_2D", 0, 1, 8, 16, 1e-10, 2, 400, 1)
        else:
            P = H.HSS_tree_parameters("SL_H_2D", 20.0, 1, 8, 16, 1e-10, 3, 800, 1,
                                      proxy_layers=4, preset_all=True)
            p = o.tree_parameters("SL_H_2D", 20.0, 1, 8, 16, 1e-10, 3, 800, 1,
                                  corr0=corr0(2 / 128, 20.0))
            p["proxy_layers"] = 4
        tree = H.HSS2D_build_tree_fsym(P)
        otree = o.build_tree(p, interform="dense", preset_all=True)
        _cache[kind] = dict(P=P, p=p, tree=tree, otree=otree)
    return _cache[kind]


@pytest.mark.gpu
@pytest.mark.parametrize("kind", CASES)
def test_tall_panel_least_squares_residual(kind):
    c = case(kind)
    tree, ot, p = c["tree"], c["otree"], c["p"]
    X = p["X_source"]
    tall = 0
    for lev in range(1, ot.depth + 1):
        b = 2 ** lev
        info = tree.level_info(lev)
        tall += info["M"] > 1024
        J, k = ot.J[ot.key(b)], ot.k[ot.key(b)]
        assert np.array_equal(tree.J(b), J)
        K = o.kernel_eval(ot.Xp[ot.key(b)], X[ot.I_src[b]][J], p)
        K11, K21 = K[:, :k], K[:, k:]
        T = tree.T(b)
        res = np.linalg.norm(K11 @ T - K21) / np.linalg.norm(K21)
        res_ref = np.linalg.norm(K11 @ ot.Tdense(b) - K21) / np.linalg.norm(K21)
        assert res <= max(4 * res_ref, 1e-11), (lev, info["M"], res, res_ref)
    assert tall >= 1


@pytest.mark.gpu
@pytest.mark.parametrize("kind", CASES)
def test_tall_panel_tree_solution(kind):
    c = case(kind)
    tree, ot, p = c["tree"], c["otree"], c["p"]
    N = p["X_source"].shape[0]
    rng = np.random.default_rng(11)
    b = rng.standard_normal(N) + (1j * rng.standard_normal(N) if kind == "helmholtz" else 0)
    b /= np.linalg.norm(b)
    y = H.HSS2D_applyinv_TI(tree, H.HSS2D_invert_fsym(tree), b)
    yor = o.applyinv(ot, o.invert(ot), b).ravel()
    assert rel(y, yor) < 1e-9
    op = H.HSS2D_compress_fsym(tree)
    assert np.linalg.norm(b - H.HSS2D_apply_TI(tree, op, y)) < 1e-10
