"""HSS1D path: oracle against dense linear algebra (CPU) and the GPU library against the
oracle and the dense operator (Test_HSS1D_compress.m:79-109 semantics)."""
import numpy as np
import pytest

import hssc_directsolver_amd as H
from oracle import hss1d as o1


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_double_layer_gauss_identity():
    # D 1 = -1/2 inside the contour (Gauss), so A 1 = -1 for A = -1/2 I + D: pins the
    # kernel including its curvature diagonal
    C = o1.star_contour(512)
    A = o1.dense_matrix(C)
    assert np.max(np.abs(A @ np.ones(512) + 1.0)) < 1e-12


def test_geometry_matches_oracle():
    C = o1.star_contour(300)
    G = H.star_contour(300)
    for a, b in ((C.X, G.X), (C.nrm, G.nrm), (C.w, G.w), (C.kappa, G.kappa)):
        assert np.array_equal(a, b)


def test_tree_matches_reference_rule():
    # LOCAL_get_tree: halve (first half = ceil) until <= nbox_max, BFS numbering
    boxes = o1.get_tree(1000, 64)
    leaves = [b for b in range(1, len(boxes)) if boxes[b].c1 < 0]
    assert sum(boxes[b].count for b in leaves) == 1000
    assert all(boxes[b].count <= 64 for b in leaves)
    assert boxes[2].count == 500 and boxes[4].count == 250 and boxes[8].count == 125
    assert boxes[16].count == 63 and boxes[17].count == 62


@pytest.mark.parametrize("N", [1024])
def test_oracle_hss1d_vs_dense(N):
    C = o1.star_contour(N)
    A = o1.dense_matrix(C)
    Hs = o1.compress(C, 64, 1e-10)
    x = np.random.default_rng(1).standard_normal(N)
    assert rel(o1.apply(Hs, x)[:, 0], A @ x) < 1e-10
    y = o1.applyinv(Hs, o1.invert(Hs), x)[:, 0]
    assert rel(y, np.linalg.solve(A, x)) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("N", [1024, 4096])
def test_gpu_hss1d_vs_dense_and_oracle(N):
    C = H.star_contour(N)
    Co = o1.star_contour(N)
    A = o1.dense_matrix(Co)
    K = H.HSS1D_compress_nsym("green", C, 64, 1e-10)
    x = np.random.default_rng(2).standard_normal(N)
    assert rel(H.HSS1D_apply_nsym(K, x), A @ x) < 1e-10
    H.HSS1D_invert_nsym(K)
    y = H.HSS1D_applyinv_nsym(K, x)
    ytrue = np.linalg.solve(A, x)
    assert rel(y, ytrue) < 1e-10
    # exact inverse of its own compressed operator
    assert rel(H.HSS1D_apply_nsym(K, y), x) < 1e-12
    # against the oracle's HSS1D solution and ranks
    Ho = o1.compress(Co, 64, 1e-10)
    yo = o1.applyinv(Ho, o1.invert(Ho), x)[:, 0]
    assert rel(y, yo) < 1e-9
    for b in range(2, K.nboxes + 1):
        assert abs(K.box(b)["kskel"] - Ho.kskel[b]) <= 3, b


@pytest.mark.gpu
def test_gpu_hss1d_batched_matches_single_problems():
    # configuration 3 in small: several star contours of different sizes in one operator
    from hssc_directsolver_amd.rank_mlp import random_contour
    rng = np.random.default_rng(5)
    curves = [random_contour(rng, N) for N in (1024, 1500, 2048, 3000, 1111)]
    K = H.HSS1D_compress_batched(curves, 64, 1e-10)
    H.HSS1D_invert_nsym(K)
    fs = [np.random.default_rng(10 + i).standard_normal(c.N) for i, c in enumerate(curves)]
    f = np.concatenate(fs)
    x = H.HSS1D_applyinv_nsym(K, f)
    assert rel(H.HSS1D_apply_nsym(K, x), f) < 1e-12
    for c, fi, xi in zip(curves, fs, H.split(K, x)):
        Ki = H.HSS1D_compress_nsym("green", c, 64, 1e-10)
        H.HSS1D_invert_nsym(Ki)
        assert rel(xi, H.HSS1D_applyinv_nsym(Ki, fi)) < 1e-10
    # the smallest against the dense solve
    A = o1.dense_matrix(o1.Curve(curves[0].X, curves[0].nrm, curves[0].w, curves[0].kappa))
    assert rel(H.split(K, x)[0], np.linalg.solve(A, fs[0])) < 1e-10


@pytest.mark.gpu
def test_gpu_hss1d_fp32_factors_with_refinement():
    # configuration 3's mixed path: FP32 inverse factors, FP64 refinement against the FP64
    # compressed operator (Common/Iterative_Refinement.m:20-35), single and batched
    from hssc_directsolver_amd.rank_mlp import random_contour
    N = 4096
    C = H.star_contour(N)
    A = o1.dense_matrix(o1.star_contour(N))
    f = np.random.default_rng(3).standard_normal(N)
    K = H.HSS1D_compress_nsym("green", C, 64, 1e-10)
    H.HSS1D_invert_nsym(K)
    y64 = H.HSS1D_applyinv_nsym(K, f)
    H.HSS1D_invert_nsym(K, precision="fp32")
    x, it, relres = H.HSS1D_solve(K, f, tol=1e-10, maxit=10)
    assert relres <= 1e-10 and 2 <= it <= 6
    assert rel(x, y64) < 1e-9
    assert rel(x, np.linalg.solve(A, f)) < 1e-9
    rng = np.random.default_rng(8)
    curves = [random_contour(rng, n) for n in (1024, 2500, 1800)]
    Kb = H.HSS1D_compress_batched(curves, 64, 1e-10)
    H.HSS1D_invert_nsym(Kb, precision="fp32")
    fb = np.random.default_rng(9).standard_normal(sum(c.N for c in curves))
    xb, itb, relb = H.HSS1D_solve(Kb, fb, tol=1e-10, maxit=10)
    assert relb <= 1e-10 and itb <= 6
    assert rel(H.HSS1D_apply_nsym(Kb, xb), fb) <= 1e-10
    H.HSS1D_invert_nsym(Kb)
    assert rel(xb, H.HSS1D_applyinv_nsym(Kb, fb)) < 1e-9
