"""The exact grid operator (FFT convolution) against dense kernel matrices, and the solves
refined against it: FP64 trees and inverses, and the mixed path whose interpolation operators
and factorization are FP32 / complex64 (a preconditioner only; the residual is the true one).

Tolerances: the FFT product carries rounding of order eps log2(L^2) relative to ||K|| ||x||,
far below the solver tolerance, so it is checked to 1e-12 against the dense product; the
refined solutions are checked by their true residual (tol 1e-10) and against A\\b."""
import math

import numpy as np
import pytest

import hssc_directsolver_amd as H
from oracle import hss2d as o
from oracle.dr_weights import corr0


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _params(kind, Ng, np_, TI, **kw):
    if kind == "helmholtz":
        h = 2 / (Ng * np_)
        P = H.HSS_tree_parameters("SL_H_2D", 20.0, 1, Ng, np_, 1e-10, 3, 800, TI, proxy_layers=4,
                                  **kw)
        p = o.tree_parameters("SL_H_2D", 20.0, 1, Ng, np_, 1e-10, 3, 800, TI, corr0=corr0(h, 20.0))
        p["proxy_layers"] = 4
    else:
        P = H.HSS_tree_parameters("SL_L_2D", 0, 1, Ng, np_, 1e-10, 2, 400, TI, **kw)
        p = o.tree_parameters("SL_L_2D", 0, 1, Ng, np_, 1e-10, 2, 400, TI)
    if "contrast" in kw:
        p["contrast"] = kw["contrast"]
    if "kernel_scale" in kw:
        p["kernel_scale"] = kw["kernel_scale"]
    if "length_scale" in kw:
        p["length_scale"] = kw["length_scale"]
    return P, p


def _rhs(P, m=1, seed=0):
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((P.N, m))
    if P.is_complex:
        b = b + 1j * rng.standard_normal((P.N, m))
    return b


CASES = [
    ("laplace", 8, 8, 1, {}),
    ("laplace", 8, 8, 0, {}),
    ("laplace", 9, 4, 0, {}),  # G = 36: FFT length 128, non-dyadic grid
    ("laplace", 8, 8, 1, dict(kernel_scale=1 / (2 * math.pi), length_scale=0.5)),
    ("helmholtz", 8, 8, 0, {}),
    ("helmholtz", 8, 4, 1, {}),
    ("helmholtz", 8, 8, 0, dict(contrast=("gaussian", 0.5, 0.15))),
]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,Ng,np_,TI,kw", CASES)
def test_grid_operator_matches_dense(kind, Ng, np_, TI, kw):
    print(f"grid op {kind} Ng={Ng} np={np_} TI={TI} {kw}", flush=True)
    P, p = _params(kind, Ng, np_, TI, **kw)
    A = o.dense_matrix(p)
    x = _rhs(P, m=2, seed=1)
    y = H.GridOperator(P).apply(x)
    assert rel(y, A @ x) < 1e-12


@pytest.mark.gpu
def test_grid_operator_exact_columns_512():
    # the metric's grid: A e_j against the kernel columns K(X, x_j)
    print("grid op exact columns 512^2", flush=True)
    P, p = _params("helmholtz", 8, 64, 0)
    N = P.N
    cols = np.random.default_rng(4).choice(N, 4, replace=False)
    E = np.zeros((N, 4), dtype=complex)
    E[cols, np.arange(4)] = 1.0
    Y = H.GridOperator(P).apply(E)
    X = p["X_source"]
    Yt = o.kernel_eval(X, X[cols], p)
    assert np.max(np.linalg.norm(Y - Yt, axis=0) / np.linalg.norm(Yt, axis=0)) < 1e-12


def _solve_exact(P, tree_prec, fact_prec, b, tol=1e-10, maxit=12):
    import torch
    if tree_prec is not None:
        P.raw.tree_precision = {"fp64": 0, "fp32": 1}[tree_prec]
    tree = H.HSS2D_build_tree_fsym(P)
    inv = H.HSS2D_invert_fsym(tree, precision=fact_prec)
    A = H.GridOperator(P)
    B = torch.from_numpy(np.ascontiguousarray(b.T)).to("cuda:0")
    Y = torch.empty_like(B)
    it, res = H.solve_exact_device(inv, A, B, Y, tol=tol, maxit=maxit)
    return Y.cpu().numpy().T, it, res


@pytest.mark.gpu
@pytest.mark.parametrize("kind,TI", [("laplace", 1), ("laplace", 0), ("helmholtz", 0)])
@pytest.mark.parametrize("tree_prec,fact_prec", [("fp64", "fp64"), ("fp64", "fp32"),
                                                 ("fp32", "fp32"), ("fp32", "fp64")])
def test_solve_refined_against_exact_operator(kind, TI, tree_prec, fact_prec):
    print(f"solve exact {kind} TI={TI} tree {tree_prec} factors {fact_prec}", flush=True)
    P, p = _params(kind, 8, 8, TI)
    A = o.dense_matrix(p)
    b = _rhs(P, seed=2)
    b /= np.linalg.norm(b)
    y, it, res = _solve_exact(P, tree_prec, fact_prec, b)
    # the stopping residual is the true one
    assert res <= 1e-10
    assert np.linalg.norm(b - A @ y) <= 1.01e-10 * np.linalg.norm(b)
    x = np.linalg.solve(A, b)
    assert rel(y, x) < 1e-9 * np.linalg.cond(A)
    if tree_prec == "fp64" and fact_prec == "fp64":
        assert it <= 2  # one correction removes the compression error
    else:
        assert it <= 6


@pytest.mark.gpu
def test_mixed_path_at_512():
    # the metric's workload: FP32 tree and complex64 factors, refined against the exact operator
    import torch
    print("mixed path 512^2 Helmholtz", flush=True)
    P, _ = _params("helmholtz", 8, 64, 0, tree_precision="fp32")
    rng = np.random.default_rng(0)
    b = rng.standard_normal(P.N) + 1j * rng.standard_normal(P.N)
    b /= np.linalg.norm(b)
    y, it, res = _solve_exact(P, None, "fp32", b[:, None])
    assert res <= 1e-10 and it <= 6, (it, res)
    # independent check of the residual with a second operator handle
    r = b - H.GridOperator(P).apply(y[:, 0])
    assert np.linalg.norm(r) <= 1.01e-10
