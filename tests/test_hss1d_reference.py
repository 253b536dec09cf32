"""HSS1D on the reference's own semantics (HSS1D/Test/Run_HSS1D_tests.m:15-27,
HSS1D/Test/Test_HSS1D_compress.m:24-110): the skeleton points of box 2 of a 2D tree
(SL_L_2D, kh = 0, sym = 0, Ng = 7, np = 16, acc = 1e-10, layers 2, params.proxy = 1),
A = Kernel_Eval(X_sk, X_sk, params), compressed with leaf size 56 by the 'green' trick
(Kernel_Eval on points, 50-point proxy circles) and by 'brute' force on the dense A, applied
to 10 random right-hand sides and inverted, against the dense A @ q and A \\ q.

CPU: the oracle's restatement of both compressions against dense linear algebra, and the
params.proxy option of Kernel_Eval's factors.  GPU: the library's compressions through the
C ABI against the dense operator and the oracle."""
import numpy as np
import pytest

import hssc_directsolver_amd as H
from oracle import hss1d as o1
from oracle import hss2d as o

NMAX, ACC = 56, 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def reference_case():
    """Run_HSS1D_tests.m:15-19 + Test_HSS1D_compress.m:24-38: X_sk = X(I_sk of box 2)."""
    P = o.tree_parameters("SL_L_2D", 0, 0, 7, 16, ACC, 2, 432, 0)
    P["proxy"] = 1
    T = o.build_index_sets(dict(P, sym=1))  # index sets do not depend on sym
    X = P["X_source"][np.asarray(T.I_sk[2])]
    return P, X


@pytest.fixture(scope="module")
def case():
    P, X = reference_case()
    A = o.kernel_eval(X, X, P)
    q = np.random.default_rng(0).standard_normal((X.shape[0], 10))
    return P, X, A, q


def test_proxy_option_drops_row_factor():
    # Kernel_Eval.m:129-154 with proxy = 1, kh = 0, sym = 0: b = 1, c = b_func(y; 0.2, -0.6)
    P, X = reference_case()
    X1, X2 = X[:7], X[100:105]
    A = o.kernel_eval(X1, X2, P)
    d2 = ((X1[:, None, :] - X2[None, :, :]) ** 2).sum(-1)
    c = 1 + 0.25 * np.exp(-(X2[:, 0] - 0.2) ** 2 - (X2[:, 1] + 0.6) ** 2)
    assert np.allclose(A, P["h"] ** 2 * 0.5 * np.log(d2) * c[None, :], rtol=1e-14, atol=0)
    P0 = dict(P, proxy=0)
    b = 1 + 0.25 * np.exp(-(X1[:, 0] - 0.3) ** 2 - (X1[:, 1] - 0.6) ** 2)
    assert np.allclose(o.kernel_eval(X1, X2, P0), A * b[:, None], rtol=1e-14, atol=0)
    P2 = dict(P, proxy=-1)  # c dropped, b kept
    assert np.allclose(o.kernel_eval(X1, X2, P2), P["h"] ** 2 * 0.5 * np.log(d2) * b[:, None],
                       rtol=1e-14, atol=0)


def test_reference_case_shape(case):
    P, X, A, q = case
    # box 2 is the left half of [-1,1]^2: two layers along its three outer sides
    assert X.shape[0] == 656 and np.all(X[:, 0] < 0)
    assert np.linalg.cond(A) < 2


@pytest.mark.parametrize("mode", ["green", "brute"])
def test_oracle_hss1d_reference_semantics(case, mode):
    P, X, A, q = case
    Hs = (o1.compress(o1.PointKernel(X, P), NMAX, ACC, nproxy=50) if mode == "green"
          else o1.compress_brute(A, NMAX, ACC))
    assert rel(o1.apply(Hs, q), A @ q) < 1e-11
    y = o1.applyinv(Hs, o1.invert(Hs), q)
    assert rel(y, np.linalg.solve(A, q)) < 1e-11
    # every parent's skeletons are drawn from its children's
    for b in range(2, len(Hs.boxes)):
        B = Hs.boxes[b]
        if not Hs.leaf(b):
            kids = np.concatenate([Hs.isk_up[B.c1], Hs.isk_up[B.c2]])
            assert np.isin(Hs.isk_up[b], kids).all()


def test_oracle_brute_rows_are_offdiagonal(case):
    # 'brute' compresses the true off-diagonal block rows: planting a change inside a leaf's
    # own diagonal block leaves that leaf's skeletons unchanged, one outside changes A_out
    P, X, A, q = case
    Hs = o1.compress_brute(A, NMAX, ACC)
    leaf = max(b for b in range(2, len(Hs.boxes)) if Hs.leaf(b))
    lo = Hs.boxes[leaf].start
    A2 = A.copy()
    A2[lo, lo + 1] += 5.0
    H2 = o1.compress_brute(A2, NMAX, ACC)
    assert np.array_equal(H2.isk_up[leaf], Hs.isk_up[leaf])
    assert rel(o1.apply(H2, q), A2 @ q) < 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["green", "brute"])
def test_gpu_hss1d_reference_semantics(case, mode):
    P, X, A, q = case
    Pg = H.HSS_tree_parameters("SL_L_2D", 0, 0, 7, 16, ACC, 2, 432, 0, proxy=1)
    # the library's Kernel_Eval with the same params is the oracle's matrix
    assert rel(H.Kernel_Eval(X, X, Pg), A) < 1e-14
    if mode == "green":
        K = H.HSS1D_compress_nsym("green", X, NMAX, ACC, Pg)
        Ho = o1.compress(o1.PointKernel(X, P), NMAX, ACC, nproxy=50)
    else:
        K = H.HSS1D_compress_nsym("brute", A, NMAX, ACC)
        Ho = o1.compress_brute(A, NMAX, ACC)
    u = H.HSS1D_apply_nsym(K, q)
    assert u.shape == q.shape
    assert rel(u, A @ q) < 1e-11
    H.HSS1D_invert_nsym(K)
    y = H.HSS1D_applyinv_nsym(K, q)
    assert rel(y, np.linalg.solve(A, q)) < 1e-11
    # one column at a time gives the same columns
    assert rel(H.HSS1D_applyinv_nsym(K, q[:, 3]), y[:, 3]) < 1e-14
    # the oracle's compression of the same matrix: solutions and ranks
    yo = o1.applyinv(Ho, o1.invert(Ho), q)
    assert rel(y, yo) < 1e-10
    for b in range(2, K.nboxes + 1):
        assert abs(K.box(b)["kskel"] - Ho.kskel[b]) <= 3, b


@pytest.mark.gpu
def test_gpu_hss1d_kernel_rejects_complex_kernel():
    Pg = H.HSS_tree_parameters("SL_H_2D", 10.0, 0, 4, 4, ACC, 2, 432, 0)
    X = Pg.X_source[:200]
    with pytest.raises(H.HsscError):
        H.HSS1D_compress_nsym("green", X, NMAX, ACC, Pg)
