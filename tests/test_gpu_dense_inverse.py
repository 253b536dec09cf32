"""The dense block inverse (MATLAB's inv on F and R F^-1 R^T: HSS2D_invert_fsym.m:108,
HSS2D_build_Finv_fsym.m:101, HSS2D_build_E_fsym.m:51,64) through hssc_dense_invert_device:
the register kernel for one block of up to 64 columns (128 for real entries) and the blocked
sweeps above (for complex symmetric blocks of 2048 columns or more, 128-wide steps whose
diagonal blocks take the 64-wide sweep), against LAPACK's inverse (numpy), on matrices that
need row pivoting, complex symmetric blocks, and the singular-pivot flag."""
import numpy as np
import pytest

import hssc_directsolver_amd as H

torch = pytest.importorskip("torch")

DT = {"f64": (np.float64, "float64"), "f32": (np.float32, "float32"),
      "c128": (np.complex128, "complex128"), "c64": (np.complex64, "complex64")}
TOL = {"f64": 1e-10, "c128": 1e-10, "f32": 2e-3, "c64": 2e-3}


def rand(rng, shape, kind):
    x = rng.standard_normal(shape)
    if kind.startswith("c"):
        x = x + 1j * rng.standard_normal(shape)
    return x


def run(blocks, kind, sym=False):
    """blocks: (batch, n, n) numpy matrices -> their inverses from the library."""
    npd, td = DT[kind]
    # column-major blocks: the tensor holds each block transposed
    t = torch.from_numpy(np.ascontiguousarray(np.swapaxes(blocks, 1, 2)).astype(npd)).to("cuda:0")
    H.dense_inv_device(t, sym=sym)
    return np.swapaxes(t.cpu().numpy(), 1, 2).astype(np.complex128 if kind.startswith("c")
                                                       else np.float64)


def check(blocks, X, kind):
    for A, Xi in zip(blocks, X):
        ref = np.linalg.inv(A)
        err = np.linalg.norm(Xi - ref) / np.linalg.norm(ref)
        assert err < TOL[kind] * np.linalg.cond(A) / 10, (kind, A.shape, err)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["c64", "c128", "f32", "f64"])
@pytest.mark.parametrize("n", [1, 5, 37, 64])
def test_register_kernel_pivots(kind, n):
    print(f"dense inverse {kind} n={n}", flush=True)
    rng = np.random.default_rng(n)
    blocks = []
    for b in range(3):
        # a row permutation of a unit upper triangular matrix: every column's largest entry
        # sits off the diagonal, so the sweep must pick its pivots among the rows
        T = np.triu(0.1 * rand(rng, (n, n), kind), 1) + np.eye(n)
        blocks.append(T[rng.permutation(n)] if b < 2 else rand(rng, (n, n), kind))
    blocks = np.array(blocks)
    check(blocks, run(blocks, kind), kind)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["c64", "c128", "f64"])
@pytest.mark.parametrize("n", [100, 200])
def test_blocked_sweep(kind, n):
    print(f"dense inverse blocked {kind} n={n}", flush=True)
    rng = np.random.default_rng(7 + n)
    blocks = np.array([rand(rng, (n, n), kind) + 4 * np.sqrt(n) * np.eye(n) for _ in range(2)])
    check(blocks, run(blocks, kind), kind)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["c64", "c128"])
@pytest.mark.parametrize("n", [48, 64, 200, 2240])
def test_complex_symmetric_blocks(kind, n):
    print(f"dense inverse symmetric {kind} n={n}", flush=True)
    rng = np.random.default_rng(n)
    B = rand(rng, (2 if n < 1000 else 1, n, n), kind)
    blocks = B + np.swapaxes(B, 1, 2) + 3 * np.sqrt(n) * np.eye(n)
    X = run(blocks, kind, sym=True)
    check(blocks, X, kind)
    assert np.array_equal(X, np.swapaxes(X, 1, 2))  # symmetric to the bit


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["c64", "f64"])
def test_singular_pivot_is_reported(kind):
    print(f"dense inverse singular {kind}", flush=True)
    rng = np.random.default_rng(3)
    A = rand(rng, (1, 20, 20), kind)
    A[0, :, 7] = 0  # a zero column
    with pytest.raises(H.HsscError) as e:
        run(A, kind)
    assert e.value.status == 3  # HSSC_ERR_SINGULAR
    A = rand(rng, (1, 20, 20), kind)
    A[0, 4, 4] = np.nan
    with pytest.raises(H.HsscError):
        run(A, kind)
