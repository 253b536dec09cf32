"""The preconditioned iterative solve of HSS2D/Test/Test_HSS2D_InverseCompression.m:201-228:
bicgstab on the compressed variable-coefficient (NTI) Lippmann-Schwinger operator with the HSS
inverse of the translation-invariant (TI) operator, two boundary layers, as the
preconditioner, to tol = acc / 10 within 200 iterations."""
import numpy as np
import pytest

import hssc_directsolver_amd as H
from oracle import hss2d as o

torch = pytest.importorskip("torch")


@pytest.mark.gpu
def test_ti_preconditioned_bicgstab():
    print("TI-preconditioned bicgstab, Lippmann-Schwinger Ng=4 np=16 k=10", flush=True)
    torch.cuda.set_device(0)
    Ng, np_, kh, acc = 4, 16, 10.0, 1e-10
    P = H.HSS_tree_parameters("SL_H_2D", kh, 1, Ng, np_, acc, 3, 400, 0)
    tree = H.HSS2D_build_tree_fsym(P)
    op = H.HSS2D_compress_fsym(tree)
    inv = H.HSS2D_invert_fsym(tree)
    Pti = H.HSS_tree_parameters("SL_H_2D", kh, 1, Ng, np_, acc, 2, 400, 1)
    tti = H.HSS2D_build_tree_fsym(Pti)
    inv_ti = H.HSS2D_invert_fsym(tti)
    X = P.X_source
    bump = o.bump_function(X[:, 0], X[:, 1], P.h)
    vec = -np.sqrt(bump) * np.exp(1j * kh * X[:, 0])
    b = torch.from_numpy(vec).to("cuda:0")

    def A(x):
        y = torch.empty_like(x)
        H.apply_device(op, x.contiguous(), y)
        return y

    def M(x):
        y = torch.empty_like(x)
        H.applyinv_device(inv_ti, x.contiguous(), y)
        return y

    x, flag, relres, it = H.bicgstab(A, b, acc / 10, 200, M)
    assert flag == 0 and relres <= acc / 10, (flag, relres, it)
    # the true residual, and the direct solve with the NTI inverse
    r = torch.linalg.vector_norm(b - A(x)).item() / torch.linalg.vector_norm(b).item()
    assert r <= acc / 10, r
    yd = torch.empty_like(b)
    H.applyinv_device(inv, b, yd)
    err = (torch.linalg.vector_norm(x - yd) / torch.linalg.vector_norm(yd)).item()
    assert err < 1e-8, err
    # the TI preconditioner is what makes the iteration short
    _, flag0, _, it0 = H.bicgstab(A, b, acc / 10, 200, None)
    print(f"preconditioned: {it} iterations; unpreconditioned: {it0} (flag {flag0})", flush=True)
    assert it < it0 or flag0 != 0
    assert it <= 40, it
