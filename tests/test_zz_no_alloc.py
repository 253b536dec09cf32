"""hssc_hss2d_applyinv_device_noalloc (the applyinv variant with the library's fixed device
workspace) against hssc_hss2d_applyinv_device on the same inverse: two calls in a row, the
second on a smaller tree with fewer right-hand sides (the fixed buffer then holds the first
call's contents), FP64 and FP32 factorizations, real and complex; and the size limit."""
import ctypes as C

import numpy as np
import pytest

import hssc_directsolver_amd as H
from hssc_directsolver_amd._lib import HsscError, lib


def _tree(kernel, np_, TI):
    if kernel == "helmholtz":
        P = H.HSS_tree_parameters("SL_H_2D", 20.0, 1, 8, np_, 1e-10, 3, 800, TI, proxy_layers=4)
    else:
        P = H.HSS_tree_parameters("SL_L_2D", 0, 1, 8, np_, 1e-10, 2, 400, TI)
    return P, H.HSS2D_build_tree_fsym(P)


def _rhs(torch, P, m, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    if P.is_complex:
        X = torch.complex(torch.randn(m, P.N, generator=g, dtype=torch.float64),
                          torch.randn(m, P.N, generator=g, dtype=torch.float64))
    else:
        X = torch.randn(m, P.N, generator=g, dtype=torch.float64)
    return X.to("cuda:0")


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,TI", [("helmholtz", 0), ("laplace", 1), ("laplace", 0)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_matches_applyinv_device(kernel, TI, prec):
    import torch
    print(f"noalloc {kernel} TI={TI} {prec}", flush=True)
    big = _tree(kernel, 8, TI)
    small = _tree(kernel, 4, TI)
    for (P, tree), m, seed in ((big, 3, 1), (small, 1, 2)):
        inv = H.HSS2D_invert_fsym(tree, precision=prec)
        X = _rhs(torch, P, m, seed)
        Y0 = torch.empty_like(X)
        Y1 = torch.full_like(X, float("nan"))
        H.applyinv_device(inv, X, Y0)
        H.applyinv_device_noalloc(inv, X, Y1)
        torch.cuda.synchronize()
        assert torch.isfinite(Y1).all()
        err = (torch.linalg.vector_norm(Y1 - Y0) / torch.linalg.vector_norm(Y0)).item()
        assert err <= 1e-14, (m, err)


@pytest.mark.gpu
def test_noalloc_rejects_oversized_workspace():
    import torch
    print("noalloc size limit", flush=True)
    P, tree = _tree("laplace", 32, 1)  # N = 65536
    inv = H.HSS2D_invert_fsym(tree)
    m = 64  # about 0.9 GB of per-level vectors: beyond the fixed buffer
    X = torch.zeros(m, P.N, dtype=torch.float64, device="cuda:0")
    Y = torch.zeros_like(X)
    with pytest.raises(HsscError, match="fixed buffer"):
        H.applyinv_device_noalloc(inv, X, Y)
    # the library keeps working after the refusal, with the same inverse
    Y0 = torch.empty(P.N, dtype=torch.float64, device="cuda:0")
    Y1 = torch.empty_like(Y0)
    x = torch.ones(P.N, dtype=torch.float64, device="cuda:0")
    H.applyinv_device(inv, x, Y0)
    H.applyinv_device_noalloc(inv, x, Y1)
    torch.cuda.synchronize()
    assert torch.equal(Y0, Y1)
