"""Repeated builds in one process give identical factors.

The 512^2 Lippmann-Schwinger workload is built alternately with the mixed path (complex64
tree and factors) and the plain complex128 path, each call sequence as a caller makes it: the
tree, the exact grid operator and the inversion back to back, then the solve refined against
the exact operator.  Every plain build must solve in one step to the same residual bit for
bit, every mixed build in the same number of steps to the same residual.  (An inversion that
overlapped the tree build's streams once gave plain factors that varied between such runs.)"""
import ctypes as C

import pytest

import hssc_directsolver_amd as H

torch = pytest.importorskip("torch")


@pytest.mark.gpu
def test_alternating_builds_are_identical():
    print("alternating mixed / complex128 builds at 512^2", flush=True)
    torch.cuda.set_device(0)
    P = H.HSS_tree_parameters("SL_H_2D", 20.0, 1, 8, 64, 1e-10, 3, 800, 0, proxy_layers=4)
    gen = torch.Generator(device="cpu").manual_seed(0)
    b = torch.complex(torch.randn(P.N, generator=gen, dtype=torch.float64),
                      torch.randn(P.N, generator=gen, dtype=torch.float64))
    b = (b / b.norm()).to("cuda:0")
    seen = {}
    for trial in range(4):
        for prec in ("fp32", "fp64"):
            P.raw.tree_precision = 1 if prec == "fp32" else 0
            torch.cuda.synchronize()
            tree = H.HSS2D_build_tree_fsym(P)
            A = H.GridOperator(P)
            inv = H.HSS2D_invert_fsym(tree, precision=prec)
            y = torch.empty_like(b)
            it, rel = H.solve_exact_device(inv, A, b, y, tol=1e-10, maxit=12)
            torch.cuda.synchronize()
            print(trial, prec, it, rel, flush=True)
            assert rel <= 1e-10
            if prec == "fp64":
                assert it == 1
            assert seen.setdefault(prec, (it, rel)) == (it, rel), (trial, prec)
            del inv, A, tree
    P.raw.tree_precision = 0
