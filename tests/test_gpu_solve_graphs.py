"""Repeated solves with one inverse and one exact grid operator replay HIP graphs of the
refinement (captured on the second solve with that operator and m).  Every replay must give
what the directly launched solve gives on the same right-hand side, bit for bit (the same
kernels on other buffers), also with other work on the same inverse between the replays, and
a change of m or of the operator must capture again."""
import numpy as np
import pytest

import hssc_directsolver_amd as H


def _problem(prec):
    P = H.HSS_tree_parameters("SL_H_2D", 20.0, 1, 8, 8, 1e-10, 3, 800, 0, proxy_layers=4)
    tree = H.HSS2D_build_tree_fsym(P)
    inv = H.HSS2D_invert_fsym(tree, precision=prec)
    return P, tree, inv


def _rhs(torch, P, m, seed):
    g = torch.Generator().manual_seed(seed)
    shape = (P.N,) if m == 1 else (m, P.N)
    b = torch.complex(torch.randn(shape, generator=g, dtype=torch.float64),
                      torch.randn(shape, generator=g, dtype=torch.float64))
    return b.to("cuda:0")


def _solve(torch, inv, A, b):
    y = torch.empty_like(b)
    it, rel = H.solve_exact_device(inv, A, b, y, tol=1e-10, maxit=12)
    torch.cuda.synchronize()
    return y.cpu().numpy(), it, rel


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_graph_replays_match_direct_solves(prec):
    import torch
    print(f"solve graphs {prec}", flush=True)
    P, tree, inv = _problem(prec)
    A = H.GridOperator(P)
    bs = [_rhs(torch, P, 1, s) for s in range(3)]
    # first solve with this operator: direct launches (the reference results)
    direct = []
    for b in bs:
        # a fresh inverse per right-hand side keeps every one of these solves direct
        inv_d = H.HSS2D_invert_fsym(tree, precision=prec)
        direct.append(_solve(torch, inv_d, A, b))
        del inv_d
    # the same inverse: solve 1 direct, solve 2 captures, the rest replay
    for rep in range(2):
        for i, b in enumerate(bs):
            y, it, rel = _solve(torch, inv, A, b)
            y0, it0, rel0 = direct[i]
            assert it == it0 and rel == rel0, (rep, i, it, it0, rel, rel0)
            assert np.array_equal(y, y0), (rep, i, np.abs(y - y0).max())
            # other work on the same inverse between replays
            z = torch.empty_like(b)
            H.applyinv_device(inv, b, z)
            torch.cuda.synchronize()
    # another m, then m = 1 again, then another operator: each captures anew
    b2 = _rhs(torch, P, 2, 7)
    for _ in range(3):
        y2, it2, rel2 = _solve(torch, inv, A, b2)
        assert rel2 <= 1e-10
        for j in range(2):
            yj, _, _ = _solve(torch, H.HSS2D_invert_fsym(tree, precision=prec), A, b2[j].clone())
            assert np.abs(y2[j] - yj).max() <= 1e-8 * np.abs(yj).max()
    A2 = H.GridOperator(P)
    for _ in range(3):
        y, it, rel = _solve(torch, inv, A2, bs[0])
        assert np.array_equal(y, direct[0][0])
