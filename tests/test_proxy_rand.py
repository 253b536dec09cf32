"""The 'rand' proxy rule (HSS2D/HSS2D_build_tree_fsym.m:899-921, LOCAL_random_sampler
:979-1006): the random stream, the density of the kept points by dyadic shell, the library's
index sets and proxies bit-exact against the oracle, and (GPU) solutions with thinned proxies
against the dense solve."""
import math

import numpy as np
import pytest

import hssc_directsolver_amd as H
from oracle import hss2d as o
from oracle.dr_weights import corr0


def test_default_stream_is_matlabs():
    # MATLAB's documented rand(3) right after rng('default') (column-major order)
    doc = [0.8147, 0.9058, 0.1270, 0.9134, 0.6324, 0.0975, 0.2785, 0.5469, 0.9575]
    assert np.allclose(o.default_rand(9), doc, atol=5e-5)


def test_sampler_density_by_dyadic_shell():
    # a thick ring (32 layers) around a 2h x 2h box: every point within lay*h of the box is
    # kept; beyond, the kept fraction of shell j (distances in (rad/2^(j+1), rad/2^j]) is
    # 2^(-lambda * (sz - 1 - j)), sz = 1 - floor(log2(lay h / rad)).  Shells are counted here
    # from the integer layer index t (distance (t - 1/2) h), not from the sampler's formulas.
    h, lay, L, lam = 1 / 64, 2, 32, 2.0
    cB = (0.5 * h, 0.5 * h)
    X = o.box_proxy(cB, h, h, L, h)
    t = np.rint(np.max(np.abs(X - cB), axis=1) / h - 1 + 0.5).astype(int)  # layer 1..L
    assert t.min() == 1 and t.max() == L
    Xk = o.random_sampler(X, cB, h, h, lam, lay, h)
    kept = {tuple(p) for p in Xk}
    keep = np.array([tuple(p) in kept for p in X])
    assert keep[t <= lay].all()
    rad = (L - 0.5) * h
    sz = 1 - math.floor(math.log2(lay * h / rad))
    dist = (t - 0.5) * h
    for j in range(sz - 1):
        sel = (t > lay) & (dist <= rad / 2 ** j) & (dist > rad / 2 ** (j + 1))
        if j == sz - 2:
            sel = (t > lay) & (dist <= rad / 2 ** j)
        if sel.sum() < 50:
            continue
        p = 2.0 ** (-lam * (sz - 1 - j))
        frac, sd = keep[sel].mean(), math.sqrt(p * (1 - p) / sel.sum())
        assert abs(frac - p) < 5 * sd + 1e-12, (j, frac, p)


def _rand_params(pot, Ng, np_, TI, lib):
    kh = 0.0 if pot == "SL_L_2D" else 8.0
    if lib:
        return H.HSS_tree_parameters(pot, kh, 1, Ng, np_, 1e-10, 2 if kh == 0 else 3,
                                     400, TI, proxy_rule="rand", preset_all=kh > 0)
    p = o.tree_parameters(pot, kh, 1, Ng, np_, 1e-10, 2 if kh == 0 else 3, 400, TI,
                          corr0=corr0(2 / (Ng * np_), kh) if kh > 0 else None)
    p["proxy_rule"] = "rand"
    return p


@pytest.mark.parametrize("pot,Ng,np_,TI", [("SL_L_2D", 8, 4, 1), ("SL_L_2D", 8, 8, 0),
                                          ("SL_L_2D", 8, 16, 1), ("SL_H_2D", 8, 8, 0)])
def test_rand_proxy_index_sets_bit_exact(pot, Ng, np_, TI):
    P = _rand_params(pot, Ng, np_, TI, lib=True)
    p = _rand_params(pot, Ng, np_, TI, lib=False)
    t = H.HSS2D_build_index_sets(P)
    ref = o.build_tree(p, with_T=False, preset_all=True)
    thinned = False
    for lev in range(1, ref.depth + 1):
        b = 2 ** lev
        key = ref.key(b)
        full = o.box_proxy(ref.cent[b], *o.level_radii(lev), o.proxy_layers(lev, p), p["h"])
        thinned |= ref.Xp[key].shape[0] < full.shape[0]
        for bb in ([b] if TI else range(2 ** lev, 2 ** (lev + 1))):
            assert np.array_equal(t.I_sk(bb), ref.I_sk[bb]), bb
            assert np.array_equal(t.J(bb), ref.J[ref.key(bb)]), bb
            assert np.array_equal(t.proxy(bb), ref.Xp[ref.key(bb)]), bb
    assert thinned


@pytest.mark.gpu
@pytest.mark.parametrize("pot,TI", [("SL_L_2D", 1), ("SL_L_2D", 0), ("SL_H_2D", 0)])
def test_gpu_solve_with_rand_proxies(pot, TI):
    # the compressed inverse built on thinned proxies solves A x = b like the dense solve
    print(f"rand proxies {pot} TI={TI}", flush=True)
    kh = 0.0 if pot == "SL_L_2D" else 8.0
    P = H.HSS_tree_parameters(pot, kh, 1, 8, 8, 1e-10, 2 if kh == 0 else 3, 400, TI,
                              proxy_rule="rand")
    X = P.X_source
    A = H.Kernel_Eval(X, X, P)
    b = np.random.default_rng(4).standard_normal(P.N).astype(P.dtype)
    b /= np.linalg.norm(b)
    tree = H.HSS2D_build_tree_fsym(P)
    inv = H.HSS2D_invert_fsym(tree)
    y = H.HSS2D_applyinv_TI(tree, inv, b)
    ytrue = np.linalg.solve(A, b)
    assert np.linalg.norm(y - ytrue) / np.linalg.norm(ytrue) < 1e-8
    assert np.linalg.norm(A @ y - b) < 1e-9
