"""The C-ABI library loads, exports every declared entry point, and its host-side index
sets agree exactly with the oracle (no GPU needed)."""
import re
from pathlib import Path

import ctypes

import numpy as np
import pytest

import hssc_directsolver_amd as H
from hssc_directsolver_amd import _lib
from oracle import hss2d as o

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "hssc.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hssc_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name
    bound = {s[0] for s in _lib.SIGNATURES}
    assert set(names) == bound


def test_params_match_reference_defaults():
    # Common/HSS_tree_parameters.m:108-139
    p = H.HSS_tree_parameters("SL_L_2D", 0, 1, 8, 32, 1e-10, 2, 400, 1)
    assert p.depth == 10 and p.N == 65536 and p.h == 2 / 256
    assert p.max_particles == 64 and p.k_cut == 2500 and p.transinv == 1
    p = H.HSS_tree_parameters("SL_L_2D", 0, 1, 8, 64, 1e-10, 2, 400, 0)
    assert p.depth == 12 and p.N == 262144


def test_invalid_arguments_raise():
    with pytest.raises(H.HsscError):
        H.HSS_tree_parameters("SL_L_2D", 0, 1, 8, 6, 1e-10, 2, 400, 1)  # np not a power of 2
    with pytest.raises(H.HsscError):
        H.HSS_tree_parameters("SL_L_2D", 0, 1, 8, 8, 2.0, 2, 400, 1)    # acc out of range


@pytest.mark.parametrize("TI", [1, 0])
@pytest.mark.parametrize("Ng,np_", [(8, 4), (8, 8), (16, 8), (8, 16)])
def test_index_sets_bit_exact(TI, Ng, np_):
    P = H.HSS_tree_parameters("SL_L_2D", 0, 1, Ng, np_, 1e-10, 2, 400, TI)
    p = o.tree_parameters("SL_L_2D", 0, 1, Ng, np_, 1e-10, 2, 400, TI)
    assert np.array_equal(P.X_source, p["X_source"])
    t = H.HSS2D_build_index_sets(P)
    ref = o.build_index_sets(p)
    for b in range(1, 2 ** (ref.depth + 1)):
        assert np.array_equal(t.I_src(b), ref.I_src[b]), b
        assert np.array_equal(t.I_sk(b), ref.I_sk[b]), b
        assert np.array_equal(t.J(b), ref.J[ref.key(b)]), b
        assert np.array_equal(t.proxy(b), ref.Xp[ref.key(b)]), b


def test_appendix_level_table_512():
    # per-level (n, k, M) of the reference's 'bdry' rules at N = 512^2 (Ng=8, np=64,
    # layers 2): M is the full proxy ring (corner points kept, LOCAL_Build_Proxy), i.e.
    # (w + 2L)(h + 2L) - w h points for a w x h box and L = layers + 1 (L = layers at the root).
    want = {12: (64, 48, 132), 11: (96, 80, 180), 10: (160, 112, 228), 9: (224, 176, 324),
            8: (352, 240, 420), 7: (480, 368, 612), 6: (736, 496, 804), 5: (992, 752, 1188),
            4: (1504, 1008, 1572), 3: (2016, 1520, 2340), 2: (3040, 2032, 3108),
            1: (4064, 3056, 4644), 0: (6112, 4080, 4112)}
    t = H.HSS2D_build_index_sets(H.HSS_tree_parameters("SL_L_2D", 0, 1, 8, 64, 1e-10, 2, 400, 1))
    for lev, (n, k, M) in want.items():
        info = t.level_info(lev)
        assert (info["n"], info["k"], info["M"]) == (n, k, M), lev


def test_helmholtz_level_table_512():
    # Lippmann-Schwinger Helmholtz at N = 512^2 (k = 20, layers 3, NTI) with 4-layer proxy
    # rings: the same geometry rules with layers = 3 (every box of a level equal)
    want = {12: (64, 60, 192), 11: (120, 108, 256), 10: (216, 156, 320), 9: (312, 252, 448),
            8: (504, 348, 576), 7: (696, 540, 832), 6: (1080, 732, 1088), 5: (1464, 1116, 1600),
            4: (2232, 1500, 2112), 3: (3000, 2268, 3136), 2: (4536, 3036, 4160),
            1: (6072, 4572, 6208), 0: (9144, 6108, 6180)}
    P = H.HSS_tree_parameters("SL_H_2D", 20.0, 1, 8, 64, 1e-10, 3, 800, 0, proxy_layers=4,
                              preset_all=True)
    t = H.HSS2D_build_index_sets(P)
    for lev, (n, k, M) in want.items():
        info = t.level_info(lev)
        assert (info["n"], info["k"], info["M"]) == (n, k, M), lev


def test_dr_corr0_known_answer_and_oracle():
    from oracle.dr_weights import corr0
    # INSTALL.txt:96-115 (the reference's Fortran side-car, h = 1/32, rk = 1)
    ref = complex(0.976562500000008e-03, -0.304146129302758e-02)
    c = corr0(1 / 32, 1.0)
    assert abs(c - ref) < 2e-17
    assert abs(c.real - (1 / 32) ** 2) == 0
    L = _lib.lib()
    cr, ci = ctypes.c_double(), ctypes.c_double()
    for h, k in [(1 / 32, 1.0), (1 / 256, 20.0), (2 / 512, 20.0), (1 / 64, 50.0), (0.01, 100.0)]:
        _lib.check(L.hssc_dr_corr0(h, k, ctypes.byref(cr), ctypes.byref(ci)))
        want = corr0(h, k)
        assert abs(complex(cr.value, ci.value) - want) <= 1e-15 * abs(want), (h, k)
    # beyond the side-car's tables (hk <= 1.01) the series still holds: hk = 1.25, 2
    for h, k in [(1 / 16, 20.0), (1 / 10, 20.0)]:
        _lib.check(L.hssc_dr_corr0(h, k, ctypes.byref(cr), ctypes.byref(ci)))
        want = corr0(h, k, terms=40)
        assert abs(complex(cr.value, ci.value) - want) <= 1e-15 * abs(want), (h, k)
    assert L.hssc_dr_corr0(1 / 32, 100.0, ctypes.byref(cr), ctypes.byref(ci)) != 0
    # params_init sets corr0 for kh > 0 (HSS_tree_parameters.m:167-175)
    P = H.HSS_tree_parameters("SL_H_2D", 20.0, 1, 8, 64, 1e-10, 3, 800, 0)
    assert abs(complex(P.dr_weights_re, P.dr_weights_im) - corr0(1 / 256, 20.0)) < 1e-18


def test_gpu_entry_points_fail_loudly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    P = H.HSS_tree_parameters("SL_L_2D", 0, 1, 8, 4, 1e-10, 2, 400, 1)
    with pytest.raises(H.HsscError):
        H.HSS2D_build_tree_fsym(P)
    # kh > 0 index sets need the GPU for the ID-selected bottom levels, and not otherwise
    Ph = H.HSS_tree_parameters("SL_H_2D", 20.0, 1, 8, 4, 1e-10, 3, 800, 0)
    with pytest.raises(H.HsscError):
        H.HSS2D_build_index_sets(Ph)
    Ph = H.HSS_tree_parameters("SL_H_2D", 20.0, 1, 8, 4, 1e-10, 3, 800, 0, preset_all=True)
    assert H.HSS2D_build_index_sets(Ph).level_info(4)["k"] == 60
