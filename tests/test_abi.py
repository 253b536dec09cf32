"""The C-ABI library loads, exports every declared entry point, and its host-side index
sets agree exactly with the oracle (no GPU needed)."""
import re
from pathlib import Path

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
    # per-level (n, k) of the reference's 'bdry' rule at N = 512^2 (Ng=8, np=64, layers 2)
    want = {12: (64, 48), 11: (96, 80), 10: (160, 112), 9: (224, 176), 8: (352, 240),
            7: (480, 368), 6: (736, 496), 5: (992, 752), 4: (1504, 1008), 3: (2016, 1520),
            2: (3040, 2032), 1: (4064, 3056), 0: (6112, 4080)}
    t = H.HSS2D_build_index_sets(H.HSS_tree_parameters("SL_L_2D", 0, 1, 8, 64, 1e-10, 2, 400, 1))
    for lev, (n, k) in want.items():
        info = t.level_info(lev)
        assert (info["n"], info["k"]) == (n, k), lev


def test_gpu_entry_points_fail_loudly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    P = H.HSS_tree_parameters("SL_L_2D", 0, 1, 8, 4, 1e-10, 2, 400, 1)
    with pytest.raises(H.HsscError):
        H.HSS2D_build_tree_fsym(P)
