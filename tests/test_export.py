"""MATLAB round trip of INFO_TREE / A_HSS (export.py): MAT files with the reference's field
names and 1-based indices, read back with scipy.io.loadmat and checked against the library's
getters and against the tree's own geometry."""
import numpy as np
import pytest

import hssc_directsolver_amd as H
from hssc_directsolver_amd.export import export_mat

loadmat = pytest.importorskip("scipy.io").loadmat


def _box(m, b):
    return m["INFO_TREE"]["BOX"][0, 0][0, b - 1]


@pytest.mark.parametrize("TI", [0, 1])
def test_index_sets_round_trip(tmp_path, TI):
    # index sets only (no device needed for a kh = 0 tree)
    P = H.HSS_tree_parameters("SL_L_2D", 0, 1, 4, 8, 1e-10, 2, 400, TI)
    tree = H.HSS2D_build_index_sets(P)
    path = tmp_path / "tree.mat"
    export_mat(str(path), tree, with_T=False)
    m = loadmat(str(path))
    T = m["INFO_TREE"][0, 0]
    D = tree.depth
    assert int(T["depth"][0, 0]) == D
    assert np.array_equal(T["numlev"].ravel(), 2.0 ** np.arange(D + 1))
    assert T["BOX"].shape == (1, 2 ** (D + 1) - 1)
    assert np.array_equal(T["leaves"][:, 0], np.arange(2 ** D, 2 ** (D + 1)))
    X = P.X_source
    for b in range(1, 2 ** (D + 1)):
        e = _box(m, b)
        lev = b.bit_length() - 1
        assert int(e["levbox"][0, 0]) == lev and int(e["parent"][0, 0]) == b // 2
        if lev < D:
            assert np.array_equal(e["child"].ravel(), [2 * b, 2 * b + 1])
        isrc = e["I_src"].ravel().astype(int)
        assert np.array_equal(isrc - 1, tree.I_src(b))
        assert np.array_equal(e["I_sk"].ravel().astype(int) - 1, tree.I_sk(b))
        assert int(e["n_skel"][0, 0]) == isrc.size
        if not TI:
            J = e["J"].ravel().astype(int)
            # J puts the skeleton first: I_src(J(1:k)) = I_sk
            k = int(e["k"][0, 0])
            assert np.array_equal(isrc[J[:k] - 1], e["I_sk"].ravel().astype(int))
        if lev == D:  # a leaf holds its slab of the uniform grid, centred on the box centre
            assert np.allclose(X[isrc - 1].mean(axis=0), e["cent"].ravel(), atol=1e-14)
    if TI:
        assert T["LEV"].shape == (1, D + 1)
        for lev in range(D + 1):
            assert np.array_equal(T["LEV"][0, lev]["J"].ravel().astype(int) - 1, tree.J(2 ** lev))


@pytest.mark.gpu
def test_tree_and_inverse_round_trip(tmp_path):
    P = H.HSS_tree_parameters("SL_H_2D", 10.0, 1, 4, 16, 1e-10, 2, 400, 0)
    tree = H.HSS2D_build_tree_fsym(P)
    inv = H.HSS2D_invert_fsym(tree)
    path = tmp_path / "hss.mat"
    export_mat(str(path), tree, inv)
    m = loadmat(str(path))
    D = tree.depth
    for b in (2, 3, 2 ** D, 2 ** (D + 1) - 1):
        e = _box(m, b)
        assert np.array_equal(e["T"], tree.T(b))
        assert np.array_equal(e["J"].ravel().astype(int) - 1, tree.J(b))
    A = m["A_HSS"]
    assert A.shape == (1, 2 ** (D + 1) - 1)
    for b in (1, 2, 2 ** D):
        assert np.array_equal(A[0, b - 1]["FI"], inv.FI(b))
        if b > 1:
            assert np.array_equal(A[0, b - 1]["E"], inv.E(b))
    # Helmholtz blocks are written as complex MAT arrays
    assert np.iscomplexobj(A[0, 0]["FI"]) and np.iscomplexobj(_box(m, 2)["T"])
