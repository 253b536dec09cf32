"""MATLAB round trip of the compressed representation: INFO_TREE and A_HSS written as MAT
files (scipy.io.savemat, MAT v5) with the field names and 1-based indices of the reference's
structures, so a MATLAB session can load them and call the reference's apply functions on
them, and loadmat reads them back.

INFO_TREE (HSSC_DirectSolver/DOCUMENTATION_HSS2D.txt:262-292, HSS2D_build_tree_fsym.m):
  .depth, .numlev (boxes per level 0..depth), .leaves (index and level of each leaf box),
  .BOX(b), b = 1 .. 2^(depth+1) - 1 heap-numbered: .cent, .levbox, .parent (0 for the root),
  .child, .I_src, .I_sk, .n_skel, .k, .J (= [I_sk; I_rs] as positions in I_src), .T (k x
  (n_skel - k), dense); for TI kernels (params.transinv) J and T are stored once per level in
  .LEV(lev + 1) instead of per box.
A_HSS (DOCUMENTATION_HSS2D.txt:808-817, HSS2D_invert_fsym.m): A_HSS(b).FI (n x n, dense
F^-1) and A_HSS(b).E (k x k; the root has none); for TI kernels one entry per level.
"""
from __future__ import annotations

import numpy as np


def _centres(depth):
    """Box centres of the heap-numbered tree: the root at the origin; the children of a box
    on level lev at (-s, +s) along x when lev is even and (+s, -s) along y when it is odd,
    s = 2^-(lev // 2) / 2 (the bisection of HSS2D_build_tree_fsym.m:1289-1344, as the library
    builds it)."""
    nb = 2 ** (depth + 1)
    c = np.zeros((nb, 2))
    for lev in range(depth):
        s = 0.5 * 2.0 ** (-(lev // 2))
        for b in range(2 ** lev, 2 ** (lev + 1)):
            if lev % 2 == 0:
                c[2 * b] = c[b] + (-s, 0.0)
                c[2 * b + 1] = c[b] + (s, 0.0)
            else:
                c[2 * b] = c[b] + (0.0, s)
                c[2 * b + 1] = c[b] + (0.0, -s)
    return c


def info_tree(tree, with_T=True):
    """INFO_TREE of a built tree as nested dicts / object arrays in MATLAB's layout."""
    D = tree.depth
    TI = bool(tree.params.transinv)
    nbox = 2 ** (D + 1) - 1
    cent = _centres(D)
    fields = ["cent", "levbox", "parent", "child", "I_src", "I_sk", "n_skel", "k", "J"]
    if with_T:
        fields.append("T")
    box = np.empty((1, nbox), dtype=[(f, "O") for f in fields])
    for b in range(1, nbox + 1):
        lev = b.bit_length() - 1
        e = box[0, b - 1]
        e["cent"] = cent[b].reshape(1, 2)
        e["levbox"] = float(lev)
        e["parent"] = float(b // 2)
        e["child"] = (np.array([[2 * b, 2 * b + 1]], dtype=float) if lev < D
                      else np.zeros((1, 0)))
        isrc = tree.I_src(b)
        e["I_src"] = (isrc + 1).astype(float).reshape(-1, 1)
        e["I_sk"] = (tree.I_sk(b) + 1).astype(float).reshape(-1, 1)
        e["n_skel"] = float(isrc.size)
        e["k"] = float(tree.level_info(lev)["k"])
        if TI:
            e["J"] = np.zeros((0, 1))
            if with_T:
                e["T"] = np.zeros((0, 0))
        else:
            e["J"] = (tree.J(b) + 1).astype(float).reshape(-1, 1)
            if with_T:
                e["T"] = tree.T(b) if lev > 0 else np.zeros((0, 0))
    out = dict(depth=float(D), numlev=np.array([[2.0 ** l] for l in range(D + 1)]),
               leaves=np.array([[float(b), float(D)] for b in range(2 ** D, 2 ** (D + 1))]),
               BOX=box)
    if TI:
        lf = ["J", "T"] if with_T else ["J"]
        lv = np.empty((1, D + 1), dtype=[(f, "O") for f in lf])
        for lev in range(D + 1):
            b = 2 ** lev
            lv[0, lev]["J"] = (tree.J(b) + 1).astype(float).reshape(-1, 1)
            if with_T:
                lv[0, lev]["T"] = tree.T(b) if lev > 0 else np.zeros((0, 0))
        out["LEV"] = lv
    return out


def a_hss(inv):
    """A_HSS of an inverse (HSS2D_invert_fsym): FI and E per box (per level for TI)."""
    tree = inv.tree
    D = tree.depth
    TI = bool(tree.params.transinv)
    n = D + 1 if TI else 2 ** (D + 1) - 1
    arr = np.empty((1, n), dtype=[("FI", "O"), ("E", "O")])
    for i in range(n):
        b = 2 ** i if TI else i + 1
        arr[0, i]["FI"] = inv.FI(b)
        arr[0, i]["E"] = inv.E(b) if b > 1 else np.zeros((0, 0))
    return arr


def export_mat(path, tree, inv=None, with_T=True):
    """Writes INFO_TREE (and A_HSS when an inverse is given) to a MAT file."""
    from scipy.io import savemat
    d = {"INFO_TREE": info_tree(tree, with_T=with_T)}
    if inv is not None:
        d["A_HSS"] = a_hss(inv)
    savemat(path, d, do_compression=False, oned_as="column")
