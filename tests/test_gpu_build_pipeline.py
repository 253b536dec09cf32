"""The tree build queues each level's least squares while the host still builds the levels
above (HSS2D_build_tree_fsym's index sets and the interpolation operators overlap).  A level
that cannot be built stops the build with an error while the levels below it are already
running on the GPU; the library drains them, and the process keeps working."""
import numpy as np
import pytest

import hssc_directsolver_amd as H
from hssc_directsolver_amd._lib import STATUS

torch = pytest.importorskip("torch")


@pytest.mark.gpu
def test_failing_level_after_queued_levels():
    # Ng = 4, np = 32, two boundary layers, one-layer proxy rings: the leaf level (k = 16,
    # M = 20) is queued, the level above (k = 32, M = 28) has fewer proxy points than skeleton
    # points and stops the build.
    print("pipelined build: failing level 9 after the queued leaf level", flush=True)
    P = H.HSS_tree_parameters("SL_L_2D", 0, 1, 4, 32, 1e-10, 2, 400, 0, proxy_layers=1)
    with pytest.raises(H.HsscError) as e:
        H.HSS2D_build_tree_fsym(P)
    assert STATUS[e.value.status] == "UNSUPPORTED"
    assert "proxy smaller than skeleton" in str(e.value)
    # the same process builds and solves a valid tree afterwards
    Q = H.HSS_tree_parameters("SL_L_2D", 0, 1, 4, 32, 1e-10, 2, 400, 0)
    tree = H.HSS2D_build_tree_fsym(Q)
    inv = H.HSS2D_invert_fsym(tree)
    op = H.HSS2D_compress_fsym(tree)
    b = np.random.default_rng(0).standard_normal(Q.N)
    y = H.HSS2D_applyinv_fsym(tree, inv, b)
    res = np.linalg.norm(b - H.HSS2D_apply_fsym(tree, op, y)) / np.linalg.norm(b)
    assert res < 1e-10, res
