"""The skeleton IDs of the HSS2D ID-selected levels are deterministic: repeated builds of the
same tree give the same pivot orders and ranks, box by box.  The register-resident CPQR ID
(the kernel of the complex skeleton levels) once let thread 0 swap the pivot positions while
other waves still read them, so a wave could take another pivot column than the rest and the
box got a wrong ID now and then; repeated builds of a 4096-box level show such a race."""
import numpy as np
import pytest

import hssc_directsolver_amd as H


@pytest.mark.gpu
def test_repeated_index_sets_are_identical():
    P = H.HSS_tree_parameters("SL_H_2D", 20, 1, 8, 64, 1e-10, 3, 800, 0, proxy_layers=4)
    ref = None
    for rep in range(8):
        print(f"512^2 index sets, build {rep}", flush=True)
        t = H.HSS2D_build_index_sets(P)
        D = t.depth
        got = {lev: t.box_ranks(lev)[0].copy() for lev in (D, D - 1)}
        # the key boxes' skeleton orders of the two ID levels
        for lev in (D, D - 1):
            got[(lev, "J")] = np.concatenate([t.J(b) for b in range(2 ** lev, 2 ** (lev + 1))])
        if ref is None:
            ref = got
        else:
            for key in ref:
                assert np.array_equal(ref[key], got[key]), (rep, key)
        del t
