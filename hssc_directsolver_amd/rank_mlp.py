"""Learned rank prediction for the randomized IDs (the part that replaces HSS2D/predict_rank.m).

A RankMLP (8 -> 64 -> 64 -> 1, GELU, FP32) is trained on the GPU by the library's own
kernels (hssc_mlp_train) to predict log2 of a box's ID rank from geometric features
(HSS1D.features).  Training data are the ranks that the plain ID produces on seeded random
star contours: mode m in {3..8}, amplitude in [0.1, 0.4], N log-uniform.  At compression
time (HSS1D_compress_nsym(..., rank='mlp', mlp=...)) the prediction sizes each block's
sketch; sketches that fail fall back to the plain ID.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .hss1d import HSS1D_compress_nsym, star_contour

_FP = C.POINTER(C.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class RankMLP:
    NFEAT = 8

    def __init__(self, seed=0):
        h = C.c_void_p()
        check(lib().hssc_mlp_create(C.c_uint64(seed), C.byref(h)))
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                lib().hssc_mlp_destroy(h)
            except TypeError:  # interpreter shutdown: module globals already cleared
                pass
            self.handle = None

    @property
    def nparams(self):
        return lib().hssc_mlp_num_params()

    def get_params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        check(lib().hssc_mlp_get_params(self.handle, p.ctypes.data_as(_FP)))
        return p

    def set_params(self, p):
        p = _f32(p)
        check(lib().hssc_mlp_set_params(self.handle, p.ctypes.data_as(_FP)))

    def set_scaling(self, mu, sd):
        mu, sd = _f32(mu), _f32(sd)
        check(lib().hssc_mlp_set_scaling(self.handle, mu.ctypes.data_as(_FP), sd.ctypes.data_as(_FP)))
        self.scaling = (mu.copy(), sd.copy())

    def save(self, path):
        """Writes the weights and the feature scaling to a .npz file (plain arrays, no pickles)."""
        mu, sd = self.scaling
        np.savez(path, params=self.get_params(), mu=mu, sd=sd)

    @classmethod
    def load(cls, path):
        """A network with the weights and scaling that save() wrote."""
        m = cls()
        with np.load(path, allow_pickle=False) as f:
            m.set_params(f["params"])
            m.set_scaling(f["mu"], f["sd"])
        return m

    def fit_scaling(self, X, t):
        mu = np.concatenate([X.mean(0), [t.mean()]])
        sd = np.concatenate([X.std(0), [t.std()]])
        sd[sd < 1e-6] = 1.0
        self.set_scaling(mu, sd)
        return mu, sd

    def predict(self, X):
        """X: B x 8 raw features -> B predicted targets (log2 rank)."""
        Xf = _f32(np.asfortranarray(X, dtype=np.float32).T.reshape(-1))  # column-major B x 8
        y = np.empty(X.shape[0], dtype=np.float32)
        check(lib().hssc_mlp_predict(self.handle, Xf.ctypes.data_as(_FP), X.shape[0],
                                     y.ctypes.data_as(_FP)))
        return y

    def grads(self, X, t):
        Xf = _f32(np.asfortranarray(X, dtype=np.float32).T.reshape(-1))
        tt = _f32(t)
        g = np.empty(self.nparams, dtype=np.float32)
        loss = C.c_double()
        check(lib().hssc_mlp_grads(self.handle, Xf.ctypes.data_as(_FP), tt.ctypes.data_as(_FP),
                                   X.shape[0], C.byref(loss), g.ctypes.data_as(_FP)))
        return loss.value, g

    def train(self, X, t, epochs=2000, lr=3e-3):
        Xf = _f32(np.asfortranarray(X, dtype=np.float32).T.reshape(-1))
        tt = _f32(t)
        loss = C.c_double()
        check(lib().hssc_mlp_train(self.handle, Xf.ctypes.data_as(_FP), tt.ctypes.data_as(_FP),
                                   X.shape[0], int(epochs), C.c_float(lr), C.byref(loss)))
        return loss.value


def random_contour(rng, N):
    """A configuration-3 geometry: star contour r(t) = 1 + a cos(m t), m in {3..8},
    a in [0.1, 0.4]."""
    return star_contour(N, a=float(rng.uniform(0.1, 0.4)), m=int(rng.integers(3, 9)))


def box_dataset(hss):
    """(features, log2 rank) of every non-root box of a compressed HSS1D operator."""
    X, t = [], []
    for b in range(1, hss.nboxes + 1):
        info = hss.box(b)
        if info["level"] == 0:  # roots have no ID
            continue
        X.append(hss.features(b))
        t.append(np.log2(max(info["kskel"], 1)))
    return np.array(X, dtype=np.float32), np.array(t, dtype=np.float32)


def collect(ngeom, seed=0, Nmin=1024, Nmax=16384, nbox_max=64, acc=1e-10, nproxy=64):
    """Ranks produced by the plain ID on `ngeom` seeded geometries (N log-uniform)."""
    rng = np.random.default_rng(seed)
    Xs, ts = [], []
    for _ in range(ngeom):
        N = int(np.exp(rng.uniform(np.log(Nmin), np.log(Nmax))))
        K = HSS1D_compress_nsym("green", random_contour(rng, N), nbox_max, acc, nproxy=nproxy)
        X, t = box_dataset(K)
        Xs.append(X), ts.append(t)
    return np.concatenate(Xs), np.concatenate(ts)


def train_rank_mlp(X, t, seed=0, epochs=3000, lr=3e-3):
    m = RankMLP(seed)
    m.fit_scaling(X, t)
    loss = m.train(X, t, epochs=epochs, lr=lr)
    return m, loss


# ---- HSS2D: ranks of the ID-selected skeleton levels (kh > 0) -----------------------------

def hss2d_geometry(rng, Ng_choices=(6, 8, 10), np_choices=(4, 8, 16, 32, 64),
                   kh_range=(4.0, 40.0)):
    """A seeded Lippmann-Schwinger configuration: grid (Ng, np), wavenumber log-uniform in
    kh_range, contrast the tanh plateau or a Gaussian of random amplitude and width."""
    from .hss2d import HSS_tree_parameters
    Ng = int(rng.choice(Ng_choices))
    np_ = int(rng.choice(np_choices))
    kh = float(np.exp(rng.uniform(np.log(kh_range[0]), np.log(kh_range[1]))))
    contrast = None
    if rng.uniform() < 0.5:
        contrast = ("gaussian", float(rng.uniform(0.2, 1.0)), float(rng.uniform(0.1, 0.5)))
    # the Duan-Rokhlin weight needs h k <= 2
    kh = min(kh, 1.9 * Ng * np_ / 2)
    return HSS_tree_parameters("SL_H_2D", kh, 1, Ng, np_, 1e-10, 3, 800, 0, proxy_layers=4,
                               contrast=contrast)


def hss2d_box_dataset(tree):
    """(features, log2 rank) of the boxes of the ID-selected levels of a tree."""
    Xs, ts = [], []
    for lev in range(tree.depth + 1):
        ranks, idl = tree.box_ranks(lev)
        if not idl:
            continue
        Xs.append(tree.box_features(lev))
        ts.append(np.log2(np.maximum(ranks, 1)).astype(np.float32))
    return np.concatenate(Xs), np.concatenate(ts)


def collect_hss2d(ngeom, seed=0, max_boxes=384, **kw):
    """Ranks the plain ID produces at the ID-selected levels of `ngeom` seeded geometries, at
    most `max_boxes` boxes drawn from each (large grids would outweigh the others)."""
    from .hss2d import HSS2D_build_index_sets
    rng = np.random.default_rng(seed)
    Xs, ts = [], []
    for _ in range(ngeom):
        X, t = hss2d_box_dataset(HSS2D_build_index_sets(hss2d_geometry(rng, **kw)))
        if len(t) > max_boxes:
            pick = rng.choice(len(t), max_boxes, replace=False)
            X, t = X[pick], t[pick]
        Xs.append(X), ts.append(t)
    return np.concatenate(Xs), np.concatenate(ts)
