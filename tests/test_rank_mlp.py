"""Rank prediction (the learned replacement of HSS2D/predict_rank.m): the network's GPU
forward/backward against the CPU oracle, training, and what the trained network gains over
the closed-form guess when it sizes the randomized IDs of the HSS1D compression."""
import numpy as np
import pytest

from oracle import mlp as om

KNOWN_TABLE = [  # HSS2D/predict_rank.m:13-23 evaluated by hand: (n, acc) -> (rank, C)
    ((100, 1e-10), (100, 42)),    # n < 150 -> min(n, r0(10) = 130)
    ((140, 1e-10), (130, 42)),
    ((1000, 1e-10), (329, 42)),   # ceil(275*3 - 496)
    ((5000, 1e-10), (522, 42)),   # ceil(275*log10(5000) - 496)
    ((1000, 1e-8), (199, 21)),    # p = 8: ceil(138*3 - 215), C = ceil(138*log10(2)/2)
    ((1000, 1e-16), (507, 74)),   # p clamped to 12: ceil(487*3 - 954)
]


@pytest.mark.parametrize("args,want", KNOWN_TABLE)
def test_predict_rank_table(args, want):
    assert om.predict_rank(*args) == want


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    W1 = rng.standard_normal((8, 64)) * 0.3
    W2 = rng.standard_normal((64, 64)) * 0.1
    w3 = rng.standard_normal(64) * 0.1
    p = om.pack(W1, rng.standard_normal(64) * 0.1, W2, rng.standard_normal(64) * 0.1, w3, 0.3)
    X = rng.standard_normal((11, 8))
    t = rng.standard_normal(11)
    _, g = om.loss_and_grad(p, X, t)
    for i in rng.choice(p.size, 40, replace=False):
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (om.loss_and_grad(p + e, X, t)[0] - om.loss_and_grad(p - e, X, t)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), i


def _data(seed, B=300):
    rng = np.random.default_rng(seed)
    X = np.column_stack([rng.integers(1, 7, B), rng.uniform(4, 6, B), rng.uniform(4, 7, B),
                         rng.uniform(-6, -1, B), rng.uniform(0, 3, B), rng.uniform(0, 1, B),
                         rng.uniform(6, 8, B), np.full(B, 10.0)]).astype(np.float32)
    t = (0.4 * X[:, 2] + 0.3 * np.sin(X[:, 4]) + 0.1 * X[:, 0]).astype(np.float32)
    return X, t


@pytest.mark.gpu
def test_gpu_forward_and_gradients_match_oracle():
    from hssc_directsolver_amd.rank_mlp import RankMLP
    X, t = _data(1)
    m = RankMLP(seed=3)
    mu, sd = m.fit_scaling(X, t)
    p = m.get_params()
    Xs = om.standardise(X, mu, sd)
    ts = (t - mu[8]) / sd[8]
    y_or, _ = om.forward(p, Xs)
    y = m.predict(X)
    assert np.max(np.abs((y - mu[8]) / sd[8] - y_or)) < 2e-5 * max(1, np.max(np.abs(y_or)))
    loss, g = m.grads(X, t)
    loss_or, g_or = om.loss_and_grad(p, Xs, ts)
    assert abs(loss - loss_or) <= 1e-5 * loss_or
    assert np.linalg.norm(g - g_or) <= 1e-4 * np.linalg.norm(g_or)


@pytest.mark.gpu
def test_training_fits_a_smooth_target(tmp_path):
    from hssc_directsolver_amd.rank_mlp import RankMLP
    X, t = _data(2, B=2000)
    m = RankMLP(seed=5)
    m.fit_scaling(X, t)
    l0, _ = m.grads(X, t)
    l1 = m.train(X, t, epochs=1500, lr=3e-3)
    assert l1 < 0.02 * l0
    Xv, tv = _data(9, B=500)
    assert np.sqrt(np.mean((m.predict(Xv) - tv) ** 2)) < 0.1 * np.std(tv)
    # the saved network predicts the same bits once loaded into a fresh handle
    path = tmp_path / "rank_mlp.npz"
    m.save(path)
    m2 = RankMLP.load(path)
    assert np.array_equal(m2.get_params(), m.get_params())
    assert np.array_equal(m2.predict(Xv), m.predict(Xv))


@pytest.mark.gpu
def test_learned_ranks_size_the_randomized_ids():
    import hssc_directsolver_amd as H
    from hssc_directsolver_amd.rank_mlp import collect, random_contour, train_rank_mlp
    # train on ranks of the plain ID over seeded configuration-3 geometries
    X, t = collect(40, seed=11, Nmin=1024, Nmax=8192)
    m, loss = train_rank_mlp(X, t, seed=1, epochs=3000)
    assert loss < 0.1
    rng = np.random.default_rng(123)
    rows = {}
    for mode in ("id", "table", "mlp"):
        rows[mode] = dict(blocks=0, fallbacks=0, rows=0)
    for g in range(4):
        N = 2048
        curve = random_contour(rng, N)
        f = np.random.default_rng(g).standard_normal(N)
        xs = {}
        for mode in ("id", "table", "mlp"):
            K = H.HSS1D_compress_nsym("green", curve, 64, 1e-10, nproxy=64, rank=mode,
                                      mlp=m if mode == "mlp" else None)
            H.HSS1D_invert_nsym(K)
            xs[mode] = H.HSS1D_applyinv_nsym(K, f)
            st = K.id_stats()
            rows[mode]["blocks"] += st["blocks"]
            rows[mode]["fallbacks"] += st["fallbacks"]
            rows[mode]["rows"] += st["sketch_rows"] + st["full_rows"]
            # the plain ID judges: every mode solves the same system to tolerance
            resid = np.linalg.norm(H.HSS1D_apply_nsym(K, xs[mode]) - f) / np.linalg.norm(f)
            assert resid < 1e-12, (mode, resid)
        for mode in ("table", "mlp"):
            assert np.linalg.norm(xs[mode] - xs["id"]) < 1e-9 * np.linalg.norm(xs["id"])
    # what the learned ranks gain: fewer rows factored than the closed-form guess and
    # the plain ID, with few fallbacks
    assert rows["mlp"]["rows"] < rows["table"]["rows"]
    assert rows["mlp"]["rows"] < 0.8 * rows["id"]["rows"]
    assert rows["mlp"]["fallbacks"] <= 0.2 * rows["mlp"]["blocks"]
