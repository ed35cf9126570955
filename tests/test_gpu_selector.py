"""The learned accumulator selector on the GPU: the network's forward pass
and gradients against the FP64 oracle (oracle/mlp.py, pinned by finite
differences), training, the timing samples the product produces, and the
learned selector's results (bit-exact counts) and timing beside the
fixed-threshold heuristic."""
import numpy as np
import pytest

from graphkokkos_amd import SEL_LEARNED, HostGraph, default_opts, selector
from oracle import mlp
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_forward_matches_fp64_oracle(gpu):
    rng = np.random.default_rng(1)
    p = mlp.random_params(rng)
    selector.set_params(gpu, p)
    X = rng.normal(0, 2, (1000, 16))
    want, _ = mlp.forward(p, X)
    got = selector.forward(gpu, X)
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-4 * np.abs(want).max())


def test_gradient_matches_fp64_oracle(gpu):
    """FP32 gradients accumulated with atomics: |g - g64| <= 1e-3 max|g64|."""
    rng = np.random.default_rng(2)
    p = mlp.random_params(rng)
    selector.set_params(gpu, p)
    X = rng.normal(0, 2, (3000, 16))
    T = rng.normal(0, 1, (3000, 4))
    V = (rng.random((3000, 4)) < 0.8).astype(np.uint8)
    g, loss = selector.grad(gpu, X, T, V)
    l64, g64 = mlp.loss_grad(p, X, T, V)
    assert abs(loss - l64) <= 1e-4 * abs(l64)
    err = np.abs(g - g64).max()
    assert err <= 1e-3 * np.abs(g64).max(), err


def test_training_fits_a_known_cost_function(gpu):
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (20000, 16)).astype(np.float32)
    W = rng.normal(0, 1, (16, 4))
    T = np.tanh(X @ W) + 0.5 * X[:, :4] ** 2
    V = np.ones_like(T, np.uint8)
    loss0 = float(((T - T.mean(0)) ** 2).mean())
    loss = selector.train(gpu, X, T, V, epochs=30, batch=1024, lr=3e-3, seed=4)
    assert loss < 0.1 * loss0, (loss, loss0)
    pred = selector.forward(gpu, X[:2000])
    assert np.corrcoef(pred.ravel(), T[:2000].ravel())[0, 1] > 0.95


def test_samples_learned_selection_and_gain(gpu):
    train_graphs = [HostGraph.rmat(14, 16, seed=5), HostGraph.grid(128), HostGraph.chung_lu(20000, 2.1, 16, seed=5)]
    rep = selector.train_from_graphs(gpu, train_graphs, epochs=20, seed=1)
    assert rep["samples"] > 10000
    print("selector training:", rep, flush=True)
    G = HostGraph.rmat(16, 16, seed=9)
    d = gpu.upload(G.L)
    want = O.tc_sandia(G.L)
    assert gpu.tc_count(d, {"selector": SEL_LEARNED}) == want
    assert gpu.tc_count(d) == want
    # a margin no prediction can exceed keeps every fixed-threshold choice
    assert gpu.tc_count(d, {"selector": SEL_LEARNED, "learned_margin": 10 ** 8}) == want
    rows_l = gpu.last_timing()["rows_per_acc"]
    gpu.tc_count(d)
    assert rows_l == gpu.last_timing()["rows_per_acc"]
    # The measured gain of the trained network over fixed thresholds that send
    # mask rows up to 16 columns to MCA-wave (the support product's default),
    # on a graph large enough to be throughput-bound: the labels are
    # wave-time, a throughput cost, and products of half a millisecond are
    # bound by latency instead.  Trained on these small graphs, the network
    # moves the mask rows of 17-48 columns from MSA-LDS to MCA-wave, as the
    # autotuner's search over thresholds does (triangle counting's default
    # threshold, 48, came from that search).
    G = HostGraph.rmat(19, 16, seed=9)
    d = gpu.upload(G.L)
    want = gpu.tc_count(d)
    assert gpu.tc_count(d, {"selector": SEL_LEARNED}) == want
    t_h = selector.time_tc(gpu, d, {"mca_max_mask": 16}, reps=7)
    rows_h = gpu.last_timing()["rows_per_acc"]
    t_l = selector.time_tc(gpu, d, {"selector": SEL_LEARNED, "mca_max_mask": 16}, reps=7)
    rows_l = gpu.last_timing()["rows_per_acc"]
    print(f"R-MAT s19 triangle counting: heuristic {t_h * 1e3:.3f} ms (rows MSA-LDS/MCA-wave {rows_h[:2]}), "
          f"learned {t_l * 1e3:.3f} ms ({rows_l[:2]}), learned/heuristic speedup {t_h / t_l:.3f}", flush=True)
    assert rows_l[1] > rows_h[1]
    assert t_h / t_l > 1.1


def _heuristic_choice(X, V):
    """The fixed-threshold rule on the features' mask length, at the
    library's default thresholds (MCA-wave, MSA-LDS, dense-global)."""
    o = default_opts()
    # -1: the product mode's default, 48 for triangle counting
    mca = o.mca_max_mask if o.mca_max_mask >= 0 else 48
    mlen = np.round(2 ** X[:, 0] - 1).astype(np.int64)
    c = np.where(mlen <= mca, 1,
                 np.where(mlen <= o.lds_max_mask, 0, np.where(mlen >= o.dense_min_mask, 3, 2)))
    c = np.where(V[np.arange(len(c)), c].astype(bool), c, 0)  # capacity fallback
    return c


def test_network_gain_on_held_out_rows(gpu):
    """What the trained network gains: train on some graphs, then on the rows
    of a held-out graph compare the measured cost (the labels) of the
    accumulators picked by the network, by the fixed-threshold heuristic and
    by a trivial predictor (each accumulator's mean label, one choice for
    every row), all among the selector's candidates.  The network's
    predictions must explain most of the label variance and its picks must
    cost less than the trivial predictor's."""
    train_graphs = [HostGraph.rmat(14, 16, seed=5), HostGraph.grid(128), HostGraph.chung_lu(20000, 2.1, 16, seed=5)]
    selector.train_from_graphs(gpu, train_graphs, epochs=30, seed=1)
    G = HostGraph.rmat(15, 16, seed=11)  # held out
    d = gpu.upload(G.L)
    X, T, V = selector.samples(gpu, d)
    Vb = V.astype(bool)
    Y = selector.forward(gpu, X)
    resid = (Y - T)[Vb]
    r2 = 1.0 - np.mean(resid ** 2) / np.var(T[Vb])
    rows = np.arange(len(X))
    # the selector's candidates: MSA-LDS and MCA-wave where the row fits LDS,
    # the global accumulators only beyond it
    mlen = np.round(2 ** X[:, 0] - 1).astype(np.int64)
    fits = (mlen <= default_opts().lds_max_mask)[:, None]
    cand = Vb & np.where(fits, np.array([True, True, False, False]), np.array([False, False, True, True]))
    net = np.where(cand, Y, np.inf).argmin(axis=1)
    heur = _heuristic_choice(X, V)
    mean_label = np.array([T[Vb[:, a], a].mean() if Vb[:, a].any() else np.inf for a in range(4)])
    triv = np.where(cand, mean_label[None, :], np.inf).argmin(axis=1)
    cost = lambda c: float(np.exp2(T[rows, c]).sum())
    c_net, c_heur, c_triv = cost(net), cost(heur), cost(triv)
    best = float(np.exp2(np.where(cand, T, np.inf)).min(axis=1).sum())
    print(f"held-out R-MAT s15 rows {len(X)}: R^2 {r2:.3f}; summed label ns of the picks: network {c_net:.4g}, "
          f"heuristic {c_heur:.4g}, trivial {c_triv:.4g}, per-row best {best:.4g}; "
          f"network/heuristic {c_heur / c_net:.3f}x", flush=True)
    assert r2 > 0.5
    assert c_net <= c_triv
    assert c_net >= best * (1 - 1e-6)


def test_sample_labels_follow_capacities(gpu):
    G = HostGraph.rmat(12, 16, seed=6)
    d = gpu.upload(G.L)
    X, T, V = selector.samples(gpu, d)
    mlen = np.round(2 ** X[:, 0] - 1).astype(int)
    assert not V[mlen > 64, 1].any()      # MCA-wave holds mask rows up to 64
    assert V[:, 0].all()                  # MSA-LDS holds every row here
    assert np.isfinite(T[V.astype(bool)]).all()


def test_autotuner_returns_correct_configuration(gpu):
    G = HostGraph.rmat(13, 16, seed=2)
    d = gpu.upload(G.L)
    space = {"mca_max_mask": [0, 16], "small_max_mask": [64, 128]}
    best, t, log = selector.autotune_tc(gpu, d, rounds=1, reps=2, space=space)
    assert len(log) >= 3 and t > 0
    assert gpu.tc_count(d, best) == O.tc_sandia(G.L)
