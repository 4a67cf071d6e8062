"""The learned tuner on the committed sweep data (CPU): the committed network,
evaluated with the FP64 oracle forward pass, is within 5% of the exhaustive
best on held-out matrices and beats the fixed default tile."""
import json
import os

import numpy as np

import oracle
from benchmark_spmv_using_csr_amd import tuner

DATA = os.path.join(tuner.DATA_DIR, "tuner_sweep.npz")


def _data():
    return tuner.load_data([DATA])


def _scores(net, data):
    X = tuner.transform(data["raw"]).astype(np.float64)
    norm = np.r_[np.array(net["mean"]), np.array(net["std"])].astype(np.float32).astype(np.float64)
    p = np.array(net["params"], np.float32).astype(np.float64)
    T = np.zeros((len(X), tuner.K))
    _, out, _ = oracle.mlp_loss_grad(tuner.N_FEATURES, net["hidden"], tuner.K, p, norm, X, T,
                                     np.ones_like(T))
    return out


def test_dataset_shape_and_split():
    d = _data()
    assert len(d["seed"]) == 2000 and np.array_equal(d["seed"], np.arange(2000))
    assert d["us"].shape == (2000, tuner.K) and (d["us"] > 0).all()
    assert np.array_equal(d["classes"], np.array(tuner.CLASSES))
    assert set(np.unique(d["family"])) == {0, 1, 2, 3}
    assert abs(d["fp64"].mean() - 0.5) < 0.01                      # precision split evenly
    lm = np.log10(d["m"])
    assert lm.min() >= 3 and lm.max() <= np.log10(4e6)
    tr, te = tuner.split(d)
    assert tr.sum() == 1500 and te.sum() == 500


def test_targets():
    us = np.array([[10.0, 20.0, -1.0, 40.0]])
    t, m = tuner.targets(us)
    assert np.allclose(t, [[0, 1, 0, 2]]) and np.array_equal(m, [[1, 1, 0, 1]])


def test_committed_network_regret_and_gain():
    d = _data()
    net = tuner.load_net()
    assert net["classes"] == tuner.CLASSES
    tr, te = tuner.split(d)
    sc = _scores(net, d)
    held = tuner.evaluate(d, sc[te], te)
    # within 5% of the exhaustive best on held-out matrices ...
    assert held["mean_regret"] <= 0.05
    assert held["geomean_time_ratio_to_best"] <= 1.05
    # ... and a real gain over the method without it (one fixed tile) and over
    # the best single fixed class
    assert held["geomean_speedup_vs_default"] >= 1.3
    us = d["us"][te]
    best = us.min(axis=1)
    fixed = min(np.exp(np.log(us[:, k] / best).mean()) for k in range(tuner.K))
    assert held["geomean_time_ratio_to_best"] < fixed - 0.1
    # the recorded metadata matches what the oracle forward pass gives
    assert abs(net["meta"]["held_out"]["mean_regret"] - held["mean_regret"]) < 0.01


def test_transform_matches_formula():
    raw = np.array([1e6, 1e7, 10.0, 400.0, 1e5, 0.1, 5e5, 1.0])
    f = tuner.transform(raw)[0]
    want = [np.log10(1e6) / 7, np.log10(1e7 + 1) / 9, np.log2(11) / 8, np.log2(21) / 10,
            np.log2(1e5 + 1) / 17, 0.1, np.log2((5e5 + 1) / (1e6 + 1)) / 23, 1.0]
    np.testing.assert_allclose(f, want, rtol=1e-6)
