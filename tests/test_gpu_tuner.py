"""The tuner's HIP kernels against the CPU FP64 oracle, and the on-device
inference and autotune paths through the C ABI."""
import numpy as np
import pytest

import oracle
from benchmark_spmv_using_csr_amd import _lib, matgen, tuner
from benchmark_spmv_using_csr_amd.spmv import bhsparse_spmv

pytestmark = pytest.mark.gpu

H, K, NF = tuner.HIDDEN, tuner.K, tuner.N_FEATURES


def _problem(N=60, seed=0):
    rng = np.random.default_rng(seed)
    params = tuner.init_params(H, K, seed).astype(np.float64)
    params[H * NF: H * NF + H] = rng.normal(size=H) * 0.1          # nonzero biases
    norm = np.r_[rng.normal(size=NF) * 0.2, rng.uniform(0.5, 2.0, NF)]
    X = rng.normal(size=(N, NF))
    T = rng.uniform(0, 3, size=(N, K))
    M = (rng.random((N, K)) > 0.15).astype(np.float64)
    return params, norm, X, T, M


def test_mlp_gradient_f64_matches_oracle():
    params, norm, X, T, M = _problem()
    lo, _, go = oracle.mlp_loss_grad(NF, H, K, params, norm, X, T, M)
    lg, gg = tuner.loss_grad_f64(params, norm, X, T, M)
    assert abs(lg - lo) <= 1e-12 * abs(lo)
    assert np.max(np.abs(gg - go)) <= 1e-11 * np.max(np.abs(go))


def test_mlp_float_kernels_match_oracle():
    params, norm, X, T, M = _problem(seed=1)
    p32 = params.astype(np.float32)
    net = tuner.Mlp(H, K)
    net.set(p32, norm[:NF].astype(np.float32), norm[NF:].astype(np.float32))
    n32 = np.r_[norm[:NF].astype(np.float32), norm[NF:].astype(np.float32)].astype(np.float64)
    X32 = X.astype(np.float32)
    lo, oo, go = oracle.mlp_loss_grad(NF, H, K, p32.astype(np.float64), n32,
                                      X32.astype(np.float64), T.astype(np.float32), M)
    out = net.forward(X32)
    assert np.max(np.abs(out - oo)) <= 1e-4 * max(1.0, np.max(np.abs(oo)))
    lg, gg = net.loss_grad(X32, T, M)
    assert abs(lg - lo) <= 1e-4 * abs(lo)
    assert np.max(np.abs(gg - go)) <= 1e-3 * np.max(np.abs(go))
    assert np.array_equal(net.params(), p32)


def test_training_reduces_loss_and_is_reproducible():
    rng = np.random.default_rng(3)
    X = rng.normal(size=(200, NF)).astype(np.float32)
    W = rng.normal(size=(NF, K))
    T = np.abs(X @ W).astype(np.float32) / 4       # a learnable target
    M = np.ones_like(T)
    hs = []
    for _ in range(2):
        net = tuner.Mlp(H, K)
        net.set(tuner.init_params(H, K, 5), np.zeros(NF, np.float32), np.ones(NF, np.float32))
        h = net.train(X, T, M, epochs=400, lr=3e-3)
        hs.append((h, net.params()))
    assert hs[0][0][-1] < 0.1 * hs[0][0][0]
    assert np.array_equal(hs[0][0], hs[1][0]) and np.array_equal(hs[0][1], hs[1][1])


def test_apply_tuner_matches_host_forward():
    A = matgen.power_law(31, 20000, 20000, 1.8, 5000, 0.1, 0)
    net = tuner.Mlp(H, K)
    rng = np.random.default_rng(7)
    mean = rng.normal(size=NF).astype(np.float32) * 0.1
    std = rng.uniform(0.5, 1.5, NF).astype(np.float32)
    net.set(tuner.init_params(H, K, 11), mean, std)
    for dtype in (np.float64, np.float32):
        op = bhsparse_spmv(dtype)
        _lib.check(op.init_platform(0), "init_platform")
        y = np.zeros(A.m, dtype)
        _lib.check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val.astype(dtype),
                                  np.ones(A.n, dtype), y), "prepare_mem")
        chosen = op.apply_tuner(net, tuner.CLASSES)
        sc = net.forward(tuner.transform(op.row_features()))[0]
        assert chosen == int(np.argmin(sc))
        assert op.get_config().as_tuple() == tuner.CLASSES[chosen]
        op.close()


def test_learned_shortlist_times_the_best_predicted_classes():
    """apply_tuner_topk: topk = 1 is the network's pick; topk = 3 installs one of
    the three best-predicted runnable classes (host forward pass), the result
    stays within the error bound; bad arguments are refused."""
    A = matgen.power_law(33, 40000, 40000, 1.9, 8000, 0.1, 0)
    x = matgen.vector(34, A.n)
    net = tuner.make_net(tuner.load_net())
    op = bhsparse_spmv(np.float64)
    _lib.check(op.init_platform(0), "init_platform")
    y = np.zeros(A.m)
    _lib.check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val, x, y), "prepare_mem")
    print("[gpu] learned shortlist", flush=True)
    top1 = op.apply_tuner(net, tuner.CLASSES)
    assert op.apply_tuner(net, tuner.CLASSES, topk=1) == top1
    sc = net.forward(tuner.transform(op.row_features()))[0]
    assert int(np.argmin(sc)) == top1
    c3 = op.apply_tuner(net, tuner.CLASSES, topk=3, reps=3, flush_bytes=0)
    # among the three lowest predictions (host and device forward passes may
    # differ in the last bits, hence the small margin)
    assert sc[c3] <= np.sort(sc)[2] + 1e-4
    assert op.get_config().as_tuple() == tuner.CLASSES[c3]
    _lib.check(op.spmv(1.0), "spmv")
    yr = oracle.spmv_serial(A.row_ptr, A.col, A.val, x)
    assert oracle.rel_error(A.row_ptr, A.col, A.val, x, op.y(), yr) <= 1e-13
    with pytest.raises(_lib.BhsError):
        op.apply_tuner(net, tuner.CLASSES, topk=0)
    with pytest.raises(_lib.BhsError):
        op.apply_tuner(net, tuner.CLASSES, topk=3, reps=0)
    op.close()


def test_autotune_installs_fastest_and_stays_correct():
    A = matgen.uniform_rows(41, 30000, 30000, 0, 40)
    x = matgen.vector(42, A.n)
    op = bhsparse_spmv(np.float64)
    _lib.check(op.init_platform(0), "init_platform")
    y = np.zeros(A.m)
    _lib.check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val, x, y), "prepare_mem")
    us, best = op.autotune(tuner.CLASSES, reps=3, flush_bytes=0)
    assert best == int(np.argmin(np.where(us > 0, us, np.inf)))
    assert op.get_config().as_tuple() == tuner.CLASSES[best]
    _lib.check(op.spmv(1.0), "spmv")
    yr = oracle.spmv_serial(A.row_ptr, A.col, A.val, x)
    assert oracle.rel_error(A.row_ptr, A.col, A.val, x, op.y(), yr) <= 1e-13


def test_learned_tuner_live_on_held_out_matrices():
    """Live on the GPU: the committed network's pick (and its timed shortlist
    of five) vs the exhaustive best and the fixed default tile, on every
    tenth held-out matrix of the tuner set.  The bound is the target's: the
    tuned time within 5% of the exhaustive best (geometric mean)."""
    net = tuner.make_net(tuner.load_net())
    to_best, vs_default, short_best = [], [], []
    for seed in range(1500, 2000, 10):
        A, dtype = tuner.dataset_matrix(seed)
        if A.nnz > 3e7:
            continue
        op = bhsparse_spmv(dtype)
        _lib.check(op.init_platform(0), "init_platform")
        y = np.zeros(A.m, dtype)
        _lib.check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val.astype(dtype),
                                  matgen.vector(seed, A.n).astype(dtype), y), "prepare_mem")
        chosen = op.apply_tuner(net, tuner.CLASSES)
        short = op.apply_tuner(net, tuner.CLASSES, topk=5, reps=5, flush_bytes=1 << 30)
        us, best = op.autotune(tuner.CLASSES, reps=7, flush_bytes=1 << 30)
        to_best.append(us[chosen] / us[best])
        short_best.append(us[short] / us[best])
        vs_default.append(us[tuner.DEFAULT_CLASS] / us[chosen])
        op.close()
    gm = lambda a: float(np.exp(np.mean(np.log(a))))
    print(f"[gpu] live held-out: n={len(to_best)} tuned/best {gm(to_best):.4f} "
          f"(p90 {np.percentile(to_best, 90):.4f}) shortlist/best {gm(short_best):.4f} "
          f"default/tuned {gm(vs_default):.4f}", flush=True)
    assert len(to_best) >= 40
    assert gm(to_best) <= 1.05
    assert gm(short_best) <= 1.05
    assert gm(vs_default) >= 1.2
