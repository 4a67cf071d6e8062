"""The oracle pinned against the reference's fixture, closed forms, an
independent implementation (scipy) and brute force.  CPU only."""
import bisect

import numpy as np
import pytest
import scipy.sparse as sp

import cases
import oracle
from benchmark_spmv_using_csr_amd import matgen


def test_six_by_six_known_answer():
    # spmv_cuda/main.cu:441-474: the only known-answer test in the reference
    c = cases.six_by_six()
    assert np.array_equal(oracle.spmv_serial(c.row_ptr, c.col, c.val, c.x), cases.SIX_BY_SIX_Y)
    y32 = oracle.spmv_serial_f32(c.row_ptr, c.col, c.val, c.x)
    assert np.array_equal(y32, cases.SIX_BY_SIX_Y.astype(np.float32))


@pytest.mark.parametrize("seed", range(4))
def test_serial_matches_scipy(seed):
    c = cases.from_lengths("r", np.random.default_rng(seed).integers(0, 40, 300), n=500,
                           seed=seed)
    A = sp.csr_matrix((c.val, c.col, c.row_ptr), shape=(c.m, c.n))
    y = oracle.spmv_serial(c.row_ptr, c.col, c.val, c.x)
    np.testing.assert_allclose(y, A @ c.x, rtol=0, atol=1e-13)


def test_laplacian_closed_form():
    # x = 1: y_i = 6 - (number of grid neighbours) = number of missing neighbours
    nx, ny, nz = 7, 5, 4
    A = matgen.laplacian7(nx, ny, nz)
    y = oracle.spmv_serial(A.row_ptr, A.col, A.val, np.ones(A.n))
    i = np.arange(A.m)
    x_, y_, z_ = i % nx, (i // nx) % ny, i // (nx * ny)
    missing = ((x_ == 0).astype(int) + (x_ == nx - 1) + (y_ == 0) + (y_ == ny - 1)
               + (z_ == 0) + (z_ == nz - 1))
    assert np.array_equal(y, missing.astype(np.float64))
    assert A.nnz == 7 * A.m - 2 * (nx * ny + ny * nz + nx * nz)


def test_block3_against_dense():
    A = matgen.block3(5, 31, 4, 12)
    x = matgen.vector(9, A.n)
    D = A.dense()
    np.testing.assert_allclose(oracle.spmv_serial(A.row_ptr, A.col, A.val, x), D @ x,
                               rtol=0, atol=1e-13)
    # block structure: rows of one block row share their column set
    for I in range(A.m // 3):
        cols = [tuple(A.col[A.row_ptr[r]:A.row_ptr[r + 1]]) for r in range(3 * I, 3 * I + 3)]
        assert cols[0] == cols[1] == cols[2]


def test_upper_bound_matches_bisect():
    rng = np.random.default_rng(1)
    rp = np.concatenate([[0], np.cumsum(rng.integers(0, 4, 200))]).astype(np.int32)
    for key in range(-2, int(rp[-1]) + 3):
        assert oracle.upper_bound(rp, key) == bisect.bisect_right(list(rp), key)


@pytest.mark.parametrize("P", [4, 8, 64, 256])
def test_partition_brute_force(P):
    c = cases.from_lengths("r", [0, 0, 3, 0, 9, 1, 0, 0, 17, 2, 0, 5, 0], seed=3)
    tr = oracle.partition(c.row_ptr, c.nnz, P)
    NP = (c.nnz + P - 1) // P
    assert len(tr) == NP + 1
    for p in range(NP + 1):
        k = min(p * P, c.nnz)
        brute = max(r for r in range(c.m + 1) if c.row_ptr[r] <= k)
        assert tr[p] == brute
    # rowOf(B) for B < nnz is a nonempty row holding item B
    for p in range(NP):
        r = tr[p]
        assert c.row_ptr[r] <= p * P < c.row_ptr[r + 1]
    assert tr[NP] == c.m


REF_SHAPES = [(32, 4, 7), (32, 8, 7), (64, 8, 5), (64, 16, 2), (8, 16, 6), (64, 4, 1)]


@pytest.mark.parametrize("W,S,T", REF_SHAPES)
def test_reference_restatement_exact_on_integers(W, S, T):
    for c in cases.adversarial(W * S, T, seed=W + S + T, integer=True):
        yr = oracle.spmv_serial(c.row_ptr, c.col, c.val, c.x)
        y, st = oracle.segsum_ref(c.row_ptr, c.col, c.val, c.x, W=W, S=S, T=T)
        assert np.array_equal(y, yr), c.name


@pytest.mark.parametrize("W,S,T", REF_SHAPES[:3])
def test_reference_restatement_error_bound(W, S, T):
    for c in cases.adversarial(W * S, T, seed=7):
        yr = oracle.spmv_serial(c.row_ptr, c.col, c.val, c.x)
        y, _ = oracle.segsum_ref(c.row_ptr, c.col, c.val, c.x, W=W, S=S, T=T)
        assert oracle.rel_error(c.row_ptr, c.col, c.val, c.x, y, yr) <= 1e-13, c.name


def test_restatement_counts_match_tile_stats():
    c = cases.from_lengths("pl", np.r_[[0] * 5, [2900] * 3, [2, 0, 1] * 300], n=3000, seed=2)
    for W, S, T in REF_SHAPES:
        _, st = oracle.segsum_ref(c.row_ptr, c.col, c.val, c.x, W=W, S=S, T=T)
        ts = oracle.tile_stats(c.row_ptr, c.nnz, W * S, T)
        assert (st["n_tiles"], st["n_fast"], st["n_dirty"]) == \
            (ts["n_tiles"], ts["n_fast"], ts["n_dirty"])
        assert st["n_fast"] > 0 and st["n_dirty"] > 0


def test_tile_stats_brute_force():
    c = cases.from_lengths("x", [0, 0, 40, 0, 3, 130, 0, 1, 1, 0, 0, 64, 9], seed=4)
    P, T = 16, 2
    ts = oracle.tile_stats(c.row_ptr, c.nnz, P, T)
    tr = oracle.partition(c.row_ptr, c.nnz, P)
    NP = len(tr) - 1
    fast = sum(tr[p] == tr[p + 1] for p in range(NP))
    dirty = 0
    for p in range(NP):
        if tr[p] == tr[p + 1]:
            continue
        stop = min(tr[p + 1], c.m - 1)
        dirty += any(c.row_ptr[r] == c.row_ptr[r + 1] for r in range(tr[p], stop + 1))
    sync = sum(c.row_ptr[tr[w * T]] < w * T * P for w in range((NP + T - 1) // T))
    assert (ts["n_fast"], ts["n_dirty"], ts["n_sync"], ts["n_lead_empty"]) == (fast, dirty, sync, 2)


def test_error_metric():
    rp = np.array([0, 2, 2, 3], np.int32)
    col = np.array([0, 1, 1], np.int32)
    val = np.array([1.0, -2.0, 4.0])
    x = np.array([3.0, 0.5])
    yr = oracle.spmv_serial(rp, col, val, x)
    assert np.array_equal(yr, [2.0, 0.0, 2.0])
    assert oracle.rel_error(rp, col, val, x, yr, yr) == 0
    e, per = oracle.rel_error(rp, col, val, x, yr + np.array([0.4, 0, 0]), yr, per_row=True)
    assert per[0] == pytest.approx(0.4 / 4.0) and per[1] == 0 and e == per[0]
    # a nonzero in an empty row is infinitely wrong; so is NaN
    assert oracle.rel_error(rp, col, val, x, yr + np.array([0, 1e-300, 0]), yr) == np.inf
    assert oracle.rel_error(rp, col, val, x, np.array([np.nan, 0, 2.0]), yr) == np.inf


def test_planted_mistakes_are_caught():
    """Plausible slips in a segmented sum (a dropped item, a product shifted to
    the next row, a sign flip) all exceed the FP64 bound."""
    c = cases.from_lengths("r", np.random.default_rng(5).integers(1, 30, 400), n=800, seed=5)
    yr = oracle.spmv_serial(c.row_ptr, c.col, c.val, c.x)
    prods = c.val * c.x[c.col]
    r = 17
    j = c.row_ptr[r + 1] - 1
    dropped = yr.copy(); dropped[r] -= prods[j]
    shifted = dropped.copy(); shifted[r + 1] += prods[j]
    flipped = yr.copy(); flipped[r] -= 2 * prods[c.row_ptr[r]]
    for bad in (dropped, shifted, flipped):
        assert oracle.rel_error(c.row_ptr, c.col, c.val, c.x, bad, yr) > 1e-6


def test_openmp_equals_serial():
    A = matgen.uniform_rows(3, 5000, 5000, 0, 32)
    x = matgen.vector(4, A.n)
    y1 = oracle.spmv_serial(A.row_ptr, A.col, A.val, x)
    y2, t = oracle.spmv_omp(A.row_ptr, A.col, A.val, x, 4)
    assert t == 4 and np.array_equal(y1, y2)
    s, used = oracle.time_spmv(A.row_ptr, A.col, A.val, x, 1, 3)
    assert s > 0 and used == 1


# ---- the tuner's MLP backprop (oracle_mlp_loss_grad; no reference counterpart) ----
def _mlp_case(seed=0, n_in=8, h=6, n_out=5, N=7):
    rng = np.random.default_rng(seed)
    n_par = h * n_in + h + h * h + h + n_out * h + n_out
    params = rng.normal(0, 0.7, n_par)
    X = rng.normal(2.0, 3.0, (N, n_in))
    norm = np.concatenate([X.mean(0) + 0.1, X.std(0) + 0.5])
    T = rng.normal(0, 1, (N, n_out))
    M = (rng.random((N, n_out)) < 0.6).astype(np.float64)   # a few masked targets
    M[0, 0] = 1.0
    return (n_in, h, n_out), params, norm, X, T, M


def _mlp_forward_numpy(shape, params, norm, X):
    # the same network as three dense matrix products (numpy), not the oracle's loops
    n_in, h, n_out = shape
    o = 0
    W1 = params[o:o + h * n_in].reshape(h, n_in); o += h * n_in
    b1 = params[o:o + h]; o += h
    W2 = params[o:o + h * h].reshape(h, h); o += h * h
    b2 = params[o:o + h]; o += h
    W3 = params[o:o + n_out * h].reshape(n_out, h); o += n_out * h
    b3 = params[o:o + n_out]
    xs = (X - norm[:n_in]) / norm[n_in:]
    a1 = np.maximum(xs @ W1.T + b1, 0)
    a2 = np.maximum(a1 @ W2.T + b2, 0)
    return a2 @ W3.T + b3, (xs @ W1.T + b1, a1 @ W2.T + b2)


def _mlp_pins_hold(seed=0):
    shape, params, norm, X, T, M = _mlp_case(seed)
    loss, out, grad = oracle.mlp_loss_grad(*shape, params, norm, X, T, M)
    ref_out, pre = _mlp_forward_numpy(shape, params, norm, X)
    if np.max(np.abs(out - ref_out)) > 1e-12 * max(1.0, np.max(np.abs(ref_out))):
        return False
    if abs(loss - np.sum(M * (ref_out - T) ** 2) / np.sum(M)) > 1e-12 * max(1.0, loss):
        return False
    # central differences of the oracle's loss over every parameter; the step
    # stays away from every ReLU kink (checked below), so the loss is smooth there
    eps = 1e-6
    assert min(np.min(np.abs(p)) for p in pre) > 1e-3, "case sits on a ReLU kink"
    fd = np.empty_like(params)
    for i in range(len(params)):
        pp, pm = params.copy(), params.copy()
        pp[i] += eps
        pm[i] -= eps
        lp = oracle.mlp_loss_grad(*shape, pp, norm, X, T, M)[0]
        lm = oracle.mlp_loss_grad(*shape, pm, norm, X, T, M)[0]
        fd[i] = (lp - lm) / (2 * eps)
    scale = max(1e-3, np.max(np.abs(fd)))
    return bool(np.max(np.abs(grad - fd)) <= 1e-6 * scale)


@pytest.mark.parametrize("seed", range(3))
def test_mlp_backprop_matches_finite_differences(seed):
    # the pin of oracle_mlp_loss_grad: its outputs against numpy matrix products,
    # its gradient against central differences of its loss (8 -> 6 -> 6 -> 5,
    # masked targets, relative tolerance 1e-6 of the largest gradient)
    assert _mlp_pins_hold(seed)


@pytest.mark.parametrize("mistake", sorted(oracle.MLP_MUTATIONS))
def test_planted_mlp_mistake_breaks_the_pin(mistake):
    with oracle.mutated(mistake):
        assert not _mlp_pins_hold(0), mistake
    assert _mlp_pins_hold(0)
