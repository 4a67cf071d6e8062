"""Batched persistent kernel (hipGraph) against the per-matrix plain baseline
(bit for bit: same tile shape, same per-tile arithmetic) and against the FP64
oracle per matrix and right-hand side."""
import numpy as np
import pytest

import cases
import oracle
from benchmark_spmv_using_csr_amd import batched, matgen

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-13, np.float32: 1e-5}


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def small_batch(seed, dtype, n=20, k_choices=(1, 4, 16)):
    rng = np.random.default_rng(seed)
    mats, ks = [], []
    for i in range(n):
        A = matgen.family_matrix(seed * 1000 + i, 100, 4000, i % 4)
        mats.append(A)
        ks.append(int(rng.choice(k_choices)))
    # edge cases: no nonzeros, leading/trailing empty rows, one long row
    for c in (cases.from_lengths("empty", [0] * 7, seed=1),
              cases.from_lengths("lead/trail empty", [0] * 70 + [3] * 50 + [0] * 9, seed=2),
              cases.from_lengths("long row", [2, 5000, 3], n=5100, seed=3),
              cases.six_by_six()):
        mats.append(matgen.CSR(c.m, c.n, c.row_ptr, c.col, c.val, c.name))
        ks.append(int(rng.choice(k_choices)))
    Xs = [rng.uniform(-1, 1, (A.n, k)).astype(dtype) for A, k in zip(mats, ks)]
    return mats, ks, Xs


def check_oracle(mats, ks, Xs, Ys, dtype, every=1):
    for i in range(0, len(mats), every):
        A = mats[i]
        v = A.val.astype(dtype).astype(np.float64)
        for r in range(ks[i]):
            x = Xs[i][:, r].astype(np.float64)
            yr = oracle.spmv_serial(A.row_ptr, A.col, v, x)
            e = oracle.rel_error(A.row_ptr, A.col, v, x, Ys[i][:, r], yr)
            assert e <= TOL[dtype], (A.name, r, e)


@pytest.mark.parametrize("dtype,sched", [
    (dt, sc) for dt in (np.float64, np.float32)
    for sc in ((8, 4, 4, 0, 1), (4, 1, 1, 0, 0), (16, 16, 2, 37, 1))] + [
    (np.float64, (4, 4, 1, 0, 1, 2)), (np.float64, (4, 16, 1, 0, 1, 4)),
    (np.float64, (4, 1, 2, 0, 0, 4))])
def test_batched_equals_plain_and_oracle(dtype, sched):
    mats, ks, Xs = small_batch(7, dtype)
    say(f"small batch {np.dtype(dtype).name} schedule={sched}")
    b = batched.SpmvBatch(mats, ks, dtype)
    b.set_schedule(*sched)
    b.set_x(Xs)
    b.run(1.0)
    Yb = b.y()
    b.run_plain(1.0)
    Yp = b.y(plain=True)
    for i, (yb, yp) in enumerate(zip(Yb, Yp)):
        assert np.array_equal(yb, yp), (i, mats[i].name)
    check_oracle(mats, ks, Xs, Yb, dtype)
    b.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_odd_and_even_rhs_counts(dtype):
    """Even right-hand-side counts are reduced in pairs (2-wide x loads from
    even x offsets, 2-wide y stores to even y offsets) where the tile allows
    (S=4 in fp64, S<=8 in fp32), odd ones one RHS at a time; matrices of both
    kinds share one batch.  Bit for bit against the plain launches, and
    within the oracle's bound."""
    mats, ks, Xs = small_batch(13, dtype, n=16, k_choices=(2, 3, 5, 6, 7, 16))
    say(f"odd/even k batch {np.dtype(dtype).name} ks={ks}")
    b = batched.SpmvBatch(mats, ks, dtype)
    scheds = [(8, 4, 4, 0, 1), (4, 2, 2, 0, 0)]
    if dtype == np.float64:                 # x staged in LDS: one to four passes per tile
        scheds += [(4, 4, 1, 0, 1, 4), (4, 2, 1, 0, 0, 4), (4, 4, 2, 0, 1, 2),
                   (4, 2, 1, 0, 0, 2)]
    for sched in scheds:
        b.set_schedule(*sched)
        b.set_x(Xs)
        b.run(1.0)
        Yb = b.y()
        b.run_plain(1.0)
        Yp = b.y(plain=True)
        for i, (yb, yp) in enumerate(zip(Yb, Yp)):
            assert np.array_equal(yb, yp), (sched, i, ks[i], mats[i].name)
        check_oracle(mats, ks, Xs, Yb, dtype)
    b.close()


def test_x_stage_schedules_are_checked():
    """x staging runs on fp64 tiles of 4 items per lane only, in passes of 2 or
    4 chunks, within the block's LDS; anything else is refused with an
    error."""
    from benchmark_spmv_using_csr_amd._lib import BhsError
    mats, ks, Xs = small_batch(5, np.float64, n=4)
    b = batched.SpmvBatch(mats, ks, np.float64)
    for bad in ((8, 4, 1, 0, 1, 2), (4, 4, 1, 0, 1, 8), (4, 4, 8, 0, 1, 4)):
        with pytest.raises(BhsError):
            b.set_schedule(*bad)
    b.set_schedule(4, 4, 1, 0, 1, 4)
    assert b.schedule() == (4, 4, 1, 0, 1, 4)
    b.close()
    mats, ks, Xs = small_batch(5, np.float32, n=4)
    b = batched.SpmvBatch(mats, ks, np.float32)
    with pytest.raises(BhsError):
        b.set_schedule(4, 4, 1, 0, 1, 4)
    b.close()


def test_graph_replays_and_alpha():
    mats, ks, Xs = small_batch(9, np.float64, n=12)
    b = batched.SpmvBatch(mats, ks, np.float64)
    b.set_x(Xs)
    b.run_plain(1.0)
    ref = b.y(plain=True)
    for _ in range(3):                    # replays with eager work in between
        b.run(1.0)
        got = b.y()
        b.run_plain(1.0)
        assert all(np.array_equal(g, r) for g, r in zip(got, ref))
    b.run(-0.5)                           # new alpha: graph re-captured
    got = b.y()
    assert all(np.array_equal(g, -0.5 * r) for g, r in zip(got, ref))
    Xs2 = [2.0 * X for X in Xs]
    b.set_x(Xs2)
    b.run(1.0)
    got = b.y()
    assert all(np.array_equal(g, 2.0 * r) for g, r in zip(got, ref))
    b.close()


def test_batch_autotune_keeps_results():
    mats, ks, Xs = small_batch(11, np.float64, n=10)
    b = batched.SpmvBatch(mats, ks, np.float64)
    b.set_x(Xs)
    ms, best = b.autotune(batched.SCHEDULES[:5], reps=3)
    assert best == int(np.argmin(np.where(ms > 0, ms, np.inf)))
    assert b.schedule() == batched.SCHEDULES[best]
    b.run(1.0)
    b2 = b.y()
    b.run_plain(1.0)
    assert all(np.array_equal(g, r) for g, r in zip(b2, b.y(plain=True)))
    b.close()


def test_full_256_matrix_batch():
    mats, ks, Xs = batched.batch_set()
    say(f"256-matrix batch nnz={sum(A.nnz for A in mats)}")
    b = batched.SpmvBatch(mats, ks, np.float64)
    b.set_x(Xs)
    b.run(1.0)
    Yb = b.y()
    b.run_plain(1.0)
    Yp = b.y(plain=True)
    assert all(np.array_equal(g, r) for g, r in zip(Yb, Yp))
    check_oracle(mats, ks, Xs, Yb, np.float64, every=16)
    b.set_schedule(4, 16, 1, 0, 1, 0)
    b.run_plain(1.0)
    Yp4 = b.y(plain=True)
    for xs in (4, 2):                       # x staged in LDS: the plain launches' bits
        say(f"256-matrix batch x_stage={xs}")
        b.set_schedule(4, 16, 1, 0, 1, xs)
        b.run(1.0)
        assert all(np.array_equal(g, r) for g, r in zip(b.y(), Yp4))
    b.close()


def test_plain_graph_replay_matches_plain_launches():
    """The baseline's launches replayed as one hipGraph (bench mode 2) give the
    plain launches' y, every replay."""
    mats, ks, Xs = small_batch(21, np.float64, n=10)
    b = batched.SpmvBatch(mats, ks, np.float64)
    b.set_schedule(4, 4, 2, 0, 1)
    b.set_x(Xs)
    b.run_plain(1.0)
    Yp = b.y(plain=True)
    for _ in range(2):
        b.bench(plain=2, warmup=0, reps=2)
        for i, (u, v) in enumerate(zip(b.y(plain=True), Yp)):
            assert np.array_equal(u, v), i
    b.close()
