"""The lane-exact kernel model against the FP64 oracle (CPU only): the
design of the HIP kernel -- row ownership, carries, wave boundary partials,
in-wave fix-up of dirty tiles -- is checked here before any GPU run."""
import numpy as np
import pytest

import cases
import kernel_model as km
import oracle

SHAPES = [(8, 3), (4, 1), (4, 2), (16, 1), (16, 4), (32, 2), (8, 7)]


@pytest.mark.parametrize("S,T", SHAPES)
def test_model_exact_on_integers(S, T):
    for c in cases.adversarial(64 * S, T, seed=S * 10 + T, integer=True):
        yr = oracle.spmv_serial(c.row_ptr, c.col, c.val, c.x)
        for dt in (np.float64, np.float32):
            y = km.segsum(c.row_ptr, c.col, c.val.astype(dt), c.x.astype(dt), S=S, TPW=T,
                          dtype=dt)
            assert not np.isnan(y).any(), c.name      # every row written
            assert np.array_equal(y.astype(np.float64), yr), (c.name, dt)


@pytest.mark.parametrize("S,T", SHAPES[:4])
def test_model_error_bounds(S, T):
    for c in cases.adversarial(64 * S, T, seed=S + T):
        for dt, tol in ((np.float64, 1e-13), (np.float32, 1e-5)):
            v, x = c.val.astype(dt), c.x.astype(dt)
            yr = oracle.spmv_serial(c.row_ptr, c.col, v.astype(np.float64), x.astype(np.float64))
            y = km.segsum(c.row_ptr, c.col, v, x, S=S, TPW=T, dtype=dt)
            e = oracle.rel_error(c.row_ptr, c.col, v.astype(np.float64), x.astype(np.float64),
                                 y, yr)
            assert e <= tol, (c.name, dt, e)


@pytest.mark.parametrize("S,T", SHAPES)
def test_model_counts_match_oracle(S, T):
    for c in cases.adversarial(64 * S, T, seed=3):
        cnt = {}
        km.segsum(c.row_ptr, c.col, c.val, c.x, S=S, TPW=T, counters=cnt)
        ts = oracle.tile_stats(c.row_ptr, c.nnz, 64 * S, T)
        assert (cnt["n_fast"], cnt["n_dirty"], cnt["n_sync"]) == \
            (ts["n_fast"], ts["n_dirty"], ts["n_sync"]), c.name


def test_model_alpha():
    c = cases.from_lengths("r", np.random.default_rng(0).integers(0, 20, 500), n=700)
    y1 = km.segsum(c.row_ptr, c.col, c.val, c.x, S=8, TPW=3, alpha=1.0)
    y2 = km.segsum(c.row_ptr, c.col, c.val, c.x, S=8, TPW=3, alpha=-0.25)
    assert np.array_equal(y2, -0.25 * y1)


@pytest.mark.parametrize("npan", [2, 4, 8])
def test_panel_combine_positions(npan):
    """The combine step's addressing (one wave per 64 rows: bit p of a per-row
    mask, a per-(group, panel) base, and the rank of the lane among the
    group's rows with bit p) finds exactly the stacked row of every nonempty
    (panel, row) pair, and nothing past the stacked matrix."""
    for c in cases.adversarial(512, 3, seed=npan):
        st, idx = cases.stack_panels(c, npan)
        m = c.m
        G = (m + 63) // 64
        mask = np.zeros(m, np.int64)
        for p in range(npan):
            mask |= (idx[p] >= 0).astype(np.int64) << p
        for p in range(npan):
            for g in range(G):
                rows = np.arange(64 * g, min(m, 64 * g + 64))
                has = (mask[rows] >> p) & 1 == 1
                if not has.any():
                    continue
                base = idx[p][rows[has]].min()          # first stacked row of the group
                rank = np.cumsum(has) - has              # lower lanes with the bit
                pos = base + rank[has]
                assert np.array_equal(pos, idx[p][rows[has]]), (c.name, p, g)
                assert pos.max() < st.m
