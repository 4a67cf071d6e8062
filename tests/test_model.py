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


def chain_rule(rp, PW, NW):
    """Which waves leave a wave-boundary partial, from the stacked row_ptr and
    the wave size alone (the rule the library's analysis lists the combine's
    chain records by): wave w > 0 does when the row holding nonzero w*PW began
    before it; the entry is that row."""
    rp = np.asarray(rp, np.int64)
    out = np.full(NW, -1, np.int64)
    if NW > 1:
        s = np.arange(1, NW, dtype=np.int64) * PW
        R = np.searchsorted(rp, s, side="right") - 1
        out[1:] = np.where(rp[R] < s, R, -1)
    return out


@pytest.mark.parametrize("npan", [2, 4, 8])
@pytest.mark.parametrize("S,T", [(4, 1), (8, 3), (16, 2)])
def test_panel_chain_rule_matches_the_sync_entries(npan, S, T):
    """The structural rule gives exactly the sync entries the segmented sum
    leaves on every stacked adversarial case (no empty stacked rows)."""
    for c in cases.adversarial(64 * S, T, seed=npan + S):
        st, _ = cases.stack_panels(c, npan)
        if st.m == 0:
            continue
        _, sr, _ = km.segsum(st.row_ptr, st.col, st.val, c.x, S=S, TPW=T, calibrate=False)
        assert np.array_equal(sr, chain_rule(st.row_ptr, 64 * S * T, len(sr))), c.name


@pytest.mark.parametrize("npan", [2, 4])
@pytest.mark.parametrize("S,T", [(4, 1), (8, 3)])
def test_panel_model_error_bounds(npan, S, T):
    """Chain calibration + panel sum: exact on integer data, within the FP64
    and FP32 bounds otherwise."""
    for integer in (True, False):
        for c in cases.adversarial(64 * S, T, seed=7 * npan + S, integer=integer):
            st, idx = cases.stack_panels(c, npan)
            yr = oracle.spmv_serial(c.row_ptr, c.col, c.val, c.x)
            y = km.segsum_panels(st.row_ptr, st.col, st.val, idx, c.x, S=S, TPW=T)
            if integer:
                assert np.array_equal(y, yr), c.name
            assert oracle.rel_error(c.row_ptr, c.col, c.val, c.x, y, yr) <= 1e-13, c.name
            v32 = c.val.astype(np.float32).astype(np.float64)
            x32 = c.x.astype(np.float32).astype(np.float64)
            y32 = km.segsum_panels(st.row_ptr, st.col, st.val, idx, c.x, S=S, TPW=T,
                                   dtype=np.float32)
            yr32 = oracle.spmv_serial(c.row_ptr, c.col, v32, x32)
            assert oracle.rel_error(c.row_ptr, c.col, v32, x32, y32, yr32) <= 1e-5, c.name


def test_chain_calibration_chunks():
    """A chain of 150 partials is summed as three zero-padded 64-entry chunks
    (wave_sum each) added in order; a chain of one partial adds it unchanged;
    sync entries of other rows and -1 entries are left alone."""
    rng = np.random.default_rng(5)
    sv = rng.standard_normal(160)
    sr = np.full(160, -1, np.int64)
    sr[3:153] = 7
    sr[155] = 2
    y = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0])
    y0 = y.copy()
    km.calibrate_chains(sr, sv, y)
    assert y[2] == y0[2] + sv[155]
    ch = [np.pad(sv[j:min(j + 64, 153)], (0, 64 - (min(j + 64, 153) - j))) for j in (3, 67, 131)]
    assert np.isclose(y[7], y0[7] + sv[3:153].sum(), rtol=0, atol=1e-12)
    assert y[7] == y0[7] + ((km._wave_sum(ch[0]) + km._wave_sum(ch[1])) + km._wave_sum(ch[2]))
    assert np.array_equal(np.delete(y, [2, 7]), np.delete(y0, [2, 7]))
