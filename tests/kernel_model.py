"""Lane-exact CPU model of the HIP segmented-sum kernel (test helper).

Mirrors benchmark_spmv_using_csr_amd/csrc/bhs_kernels.h::segsum_kernel and
calibrate_kernel operation by operation, in the kernel's own precision and
summation order (products rounded before they are summed; the library is built
with -ffp-contract=off), so the GPU result must equal this model bit for bit.
The model itself is checked against the FP64 oracle in test_model.py; it is a
description of the kernel, not the parity oracle.
"""
from __future__ import annotations

import numpy as np

W = 64
FULL = (1 << 64) - 1


def partition(row_ptr, nnz, P):
    NP = (nnz + P - 1) // P
    keys = np.minimum(np.arange(NP + 1, dtype=np.int64) * P, nnz)
    return (np.searchsorted(row_ptr, keys, side="right") - 1).astype(np.int32)


def _wave_sum(v):
    """wave_sum: quads ((v0+v1)+(v2+v3)), half rows, rows (each step adds the
    mirror-image partner), then (R0+R1)+(R2+R3); returns the scalar."""
    v = v.copy()
    lanes = np.arange(W)
    v = v + v[lanes ^ 1]
    v = v + v[lanes ^ 2]
    v = v + v[(lanes & ~7) | (7 - (lanes & 7))]
    v = v + v[(lanes & ~15) | (15 - (lanes & 15))]
    return (v[0] + v[16]) + (v[32] + v[48])


def _segscan(v, present):
    """Hillis-Steele segmented inclusive scan with the kernel's early exit."""
    cov = 0
    for l in range(W):
        if present[l]:
            cov |= 1 << l
    v = v.copy()
    d = 1
    while d < W:
        if cov == FULL:
            break
        up = v.copy()
        up[d:] = v[:-d]              # __shfl_up: lanes below d keep their own value
        for l in range(W):
            if not (cov >> l) & 1:
                v[l] = up[l] + v[l]
        cov = (cov | (cov << d)) & FULL
        d <<= 1
    return v


def _calibrate(sync_row, sync_val, NW, y, T):
    """calibrate_kernel: one wave per 64 consecutive sync entries."""
    for base in range(0, NW, W):
        idx = np.arange(base, base + W)
        r = np.where(idx < NW, sync_row[np.minimum(idx, NW - 1)], -1)
        v = np.where((idx < NW) & (r >= 0), sync_val[np.minimum(idx, NW - 1)], T(0)).astype(y.dtype)
        before = sync_row[base - 1] if base > 0 else -1
        rprev = np.concatenate([[before], r[:-1]])
        starts = r != rprev
        starts[0] = True
        s = _segscan(v, starts)
        run_end = np.concatenate([r[1:] != r[:-1], [True]])
        lead_cont = r[0] >= 0 and before == r[0]
        ext = T(0)
        if r[W - 1] >= 0:
            nb = base + W
            while nb < NW:
                i2 = np.arange(nb, nb + W)
                r2 = np.where(i2 < NW, sync_row[np.minimum(i2, NW - 1)], -1)
                same = r2 == r[W - 1]
                ln = W if same.all() else int(np.argmin(same))
                if ln == 0:
                    break
                v2 = np.zeros(W, y.dtype)
                v2[:ln] = sync_val[nb:nb + ln]
                ext = T(ext + _wave_sum(v2))
                if ln < W:
                    break
                nb += W
        run_start = 0
        for i in range(W):
            if starts[i]:
                run_start = i
            if run_end[i] and r[i] >= 0 and not (run_start == 0 and lead_cont):
                tot = T(s[i] + ext) if i == W - 1 else s[i]
                y[r[i]] = T(y[r[i]] + tot)


def calibrate_chains(sync_row, sync_val, y):
    """The column-panel combine's calibration: every run of consecutive sync
    entries with the same row R (a chain) is summed 64 entries at a time --
    _wave_sum of each zero-padded chunk, the chunk sums added in order -- and
    y[R] becomes y[R] + total."""
    T = y.dtype.type
    NW = len(sync_row)
    w = 0
    while w < NW:
        r = int(sync_row[w])
        if r < 0:
            w += 1
            continue
        e = w
        while e < NW and sync_row[e] == r:
            e += 1
        tot = None
        for j in range(w, e, W):
            chunk = np.zeros(W, y.dtype)
            chunk[:min(W, e - j)] = sync_val[j:min(j + W, e)]
            s = T(_wave_sum(chunk))
            tot = s if tot is None else T(tot + s)
        y[r] = T(y[r] + tot)
        w = e


def segsum(row_ptr, col, val, x, S=8, TPW=3, alpha=1.0, dtype=np.float64, counters=None,
           calibrate=True):
    """y = alpha*A*x exactly as segsum_kernel + calibrate_kernel compute it.
    With calibrate=False: (y before calibration, sync_row, sync_val)."""
    T = np.dtype(dtype).type
    row_ptr = np.asarray(row_ptr, np.int64)
    m = len(row_ptr) - 1
    nnz = int(row_ptr[-1]) if m >= 0 else 0
    P = W * S
    NP = (nnz + P - 1) // P
    NW = (NP + TPW - 1) // TPW
    y = np.full(max(m, 0), np.nan, dtype)
    alpha = T(alpha)
    st = {"n_fast": 0, "n_dirty": 0, "n_sync": 0}
    if counters is not None:
        counters.update(st)
    none = (np.zeros(0, np.int64), np.zeros(0, dtype))
    if m == 0:
        return y if calibrate else (y,) + none
    if NP == 0:
        y[:] = 0
        return y if calibrate else (y,) + none
    tile_row = partition(row_ptr, nnz, P)
    lead_empty = int(tile_row[0])
    pad = NP * P
    colp = np.zeros(pad, np.int64)
    colp[:nnz] = col
    valp = np.zeros(pad, dtype)
    valp[:nnz] = np.asarray(val, dtype)
    xx = np.asarray(x, dtype)
    if len(xx) == 0:
        xx = np.zeros(1, dtype)
    idx = np.arange(pad)
    prod_all = np.where(idx < nnz, valp * xx[colp], T(0)).astype(dtype)

    sync_row = np.full(NW, -2, np.int64)
    sync_val = np.zeros(NW, dtype)
    lanes = np.arange(W)
    G = 16 // np.dtype(dtype).itemsize
    kk = np.arange(S)
    coal = (kk[None, :] // G) * (W * G) + lanes[:, None] * G + (kk[None, :] % G)
    for w in range(NW):
        if lead_empty > 0:
            lo = lead_empty * w // NW
            hi = lead_empty * (w + 1) // NW
            y[lo:hi] = 0
        p_begin = w * TPW
        p_end = min(p_begin + TPW, NP)
        E0 = p_begin * P
        r_first = int(tile_row[p_begin])
        syncing = row_ptr[r_first] < E0
        did_sync = False
        is_open = False
        carry = T(0)
        for p in range(p_begin, p_end):
            B = p * P
            r0 = int(tile_row[p])
            r1 = int(tile_row[p + 1])
            cont = row_ptr[r0] < B
            prod = prod_all[B:B + P].reshape(W, S).copy()
            if r0 == r1:
                # coalesced item order: lane l's item k is nonzero
                # B + (k // G) * 64 * G + l * G + k % G
                pc = prod_all[B + coal].copy()
                s = pc[:, 0].copy()
                for k in range(1, S):
                    s = s + pc[:, k]
                s = _wave_sum(s)
                carry = T(carry + s) if cont else T(s)
                is_open = True
                st["n_fast"] += 1
                continue
            stop = m - 1 if r1 == m else r1
            hb = np.zeros(W, np.int64)
            dirty = False
            for r in range(r0, stop + 1):
                if row_ptr[r + 1] > row_ptr[r]:
                    o = int(row_ptr[r] - B)
                    if 0 <= o < P:
                        hb[o // S] |= 1 << (o % S)
                else:
                    dirty = True
                    y[r] = 0
            hb[0] |= 1
            if cont:
                prod[0, 0] = T(prod[0, 0] + carry)
            nh = np.array([bin(int(h)).count("1") for h in hb])
            incl = np.cumsum(nh)
            base = incl - nh
            K = int(incl[-1])
            s_seg = np.full(P, np.nan, dtype)
            acc = np.zeros(W, dtype)
            lead = np.zeros(W, dtype)
            seen = np.zeros(W, bool)
            sidx = base.copy()
            for k in range(S):
                h = ((hb >> k) & 1).astype(bool)
                for l in lanes[h]:
                    if seen[l]:
                        s_seg[sidx[l]] = acc[l]
                        sidx[l] += 1
                    else:
                        lead[l] = acc[l]
                        seen[l] = True
                acc[h] = 0
                acc = acc + prod[:, k]
            I = _segscan(acc, seen)
            for l in range(1, W):
                if seen[l]:
                    s_seg[base[l] - 1] = I[l - 1] + lead[l]
            s_seg[K - 1] = I[W - 1]
            partial_tail = r1 < m and row_ptr[r1] < B + P
            kb = 1 if syncing else 0
            ke = K - 1 if partial_tail else K
            if not dirty:
                for k in range(kb, ke):
                    y[r0 + k] = alpha * s_seg[k]
            else:
                st["n_dirty"] += 1
                running = 0
                for r in range(r0, stop + 1):
                    if row_ptr[r + 1] > row_ptr[r]:
                        if kb <= running < ke:
                            y[r] = alpha * s_seg[running]
                        running += 1
            if syncing:
                sync_row[w] = r0
                sync_val[w] = alpha * s_seg[0]
                syncing = False
                did_sync = True
            carry = T(s_seg[K - 1])
            is_open = partial_tail
        if is_open:
            r_open = int(tile_row[p_end])
            if syncing:
                sync_row[w] = r_open
                sync_val[w] = alpha * carry
                did_sync = True
            else:
                y[r_open] = alpha * carry
        if not did_sync:
            sync_row[w] = -1
        if did_sync:
            st["n_sync"] += 1
    if counters is not None:
        counters.update(st)
    if not calibrate:
        return y, sync_row, sync_val
    _calibrate(sync_row, sync_val, NW, y, T)
    return y


def segsum_panels(st_row_ptr, st_col, st_val, idx, x, S=8, TPW=3, dtype=np.float64):
    """The column-panel variant: segsum_kernel over the stacked matrix (no
    calibrate_kernel), then combine_panels_kernel -- its chain calibration and
    the sum over panels in panel order.  idx (npan x m) holds the stacked row
    of every (panel, row) pair, -1 where the row has no entry in the panel."""
    T = np.dtype(dtype).type
    npan, m = idx.shape
    if len(st_row_ptr) > 1:
        ys, sr, sv = segsum(st_row_ptr, st_col, np.asarray(st_val, dtype), np.asarray(x, dtype),
                            S=S, TPW=TPW, dtype=dtype, calibrate=False)
        calibrate_chains(sr, sv, ys)
    else:
        ys = np.zeros(1, dtype)
    part = lambda p: np.where(idx[p] >= 0, ys[np.maximum(idx[p], 0)], T(0)).astype(dtype)
    out = part(0)
    for p in range(1, npan):
        out = (out + part(p)).astype(dtype)
    return out
