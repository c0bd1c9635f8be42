"""Seeded synthetic LPs for the benchmark workloads and the parity tests.

Every value is dyadic: coefficients are multiples of 1/8 in +-[0.5, 2], all
other data multiples of 2^-10, so every product and sum formed by presolve on
the original data is exact in FP64.  Planted reductions (fractions of rows or
columns): singleton rows, singleton columns, doubleton equalities, scaled
duplicate rows (factor +-2^k, k in -2..2), fixed columns (l = u) and empty
columns.  u_j = +inf only where c_j >= 0, so that empty and dual-fixable
columns never make the LP unbounded.

An LP is a dict of numpy arrays: m, n, A_ptr, A_idx, A_val (row-sorted CSR),
AT_ptr, AT_idx, AT_val (its transpose, rows sorted), c, l, u, AL, AU, obj.
"""
from __future__ import annotations

import numpy as np

GRID = 2.0 ** -10


def _dyadic(x, step=GRID):
    return np.round(np.asarray(x, dtype=np.float64) / step) * step


def _coeffs(rng, k):
    mag = np.round(rng.uniform(0.5, 2.0, size=k) * 8.0) / 8.0
    sign = np.where(rng.random(k) < 0.5, -1.0, 1.0)
    return mag * sign


def csr_from_coo(m, n, rows, cols, vals):
    """Row-sorted CSR from COO; duplicate (row, col) entries keep the first."""
    key = rows.astype(np.int64) * n + cols.astype(np.int64)
    order = np.argsort(key, kind="stable")
    key = key[order]
    keep = np.ones(len(key), dtype=bool)
    if len(key) > 1:
        keep[1:] = key[1:] != key[:-1]
    order = order[keep]
    r = rows[order].astype(np.int64)
    ptr = np.zeros(m + 1, dtype=np.int64)
    np.add.at(ptr, r + 1, 1)
    ptr = np.cumsum(ptr).astype(np.int32)
    return ptr, cols[order].astype(np.int32), vals[order].astype(np.float64)


def transpose_csr(m, n, ptr, idx, val):
    rows = np.repeat(np.arange(m, dtype=np.int32), np.diff(ptr))
    order = np.argsort(idx, kind="stable")
    tptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(tptr, idx.astype(np.int64) + 1, 1)
    return np.cumsum(tptr).astype(np.int32), rows[order].astype(np.int32), val[order].astype(np.float64)


def _finish(rng, m, n, rows, cols, vals, c, l, u, x0, row_kind, dup_src, dup_fac):
    """Builds CSR/AT and row bounds. row_kind: 0 general, 1 equality, 2 one-sided,
    3 duplicate (scaled copy of dup_src)."""
    ptr, idx, val = csr_from_coo(m, n, rows, cols, vals)
    ax = np.zeros(m)
    rr = np.repeat(np.arange(m), np.diff(ptr))
    np.add.at(ax, rr, val * x0[idx])
    lo = ax - _dyadic(rng.uniform(0.0, 1.0, size=m))
    hi = ax + _dyadic(rng.uniform(0.0, 1.0, size=m))
    AL = lo.copy()
    AU = hi.copy()
    eq = row_kind == 1
    AL[eq] = ax[eq]
    AU[eq] = ax[eq]
    one = row_kind == 2
    side = rng.random(m) < 0.5
    AL[one & side] = -np.inf
    AU[one & ~side] = np.inf
    dup = np.nonzero(row_kind == 3)[0]
    if len(dup):
        src = dup_src[dup]
        f = dup_fac[dup]
        a = AL[src] * f
        b = AU[src] * f
        AL[dup] = np.where(f > 0, a, b)
        AU[dup] = np.where(f > 0, b, a)
    tptr, tidx, tval = transpose_csr(m, n, ptr, idx, val)
    return {"m": int(m), "n": int(n), "A_ptr": ptr, "A_idx": idx, "A_val": val,
            "AT_ptr": tptr, "AT_idx": tidx, "AT_val": tval,
            "c": c.astype(np.float64), "l": l.astype(np.float64), "u": u.astype(np.float64),
            "AL": AL.astype(np.float64), "AU": AU.astype(np.float64), "obj": 0.0}


def generate(m, n, seed, row_len=(2, 10), mean_row_len=None, min_row_len=2, zipf=None, zipf_cap=5000,
             frac_singleton_rows=0.05, frac_singleton_cols=0.05, frac_doubleton=0.03,
             frac_dup_rows=0.02, frac_fixed_cols=0.02, frac_empty_cols=0.01,
             frac_eq=0.30, frac_one_sided=0.10, frac_inf_u=0.20):
    """General planted LP.  Row lengths of general rows: Poisson(mean_row_len)
    clipped to >= min_row_len when mean_row_len is given, else uniform in
    row_len.
    Column choice is uniform, or Zipf(zipf) popularity with a per-column cap."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = _dyadic(rng.normal(0.0, 1.0, size=n))
    perm = rng.permutation(n)
    n_empty = int(round(frac_empty_cols * n))
    n_fixed = int(round(frac_fixed_cols * n))
    n_scol = int(round(frac_singleton_cols * n))
    empty_cols = perm[:n_empty]
    fixed_cols = perm[n_empty:n_empty + n_fixed]
    scols = perm[n_empty + n_fixed:n_empty + n_fixed + n_scol]
    normal = np.setdiff1d(np.arange(n), np.concatenate([empty_cols, scols]))
    l = np.zeros(n)
    u = _dyadic(rng.uniform(1.0, 10.0, size=n))
    inf_cand = np.nonzero(c >= 0)[0]
    n_inf = min(len(inf_cand), int(round(frac_inf_u * n)))
    u[rng.choice(inf_cand, size=n_inf, replace=False)] = np.inf
    x0 = _dyadic(rng.uniform(0.0, 1.0, size=n) * np.where(np.isfinite(u), u, 10.0))
    x0 = np.minimum(np.maximum(x0, l), np.where(np.isfinite(u), u, np.inf))
    fu = np.where(np.isfinite(u[fixed_cols]), u[fixed_cols], 10.0)
    x0[fixed_cols] = _dyadic(rng.uniform(0.0, 1.0, size=n_fixed) * fu)
    l[fixed_cols] = x0[fixed_cols]
    u[fixed_cols] = x0[fixed_cols]

    rperm = rng.permutation(m)
    n_srow = int(round(frac_singleton_rows * m))
    n_dbl = int(round(frac_doubleton * m))
    n_dup = int(round(frac_dup_rows * m))
    srows = rperm[:n_srow]
    drows = rperm[n_srow:n_srow + n_dbl]
    duprows = rperm[n_srow + n_dbl:n_srow + n_dbl + n_dup]
    grows = np.sort(rperm[n_srow + n_dbl + n_dup:])
    row_kind = np.zeros(m, dtype=np.int8)
    kind_draw = rng.random(m)
    row_kind[kind_draw < frac_eq] = 1
    row_kind[(kind_draw >= frac_eq) & (kind_draw < frac_eq + frac_one_sided)] = 2
    row_kind[drows] = 1
    row_kind[srows] = np.where(row_kind[srows] == 2, 0, row_kind[srows])

    # General rows
    if mean_row_len is not None:
        lens = np.maximum(min_row_len, rng.poisson(mean_row_len, size=len(grows)))
    else:
        lens = rng.integers(row_len[0], row_len[1] + 1, size=len(grows))
    total = int(lens.sum())
    if zipf is None:
        gcols = normal[rng.integers(0, len(normal), size=total)]
    else:
        ranks = np.arange(1, len(normal) + 1, dtype=np.float64)
        w = ranks ** (-float(zipf))
        w /= w.sum()
        expected = w * total
        capped = np.zeros(len(normal), dtype=bool)
        for _ in range(50):  # water-filling of the per-column cap
            over = (~capped) & (expected > zipf_cap)
            if not over.any():
                break
            capped |= over
            rest = total - zipf_cap * capped.sum()
            ww = np.where(capped, 0.0, w)
            expected = np.where(capped, float(zipf_cap), ww / ww.sum() * rest)
        counts = np.floor(expected).astype(np.int64)
        short = total - int(counts.sum())
        if short > 0:
            frac = expected - counts
            extra = rng.choice(len(normal), size=short, replace=False, p=frac / frac.sum())
            counts[extra] += 1
        col_order = rng.permutation(normal)
        gcols = np.repeat(col_order, counts)
        rng.shuffle(gcols)
    grow_ids = np.repeat(grows, lens)
    gvals = _coeffs(rng, total)

    # Singleton columns: one entry each in a random general row.
    sc_rows = grows[rng.integers(0, len(grows), size=len(scols))]
    sc_vals = _coeffs(rng, len(scols))
    # Singleton rows.
    sr_cols = normal[rng.integers(0, len(normal), size=len(srows))]
    sr_vals = _coeffs(rng, len(srows))
    # Doubleton equalities: two distinct normal columns.
    d1 = normal[rng.integers(0, len(normal), size=len(drows))]
    d2 = normal[rng.integers(0, len(normal), size=len(drows))]
    same = d1 == d2
    d2[same] = normal[(np.searchsorted(normal, d2[same]) + 1) % len(normal)]
    dv = _coeffs(rng, 2 * len(drows))

    rows = np.concatenate([grow_ids, sc_rows, srows, drows, drows]).astype(np.int64)
    cols = np.concatenate([gcols, scols, sr_cols, d1, d2]).astype(np.int64)
    vals = np.concatenate([gvals, sc_vals, sr_vals, dv])

    # Duplicate rows: scaled copies of general rows.
    dup_src = np.full(m, -1, dtype=np.int64)
    dup_fac = np.ones(m)
    if len(duprows):
        ptr0, idx0, val0 = csr_from_coo(m, n, rows, cols, vals)
        src = grows[rng.integers(0, len(grows), size=len(duprows))]
        fac = (2.0 ** rng.integers(-2, 3, size=len(duprows))) * np.where(rng.random(len(duprows)) < 0.5, -1.0, 1.0)
        dup_src[duprows] = src
        dup_fac[duprows] = fac
        row_kind[duprows] = 3
        lens_src = np.diff(ptr0)[src]
        starts = ptr0[src]
        pos = np.repeat(starts, lens_src) + (np.arange(lens_src.sum()) - np.repeat(np.cumsum(lens_src) - lens_src, lens_src))
        rows = np.concatenate([rows, np.repeat(duprows, lens_src)])
        cols = np.concatenate([cols, idx0[pos]])
        vals = np.concatenate([vals, val0[pos] * np.repeat(fac, lens_src)])
    return _finish(rng, m, n, rows, cols, vals, c, l, u, x0, row_kind, dup_src, dup_fac)


def config0(seed=0):
    """m=2,000, n=3,000, ~12k nnz, mean row length 6."""
    return generate(2000, 3000, seed, mean_row_len=6.3)


def config1(seed=1, m=1_000_000, n=1_500_000, row_len=(2, 14)):
    """Large sparse LP: m=1e6, n=1.5e6, general row lengths U[2,14] (~8M nnz),
    Zipf(1.5) column popularity capped at 5,000 per column."""
    return generate(m, n, seed, row_len=row_len, zipf=1.5, zipf_cap=5000)


def config2(seed=2, blocks=2000, brows=40, bcols=60, density=0.15, link_rows=200, link_nnz=3000,
            frac_parallel_cols=0.30, frac_identical_blocks=0.10):
    """Block-diagonal LP with dense linking rows and planted parallel columns."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = blocks * brows + link_rows
    n = blocks * bcols
    rows, cols, vals = [], [], []
    proto = None
    for b in range(blocks):
        r0, c0 = b * brows, b * bcols
        if proto is not None and rng.random() < frac_identical_blocks:
            rr, cc, vv = proto
        else:
            mask = rng.random((brows, bcols)) < density
            rr, cc = np.nonzero(mask)
            vv = _coeffs(rng, len(rr))
            # parallel copies: column k := scale * column j inside the block
            npar = int(round(frac_parallel_cols * bcols))
            src = rng.integers(0, bcols, size=npar)
            dst = rng.choice(bcols, size=npar, replace=False)
            keep = ~np.isin(cc, dst)
            rr, cc, vv = rr[keep], cc[keep], vv[keep]
            add_r, add_c, add_v = [], [], []
            for s, d in zip(src, dst):
                if s == d or s in dst:
                    continue
                sel = cc == s
                f = (2.0 ** rng.integers(-1, 2)) * (1.0 if rng.random() < 0.7 else -1.0)
                add_r.append(rr[sel]); add_c.append(np.full(sel.sum(), d)); add_v.append(vv[sel] * f)
            if add_r:
                rr = np.concatenate([rr] + add_r); cc = np.concatenate([cc] + add_c)
                vv = np.concatenate([vv] + add_v)
            if proto is None:
                proto = (rr, cc, vv)
        rows.append(rr + r0); cols.append(cc + c0); vals.append(vv)
    for k in range(link_rows):
        cc = rng.choice(n, size=min(link_nnz, n), replace=False)
        rows.append(np.full(len(cc), blocks * brows + k)); cols.append(cc); vals.append(_coeffs(rng, len(cc)))
    rows = np.concatenate(rows).astype(np.int64)
    cols = np.concatenate(cols).astype(np.int64)
    vals = np.concatenate(vals)
    c = _dyadic(rng.normal(0.0, 1.0, size=n))
    l = np.zeros(n)
    u = _dyadic(rng.uniform(1.0, 10.0, size=n))
    inf_cand = np.nonzero(c >= 0)[0]
    u[rng.choice(inf_cand, size=min(len(inf_cand), int(0.2 * n)), replace=False)] = np.inf
    x0 = _dyadic(rng.uniform(0.0, 1.0, size=n) * np.where(np.isfinite(u), u, 10.0))
    row_kind = np.zeros(m, dtype=np.int8)
    d = rng.random(m)
    row_kind[d < 0.30] = 1
    row_kind[(d >= 0.30) & (d < 0.40)] = 2
    return _finish(rng, m, n, rows, cols, vals, c, l, u, x0, row_kind, np.full(m, -1), np.ones(m))


def reduced_solution(seed, m1, n1, l, u):
    """Seeded reduced-space (x, y, z) with x projected onto [l, u]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = _dyadic(rng.uniform(-1.0, 11.0, size=n1))
    x = np.minimum(np.maximum(x, l), u)
    x = np.where(np.isfinite(x), x, 0.0)
    y = _dyadic(rng.normal(0.0, 1.0, size=m1))
    z = _dyadic(rng.normal(0.0, 1.0, size=n1))
    return x, y, z
