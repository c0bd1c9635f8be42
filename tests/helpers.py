"""Shared helpers of the test suite: phase plans and statistics built on the
host exactly as the presolve driver builds them on the device, and the
field-by-field comparison of results."""
import numpy as np


def fresh_plan(lp):
    """A phase plan over `lp`: all rows/columns kept, copies of the vectors."""
    return {"keep_row": np.ones(lp["m"], np.uint8), "keep_col": np.ones(lp["n"], np.uint8),
            "c": np.array(lp["c"], np.float64), "l": np.array(lp["l"], np.float64),
            "u": np.array(lp["u"], np.float64), "AL": np.array(lp["AL"], np.float64),
            "AU": np.array(lp["AU"], np.float64), "obj_constant_delta": 0.0}


def stats(lp):
    """Row/column statistics from the CSR counts (explicit zeros included; the
    singleton entry is the first entry of its row/column)."""
    out = {}
    for side, ptr, idx, val in (("row", lp["A_ptr"], lp["A_idx"], lp["A_val"]),
                                ("col", lp["AT_ptr"], lp["AT_idx"], lp["AT_val"])):
        ptr = np.asarray(ptr)
        cnt = np.diff(ptr).astype(np.int32)
        single = cnt == 1
        first = np.minimum(ptr[:-1], max(len(idx) - 1, 0))
        sidx = np.where(single, np.asarray(idx)[first] if len(idx) else 0, -1).astype(np.int32)
        sval = np.where(single, np.asarray(val)[first] if len(val) else 0.0, 0.0)
        if side == "row":
            out.update(row_nnz=cnt, singleton_row_mask=single.astype(np.uint8),
                       singleton_row_col=sidx, singleton_row_val=sval)
        else:
            out.update(col_nnz=cnt, empty_col_mask=(cnt == 0).astype(np.uint8),
                       column_singleton_mask=single.astype(np.uint8), column_singleton_row=sidx,
                       column_singleton_val=sval)
    return out


def assert_same_array(name, a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, f"{name}: shape {a.shape} vs {b.shape}"
    if a.dtype.kind == "f":
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        if not same.all():
            i = int(np.nonzero(~same)[0][0])
            raise AssertionError(f"{name}: first difference at {i}: {a[i]!r} vs {b[i]!r} "
                                 f"({int((~same).sum())} differ)")
    else:
        if not np.array_equal(a, b):
            i = int(np.nonzero(a != b)[0][0])
            raise AssertionError(f"{name}: first difference at {i}: {a[i]!r} vs {b[i]!r}")


RECORD_KEYS = ("row_red2org", "col_red2org", "fixed_idx", "fixed_val", "types", "index_starts",
               "value_starts", "dual_modes", "indices", "vals")
LP_KEYS = ("A_ptr", "A_idx", "A_val", "AT_ptr", "AT_idx", "AT_val", "c", "l", "u", "AL", "AU")


def assert_same_record(ra, rb):
    for k in ("m0", "n0", "m1", "n1"):
        assert ra[k] == rb[k], f"record {k}: {ra[k]} vs {rb[k]}"
    for k in RECORD_KEYS:
        assert_same_array("record " + k, ra[k], rb[k])
    assert ra["obj_constant_new"] == rb["obj_constant_new"], \
        f"obj_constant_new {ra['obj_constant_new']!r} vs {rb['obj_constant_new']!r}"


def assert_same_lp(la, lb):
    assert (la["m"], la["n"]) == (lb["m"], lb["n"])
    for k in LP_KEYS:
        assert_same_array("lp " + k, la[k], lb[k])
    assert la["obj"] == lb["obj"], f"obj {la['obj']!r} vs {lb['obj']!r}"


def digests(d, keys):
    """SHA-256 of each array (dtype, shape and bytes), for comparing results
    too large to keep as fixtures."""
    import hashlib

    out = {}
    for k in keys:
        a = np.ascontiguousarray(d[k])
        h = hashlib.sha256(f"{a.dtype.str}{a.shape}".encode())
        h.update(a.tobytes())
        out[k] = h.hexdigest()
    return out


def lp_from_rows(rows, n, c, AL, AU, l, u):
    """LP dict from rows given as lists of (column, value) in column order."""
    m = len(rows)
    ptr = np.zeros(m + 1, np.int32)
    idx, val = [], []
    for i, r in enumerate(rows):
        for j, v in r:
            idx.append(j)
            val.append(v)
        ptr[i + 1] = len(idx)
    idx, val = np.array(idx, np.int32), np.array(val, np.float64)
    order = np.argsort(idx, kind="stable")
    tptr = np.zeros(n + 1, np.int32)
    np.add.at(tptr, idx + 1, 1)
    tptr = np.cumsum(tptr).astype(np.int32)
    row_of = np.repeat(np.arange(m, dtype=np.int32), np.diff(ptr))
    f = lambda a: np.array(a, np.float64)  # noqa: E731
    return {"m": m, "n": n, "A_ptr": ptr, "A_idx": idx, "A_val": val, "AT_ptr": tptr,
            "AT_idx": row_of[order].astype(np.int32), "AT_val": val[order], "c": f(c), "l": f(l), "u": f(u),
            "AL": f(AL), "AU": f(AU), "obj": 0.0}
