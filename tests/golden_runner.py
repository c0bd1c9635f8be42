"""Runs the reference known-answer cases (tests/golden/reference_kats.json)
against a backend: the CPU oracle (oracle_ffi) or the HIP library (gp_ffi).
Both expose the same functions on host numpy arrays."""
import json
import math
import os

import numpy as np

import lp_synth
from helpers import stats as csr_stats

KATS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_kats.json")
TOL = 1e-12


def load_cases(kind=None):
    with open(KATS) as f:
        cases = json.load(f)["cases"]
    return [c for c in cases if kind is None or c["kind"] == kind]


def full_lp(spec):
    """An LP dict from a partial spec: the missing one of A / AT is the
    transpose of the other (empty when both are missing); absent vectors
    stay None (null pointers), as in the reference test."""
    m, n = spec["m"], spec["n"]
    A, AT = spec.get("A"), spec.get("AT")
    if A is None and AT is None:
        A = [[0] * (m + 1), [], []]
    if A is None:
        tp, ti, tv = [np.asarray(x) for x in AT]
        A = lp_synth.transpose_csr(n, m, tp.astype(np.int32), ti.astype(np.int32), tv.astype(np.float64))
    if AT is None:
        ap, ai, av = [np.asarray(x) for x in A]
        AT = lp_synth.transpose_csr(m, n, ap.astype(np.int32), ai.astype(np.int32), av.astype(np.float64))
    lp = {"m": m, "n": n, "obj": spec.get("obj", 0.0),
          "A_ptr": np.asarray(A[0], np.int32), "A_idx": np.asarray(A[1], np.int32),
          "A_val": np.asarray(A[2], np.float64),
          "AT_ptr": np.asarray(AT[0], np.int32), "AT_idx": np.asarray(AT[1], np.int32),
          "AT_val": np.asarray(AT[2], np.float64)}
    for k in ("c", "l", "u", "AL", "AU"):
        lp[k] = None if spec.get(k) is None else np.asarray(spec[k], np.float64)
    return lp


def full_plan(lp, spec):
    m, n = lp["m"], lp["n"]
    defaults = {"keep_row": np.ones(m, np.uint8), "keep_col": np.ones(n, np.uint8), "c": np.zeros(n),
                "l": np.full(n, -np.inf), "u": np.full(n, np.inf), "AL": np.full(m, -np.inf),
                "AU": np.full(m, np.inf)}
    plan = {}
    for k, d in defaults.items():
        plan[k] = np.asarray(spec[k], d.dtype).copy() if k in spec else d
    plan["obj_constant_delta"] = 0.0
    return plan


def _get(ctx, field):
    obj = ctx
    for part in field.split("."):
        obj = obj[part]
    return obj


def _close(got, want):
    if isinstance(want, float) and math.isinf(want):
        return math.isinf(got) and (got > 0) == (want > 0)
    return abs(float(got) - float(want)) < TOL


def check(ctx, asserts, where):
    for a in asserts:
        op = a[0]
        if op == "one_of":
            errors = []
            for alt in a[1:]:
                try:
                    check(ctx, alt, where)
                    break
                except AssertionError as e:
                    errors.append(str(e))
            else:
                raise AssertionError(f"{where}: no alternative holds: {errors}")
        elif op == "eq":
            got, want = _get(ctx, a[1]), a[2]
            if np.isscalar(got) or isinstance(got, (int, float)):
                assert got == want, f"{where}: {a[1]} = {got!r}, expected {want!r}"
            else:
                got = np.asarray(got)
                assert len(got) == len(want) and all(_close(g, w) for g, w in zip(got, want)), \
                    f"{where}: {a[1]} = {got.tolist()}, expected {want}"
        elif op == "approx":
            got = _get(ctx, a[1])
            got = got if a[2] is None else got[a[2]]
            assert _close(got, a[3]), f"{where}: {a[1]}[{a[2]}] = {got!r}, expected {a[3]!r}"
        elif op == "bits":
            got = float(_get(ctx, a[1])[a[2]])
            assert np.float64(got).tobytes() == np.float64(a[3]).tobytes(), \
                f"{where}: {a[1]}[{a[2]}] = {got!r} is not bit-identical to {a[3]!r}"
        elif op == "len":
            got = _get(ctx, a[1])
            assert len(got) == a[2], f"{where}: len({a[1]}) = {len(got)}, expected {a[2]}"
        elif op == "flag":
            got = int(ctx["flags"][a[1]])
            assert got == a[2], f"{where}: flag {a[1]} = {got}, expected {a[2]}"
        elif op == "le":
            x, y = _get(ctx, a[1])[a[2]], _get(ctx, a[3])[a[4]]
            assert x <= y, f"{where}: {a[1]}[{a[2]}] = {x} > {a[3]}[{a[4]}] = {y}"
        elif op == "index_at":
            got = int(_get(ctx, a[1])[_get(ctx, a[2])[a[3]]])
            assert got == a[4], f"{where}: {a[1]}[{a[2]}[{a[3]}]] = {got}, expected {a[4]}"
        elif op == "gt":
            got = _get(ctx, a[1])
            assert got > a[2], f"{where}: {a[1]} = {got}, expected > {a[2]}"
        else:
            raise ValueError(f"unknown assertion {op}")


def run_case(backend, c):
    where = f"{c['name']} ({c['source']})"
    kind = c["kind"]
    if kind == "activity":
        A = c["A"]
        mn, mx, ni, pi = backend.row_activity(c["m"], c["n"], np.asarray(A[0], np.int32), np.asarray(A[1], np.int32),
                                              np.asarray(A[2]), np.asarray(c["l"]), np.asarray(c["u"]), c["zero_tol"])
        check({"min_fin": mn, "max_fin": mx, "neg_inf": ni, "pos_inf": pi}, c["expect"], where)
    elif kind == "col_max_abs":
        AT = c["AT"]
        out = backend.col_max_abs(c["n"], c["m"], np.asarray(AT[0], np.int32), np.asarray(AT[1], np.int32),
                                  np.asarray(AT[2]))
        check({"col_max_abs": out}, c["expect"], where)
    elif kind == "rule":
        lp = full_lp(c["lp"])
        plan = full_plan(lp, c.get("plan", {}))
        st = csr_stats(lp)
        for k, v in c.get("stats", {}).items():
            st[k] = np.asarray(v, st[k].dtype)
        for i, step in enumerate(c["steps"]):
            for k, v in step.get("set", {}).items():
                plan[k] = v
            cand = step.get("candidate")
            flags, tape, new_A = backend.rule_apply(step["rule"], plan, lp, st, backend.params(**step["params"]),
                                                    None if cand is None else np.asarray(cand, np.uint8))
            for k in ("has_change", "has_row_action", "has_col_action", "has_infeasible", "has_unbounded"):
                plan[k] = flags[k]
            plan["obj_constant_delta"] = flags["obj_constant_delta"]
            ctx = {"plan": plan, "flags": flags, "tape": tape, "obj": flags["obj_constant_delta"]}
            check(ctx, step["expect"], f"{where} step {i}")
    elif kind == "presolve":
        lp = full_lp(c["lp"])
        summary, rec, red = backend.presolve(lp, backend.params(**c["params"]), True)
        check({"summary": summary, "record": rec, "reduced": red}, c["expect"], where)
        ps = c.get("postsolve")
        if ps:
            x, y, z = backend.postsolve(rec, ps["x_red"], ps["y_red"], ps["z_red"], lp if ps["original"] else None)
            check({"x": x, "y": y, "z": z}, ps["expect"], where + " postsolve")
    elif kind == "postsolve":
        r = dict(c["record"])
        for k, dt in (("row_red2org", np.int32), ("col_red2org", np.int32), ("fixed_idx", np.int32),
                      ("fixed_val", np.float64), ("types", np.int32), ("index_starts", np.int32),
                      ("value_starts", np.int32), ("dual_modes", np.uint8), ("indices", np.int32),
                      ("vals", np.float64)):
            r[k] = np.asarray(r[k], dt)
        orig = full_lp(c["original"]) if c.get("original") else None
        x, y, z = backend.postsolve(r, c["x_red"], c["y_red"], c["z_red"], orig)
        check({"x": x, "y": y, "z": z}, c["expect"], where)
    elif kind == "progress":
        for before, after, ratio, want in c["checks"]:
            got = backend.has_good_nnz_progress(before, after, ratio)
            assert got == want, f"{where}: progress({before}, {after}, {ratio}) = {got}"
    else:
        raise ValueError(kind)
