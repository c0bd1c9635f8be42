"""The continuous-MPS reader (host only).

* tests/cpp/mps_reader_test: the reference reader's own known answers
  (cpp/mps_reader/tests/*.cpp), compiled against include/mps_reader/.
* The reference's reader itself, built from its sources into
  oracle/_ref/libmps_ref.so (where the reference tree is present), reads
  the same texts as the library's reader: a few hundred seeded random MPS
  files in free and fixed format (comments, several RHS / RANGES / BOUNDS
  sets, every bound type, repeated entries, numbers written in many forms)
  and a set of malformed ones.  Both must accept or reject each text, and
  accepted models must agree bit for bit in every array.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import gp_ffi
import gp_types as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP_TEST = os.path.join(ROOT, "tests", "cpp", "build", "mps_reader_test")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libmps_ref.so")


def test_reader_known_answers_in_cpp():
    if not os.path.exists(CPP_TEST):
        subprocess.run(["make", "-C", os.path.dirname(os.path.dirname(CPP_TEST))], check=True,
                       stdout=subprocess.DEVNULL)
    r = subprocess.run([CPP_TEST], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok ") == 15, r.stdout


def test_view_through_the_c_abi():
    m = gp_ffi.MpsModel(
        "NAME toy\nROWS\n N OBJ\n G DEMAND\nCOLUMNS\n    X OBJ 1 DEMAND 1\n    Y OBJ 2 DEMAND 1\n"
        "RHS\n    RHS1 DEMAND 3\n    RHS1 OBJ 1.5\nBOUNDS\n UP BND1 X 2\n UP BND1 Y 4\nENDATA\n")
    lp = m.lp
    assert lp["name"] == "toy" and lp["m"] == 1 and lp["n"] == 2
    assert np.array_equal(lp["A_ptr"], [0, 2]) and np.array_equal(lp["A_idx"], [0, 1])
    assert np.array_equal(lp["AT_ptr"], [0, 1, 2]) and np.array_equal(lp["AT_idx"], [0, 0])
    assert np.array_equal(lp["c"], [1.0, 2.0]) and lp["obj"] == -1.5
    assert np.array_equal(lp["AL"], [3.0]) and lp["AU"][0] == np.inf
    assert np.array_equal(lp["l"], [0.0, 0.0]) and np.array_equal(lp["u"], [2.0, 4.0])
    with pytest.raises(gp_ffi.GpError, match="OBJSENSE MAX"):
        gp_ffi.MpsModel("NAME x\nOBJSENSE MAX\nROWS\n N OBJ\nCOLUMNS\n    X OBJ 1\nENDATA\n")


# ---- differential reading against the reference reader ----------------------
class _Ref:
    def __init__(self, path):
        self.L = C.CDLL(path)
        self.L.ref_mps_parse.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]
        self.L.ref_mps_get_view.argtypes = [C.c_void_p, C.POINTER(T.GpMpsView)]
        self.L.ref_mps_get_view.restype = None
        self.L.ref_mps_free.argtypes = [C.c_void_p]
        self.L.ref_mps_free.restype = None

    def read(self, text):
        data = text.encode()
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        if self.L.ref_mps_parse(data, len(data), C.byref(h), err, 512) != 0:
            return None, err.value.decode()
        v = T.GpMpsView()
        self.L.ref_mps_get_view(h, C.byref(v))
        out = dict(T.mps_view_arrays(v), name=v.name.decode(), m=v.rows, n=v.cols, obj=v.obj_constant)
        self.L.ref_mps_free(h)
        return out, None


def _ours(text):
    try:
        return gp_ffi.MpsModel(text).lp, None
    except gp_ffi.GpError as e:
        return None, str(e)


def _number(rng, magnitude=None):
    """A number token in one of the forms MPS files use."""
    x = float(rng.choice([0.0, 1.0, 2.5, -3.0, 1e-6, 0.000001, 123456.789, 1e5, 7.0 / 3.0, -0.125, 1e-9, 4e7]))
    if magnitude is not None:
        x = magnitude
    else:
        x = x * float(rng.choice([1.0, -1.0, 10.0, 0.01])) + float(rng.integers(-5, 6)) * float(rng.random() < 0.5)
    form = int(rng.integers(0, 11))
    if form == 0:
        return repr(x)
    if form == 1:
        return f"{x:.17g}"
    if form == 2:
        return f"{x:.3f}"
    if form == 3:
        return f"{x:.6e}".replace("e", "D")
    if form == 4:
        return f"{x:e}"
    if form == 5:
        return f"{x:.20f}"[:24]  # many digits: the strtod path
    if form == 6:
        return ("+" if x >= 0 else "") + f"{x:g}"
    if form == 7:
        s = f"{abs(x):.4f}".lstrip("0") or "0"
        return ("-" if x < 0 else "") + s  # ".5"
    if form == 8:
        return f"{int(round(x))}."  # "5."
    if form == 9:
        return f"{x:.15g}E0"
    return str(int(round(x * 1000)))


def _random_mps(rng, fixed=False):
    m = int(rng.integers(1, 7))
    n = int(rng.integers(1, 8))
    rows = [f"R{i}" for i in range(m)]
    types = [str(rng.choice(list("EGL"))) for _ in range(m)]
    cols = [f"C{j}" for j in range(n)]
    lines = ["NAME          rnd" + str(int(rng.integers(0, 99)))]
    if rng.random() < 0.2:
        lines += ["OBJSENSE", "    MIN"] if rng.random() < 0.5 else ["OBJSENSE MIN"]
    lines.append("ROWS")
    lines.append(" N  OBJ")
    if rng.random() < 0.3:
        lines.append(" N  FREEROW")
    for r, t in zip(rows, types):
        lines.append(f" {t}  {r}")
    lines.append("COLUMNS")
    names = rows + ["OBJ"] + (["FREEROW"] if "FREEROW" in " ".join(lines) else [])
    for c in cols:
        k = int(rng.integers(1, 5))
        picks = [str(rng.choice(names)) for _ in range(k)]
        if rng.random() < 0.2 and picks:
            picks.append(picks[0])  # a repeated entry, summed
        for a in range(0, len(picks), 2):
            pair = picks[a:a + 2]
            if fixed:
                s = f"    {c:<8}  {pair[0]:<8}  {_number(rng):>12}"
                if len(pair) == 2:
                    s += f"   {pair[1]:<8}  {_number(rng):>12}"
            else:
                s = f"    {c} " + " ".join(f"{p} {_number(rng)}" for p in pair)
            lines.append(s)
        if rng.random() < 0.15:
            lines.append("* a comment inside COLUMNS")
    for sec, setname in (("RHS", "RHS1"), ("RANGES", "RNG1")):
        if sec == "RANGES" and rng.random() < 0.5:
            continue
        lines.append(sec)
        for sname in ([setname, setname + "B"] if rng.random() < 0.3 else [setname]):
            for r in rows + (["OBJ"] if sec == "RHS" and rng.random() < 0.4 else []):
                if rng.random() < 0.7:
                    lines.append(f"    {sname} {r} {_number(rng)}")
    if rng.random() < 0.8:
        lines.append("BOUNDS")
        for bname in (["BND1", "BND2"] if rng.random() < 0.3 else ["BND1"]):
            for c in cols:
                if rng.random() < 0.6:
                    t = str(rng.choice(["LO", "UP", "FX", "FR", "MI", "PL", "UP", "LO"]))
                    if t in ("FR", "MI", "PL"):
                        lines.append(f" {t} {bname} {c}")
                    else:
                        v = _number(rng)
                        if t == "UP" and rng.random() < 0.2:
                            v = str(rng.choice(["inf", "+Infinity", "-2", "INF"]))
                        if t == "LO" and rng.random() < 0.1:
                            v = "-inf"
                        lines.append(f" {t} {bname} {c} {v}")
    lines.append("ENDATA")
    return "\n".join(lines) + "\n"


MALFORMED = [
    "",
    "NAME x\nROWS\n N OBJ\n X BAD\nENDATA\n",
    "NAME x\nROWS\n N OBJ\n L R\nCOLUMNS\n    X NOROW 1\nENDATA\n",
    "NAME x\nROWS\n N OBJ\n L R\nCOLUMNS\n    X R 1.2.3\nENDATA\n",
    "NAME x\nROWS\n N OBJ\n L R\nCOLUMNS\n    X R 1\nRHS\n    RHS1 NOROW 1\nENDATA\n",
    "NAME x\nROWS\n N OBJ\n L R\nCOLUMNS\n    X R 1\nBOUNDS\n UP BND1 NOCOL 1\nENDATA\n",
    "NAME x\nROWS\n N OBJ\n L R\nCOLUMNS\n    X R 1\nBOUNDS\n XX BND1 X 1\nENDATA\n",
    "NAME x\nROWS\n N OBJ\n L R\nCOLUMNS\n    X R 1\nBOUNDS\n LO BND1 X 5\n UP BND1 X 1\nENDATA\n",
    "NAME x\nROWS\n N OBJ\n L R\nCOLUMNS\n    X R 1\nQUADOBJ\n    X X 1\nENDATA\n",
    "NAME x\nFOO\nROWS\n N OBJ\nENDATA\n",
    "  X R 1\nNAME x\nROWS\n N OBJ\nENDATA\n",
    "NAME x\nROWS\n N OBJ\n L R\nCOLUMNS\n    X R 1\nRANGES\n    RNG R 1 R\nENDATA\n",
    "NAME x\nROWS\n N OBJ\n L R\nCOLUMNS\n    X R 1e-400\nENDATA\n",
    "NAME x\nROWS\n N OBJ\n L R\nCOLUMNS\n    X R 1e5x\nENDATA\n",
]

# Read alike by both readers: strtod also takes hexadecimal numbers.
ODD_BUT_READ = ["NAME x\nROWS\n N OBJ\n L R\nCOLUMNS\n    X R 0x10\nENDATA\n"]

def _fixed(*fields):
    """A data line with its fields at the fixed MPS columns (2-3, 5-12,
    15-22, 25-36, 40-47, 50-61)."""
    starts = (1, 4, 14, 24, 39, 49)
    line = ""
    for start, f in zip(starts, fields):
        if f:
            line = line.ljust(start) + f
    return line


FIXED_ONLY = [
    # names with blanks are readable only in the fixed column layout
    "\n".join(["NAME          fix two", "ROWS", _fixed("N", "COST"), _fixed("L", "MY ROW"), "COLUMNS",
               _fixed("", "MY COL", "COST", "1.0", "MY ROW", "2.5"), "RHS", _fixed("", "RHS 1", "MY ROW", "3.0"),
               "BOUNDS", _fixed("UP", "BND 1", "MY COL", "4.0"), "ENDATA", ""]),
    "\n".join(["NAME          fix", "ROWS", _fixed("N", "COST"), _fixed("E", "R 1"), _fixed("G", "R 2"), "COLUMNS",
               _fixed("", "X 1", "COST", "1.0", "R 1", "1.0"), _fixed("", "X 1", "R 2", "-1.0"),
               "RHS", _fixed("", "RHS", "R 1", "2.0", "R 2", "-5.0"), "RANGES", _fixed("", "RNG", "R 1", "-1.5"),
               "ENDATA", ""]),
]


def _same(a, b):
    if (a is None) != (b is None):
        return False
    if a is None:
        return True
    for k in ("name", "m", "n"):
        if a[k] != b[k]:
            return False
    if np.float64(a["obj"]).tobytes() != np.float64(b["obj"]).tobytes():
        return False
    for k in ("A_ptr", "A_idx", "A_val", "AT_ptr", "AT_idx", "AT_val", "c", "l", "u", "AL", "AU"):
        if a[k].shape != b[k].shape or a[k].tobytes() != b[k].tobytes():
            return False
    return True


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("the reference reader is built only where the reference tree is present")
    return _Ref(REF_LIB)


def test_random_models_read_like_the_reference(ref):
    rng = np.random.default_rng(20240615)
    accepted = 0
    for k in range(400):
        text = _random_mps(rng, fixed=(k % 4 == 3))
        mine, e1 = _ours(text)
        theirs, e2 = ref.read(text)
        assert _same(mine, theirs), f"case {k}: ours {e1}, reference {e2}\n{text}"
        accepted += mine is not None
    assert accepted > 250


def test_malformed_and_fixed_only_models_read_like_the_reference(ref):
    for text in MALFORMED + FIXED_ONLY + ODD_BUT_READ:
        mine, e1 = _ours(text)
        theirs, e2 = ref.read(text)
        assert _same(mine, theirs), f"ours {e1}, reference {e2}\n{text}"
    for text in FIXED_ONLY:
        assert _ours(text)[0] is not None
    for text in MALFORMED:
        assert _ours(text)[0] is None, text
