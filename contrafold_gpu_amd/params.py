"""CONTRAfold complementary-model parameters: raw weights -> scoring tables.

``load(path=None)`` returns the same dict as the reference's ``cpu.load``
(gpu_contrafold/cpu.py:40-135): FP64 tables ``bp stack tm hc dl dr hp_cum cs
i11 b01`` and the scalars ``mb mu mp eu ep``.

The mapping from the raw weights to the tables is linear (every table entry is
a sum of raw weights, or 0 when the file lacks the key).  It is built here as
an explicit list of (table entry, raw weight) incidences, so the same object
gives the tables (``J @ raw``) and the chain rule for gradients
(``J.T @ d/d tables``).

Flat table layout (2577 values, shared with the C ABI): bp[5][5]
stack[5][5][5][5] tm[5][5][5][5] hc[5][5] dl[5][5][5] dr[5][5][5] hp_cum[31]
cs[31][31] i11[5][5] b01[5] mb mu mp eu ep.
"""
import os

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_PARAMS = os.path.join(DATA_DIR, "contrafold.params.complementary")

BASES = "ACGU"
IDX = {"A": 0, "C": 1, "G": 2, "U": 3}
N_TABLE = 2577
TABLES = (("bp", (5, 5)), ("stack", (5, 5, 5, 5)), ("tm", (5, 5, 5, 5)), ("hc", (5, 5)), ("dl", (5, 5, 5)),
          ("dr", (5, 5, 5)), ("hp_cum", (31,)), ("cs", (31, 31)), ("i11", (5, 5)), ("b01", (5,)))
SCALARS = ("mb", "mu", "mp", "eu", "ep")
SCALAR_KEYS = {"mb": "multi_base", "mu": "multi_unpaired", "mp": "multi_paired", "eu": "external_unpaired",
               "ep": "external_paired"}

OFFSETS = {}
_o = 0
for _name, _shape in TABLES:
    OFFSETS[_name] = _o
    _o += int(np.prod(_shape))
for _name in SCALARS:
    OFFSETS[_name] = _o
    _o += 1
assert _o == N_TABLE

# loop-length caps of the complementary model (cpu.py:31-37)
MAX_SINGLE, D_HP, D_BULGE, D_INT, D_ISYM, D_IASYM, D_IEXP = 30, 30, 30, 30, 15, 28, 4


def read_raw(path=None):
    """Ordered (name, value) pairs of a ``name value`` parameter file."""
    names, values = [], []
    with open(path or DEFAULT_PARAMS) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) == 2:
                names.append(parts[0])
                values.append(float(parts[1]))
    return names, np.asarray(values, np.float64)


class ParamMap:
    """Incidence form of the raw -> table map for one list of raw names."""

    def __init__(self, names):
        self.names = list(names)
        self.index = {n: k for k, n in enumerate(self.names)}
        # later duplicates in a file win, as dict assignment does in the reference
        rows, cols = [], []
        missing_scalars = [SCALAR_KEYS[s] for s in SCALARS if SCALAR_KEYS[s] not in self.index]
        if missing_scalars:
            raise KeyError(missing_scalars[0])

        def put(entry, *keys):
            for key in keys:
                if key in self.index:
                    rows.append(entry)
                    cols.append(self.index[key])

        def first(*keys):
            for key in keys:
                if key in self.index:
                    return key
            return None

        def sym_pair(table, prefix):
            # the reference writes [a][b] and [b][a] from key 'ab' (or 'ba');
            # the later loop visit (larger first index) decides the value
            for x in range(4):
                for y in range(4):
                    hi, lo = (x, y) if x >= y else (y, x)
                    key = first(f"{prefix}{BASES[hi]}{BASES[lo]}", f"{prefix}{BASES[lo]}{BASES[hi]}")
                    if key is not None:
                        put(OFFSETS[table] + x * 5 + y, key)

        sym_pair("bp", "base_pair_")
        sym_pair("i11", "internal_1x1_nucleotides_")
        for a in range(4):
            for b in range(4):
                put(OFFSETS["hc"] + a * 5 + b, f"helix_closing_{BASES[a]}{BASES[b]}")
                for c in range(4):
                    put(OFFSETS["dl"] + (a * 5 + b) * 5 + c, f"dangle_left_{BASES[a]}{BASES[b]}{BASES[c]}")
                    put(OFFSETS["dr"] + (a * 5 + b) * 5 + c, f"dangle_right_{BASES[a]}{BASES[b]}{BASES[c]}")
                    for d in range(4):
                        q = ((a * 5 + b) * 5 + c) * 5 + d
                        word = BASES[a] + BASES[b] + BASES[c] + BASES[d]
                        key = first("helix_stacking_" + word, "helix_stacking_" + word[::-1])
                        if key is not None:
                            put(OFFSETS["stack"] + q, key)
                        put(OFFSETS["tm"] + q, "terminal_mismatch_" + word)
        for a in range(4):
            put(OFFSETS["b01"] + a, f"bulge_0x1_nucleotides_{BASES[a]}")

        def cum(prefix, k, start):
            """keys whose sum is the cumulative ``prefix_at_least`` table at k"""
            return [f"{prefix}_at_least_{m}" for m in range(start, k + 1)]

        for k in range(D_HP + 1):
            put(OFFSETS["hp_cum"] + k, *cum("hairpin_length", k, 0))
        for l1 in range(MAX_SINGLE + 1):
            for l2 in range(MAX_SINGLE + 1 - l1):
                if l1 == 0 and l2 == 0:
                    continue
                e = OFFSETS["cs"] + l1 * 31 + l2
                if l1 == 0 or l2 == 0:
                    put(e, *cum("bulge_length", min(D_BULGE, l1 + l2), 1))
                    continue
                if l1 <= D_IEXP and l2 <= D_IEXP:
                    put(e, f"internal_explicit_{min(l1, l2)}_{max(l1, l2)}")
                put(e, *cum("internal_length", min(D_INT, l1 + l2), 2))
                if l1 == l2:
                    put(e, *cum("internal_symmetric_length", min(D_ISYM, l1), 1))
                put(e, *cum("internal_asymmetry", min(D_IASYM, abs(l1 - l2)), 1))
        for s in SCALARS:
            put(OFFSETS[s], SCALAR_KEYS[s])
        self.rows = np.asarray(rows, np.int64)
        self.cols = np.asarray(cols, np.int64)
        # entries outside hp_cum and cs are copies of exactly one raw weight
        lo_hp, hi_cs = OFFSETS["hp_cum"], OFFSETS["cs"] + 31 * 31
        keep = (self.rows < lo_hp) | (self.rows >= hi_cs)
        self._single = np.stack([self.rows[keep], self.cols[keep]], axis=1)

    @property
    def n_raw(self):
        return len(self.names)

    def tables(self, raw):
        """Flat FP64 table vector (2577) of a raw weight vector.

        Entries that are sums are formed in the reference's order (per-family
        running sums first, then the internal-loop composition), so the FP64
        tables -- and their FP32 casts -- are the reference's to the bit."""
        raw = np.asarray(raw, np.float64)
        out = np.zeros(N_TABLE, np.float64)
        single = self._single
        out[single[:, 0]] = raw[single[:, 1]]

        def running(prefix, dmax, start):
            al = np.zeros(dmax + 1)
            for k in range(start, dmax + 1):
                j = self.index.get(f"{prefix}_at_least_{k}")
                if j is not None:
                    al[k] = raw[j]
            return np.cumsum(al)

        hp = running("hairpin_length", D_HP, 0)
        bulge = running("bulge_length", D_BULGE, 1)
        il = running("internal_length", D_INT, 2)
        isym = running("internal_symmetric_length", D_ISYM, 1)
        iasym = running("internal_asymmetry", D_IASYM, 1)
        out[OFFSETS["hp_cum"]:OFFSETS["hp_cum"] + D_HP + 1] = hp
        for l1 in range(MAX_SINGLE + 1):
            for l2 in range(MAX_SINGLE + 1 - l1):
                if l1 == 0 and l2 == 0:
                    continue
                if l1 == 0 or l2 == 0:
                    v = bulge[min(D_BULGE, l1 + l2)]
                else:
                    v = 0.0
                    if l1 <= D_IEXP and l2 <= D_IEXP:
                        j = self.index.get(f"internal_explicit_{min(l1, l2)}_{max(l1, l2)}")
                        v += raw[j] if j is not None else 0.0
                    v += il[min(D_INT, l1 + l2)]
                    if l1 == l2:
                        v += isym[min(D_ISYM, l1)]
                    v += iasym[min(D_IASYM, abs(l1 - l2))]
                out[OFFSETS["cs"] + l1 * 31 + l2] = v
        return out

    def raw_grad(self, table_grad):
        """Chain rule: d/d raw = J^T d/d tables."""
        g = np.asarray(table_grad, np.float64)
        out = np.zeros(self.n_raw, np.float64)
        np.add.at(out, self.cols, g[self.rows])
        return out

    def dense(self):
        J = np.zeros((N_TABLE, self.n_raw), np.float64)
        np.add.at(J, (self.rows, self.cols), 1.0)
        return J


def unflatten(flat):
    """Flat vector -> dict in the layout of cpu.load."""
    flat = np.asarray(flat, np.float64)
    P = {}
    for name, shape in TABLES:
        o = OFFSETS[name]
        P[name] = flat[o:o + int(np.prod(shape))].reshape(shape).copy()
    for name in SCALARS:
        P[name] = float(flat[OFFSETS[name]])
    return P


def flatten(P):
    """Dict in the layout of cpu.load -> flat FP64 vector (2577)."""
    parts = [np.asarray(P[name], np.float64).reshape(-1) for name, _ in TABLES]
    parts.append(np.asarray([float(P[name]) for name in SCALARS], np.float64))
    v = np.ascontiguousarray(np.concatenate(parts))
    if v.size != N_TABLE:
        raise ValueError(f"parameter tables have {v.size} values, expected {N_TABLE}")
    return v


_MAPS = {}


def param_map(path=None):
    path = os.path.abspath(path or DEFAULT_PARAMS)
    m = _MAPS.get(path)
    if m is None:
        names, _ = read_raw(path)
        m = _MAPS[path] = ParamMap(names)
    return m


def load(path=None):
    """Drop-in for cpu.load (cpu.py:40-135): bundled complementary parameters."""
    names, values = read_raw(path)
    return unflatten(ParamMap(names).tables(values))


def load_raw(path=None):
    """(names, FP64 raw vector, ParamMap) for training / differentiation."""
    names, values = read_raw(path)
    return names, values, param_map(path)


def write_raw(path, names, values):
    """Write a raw vector back in the ``name value`` format, one weight per
    line in the given order.  Values are written with Python's shortest
    round-trip representation, so reading the file back gives the same FP64
    weights bit for bit (trained weights survive a save and a load)."""
    names = list(names)
    values = np.asarray(values, np.float64).reshape(-1)
    if len(names) != values.size:
        raise ValueError(f"{len(names)} names for {values.size} values")
    if not np.isfinite(values).all():
        raise ValueError("non-finite weight")
    with open(path, "w") as fh:
        for n, v in zip(names, values.tolist()):
            fh.write(f"{n} {v!r}\n")
