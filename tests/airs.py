"""AIRs as data for GenericStark (the zs_air / zo_air struct: include/zs_hip.h,
oracle/zs_oracle.h), with the reference's test AIRs and their traces:

  * FibSeededAir: FibAir (src/m31/stark.zig:877-931) with seeded boundary
    assertions (0,0) = a0, (1,0) = a1, (0,n-1) = claimed;
  * RangeCheckAir (stark.zig:957-1040), AndTableAir (:1046-1128),
    MultiplicityAir (:1135-1212): one LogUp relation each, no transition
    constraints of their own, their generateTrace / generateTable.

A transition constraint is a list of terms (coeff, factors), a factor being
("cur", j), ("next", j) or ("x",); the constraint's value is the sum of
coeff * product(factors) -- evalTransition written as data."""
import ctypes

import numpy as np

P = (1 << 31) - 1
KIND = {"cur": 0, "next": 1, "x": 2}


class Factor(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("column", ctypes.c_uint32)]


class Term(ctypes.Structure):
    _fields_ = [("constraint", ctypes.c_uint32), ("coeff", ctypes.c_uint32 * 4), ("n_factors", ctypes.c_uint32),
                ("factor", Factor * 4)]


class Boundary(ctypes.Structure):
    _fields_ = [("column", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("step", ctypes.c_uint64),
                ("value", ctypes.c_uint32 * 4)]


U32P = ctypes.POINTER(ctypes.c_uint32)


class AirStruct(ctypes.Structure):
    _fields_ = [("num_columns", ctypes.c_uint32), ("num_preprocessed", ctypes.c_uint32),
                ("num_transition_constraints", ctypes.c_uint32), ("num_terms", ctypes.c_uint32),
                ("terms", ctypes.POINTER(Term)),
                ("degree_mul", ctypes.c_uint32), ("degree_add", ctypes.c_uint32),
                ("num_boundary", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("boundary", ctypes.POINTER(Boundary)),
                ("num_lookup_columns", ctypes.c_uint32), ("num_lookup_relations", ctypes.c_uint32),
                ("selector_column", U32P), ("multiplicity_column", U32P),
                ("key_start", U32P), ("key_columns", U32P), ("table_start", U32P), ("table_columns", U32P)]


def qm(x):
    """An int (M31) or 4 words as QM31 words."""
    if isinstance(x, (int, np.integer)):
        return [int(x) % P, 0, 0, 0]
    return [int(v) for v in x]


class Air:
    """An AIR description; .struct() gives the C struct (kept alive by this object)."""

    def __init__(self, num_columns, constraints=(), boundary=(), num_preprocessed=0, lookups=(),
                 num_lookup_columns=0, degree=(0, 0)):
        self.num_columns = num_columns
        self.constraints = [list(c) for c in constraints]
        self.boundary = list(boundary)  # (column, step, value)
        self.num_preprocessed = num_preprocessed
        self.lookups = list(lookups)  # (selector, multiplicity, key columns, table columns)
        self.num_lookup_columns = num_lookup_columns
        self.degree = degree

    def struct(self):
        terms = [(k, qm(c), f) for k, cons in enumerate(self.constraints) for c, f in cons]
        self._terms = (Term * max(1, len(terms)))()
        for i, (k, c, f) in enumerate(terms):
            t = self._terms[i]
            t.constraint = k
            t.coeff[:] = c
            t.n_factors = len(f)
            for j, fac in enumerate(f):
                t.factor[j].kind = KIND[fac[0]]
                t.factor[j].column = fac[1] if len(fac) > 1 else 0
        self._bound = (Boundary * max(1, len(self.boundary)))()
        for i, (col, step, val) in enumerate(self.boundary):
            self._bound[i].column, self._bound[i].step = col, step
            self._bound[i].value[:] = qm(val)

        def arr(xs):
            a = (ctypes.c_uint32 * max(1, len(xs)))(*xs)
            self._keep.append(a)
            return ctypes.cast(a, U32P)

        self._keep = []
        ks, ts, kc, tc = [0], [0], [], []
        for _, _, keys, tabs in self.lookups:
            kc += keys
            tc += tabs
            ks.append(len(kc))
            ts.append(len(tc))
        s = AirStruct()
        s.num_columns, s.num_preprocessed = self.num_columns, self.num_preprocessed
        s.num_transition_constraints, s.num_terms = len(self.constraints), len(terms)
        s.terms = ctypes.cast(self._terms, ctypes.POINTER(Term))
        s.degree_mul, s.degree_add = self.degree
        s.num_boundary = len(self.boundary)
        s.boundary = ctypes.cast(self._bound, ctypes.POINTER(Boundary))
        s.num_lookup_columns, s.num_lookup_relations = self.num_lookup_columns, len(self.lookups)
        s.selector_column = arr([l[0] for l in self.lookups])
        s.multiplicity_column = arr([l[1] for l in self.lookups])
        s.key_start, s.key_columns = arr(ks), arr(kc)
        s.table_start, s.table_columns = arr(ts), arr(tc)
        self._struct = s
        return s


def fib_seeded_air(n, a0, a1, claimed):
    """FibAir.evalTransition (stark.zig:886-891): next[0] - cur[1], next[1] - cur[0] - cur[1]."""
    return Air(2, constraints=[[(1, [("next", 0)]), (P - 1, [("cur", 1)])],
                               [(1, [("next", 1)]), (P - 1, [("cur", 0)]), (P - 1, [("cur", 1)])]],
               boundary=[(0, 0, a0), (1, 0, a1), (0, n - 1, claimed)], degree=(1, 0))


def _col(vals):
    out = np.zeros((len(vals), 4), np.uint32)
    out[:, 0] = vals
    return out


def range_check_air():
    """[value, mult, selector], table [table]; key value against table (stark.zig:957-970)."""
    return Air(3, num_preprocessed=1, lookups=[(2, 1, [0], [0])], num_lookup_columns=1, degree=(1, 0))


def range_check_trace(n):
    """generateTrace / generateTable of stark.zig:990-1031."""
    h = n // 2
    value = [i % 4 if i < h else 0 for i in range(n)]
    mult = [1 if i < h else 0 for i in range(n)]
    sel = [1 if i < h else 0 for i in range(n)]
    table = [0 if i < h else i % 4 for i in range(n)]
    return np.stack([_col(value), _col(mult), _col(sel)]), np.stack([_col(table)])


def and_table_air():
    """[x, y, selector], table [ta, tb]; two-column key, multiplicity = the selector (stark.zig:1046-1059)."""
    return Air(3, num_preprocessed=2, lookups=[(2, 2, [0, 1], [0, 1])], num_lookup_columns=2, degree=(1, 0))


def and_table_trace(n):
    """stark.zig:1079-1117."""
    h = n // 2
    x = [((i % 4) & 1) if i < h else 0 for i in range(n)]
    y = [((i % 4) >> 1) if i < h else 0 for i in range(n)]
    sel = [1 if i < h else 0 for i in range(n)]
    ta = [0 if i < h else (i % 4) & 1 for i in range(n)]
    tb = [0 if i < h else (i % 4) >> 1 for i in range(n)]
    return np.stack([_col(x), _col(y), _col(sel)]), np.stack([_col(ta), _col(tb)])


def multiplicity_air():
    """[value, mult, selector], table [table] (stark.zig:1135-1148)."""
    return Air(3, num_preprocessed=1, lookups=[(2, 1, [0], [0])], num_lookup_columns=1, degree=(1, 0))


def multiplicity_trace(n):
    """stark.zig:1168-1203: rows 0..3 carry 0..3 with multiplicity 3."""
    value = [i if i < 4 else 0 for i in range(n)]
    mult = [3 if i < 4 else 0 for i in range(n)]
    sel = [1 if i < 4 else 0 for i in range(n)]
    table = [0 if i < 4 else i % 4 for i in range(n)]
    return np.stack([_col(value), _col(mult), _col(sel)]), np.stack([_col(table)])


LOOKUP_AIRS = {"range_check": (range_check_air, range_check_trace),
               "and_table": (and_table_air, and_table_trace),
               "multiplicity": (multiplicity_air, multiplicity_trace)}


def combined_air(n, a0, a1, claimed):
    """FibSeededAir's two columns and transitions with both lookup relations beside
    them: [a, b, value, mult, sel, x, y, sel2], tables [table, ta, tb]; relation 0 is
    RangeCheckAir's (value in table), relation 1 AndTableAir's ((x, y) in (ta, tb), the
    selector as multiplicity).  Exercises several relations over shared column
    challenges (alpha_cols), per-relation key and table offsets, and transition,
    boundary, lookup and binarity terms in one composition."""
    return Air(8, constraints=[[(1, [("next", 0)]), (P - 1, [("cur", 1)])],
                               [(1, [("next", 1)]), (P - 1, [("cur", 0)]), (P - 1, [("cur", 1)])]],
               boundary=[(0, 0, a0), (1, 0, a1), (0, n - 1, claimed)], num_preprocessed=3,
               lookups=[(4, 3, [2], [0]), (7, 7, [5, 6], [1, 2])], num_lookup_columns=2, degree=(1, 0))


def combined_trace(fib, n):
    """fib: a (2, n, 4) Fibonacci trace; the lookup halves of range_check_trace and
    and_table_trace beside it."""
    rt, rtab = range_check_trace(n)
    at, atab = and_table_trace(n)
    return np.concatenate([fib, rt, at]), np.concatenate([rtab, atab])
