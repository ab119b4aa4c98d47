"""Pure-Python reference of the aggregation and join semantics -- TEST INFRASTRUCTURE ONLY.

An independent second statement of what the C oracle (cpu_ref.cpp) computes, written the way the
reference's own kernel tests build their expected values (a map per key tuple, exact integers,
sequential double folds: kernels/src/test/java/io/vecruntime/kernels/GroupedAggregationTest.java:84-192,
201-244), for inputs small enough for Python loops.  It shares no code with the C oracle: group
identity is a dict over key tuples, decimal averages use fractions.Fraction, wide sums use Python
ints.  Values are Python objects: int / float / str / None (null); decimals as unscaled ints.
"""
import math
import struct
from fractions import Fraction

INT64_MIN, INT64_MAX = -(1 << 63), (1 << 63) - 1
INT128_MIN, INT128_MAX = -(1 << 127), (1 << 127) - 1

COUNT_STAR, COUNT, SUM_LONG, SUM_DOUBLE, SUM_WIDE, AVG_DOUBLE, AVG_WIDE, MIN, MAX = range(9)
CHECKED, RESULT = 1, 2


class ArithmeticOverflow(Exception):
    pass


class DecimalOverflow(Exception):
    pass


def key_identity(v, is_double):
    """Grouping identity: None stays None (its own group); doubles by raw bits
    (GroupKeyTable.java:2350-2354)."""
    if v is None:
        return None
    if is_double:
        return ("d", struct.pack("<d", v))
    return v


def nan_safe_compare(x, y):
    """CompareOp.nanSafeCompare (CompareOp.java:68-82): NaN == NaN, NaN greatest, -0.0 == 0.0."""
    xn, yn = math.isnan(x), math.isnan(y)
    if (xn and yn) or x == y:
        return 0
    if xn:
        return 1
    if yn:
        return -1
    return 1 if x > y else -1


def wrap64(v):
    v &= (1 << 64) - 1
    return v - (1 << 64) if v >= (1 << 63) else v


def round_half_up(fr, scale):
    """Fraction -> integer at `scale` decimal places, ties away from zero (java.math HALF_UP)."""
    scaled = fr * (10 ** scale)
    neg = scaled < 0
    mag = -scaled if neg else scaled
    q, r = divmod(mag.numerator, mag.denominator)
    if 2 * r >= mag.denominator:
        q += 1
    return -q if neg else q


class _State:
    def __init__(self):
        self.rows = 0
        self.count = 0
        self.lsum = 0
        self.exact = 0
        self.sticky = False
        self.dsum = 0.0
        self.best = None


def aggregate(batches, key_is_double, aggs):
    """Folds `batches` and returns (group key tuples in first-seen order, per-group output rows).

    batches: list of dicts {"keys": [column lists], "values": [column lists], "selection": list or None}
    key_is_double: per key column, whether it is FLOAT64 (compared by raw bits)
    aggs: list of dicts {kind, input, flags, buffer_precision, buffer_scale, result_precision,
          result_scale, in_type} (in_type 'int32' | 'int64' | 'double' | 'decimal')
    Output per group: a list with one entry per aggregate, each a tuple of that aggregate's
    emitted column values (Spark's partial buffer, or its result under RESULT).
    """
    index = {}
    order = []
    states = []
    for b in batches:
        keys, values, sel = b["keys"], b["values"], b.get("selection")
        n = len(keys[0]) if keys else len(values[0])
        for i in range(n):
            if sel is not None and not sel[i]:
                continue
            ident = tuple(key_identity(k[i], key_is_double[c]) for c, k in enumerate(keys))
            g = index.get(ident)
            if g is None:
                g = len(order)
                index[ident] = g
                order.append(tuple(k[i] for k in keys))
                states.append([_State() for _ in aggs])
            for a, st in zip(aggs, states[g]):
                kind = a["kind"]
                st.rows += 1
                if kind == COUNT_STAR:
                    st.count += 1
                    continue
                x = values[a["input"]][i]
                if x is None:
                    continue
                st.count += 1
                if kind == SUM_LONG:
                    nxt = st.lsum + x
                    if a["flags"] & CHECKED and not (INT64_MIN <= nxt <= INT64_MAX):
                        raise ArithmeticOverflow("long overflow")
                    st.lsum = wrap64(nxt)
                elif kind in (SUM_DOUBLE, AVG_DOUBLE):
                    st.dsum = st.dsum + x
                elif kind in (SUM_WIDE, AVG_WIDE):
                    st.exact += x
                    if not (INT128_MIN <= st.exact <= INT128_MAX):
                        st.sticky = True
                elif kind in (MIN, MAX):
                    if st.best is None:
                        st.best = x
                    elif a["in_type"] == "double":
                        c = nan_safe_compare(x, st.best)
                        if (c < 0) if kind == MIN else (c > 0):
                            st.best = x
                    elif (x < st.best) if kind == MIN else (x > st.best):
                        st.best = x
    out = [[_emit(a, st) for a, st in zip(aggs, sts)] for sts in states]
    return order, out


def _emit(a, st):
    kind, flags = a["kind"], a["flags"]
    result, ansi = bool(flags & RESULT), bool(flags & CHECKED)
    if kind in (COUNT_STAR, COUNT):
        return (st.count,)
    if kind == SUM_LONG:
        return (st.lsum if st.count else None,)
    if kind == SUM_DOUBLE:
        return (st.dsum if st.count else None,)
    if kind == AVG_DOUBLE:
        if result:
            return (st.dsum / st.count if st.count else None,)
        return (st.dsum, st.count)
    lim = 10 ** a.get("buffer_precision", 38)
    over = st.sticky or abs(st.exact) >= lim
    if kind == SUM_WIDE:
        if result:
            if st.count == 0:
                return (None,)
            if over:
                if ansi:
                    raise DecimalOverflow("decimal sum past its precision")
                return (None,)
            return (st.exact,)
        if st.count == 0:
            return (0, True)
        return (None if over else st.exact, False)
    if kind == AVG_WIDE:
        if not result:
            if st.count == 0:
                return (0, 0)
            return (None if over else st.exact, st.count)
        if st.count == 0:
            return (None,)
        if over:
            if ansi:
                raise ArithmeticOverflow("decimal average over an overflowed sum")
            return (None,)
        q = round_half_up(Fraction(st.exact, 10 ** a["buffer_scale"]) / st.count, a["result_scale"])
        if abs(q) >= 10 ** a["result_precision"]:
            if ansi:
                raise DecimalOverflow("decimal average past its precision")
            return (None,)
        return (q,)
    if kind in (MIN, MAX):
        return (st.best,)
    raise ValueError(kind)


def join(build_keys, probe_keys, join_type, probe_selection=None, drop_null_keys=False):
    """Nested-loop join: (probe row, build row) pairs for inner -- streamed rows in order, each
    row's build rows in build order -- or the kept streamed rows for semi (1) / anti (2).
    Keys are lists of columns; a row with any null key never matches (VectorHashJoinExec.scala:487-489,
    1046-1054).  `drop_null_keys` is the null-aware anti join (:994,1070)."""
    nb = len(build_keys[0]) if build_keys else 0
    npr = len(probe_keys[0])
    bkey = [tuple(k[r] for k in build_keys) for r in range(nb)]
    out = []
    for i in range(npr):
        if probe_selection is not None and not probe_selection[i]:
            continue
        pk = tuple(k[i] for k in probe_keys)
        null_key = any(v is None for v in pk)
        matches = [] if null_key else [r for r in range(nb) if None not in bkey[r] and bkey[r] == pk]
        if join_type == 0:
            out += [(i, r) for r in matches]
        else:
            matched = bool(matches) or (drop_null_keys and nb > 0 and null_key)
            if matched == (join_type == 1):
                out.append(i)
    return out
