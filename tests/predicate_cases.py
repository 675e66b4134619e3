"""Random general-predicate programs over tiny dimension tables, and a second
way to evaluate them: as set algebra over row-index sets (left to right,
AND = intersection, OR = union, negation = complement), with hop terms
resolved through a Python dict of the parent's codes. The oracle evaluates
row by row with booleans; the two formulations share no code."""
import numpy as np

INT8_NULL = 127


def random_program(rng, rows, parent=None, nterms=None, allow_hop=True):
    """(terms, columns) with terms in the DimProgram.of form over numpy
    columns; parent: an int8 or int32 dimension vector a hop can reach."""
    nterms = nterms or int(rng.integers(1, 9))
    terms = []
    for i in range(nterms):
        kinds = ["range", "in"] + (["hop"] if allow_hop and parent is not None else [])
        kind = kinds[int(rng.integers(0, len(kinds)))]
        prefix = []
        if i > 0:
            prefix.append("or" if rng.integers(0, 2) else "and")
        if rng.integers(0, 4) == 0:
            prefix.append("not")
        if kind == "range":
            col = rng.integers(-20, 120, rows, dtype=np.int32)
            lo = int(rng.integers(-30, 100))
            hi = lo + int(rng.integers(0, 80))
            terms.append(tuple(prefix + ["range", col, lo, hi]))
        elif kind == "in":
            col = rng.integers(0, 50, rows, dtype=np.int32)
            vals = rng.integers(-5, 55, int(rng.integers(1, 12)), dtype=np.int32)
            terms.append(tuple(prefix + ["in", col, vals]))
        else:
            bias = int(rng.integers(-1, 2))
            col = rng.integers(-2, parent.size + 2, rows, dtype=np.int32)
            terms.append(tuple(prefix + ["hop", col, parent, bias]))
    return terms


def parent_code(parent, k):
    """The parent's code at key k, None when NULL or outside."""
    if k < 0 or k >= parent.size:
        return None
    c = int(parent[k])
    if parent.dtype == np.int8 and c == INT8_NULL:
        return None
    if parent.dtype == np.int32 and c == -1:
        return None
    return c


def set_eval(terms, rows, group="none", group_col=None, group_mod=0, group_term=0, code_bytes=1):
    everything = set(range(rows))
    acc = None
    hop_codes = {}
    for i, t in enumerate(terms):
        t = list(t)
        comb, neg = "and", False
        while t[0] in ("and", "or", "not"):
            w = t.pop(0)
            if w == "not":
                neg = True
            else:
                comb = w
        if t[0] == "range":
            col, lo, hi = t[1], t[2], t[3]
            s = {r for r in range(rows) if lo <= int(col[r]) < hi}
        elif t[0] == "in":
            vals = set(int(v) for v in t[2])
            s = {r for r in range(rows) if int(t[1][r]) in vals}
        else:
            col, par, bias = t[1], t[2], t[3]
            codes = {r: parent_code(par, int(col[r]) + bias) for r in range(rows)}
            s = {r for r, c in codes.items() if c is not None}
            if i == group_term:
                hop_codes = codes
        if neg:
            s = everything - s
        acc = s if acc is None else (acc | s if comb == "or" else acc & s)
    null = INT8_NULL if code_bytes == 1 else -1
    out = np.full(rows, null, dtype=np.int8 if code_bytes == 1 else np.int32)
    for r in sorted(acc):
        if group == "none":
            code = 0
        elif group == "column":
            g = int(group_col[r])
            code = (g % (1 << 32)) % group_mod if group_mod > 0 else g
        else:
            code = hop_codes.get(r)
            if code is None:
                continue
        if code_bytes == 1 and not (0 <= code < INT8_NULL):
            continue
        out[r] = code
    return out
