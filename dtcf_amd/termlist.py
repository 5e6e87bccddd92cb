"""Front end for the UHF CCSD term lists (``_uhf_t1.eq``, ``_uhf_t2.eq``).

The files are the output of a tensor-expression generator: declarations

    range Oa, Ob = 10.0;                       # named index ranges
    index h1a, h2a = Oa;                       # index names over a range
    array vaa_oovv([Oa,Oa][Va,Va]:(0,1)(2,3)); # dims (upper, lower) + antisymmetric groups

and assignments, each followed by a ``# <flops>\\t<formula>`` line:

    X[i..] = (A[..] * B[..])                   # binary contraction
    X[i..] = (A[..] + -B[..] + (0.5 * C[..]))  # signed, scaled permuted sum

closed by ``# TOTAL COST=<sum of flops>`` (``_uhf_t1.eq:246``,
``_uhf_t2.eq:1000``).  ``parse`` turns a file into declarations and
statements; ``Statement.gen_flops`` recomputes the generator's flop count from
the declarations alone and ``check_flops`` compares it with every
``# flops`` line and the total.

The generator's count (restated from the file's own formula lines, e.g.
``_uhf_t2.eq:10-12,515-520``):

* contraction ``C = A * B``: 2 x the product over all distinct indices of the
  range size, except that a pair of same-range indices that is antisymmetric
  in every array holding both of them (the two operands for a contracted
  pair; the result and the operand holding it for an external pair) counts
  as n(n+1)/2;
* sum of t terms: t x the element count of the result plus, for each term
  written with a coefficient ``(c * X[..])``, the element count of X; element
  counts take each array's antisymmetric pairs as n(n+1)/2;
* for an outer product the formula line shows an empty contracted factor
  (``10.0*100.0*100.0*10.0**2.0``), which the evaluator reads as ``*``.

``dense_flops`` is what the engine performs: 2 x the full product of the
extents (the tensors are held dense, no antisymmetry packing).
"""
from __future__ import annotations

import itertools
import math
import os
import re
from dataclasses import dataclass, field


@dataclass
class Array:
    name: str
    ranges: tuple          # range name per dimension (upper dims first)
    nupper: int
    groups: tuple          # antisymmetric groups: tuples of dimension positions

    def antisym_pairs(self, idx):
        """Unordered index-name pairs antisymmetric in this array under the
        given index labelling."""
        out = set()
        for g in self.groups:
            for p, q in itertools.combinations(g, 2):
                out.add(frozenset((idx[p], idx[q])))
        return out


@dataclass
class Ref:
    name: str
    idx: tuple
    coef: float = 1.0
    scaled: bool = False   # written as (c * X[..])


@dataclass
class Statement:
    lhs: Ref
    kind: str              # "mul" or "sum"
    terms: list
    line: int
    file_flops: float | None = None
    formula: str | None = None
    _tl: "TermList" = field(default=None, repr=False)

    # ---------------------------------------------------------- index sets
    @property
    def external(self):
        return list(self.lhs.idx)

    @property
    def contracted(self):
        if self.kind != "mul":
            return []
        a, b = self.terms
        return [x for x in a.idx if x in b.idx and x not in self.lhs.idx]

    def extent(self, name, sizes=None):
        return self._tl.index_size(name, sizes)

    def dense_flops(self, sizes=None):
        if self.kind == "mul":
            allidx = set(self.terms[0].idx) | set(self.terms[1].idx) | set(self.lhs.idx)
            return 2.0 * math.prod(self.extent(x, sizes) for x in allidx)
        n = math.prod(self.extent(x, sizes) for x in self.lhs.idx)
        return float(n) * (len(self.terms) + sum(1 for t in self.terms if t.scaled))

    def gen_flops(self, sizes=None):
        tl = self._tl
        cdecl = tl.arrays[self.lhs.name]
        cpairs = cdecl.antisym_pairs(self.lhs.idx)
        if self.kind == "sum":
            n = self._packed_count(list(self.lhs.idx), cpairs, sizes) * len(self.terms)
            for t in self.terms:
                if t.scaled:
                    n += self._packed_count(list(t.idx), tl.arrays[t.name].antisym_pairs(t.idx),
                                            sizes)
            return n
        a, b = self.terms
        apairs = tl.arrays[a.name].antisym_pairs(a.idx)
        bpairs = tl.arrays[b.name].antisym_pairs(b.idx)
        contr = set(self.contracted)
        packed = set()
        for p in apairs | bpairs | cpairs:
            x, y = tuple(p)
            if x in contr and y in contr:
                if p in apairs and p in bpairs:
                    packed.add(p)
            elif x not in contr and y not in contr:
                holders = [s for s, r in ((apairs, a), (bpairs, b)) if x in r.idx and y in r.idx]
                if p in cpairs and holders and all(p in s for s in holders):
                    packed.add(p)
        allidx = list(dict.fromkeys(list(a.idx) + list(b.idx)))
        return 2.0 * self._packed_count(allidx, packed, sizes)

    def _packed_count(self, idx, pairs, sizes):
        used = set()
        n = 1.0
        for p in pairs:
            x, y = tuple(p)
            if x in used or y in used or x not in idx or y not in idx:
                continue
            e = self.extent(x, sizes)
            n *= e * (e + 1) / 2.0
            used |= {x, y}
        for x in idx:
            if x not in used:
                n *= self.extent(x, sizes)
        return n


class TermList:
    def __init__(self):
        self.ranges = {}       # range name -> size (as declared)
        self.indices = {}      # index name -> range name
        self.arrays = {}       # name -> Array
        self.statements = []
        self.total_cost = None
        self.path = None

    def index_size(self, name, sizes=None):
        r = self.indices[name]
        return (sizes or self.ranges)[r]

    @property
    def contractions(self):
        return [s for s in self.statements if s.kind == "mul"]

    @property
    def sums(self):
        return [s for s in self.statements if s.kind == "sum"]

    def array_elements(self, name, sizes=None):
        return math.prod((sizes or self.ranges)[r] for r in self.arrays[name].ranges)

    def inputs(self):
        """Arrays read before (or without) being assigned: the integrals,
        Fock blocks and amplitudes."""
        assigned, out = set(), []
        for s in self.statements:
            for t in s.terms:
                if t.name not in assigned and t.name not in out:
                    out.append(t.name)
            assigned.add(s.lhs.name)
        return out

    def outputs(self):
        """Arrays assigned and never read afterwards (the residuals)."""
        read = {t.name for s in self.statements for t in s.terms}
        return [s.lhs.name for s in self.statements if s.lhs.name not in read]


# ---------------------------------------------------------------- parsing
_REF = r"([A-Za-z_][A-Za-z0-9_]*)\[([A-Za-z0-9_,\s]*)\]"
_RE_REF = re.compile(_REF)
_RE_MUL = re.compile(r"^\(\s*" + _REF + r"\s*\*\s*" + _REF + r"\s*\)$")
_RE_SCALED = re.compile(r"^\(\s*([-+]?[0-9.]+(?:[eE][-+]?\d+)?)\s*\*\s*" + _REF + r"\s*\)$")
_RE_ARRAY = re.compile(r"([A-Za-z_][A-Za-z0-9_]*)\(\[([^\]]*)\]\[([^\]]*)\]((?::?\([0-9,]+\))*)\)")


def _split_top(expr, sep="+"):
    """Splits at depth-0 separators."""
    out, depth, cur = [], 0, []
    for ch in expr:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == sep and depth == 0:
            out.append("".join(cur).strip())
            cur = []
        else:
            cur.append(ch)
    out.append("".join(cur).strip())
    return [x for x in out if x]


def _idx(s):
    return tuple(x.strip() for x in s.split(",") if x.strip())


def _parse_term(t, line):
    neg = False
    while t.startswith("-"):
        neg = not neg
        t = t[1:].strip()
    m = _RE_SCALED.match(t)
    if m:
        c = float(m.group(1))
        return Ref(m.group(2), _idx(m.group(3)), -c if neg else c, True)
    m = _RE_REF.fullmatch(t)
    if m:
        return Ref(m.group(1), _idx(m.group(2)), -1.0 if neg else 1.0, False)
    raise ValueError(f"line {line}: cannot parse term {t!r}")


def eval_formula(text):
    """Evaluates a flop formula line (numbers, * / + and parentheses)."""
    s = text.replace("**", "*").replace(" ", "")
    if not re.fullmatch(r"[0-9.eE+\-*/()]+", s):
        raise ValueError(f"unexpected formula {text!r}")
    pos = 0

    def expr():
        nonlocal pos
        v = term()
        while pos < len(s) and s[pos] in "+-":
            op = s[pos]
            pos += 1
            w = term()
            v = v + w if op == "+" else v - w
        return v

    def term():
        nonlocal pos
        v = atom()
        while pos < len(s) and s[pos] in "*/":
            op = s[pos]
            pos += 1
            w = atom()
            v = v * w if op == "*" else v / w
        return v

    def atom():
        nonlocal pos
        if s[pos] == "(":
            pos += 1
            v = expr()
            assert s[pos] == ")"
            pos += 1
            return v
        m = re.match(r"[0-9.]+(?:[eE][-+]?\d+)?", s[pos:])
        pos += m.end()
        return float(m.group(0))

    v = expr()
    assert pos == len(s), text
    return v


def parse(path_or_text) -> TermList:
    if os.path.exists(str(path_or_text)):
        with open(path_or_text) as f:
            text = f.read()
        path = str(path_or_text)
    else:
        text, path = str(path_or_text), None
    tl = TermList()
    tl.path = path
    last = None
    for ln, raw in enumerate(text.splitlines(), 1):
        line = raw.strip()
        if not line or line in "{}":
            continue
        if line.startswith("#"):
            body = line[1:].strip()
            if body.startswith("TOTAL COST="):
                tl.total_cost = float(body.split("=", 1)[1])
            elif last is not None and last.file_flops is None:
                num, _, formula = body.partition("\t")
                last.file_flops = float(num)
                last.formula = formula.strip()
            continue
        if line.startswith("range "):
            names, val = line[6:].rstrip(";").split("=")
            for n in names.split(","):
                tl.ranges[n.strip()] = int(float(val))
            continue
        if line.startswith("index "):
            names, rng = line[6:].rstrip(";").split("=")
            for n in names.split(","):
                tl.indices[n.strip()] = rng.strip()
            continue
        if line.startswith("array "):
            for m in _RE_ARRAY.finditer(line[6:]):
                up, lo = _idx(m.group(2)), _idx(m.group(3))
                groups = tuple(tuple(int(x) for x in g.split(","))
                               for g in re.findall(r"\(([0-9,]+)\)", m.group(4) or ""))
                tl.arrays[m.group(1)] = Array(m.group(1), up + lo, len(up), groups)
            continue
        lhs_s, _, rhs = line.partition("=")
        m = _RE_REF.fullmatch(lhs_s.strip())
        if not m or not rhs.strip().startswith("("):
            raise ValueError(f"line {ln}: unrecognised statement {line[:80]!r}")
        lhs = Ref(m.group(1), _idx(m.group(2)))
        rhs = rhs.strip()
        mm = _RE_MUL.match(rhs)
        if mm:
            st = Statement(lhs, "mul", [Ref(mm.group(1), _idx(mm.group(2))),
                                        Ref(mm.group(3), _idx(mm.group(4)))], ln)
        else:
            assert rhs[0] == "(" and rhs[-1] == ")", line
            st = Statement(lhs, "sum", [_parse_term(t, ln) for t in _split_top(rhs[1:-1])], ln)
        st._tl = tl
        tl.statements.append(st)
        last = st
    _validate(tl)
    return tl


def _validate(tl: TermList):
    for s in tl.statements:
        for r in [s.lhs] + s.terms:
            a = tl.arrays.get(r.name)
            if a is None:
                raise ValueError(f"line {s.line}: undeclared array {r.name}")
            if len(a.ranges) != len(r.idx):
                raise ValueError(f"line {s.line}: {r.name} has {len(a.ranges)} dims, used with {len(r.idx)}")
            for x, rg in zip(r.idx, a.ranges):
                if tl.indices.get(x) != rg:
                    raise ValueError(f"line {s.line}: index {x} over {tl.indices.get(x)} used on a {rg} dim of {r.name}")
        if s.kind == "mul":
            a, b = s.terms
            for x in s.lhs.idx:
                if (x in a.idx) == (x in b.idx):
                    raise ValueError(f"line {s.line}: external index {x} must be in exactly one operand")
            for x in set(a.idx) ^ set(b.idx):
                if x not in s.lhs.idx:
                    raise ValueError(f"line {s.line}: index {x} is neither external nor contracted")
        else:
            for t in s.terms:
                if sorted(t.idx) != sorted(s.lhs.idx):
                    raise ValueError(f"line {s.line}: sum term {t.name} is not a permutation of the result")


# Statements whose flop line the rules above do not reproduce; the file's
# number is kept for them.  _uhf_t2.eq:706,960 (ta_vo x _a, antisymmetric
# external pair) are charged 4 flops per multiply-add; _uhf_t2.eq:769,996
# (raa_vvoo, rbb_vvoo) are charged as if the result's occupied pair were
# antisymmetric, which its declaration does not say.
GENERATOR_EXCEPTIONS = {"t1": set(), "t2": {"_a48054", "_a182867", "raa_vvoo", "rbb_vvoo"}}


def check_flops(tl: TermList):
    """Returns (mismatches, recomputed total, file total, sum of the file's
    per-statement numbers, formula mismatches).  A mismatch is (statement,
    recomputed, file number)."""
    bad, badf = [], []
    tot = 0.0
    ftot = 0.0
    for s in tl.statements:
        g = s.gen_flops()
        tot += g
        if s.file_flops is not None:
            ftot += s.file_flops
            if abs(g - s.file_flops) > 1e-9 * max(1.0, s.file_flops):
                bad.append((s, g, s.file_flops))
            if s.formula and abs(eval_formula(s.formula) - s.file_flops) > 1e-9 * max(1.0, s.file_flops):
                badf.append(s)
    return bad, tot, tl.total_cost, ftot, badf


def data_path(name):
    """Path of a term list shipped with the package (``uhf_t1.eq``,
    ``uhf_t2.eq``: the reference's ``_uhf_t1.eq`` / ``_uhf_t2.eq``)."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", name)


def load(which):
    """``load("t1")`` / ``load("t2")``: the shipped UHF term lists."""
    return parse(data_path(f"uhf_{which}.eq"))
