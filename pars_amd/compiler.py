"""Grammar -> scan plan compiler (host tool).

Lowers a grammar written in the pars grammar DSL to the 3177 bytes of a
`pars::scan_plan` — the input every scanner of this backend takes — without the
reference's constexpr C++ frontend.  It restates, step for step,
include/pars/compile/automaton.hpp (regular expressions in a fixed-capacity
arena with canonicalising constructors, Brzozowski derivatives, the byte-class
partition and the derivative DFA, the masking-state analysis) and
include/pars/compile/plan.hpp (recursion and bracket guards, FIRST sets, the
lexical / structural split, mode / comment classification read off the
automaton, per-rule lexer descriptors with the reset folded in, punctuation),
so the plan it emits is byte-identical to the reference compiler's.  The
arena's capacities (nodes, children, predicates, DFA states and classes) are
the reference's and behave the same way when reached.

Grammars are built from the DSL nodes below, or read from the DSL's C++ text
form (the `struct G { ... }; template<> struct G::def<G::tag> { using type =
...; };` files the reference compiles) by `parse_grammar`.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

# ---------------------------------------------------------------------------
# DSL (include/pars/dsl.hpp)
# ---------------------------------------------------------------------------


class Node:
    pass


@dataclass(frozen=True)
class Eps(Node):
    pass


@dataclass(frozen=True)
class Any(Node):
    pass


@dataclass(frozen=True)
class Lit(Node):
    c: int


@dataclass(frozen=True)
class Cls(Node):
    chars: tuple


@dataclass(frozen=True)
class Seq(Node):
    items: tuple


@dataclass(frozen=True)
class Alt(Node):
    items: tuple


@dataclass(frozen=True)
class Rep(Node):
    g: Node
    min: int = 0
    max: int = -1


@dataclass(frozen=True)
class NotP(Node):
    g: Node


@dataclass(frozen=True)
class Rule(Node):
    tag: str


def lit(c) -> Lit:
    return Lit(c if isinstance(c, int) else ord(c))


def cls(*cs) -> Cls:
    return Cls(tuple(c if isinstance(c, int) else ord(c) for c in cs))


def seq(*gs) -> Seq:
    return Seq(tuple(gs))


def alt(*gs) -> Alt:
    return Alt(tuple(gs))


def opt(g) -> Alt:
    return Alt((g, Eps()))


def star(g) -> Rep:
    return Rep(g, 0, -1)


def plus(g) -> Rep:
    return Rep(g, 1, -1)


@dataclass
class Grammar:
    rules: list              # rule tags, in declaration order
    start: str
    defs: dict               # tag -> Node
    lexical_rules: list | None = None
    fallback_rules: list | None = None
    name: str = "grammar"

    def d(self, tag: str) -> Node:
        return self.defs[tag]


class UnsupportedGrammar(ValueError):
    pass


# ---------------------------------------------------------------------------
# Regular expressions in an arena (automaton.hpp:16-239)
# ---------------------------------------------------------------------------
kArenaCap = 8192
kMaxKids = 16
kMaxPreds = 32
kMaxDfaStates = 16
kMaxDfaReps = 16

EMPTY, EPS, ANY, CH, SET, CAT, ALT, STAR = range(8)


class Arena:
    def __init__(self):
        self.k: list[int] = []
        self.c: list[int] = []
        self.s: list = []        # frozenset of bytes for SET
        self.kid: list[tuple] = []
        self.overflow = False
        self.empty_id = self.mk(EMPTY)
        self.eps_id = self.mk(EPS)
        self.any_id = self.mk(ANY)

    def mk(self, k: int, c: int = 0, s=frozenset(), kids=()) -> int:
        if len(self.k) >= kArenaCap:
            self.overflow = True
            return 0
        self.k.append(k)
        self.c.append(c)
        self.s.append(s)
        self.kid.append(tuple(kids))
        return len(self.k) - 1

    def eq(self, a: int, b: int) -> bool:
        if a == b:
            return True
        ka, kb = self.k[a], self.k[b]
        if ka != kb:
            return False
        if ka in (EMPTY, EPS, ANY):
            return True
        if ka == CH:
            return self.c[a] == self.c[b]
        if ka == SET:
            return self.s[a] == self.s[b]
        xa, xb = self.kid[a], self.kid[b]
        if ka == CAT:
            return len(xa) == len(xb) and all(self.eq(p, q) for p, q in zip(xa, xb))
        if ka == ALT:  # a greedy multiset match, as the reference does
            if len(xa) != len(xb):
                return False
            used = [False] * len(xb)
            for p in xa:
                for j, q in enumerate(xb):
                    if not used[j] and self.eq(p, q):
                        used[j] = True
                        break
                else:
                    return False
            return True
        return self.eq(xa[0], xb[0])  # STAR

    def cat(self, a: int, b: int) -> int:
        if self.k[a] == EMPTY or self.k[b] == EMPTY:
            return self.empty_id
        out: list[int] = []

        def push(x):
            if self.k[x] == EPS:
                return
            if self.k[x] == CAT:
                for y in self.kid[x]:
                    if len(out) < kMaxKids:
                        out.append(y)
            elif len(out) < kMaxKids:
                out.append(x)

        push(a)
        push(b)
        if not out:
            return self.eps_id
        if len(out) == 1:
            return out[0]
        return self.mk(CAT, kids=out)

    def alt(self, a: int, b: int) -> int:
        tmp: list[int] = []
        for x in (a, b):
            if self.k[x] == EMPTY:
                continue
            if self.k[x] == ALT:
                tmp.extend(self.kid[x])
            else:
                tmp.append(x)
        if len(tmp) > kMaxKids:
            raise UnsupportedGrammar("an alternation exceeds the compiler's 16 branches")
        out: list[int] = []
        for x in tmp:
            if not any(self.eq(y, x) for y in out):
                out.append(x)
        if not out:
            return self.empty_id
        if len(out) == 1:
            return out[0]
        return self.mk(ALT, kids=out)

    def star(self, a: int) -> int:
        if self.k[a] in (EMPTY, EPS):
            return self.eps_id
        if self.k[a] == STAR:
            return a
        return self.mk(STAR, kids=(a,))

    def nullable(self, i: int) -> bool:
        k = self.k[i]
        if k == EPS or k == STAR:
            return True
        if k in (EMPTY, ANY, CH, SET):
            return False
        if k == ALT:
            return any(self.nullable(x) for x in self.kid[i])
        return all(self.nullable(x) for x in self.kid[i])  # CAT

    def der(self, i: int, c: int) -> int:
        k = self.k[i]
        if k in (EMPTY, EPS):
            return self.empty_id
        if k == ANY:
            return self.eps_id
        if k == CH:
            return self.eps_id if c == self.c[i] else self.empty_id
        if k == SET:
            return self.eps_id if c in self.s[i] else self.empty_id
        if k == ALT:
            acc = self.empty_id
            for x in self.kid[i]:
                acc = self.alt(acc, self.der(x, c))
            return acc
        if k == CAT:
            kids = self.kid[i]
            if not kids:
                return self.empty_id
            first = kids[0]
            rest = self.eps_id
            for x in kids[1:]:
                rest = self.cat(rest, x)
            acc = self.cat(self.der(first, c), rest)
            if self.nullable(first):
                acc = self.alt(acc, self.der(rest, c))
            return acc
        return self.cat(self.der(self.kid[i][0], c), i)  # STAR

    def collect_preds(self, i: int, p: list):
        if self.k[i] == CH:
            if len(p) < kMaxPreds:
                p.append(frozenset([self.c[i]]))
        elif self.k[i] == SET:
            if len(p) < kMaxPreds:
                p.append(self.s[i])
        for x in self.kid[i]:
            self.collect_preds(x, p)


@dataclass
class Dfa:
    trans: list = field(default_factory=list)      # [state][rep] -> state or -1
    accepting: list = field(default_factory=list)
    rep_of: list = field(default_factory=lambda: [0] * 256)
    reps: list = field(default_factory=list)
    start: int = 0
    ok: bool = True

    def next(self, s: int, b: int) -> int:
        return self.trans[s][self.rep_of[b]]

    @property
    def n_states(self) -> int:
        return len(self.trans)


def build_dfa(A: Arena, root: int) -> Dfa:
    """automaton.hpp:253-308: byte classes from the predicates, derivative BFS."""
    D = Dfa()
    preds: list = []
    A.collect_preds(root, preds)
    for b in range(256):
        found = -1
        for j, r in enumerate(D.reps):
            if all((b in p) == (r in p) for p in preds):
                found = j
                break
        if found < 0:
            if len(D.reps) >= kMaxDfaReps:
                D.ok = False
                D.rep_of[b] = 0
                continue
            D.reps.append(b)
            found = len(D.reps) - 1
        D.rep_of[b] = found
    sid = [root]
    D.trans = [[-1] * kMaxDfaReps]
    D.accepting = [A.nullable(root)]
    s = 0
    while s < len(sid):
        for c, rb in enumerate(D.reps):
            nd = A.der(sid[s], rb)
            if A.k[nd] == EMPTY:
                continue
            found = -1
            for t, x in enumerate(sid):
                if A.eq(x, nd):
                    found = t
                    break
            if found < 0:
                if len(sid) >= kMaxDfaStates:
                    D.ok = False
                    continue
                sid.append(nd)
                D.trans.append([-1] * kMaxDfaReps)
                D.accepting.append(A.nullable(nd))
                found = len(sid) - 1
            D.trans[s][c] = found
        s += 1
    return D


def reach_accept(D: Dfa) -> list:
    can = list(D.accepting)
    changed = True
    while changed:
        changed = False
        for i in range(D.n_states):
            if can[i]:
                continue
            for c in range(len(D.reps)):
                k = D.trans[i][c]
                if k >= 0 and can[k]:
                    can[i] = True
                    changed = True
                    break
    return can


def masking_states(D: Dfa, structural: list) -> list:
    """automaton.hpp:335-365: non-start states with a path to accept through a structural byte."""
    can = reach_accept(D)
    mask = [False] * D.n_states
    for i in range(D.n_states):
        for b in range(256):
            if not structural[b]:
                continue
            k = D.next(i, b)
            if k >= 0 and can[k]:
                mask[i] = True
                break
    changed = True
    while changed:
        changed = False
        for i in range(D.n_states):
            if mask[i]:
                continue
            for c in range(len(D.reps)):
                k = D.trans[i][c]
                if k >= 0 and mask[k]:
                    mask[i] = True
                    changed = True
                    break
    mask[D.start] = False
    return mask


def _notp_set_byte(n: Node) -> Cls | None:
    """seq<notp<cls<S>>, any> — one byte not in S (the only supported notp form)."""
    if isinstance(n, Seq) and len(n.items) == 2 and isinstance(n.items[0], NotP) \
            and isinstance(n.items[0].g, Cls) and isinstance(n.items[1], Any):
        return n.items[0].g
    return None


def reify(G: Grammar, A: Arena, n: Node, visited: tuple = ()) -> int:
    """automaton.hpp:370-509: DSL node -> arena node."""
    if isinstance(n, Eps):
        return A.eps_id
    if isinstance(n, Any):
        return A.any_id
    if isinstance(n, Lit):
        return A.mk(CH, c=n.c & 0xFF)
    if isinstance(n, Cls):
        return A.mk(SET, s=frozenset(c & 0xFF for c in n.chars))
    if isinstance(n, Seq):
        ns = _notp_set_byte(n)
        if ns is not None:
            return A.mk(SET, s=frozenset(range(256)) - frozenset(c & 0xFF for c in ns.chars))
        ids = [reify(G, A, g, visited) for g in n.items]
        acc = A.eps_id
        for x in ids:
            acc = A.cat(acc, x)
        return acc
    if isinstance(n, Alt):
        ids = [reify(G, A, g, visited) for g in n.items]
        acc = A.empty_id
        for x in ids:
            acc = A.alt(acc, x)
        return acc
    if isinstance(n, Rep):
        acc = A.eps_id
        for _ in range(n.min):
            acc = A.cat(acc, reify(G, A, n.g, visited))
        if n.max < 0:
            acc = A.cat(acc, A.star(reify(G, A, n.g, visited)))
        else:
            for _ in range(n.min, n.max):
                acc = A.cat(acc, A.alt(reify(G, A, n.g, visited), A.eps_id))
        return acc
    if isinstance(n, Rule):
        if n.tag in visited:
            return A.any_id
        return reify(G, A, G.d(n.tag), visited + (n.tag,))
    return A.empty_id  # notp outside the supported form


# ---------------------------------------------------------------------------
# Grammar analysis (compile/plan.hpp:38-420)
# ---------------------------------------------------------------------------


def mentions(G: Grammar, n: Node, tag: str, visited: tuple = ()) -> bool:
    if isinstance(n, Rule):
        if n.tag == tag:
            return True
        if n.tag in visited:
            return False
        return mentions(G, G.d(n.tag), tag, visited + (n.tag,))
    if isinstance(n, (Seq, Alt)):
        return any(mentions(G, g, tag, visited) for g in n.items)
    if isinstance(n, (Rep, NotP)):
        return mentions(G, n.g, tag, visited)
    return False


def is_recursive(G: Grammar, tag: str) -> bool:
    return mentions(G, G.d(tag), tag)


def notp_ok(G: Grammar, n: Node, visited: tuple = ()) -> bool:
    if isinstance(n, NotP):
        return False
    if isinstance(n, Seq):
        if _notp_set_byte(n) is not None:
            return True
        return all(notp_ok(G, g, visited) for g in n.items)
    if isinstance(n, Alt):
        return all(notp_ok(G, g, visited) for g in n.items)
    if isinstance(n, Rep):
        return notp_ok(G, n.g, visited)
    if isinstance(n, Rule):
        if n.tag in visited:
            return True
        return notp_ok(G, G.d(n.tag), visited + (n.tag,))
    return True


def _is_lit(g: Node) -> bool:
    """plan.hpp:25-31: a node counts as a literal through the literal's char
    value, so lit<'\0'> does not count (in the keyword and bracket tests)."""
    return isinstance(g, Lit) and (g.c & 0xFF) != 0


def bracket_guard(n: Node) -> tuple[bool, int, int]:
    """plan.hpp:141-193: seq of >= 3 items, literal first and last, exactly two literals, distinct."""
    if not isinstance(n, Seq):
        return False, 0, 0
    items = n.items
    N = len(items)
    nlits = sum(1 for g in items if _is_lit(g))
    first_lit = N > 0 and _is_lit(items[0])
    last_lit = N > 0 and _is_lit(items[-1])
    o = items[0].c & 0xFF if first_lit else 0
    c = items[-1].c & 0xFF if last_lit else 0
    return (N >= 3 and first_lit and last_lit and nlits == 2 and o != c), o, c


def count_rule(n: Node, tag: str) -> int:
    if isinstance(n, Rule):
        return 1 if n.tag == tag else 0
    if isinstance(n, (Seq, Alt)):
        return sum(count_rule(g, tag) for g in n.items)
    if isinstance(n, (Rep, NotP)):
        return count_rule(n.g, tag)
    return 0


def is_lexical_rule(G: Grammar, tag: str) -> bool:
    if G.lexical_rules is not None:
        return tag in G.lexical_rules
    return not is_recursive(G, tag)


def rule_ok(G: Grammar, tag: str) -> bool:
    if G.fallback_rules is not None and tag in G.fallback_rules:
        return True
    d = G.d(tag)
    if is_recursive(G, tag):
        return bracket_guard(d)[0] or count_rule(d, tag) == 0
    if G.lexical_rules is not None:
        return tag not in G.lexical_rules or notp_ok(G, d)
    return notp_ok(G, d)


def nullable_node(G: Grammar, n: Node, visited: tuple = ()) -> bool:
    if isinstance(n, Eps) or isinstance(n, NotP):
        return True
    if isinstance(n, Seq):
        return all(nullable_node(G, g, visited) for g in n.items)
    if isinstance(n, Alt):
        return any(nullable_node(G, g, visited) for g in n.items)
    if isinstance(n, Rep):
        return n.min == 0 or nullable_node(G, n.g, visited)
    if isinstance(n, Rule):
        if n.tag in visited:
            return False
        return nullable_node(G, G.d(n.tag), visited + (n.tag,))
    return False


def first_terms(G: Grammar, n: Node, visited: tuple = ()) -> set:
    if isinstance(n, Any):
        return set(range(256))
    if isinstance(n, Lit):
        return {n.c & 0xFF}
    if isinstance(n, Cls):
        return {c & 0xFF for c in n.chars}
    if isinstance(n, Seq):
        out: set = set()
        for g in n.items:
            out |= first_terms(G, g, visited)
            if not nullable_node(G, g, visited):
                break
        return out
    if isinstance(n, Alt):
        out = set()
        for g in n.items:
            out |= first_terms(G, g, visited)
        return out
    if isinstance(n, Rep):
        return first_terms(G, n.g, visited)
    if isinstance(n, Rule):
        if n.tag in visited:
            return set()
        return first_terms(G, G.d(n.tag), visited + (n.tag,))
    return set()


def collect_lits(G: Grammar, n: Node, out: list, visited: tuple):
    """plan.hpp:640-694: standalone terminals of non-lexical rules (keywords excluded)."""
    if isinstance(n, Lit):
        out[n.c & 0xFF] = True
    elif isinstance(n, Seq):
        if all(_is_lit(g) for g in n.items):
            return  # a keyword
        for g in n.items:
            collect_lits(G, g, out, visited)
    elif isinstance(n, Alt):
        for g in n.items:
            collect_lits(G, g, out, visited)
    elif isinstance(n, (Rep, NotP)):
        collect_lits(G, n.g, out, visited)
    elif isinstance(n, Rule):
        if n.tag in visited or is_lexical_rule(G, n.tag):
            return
        collect_lits(G, G.d(n.tag), out, visited + (n.tag,))


# lex_kind (include/pars/plan.hpp:56)
AFFINE, COMMENT, DOUBLED, TABLE, FALLBACK = range(5)


@dataclass
class LexInfo:
    kind: int = 0  # 0 none, 1 mode, 2 comment
    delim: int = 0
    escape: int = 0


def derive_lexical(G: Grammar, tag: str, structural: list) -> LexInfo:
    """plan.hpp:416-510: a masking rule's delimiter / escape read off its automaton."""
    A = Arena()
    D = build_dfa(A, reify(G, A, G.d(tag)))
    if not D.ok:
        return LexInfo()
    if not any(masking_states(D, structural)):
        return LexInfo()
    out = LexInfo()
    if D.n_states < 1:
        return out
    starts = [a for a in range(256) if D.next(0, a) >= 0]
    if len(starts) != 1:
        return out
    D0 = starts[-1]
    S1 = D.next(0, D0)
    if S1 <= 0:
        return out
    out.delim = D0
    if D.accepting[S1]:
        if all(D.next(S1, a) in (-1, S1) for a in range(256)):
            out.kind = 2
            return out
    S2 = D.next(S1, D0)
    if S2 >= 0 and S2 != S1 and D.next(S2, D0) == S1:
        out.kind, out.escape = 1, D0
        return out
    for E in range(256):
        if E == D0:
            continue
        S2 = D.next(S1, E)
        if S2 < 0 or S2 == S1:
            continue
        if all(D.next(S2, a) in (-1, S1) for a in range(256)):
            out.kind, out.escape = 1, E
            return out
    if not D.accepting[S1]:
        q = D.next(S1, D0)
        body = any(D.next(S1, a) >= 0 for a in range(256) if a != D0)
        if q >= 0 and D.accepting[q] and body:
            out.kind, out.escape = 1, 0
            return out
    return out


kPlanStates = 8
kPlanClasses = 8


@dataclass
class LexDesc:
    kind: int = TABLE
    n_states: int = 0
    n_classes: int = 0
    start: int = 0
    ok: bool = True
    delim: int = 0
    escape: int = 0
    rep_of: list = field(default_factory=lambda: [0] * 256)
    trans: list = field(default_factory=lambda: [[0] * kPlanClasses for _ in range(kPlanStates)])
    masked: list = field(default_factory=lambda: [[0] * kPlanClasses for _ in range(kPlanStates)])


def make_desc(G: Grammar, tag: str, structural: list) -> LexDesc:
    """plan.hpp:516-574: the rule's automaton with the lexer reset folded in."""
    out = LexDesc()
    A = Arena()
    D = build_dfa(A, reify(G, A, G.d(tag)))
    if not D.ok:
        out.ok, out.kind = False, FALLBACK
        return out
    if not any(masking_states(D, structural)):
        out.ok = False
        return out
    n, nc = D.n_states, len(D.reps)
    if n > kPlanStates or nc > kPlanClasses:
        out.ok, out.kind = False, FALLBACK
        return out
    out.n_states, out.n_classes, out.start = n, nc, D.start
    for s in range(n):
        for c in range(nc):
            tok = D.trans[s][c]
            out.masked[s][c] = 1 if (s != D.start and tok >= 0) else 0
            nd = tok
            if nd < 0:
                n0 = D.trans[D.start][c]
                nd = n0 if n0 >= 0 else D.start
            out.trans[s][c] = nd
    out.rep_of = list(D.rep_of)
    info = derive_lexical(G, tag, structural)
    if info.kind == 1:
        out.delim = info.delim
        if info.escape == info.delim:
            out.kind, out.escape = DOUBLED, 0
        else:
            out.kind, out.escape = AFFINE, info.escape
    elif info.kind == 2:
        out.kind, out.delim = COMMENT, info.delim
    else:
        out.kind = TABLE
    out.ok = True
    return out


kMaxPunct, kMaxBrackets, kMaxModes, kMaxComments = 8, 8, 8, 4


@dataclass
class ScanPlan:
    punctuation: list = field(default_factory=list)
    brackets: list = field(default_factory=list)   # (open, close)
    modes: list = field(default_factory=list)      # (delim, escape)
    comments: list = field(default_factory=list)
    descs: list = field(default_factory=list)

    def to_bytes(self) -> bytes:
        """include/pars/plan.hpp:39-79 laid out as include/pars_plan.h (3177 bytes)."""
        b = bytearray()
        b += bytes(self.punctuation + [0] * (kMaxPunct - len(self.punctuation)))
        for i in range(kMaxBrackets):
            b += bytes(self.brackets[i]) if i < len(self.brackets) else b"\0\0"
        for i in range(kMaxModes):
            b += bytes(self.modes[i]) if i < len(self.modes) else b"\0\0"
        b += bytes(self.comments + [0] * (kMaxComments - len(self.comments)))
        b += bytes([len(self.punctuation), len(self.brackets), len(self.modes), len(self.comments)])
        for i in range(kMaxModes):
            d = self.descs[i] if i < len(self.descs) else LexDesc()
            b += bytes([d.kind, d.n_states, d.n_classes, d.start, 1 if d.ok else 0, d.delim & 0xFF,
                        d.escape & 0xFF])
            b += bytes(d.rep_of)
            for row in d.trans:
                b += bytes(row)
            for row in d.masked:
                b += bytes(row)
        b += bytes([len(self.descs)])
        assert len(b) == 3177
        return bytes(b)


def is_supported(G: Grammar) -> bool:
    """plan.hpp:229-238: every rule has a sound data-parallel lowering."""
    return all(rule_ok(G, t) for t in G.rules)


def compile_grammar(G: Grammar) -> ScanPlan:
    """The gate of plan.hpp:241-249 (require_supported), then make_scan_plan."""
    bad = [t for t in G.rules if not rule_ok(G, t)]
    if bad:
        raise UnsupportedGrammar(
            f"rules {bad} have no sound data-parallel lowering (unsupported recursion or notp form); "
            "declare them in fallback_rules")
    return make_scan_plan(G)


def make_scan_plan(G: Grammar) -> ScanPlan:
    """plan.hpp:711-748, without the support gate (as the reference's function)."""
    S = ScanPlan()
    for t in G.rules:  # brackets from guarded recursive rules
        if is_recursive(G, t):
            ok, o, c = bracket_guard(G.d(t))
            if ok and len(S.brackets) < kMaxBrackets and (o, c) not in S.brackets:
                S.brackets.append((o, c))
    lits = [False] * 256
    for t in G.rules:
        if not is_lexical_rule(G, t):
            collect_lits(G, G.d(t), lits, (t,))
    structural = [False] * 256
    for o, c in S.brackets:
        structural[o] = structural[c] = True
    for c in range(256):
        if lits[c]:
            structural[c] = True
    for t in G.rules:  # lexical rules: modes, comments, descriptors
        if not is_lexical_rule(G, t) or not notp_ok(G, G.d(t)):
            continue
        if len(first_terms(G, G.d(t))) != 1:
            continue
        info = derive_lexical(G, t, structural)
        if info.kind == 1 and len(S.modes) < kMaxModes:
            m = (info.delim, info.escape)
            if m not in S.modes:
                S.modes.append(m)
        elif info.kind == 2 and len(S.comments) < kMaxComments:
            if info.delim not in S.comments:
                S.comments.append(info.delim)
        if len(S.descs) < kMaxModes:
            d = make_desc(G, t, structural)
            if d.ok:
                S.descs.append(d)
    exclude = [False] * 256
    for o, c in S.brackets:
        exclude[o] = exclude[c] = True
    for d, e in S.modes:
        exclude[d] = True
        if e:
            exclude[e] = True
    for c in S.comments:
        exclude[c] = True
    for c in range(256):
        if len(S.punctuation) >= kMaxPunct:
            break
        if lits[c] and not exclude[c]:
            S.punctuation.append(c)
    return S


def compile_plan(G: Grammar):
    """The grammar's plan as a pars_amd.Plan (what Scanner takes)."""
    from pars_amd import Plan

    return Plan(compile_grammar(G).to_bytes(), G.name)


# ---------------------------------------------------------------------------
# The DSL's C++ text form
# ---------------------------------------------------------------------------
_TOKEN = re.compile(r"""
    (?P<ws>\s+|//[^\n]*|/\*.*?\*/)
  | (?P<chr>'(?:\\(?:x[0-9a-fA-F]+|[0-7]{1,3}|.)|[^\\'])')
  | (?P<num>-?\d+)
  | (?P<id>[A-Za-z_][A-Za-z_0-9]*)
  | (?P<op>::|[<>,;{}=()\[\]*&])
""", re.X | re.S)

_ESC = {"n": 10, "t": 9, "r": 13, "\\": 92, "'": 39, '"': 34, "a": 7, "b": 8, "f": 12, "v": 11, "?": 63}


def _char_value(tok: str) -> int:
    """A C++ character literal's value as the unsigned byte the DSL stores."""
    body = tok[1:-1]
    if body.startswith("\\"):
        e = body[1:]
        if e.startswith("x"):
            return int(e[1:], 16) & 0xFF
        if e[0] in "01234567":
            return int(e, 8) & 0xFF
        return _ESC[e]
    return ord(body) & 0xFF


def _tokens(text: str) -> list:
    text = re.sub(r"^[ \t]*#.*$", "", text, flags=re.M)  # preprocessor lines
    out, i = [], 0
    while i < len(text):
        m = _TOKEN.match(text, i)
        if not m:
            raise ValueError(f"cannot read the grammar text at: {text[i:i + 40]!r}")
        i = m.end()
        if m.lastgroup != "ws":
            out.append((m.lastgroup, m.group()))
    return out


class _Reader:
    def __init__(self, toks):
        self.t, self.i = toks, 0

    def peek(self) -> tuple:
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self, val=None) -> str:
        kind, v = self.peek()
        if kind is None or (val is not None and v != val):
            raise ValueError(f"grammar text: expected {val!r}, got {v!r}")
        self.i += 1
        return v

    def qname(self) -> list:
        """A possibly qualified name, as its components."""
        parts = [self.take()]
        while self.peek()[1] == "::":
            self.take("::")
            parts.append(self.take())
        return parts


_CTORS = {
    "lit": lambda a: Lit(a[0]),
    "cls": lambda a: Cls(tuple(a)),
    "seq": lambda a: Seq(tuple(a)),
    "alt": lambda a: Alt(tuple(a)),
    "opt": lambda a: Alt((a[0], Eps())),
    "star": lambda a: Rep(a[0], 0, -1),
    "plus": lambda a: Rep(a[0], 1, -1),
    "rep": lambda a: Rep(a[0], a[1] if len(a) > 1 else 0, a[2] if len(a) > 2 else -1),
    "notp": lambda a: NotP(a[0]),
}


def _node(R: _Reader, local: dict) -> Node:
    """One node expression; `local` holds the def's own `using` aliases."""
    head = R.qname()[-1]
    if head in local:
        return local[head]
    if head == "eps":
        return Eps()
    if head == "any":
        return Any()
    R.take("<")
    if head == "rule":
        tag = R.qname()[-1]
        R.take(">")
        return Rule(tag)
    args: list = []
    while R.peek()[1] != ">":
        kind, v = R.peek()
        if kind == "chr":
            R.take()
            args.append(_char_value(v))
        elif kind == "num":
            R.take()
            args.append(int(v))
        else:
            args.append(_node(R, local))
        if R.peek()[1] == ",":
            R.take(",")
    R.take(">")
    if head not in _CTORS:
        raise ValueError(f"unknown grammar node {head!r}")
    return _CTORS[head](args)


def _read_struct(R: _Reader, gname: str) -> dict:
    """Body of `struct G { struct tag {}; using rules = type_list<...>; ... };`."""
    info: dict = {"tags": []}
    R.take("{")
    while R.peek()[1] != "}":
        v = R.peek()[1]
        if v == "struct":
            R.take()
            info["tags"].append(R.take())
            R.take("{")
            R.take("}")
            R.take(";")
        elif v == "using":
            R.take()
            a = R.take()
            R.take("=")
            if R.peek()[1] == "type_list":
                R.take()
                R.take("<")
                items = []
                while R.peek()[1] != ">":
                    items.append(R.qname()[-1])
                    if R.peek()[1] == ",":
                        R.take(",")
                R.take(">")
                info[a] = items
            else:
                info[a] = R.qname()[-1]
            R.take(";")
        elif v == "template":  # template<class> struct def;
            while R.take() != ";":
                pass
        else:
            raise ValueError(f"unexpected {v!r} in grammar struct {gname}")
    R.take("}")
    R.take(";")
    return info


def parse_grammars(text: str) -> dict:
    """Every grammar in a text of the DSL's C++ form: grammar structs, their
    `def<Tag>` specialisations (with local `using` aliases) and namespace-level
    aliases of the structs.  Returns {struct name: Grammar}."""
    R = _Reader(_tokens(text))
    structs: dict = {}
    alias: dict = {}
    defs: dict = {}
    while R.peek()[0] is not None:
        v = R.peek()[1]
        if v == "namespace":
            R.take()
            if R.peek()[1] != "{":
                R.qname()
            R.take("{")
        elif v == "}":
            R.take()
            if R.peek()[1] == ";":
                R.take()
        elif v == "using":
            R.take()
            if R.peek()[1] == "namespace":
                R.take()
                R.qname()
            else:
                a = R.take()
                R.take("=")
                alias[a] = R.qname()[-1]
            R.take(";")
        elif v == "struct":
            R.take()
            gname = R.take()
            structs[gname] = _read_struct(R, gname)
        elif v == "template":
            R.take()
            R.take("<")
            while R.peek()[1] != ">":
                R.take()
            R.take(">")
            R.take("struct")
            owner = R.qname()
            if owner[-1] != "def" or len(owner) < 2:
                raise ValueError(f"expected G::def<Tag>, got {'::'.join(owner)}")
            R.take("<")
            tag = R.qname()[-1]
            R.take(">")
            R.take("{")
            local: dict = {}
            body = None
            while R.peek()[1] != "}":
                R.take("using")
                a = R.take()
                R.take("=")
                node = _node(R, local)
                R.take(";")
                if a == "type":
                    body = node
                else:
                    local[a] = node
            R.take("}")
            R.take(";")
            if body is None:
                raise ValueError(f"def<{tag}> has no `type`")
            g = alias.get(owner[-2], owner[-2])
            defs.setdefault(g, {})[tag] = body
        else:
            raise ValueError(f"unexpected {v!r} in the grammar text")
    out: dict = {}
    for gname, info in structs.items():
        if "rules" not in info or "start" not in info:
            continue
        d = defs.get(gname, {})
        missing = [t for t in info["rules"] if t not in d]
        if missing:
            raise ValueError(f"{gname}: rules without a def: {missing}")
        out[gname] = Grammar(rules=info["rules"], start=info["start"], defs=d,
                             lexical_rules=info.get("lexical_rules"),
                             fallback_rules=info.get("fallback_rules"), name=gname)
    return out


def parse_grammar(text: str, name: str | None = None) -> Grammar:
    """The one grammar of a text in the DSL's C++ form (or the struct `name`)."""
    gs = parse_grammars(text)
    if name is not None and name in gs:
        return gs[name]
    if len(gs) != 1:
        raise ValueError(f"expected one grammar, found {sorted(gs)}")
    G = next(iter(gs.values()))
    if name is not None:
        G.name = name
    return G


def compile_text(text: str, name: str | None = None):
    """Plan of a grammar given in the DSL's C++ text form."""
    return compile_plan(parse_grammar(text, name))
