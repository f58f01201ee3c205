"""The grammar -> plan compiler (pars_amd/compiler.py) against the reference
compiler's output.

Pinning:
  * the seven shipped grammars: the reference's own grammar headers
    (bench/grammars/*.hpp, read as text where the reference tree exists) and
    oracle/grammars/clike6.hpp compile to exactly the committed plans
    pars_amd/plans/*.plan that the reference's compile::make_scan_plan wrote;
  * the probe grammars of oracle/grammars/compiler_cases.hpp, which take the
    branches the shipped grammars leave untaken, compile to exactly the bytes
    and the support verdict the reference compiler gave
    (tests/golden/plans/, written by tests/golden/make_plans.py);
  * the reference's compile-time checks of tests/test_model.cpp:29-147 (plan
    counts per grammar, the support gate on three small grammars);
  * planted mistakes in the lowering each change at least one plan.

The oracle is pinned on the probe plans too (against the reference CPU
backend's outputs in tests/golden/plans/scans.json), so GPU parity on them
(tests/test_gpu_compiler.py) is parity with the reference.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from pars_amd import Plan
from pars_amd import compiler as C
from pars_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_GRAMMARS = "/root/reference/bench/grammars"
SHIPPED = ["json", "csv", "clike", "toml", "ini", "xml"]
PROBE_HDR = os.path.join(ROOT, "oracle", "grammars", "compiler_cases.hpp")
GOLD = os.path.join(ROOT, "tests", "golden", "plans")


def _read(path: str) -> str:
    with open(path) as f:
        return f.read()


def _fixture(name: str) -> bytes:
    with open(os.path.join(GOLD, name + ".plan"), "rb") as f:
        return f.read()


MANIFEST = json.loads(_read(os.path.join(GOLD, "manifest.json")))
PROBES = C.parse_grammars(_read(PROBE_HDR))


def _all_grammars() -> dict:
    """Every grammar whose reference plan is on disk: name -> (Grammar, bytes)."""
    out = {n: (PROBES[n], _fixture(n)) for n in MANIFEST}
    out["clike6"] = (C.parse_grammar(_read(os.path.join(ROOT, "oracle", "grammars", "clike6.hpp")), "clike6"),
                     Plan.load("clike6").blob)
    if os.path.isdir(REF_GRAMMARS):
        for g in SHIPPED:
            out[g] = (C.parse_grammar(_read(os.path.join(REF_GRAMMARS, g + ".hpp")), g), Plan.load(g).blob)
    return out


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("g", SHIPPED)
def test_reference_grammar_headers(g):
    path = os.path.join(REF_GRAMMARS, g + ".hpp")
    if not os.path.exists(path):
        pytest.skip("reference grammar headers absent")
    G = C.parse_grammar(_read(path), g)
    assert C.is_supported(G)
    assert C.compile_grammar(G).to_bytes() == Plan.load(g).blob


def test_clike6_header():
    G = C.parse_grammar(_read(os.path.join(ROOT, "oracle", "grammars", "clike6.hpp")), "clike6")
    assert C.compile_plan(G).blob == Plan.load("clike6").blob


@pytest.mark.parametrize("name", sorted(MANIFEST))
def test_probe_grammars(name):
    G = PROBES[name]
    blob = C.make_scan_plan(G).to_bytes()
    assert hashlib.sha256(_fixture(name)).hexdigest() == MANIFEST[name]["sha256"]
    assert blob == _fixture(name)
    assert C.is_supported(G) == MANIFEST[name]["supported"]
    if MANIFEST[name]["supported"]:
        assert C.compile_grammar(G).to_bytes() == blob
    else:
        with pytest.raises(C.UnsupportedGrammar):
            C.compile_grammar(G)


def test_probe_branch_coverage():
    """The probe plans hold the shapes they were written for."""
    p = {n: Plan(_fixture(n), n) for n in MANIFEST}
    kinds = lambda P: [P.desc(d)["kind"] for d in range(P.n_descs)]  # noqa: E731
    assert len(p["many_case"].brackets) == 8 and len(p["many_case"].punctuation) == 8  # both capped
    assert p["many_case"].n_descs == 0
    assert 0 in p["caps_case"].punctuation  # the NUL literal of a would-be keyword
    assert (ord("@"), ord(".")) in p["caps_case"].brackets  # a recursive lexical rule
    assert kinds(p["caps_case"]) == [C.AFFINE, C.AFFINE]  # the 12-state rule fell back, the wide one failed
    assert C.DOUBLED in kinds(p["sql_case"]) and C.TABLE in kinds(p["sql_case"])
    assert C.COMMENT in kinds(p["esc_case"])
    modes = {(b[24 + 2 * i], b[25 + 2 * i]) for n, b in ((n, _fixture(n)) for n in ("esc_case",))
             for i in range(b[46])}
    assert (ord("|"), ord("%")) in modes and (ord("`"), 0) in modes  # escape != delim, and escape 0
    assert _fixture("tiny_case")[:48] == bytes(48) and _fixture("tiny_case")[-1] == 0


# ---------------------------------------------------------------------------
# tests/test_model.cpp:29-147 (the reference's compile-time checks)
def _counts(blob: bytes) -> dict:
    P = Plan(blob)
    return {"brackets": blob[45], "modes": blob[46], "comments": blob[47], "punct": blob[44],
            "mode0": (blob[24], blob[25]), "kinds": [P.desc(d)["kind"] for d in range(P.n_descs)]}


def test_reference_plan_counts():
    srcs = {g: Plan.load(g).blob for g in ("json", "csv", "toml", "clike")}
    if os.path.isdir(REF_GRAMMARS):  # and the same through this compiler
        for g in list(srcs):
            srcs[g + "+compiled"] = C.compile_grammar(
                C.parse_grammar(_read(os.path.join(REF_GRAMMARS, g + ".hpp")), g)).to_bytes()
    for key, blob in srcs.items():
        c, g = _counts(blob), key.split("+")[0]
        if g == "json":
            assert c["brackets"] == 2 and c["modes"] == 1 and c["mode0"] == (ord('"'), ord("\\"))
            assert c["punct"] == 2 and c["kinds"].count(C.AFFINE) == 1
        elif g == "csv":
            assert c["brackets"] == 0 and c["modes"] == 1 and c["mode0"] == (ord('"'), ord('"'))
            assert c["punct"] == 2 and c["kinds"].count(C.DOUBLED) == 1
        elif g == "toml":
            assert c["brackets"] == 2 and c["modes"] == 2 and c["comments"] == 1 and c["punct"] == 3
            assert c["kinds"].count(C.COMMENT) == 1
        elif g == "clike":
            assert c["brackets"] == 1 and c["punct"] == 1 and C.TABLE in c["kinds"]


def test_support_gate():
    """tests/test_model.cpp:91-147: g_bad_bracket, g_bad_notp, g_fallback."""
    bad_bracket = C.Grammar(rules=["a"], start="a",
                            defs={"a": C.seq(C.lit("("), C.Rule("a"), C.lit(")"), C.lit(";"))})
    assert not C.is_supported(bad_bracket)
    bad_notp = C.Grammar(rules=["a", "b"], start="a", defs={"b": C.lit("x"), "a": C.NotP(C.Rule("b"))})
    assert not C.is_supported(bad_notp)
    fallback = C.Grammar(rules=["a"], start="a", fallback_rules=["a"],
                         defs={"a": C.seq(C.Rule("a"), C.lit("x"), C.Rule("a"))})
    assert C.is_supported(fallback)
    with pytest.raises(C.UnsupportedGrammar):
        C.compile_grammar(bad_bracket)
    # a guarded bracket rule and a cycle-glue rule are supported
    ok = C.Grammar(rules=["v", "arr"], start="v",
                   defs={"v": C.alt(C.Rule("arr"), C.lit("1")),
                         "arr": C.seq(C.lit("["), C.star(C.Rule("v")), C.lit("]"))})
    assert C.is_supported(ok)
    assert C.compile_grammar(ok).brackets == [(ord("["), ord("]"))]


# ---------------------------------------------------------------------------
def test_text_reader_forms():
    text = r"""
    #pragma once
    namespace a::b {
    using namespace pars;
    struct g {  // a grammar
        struct r {}; struct s {};
        using rules = type_list<r, s>;
        using start = r;
        template<class> struct def;
    };
    using H = g;
    /* block comment */
    template<> struct H::def<H::s> {
        using q = lit<'\x22'>;
        using type = seq<q, star<seq<notp<cls<'\042', '\\'>>, any>>, q>;
    };
    template<> struct g::def<g::r> {
        using type = seq<lit<'('>, rep<rule<g::r>, 0, 2>, opt<rule<a::b::g::s>>, lit<')'>>;
    };
    }
    """
    G = C.parse_grammar(text)
    assert G.name == "g" and G.rules == ["r", "s"] and G.start == "r"
    q = C.Lit(0x22)
    assert G.defs["s"] == C.seq(q, C.star(C.seq(C.NotP(C.cls(0x22, 0x5C)), C.Any())), q)
    assert G.defs["r"] == C.seq(C.lit("("), C.Rep(C.Rule("r"), 0, 2), C.opt(C.Rule("s")), C.lit(")"))
    with pytest.raises(ValueError):
        C.parse_grammar(text.replace("lit<'('>", "lyt<'('>"))
    with pytest.raises(ValueError):
        C.parse_grammar(text.replace("template<> struct g::def<g::r>", "template<> struct g::def<g::zz>"))


def test_python_dsl_matches_text():
    """The same grammar through the Python constructors and the text reader."""
    text = """
    struct m { struct top {}; struct str {}; struct grp {};
      using rules = type_list<top, str, grp>; using start = top; template<class> struct def; };
    template<> struct m::def<m::str> { using type = seq<lit<'"'>, star<alt<seq<lit<'\\\\'>, any>,
        seq<notp<cls<'"', '\\\\'>>, any>>>, lit<'"'>>; };
    template<> struct m::def<m::grp> { using type = seq<lit<'{'>, star<rule<m::top>>, lit<'}'>>; };
    template<> struct m::def<m::top> { using type = alt<rule<m::grp>, rule<m::str>, lit<','>>; };
    """
    G = C.Grammar(rules=["top", "str", "grp"], start="top", defs={
        "str": C.seq(C.lit('"'), C.star(C.alt(C.seq(C.lit("\\"), C.Any()),
                                              C.seq(C.NotP(C.cls('"', "\\")), C.Any()))), C.lit('"')),
        "grp": C.seq(C.lit("{"), C.star(C.Rule("top")), C.lit("}")),
        "top": C.alt(C.Rule("grp"), C.Rule("str"), C.lit(",")),
    })
    a, b = C.compile_grammar(C.parse_grammar(text)).to_bytes(), C.compile_grammar(G).to_bytes()
    assert a == b
    P = Plan(a)
    assert P.brackets == [(ord("{"), ord("}"))] and P.punctuation == b","
    assert P.n_descs == 1 and P.desc(0)["kind"] == C.AFFINE


# ---------------------------------------------------------------------------
# planted mistakes: each must change at least one plan on disk
def _no_reset_fold(G, tag, structural):
    d = _orig["make_desc"](G, tag, structural)
    if d.ok and d.n_states:
        A = C.Arena()
        D = C.build_dfa(A, C.reify(G, A, G.d(tag)))
        for s in range(d.n_states):
            for c in range(d.n_classes):
                t = D.trans[s][c]
                d.trans[s][c] = t if t >= 0 else D.start
    return d


def _no_keyword_rule(G, n, out, visited):
    if isinstance(n, C.Seq):
        for g in n.items:
            _no_keyword_rule(G, g, out, visited)
        return
    if isinstance(n, C.Lit):
        out[n.c & 0xFF] = True
        return
    if isinstance(n, C.Alt):
        for g in n.items:
            _no_keyword_rule(G, g, out, visited)
    elif isinstance(n, (C.Rep, C.NotP)):
        _no_keyword_rule(G, n.g, out, visited)
    elif isinstance(n, C.Rule):
        if n.tag in visited or C.is_lexical_rule(G, n.tag):
            return
        _no_keyword_rule(G, G.d(n.tag), out, visited + (n.tag,))


def _first_terms_no_nullable(G, n, visited=()):
    if isinstance(n, C.Seq):
        return _first_terms_no_nullable(G, n.items[0], visited) if n.items else set()
    return _orig["first_terms"](G, n, visited)


def _doubled_as_affine(G, tag, structural):
    d = _orig["make_desc"](G, tag, structural)
    if d.kind == C.DOUBLED:
        d.kind, d.escape = C.AFFINE, d.delim
    return d


def _guard_any_lits(n):
    ok, o, c = _orig["bracket_guard"](n)
    if not ok and isinstance(n, C.Seq) and len(n.items) >= 3 and C._is_lit(n.items[0]) \
            and C._is_lit(n.items[-1]) and n.items[0].c != n.items[-1].c:
        return True, n.items[0].c & 0xFF, n.items[-1].c & 0xFF
    return ok, o, c


def _mask_direct_only(D, structural):
    can = C.reach_accept(D)
    mask = [False] * D.n_states
    for i in range(D.n_states):
        mask[i] = any(structural[b] and D.next(i, b) >= 0 and can[D.next(i, b)] for b in range(256))
    mask[D.start] = False
    return mask


_orig = {k: getattr(C, k) for k in ("make_desc", "collect_lits", "first_terms", "bracket_guard",
                                    "masking_states")}
MUTANTS = {
    "no_reset_fold": ("make_desc", _no_reset_fold),
    "no_keyword_rule": ("collect_lits", _no_keyword_rule),
    "first_terms_ignores_nullable": ("first_terms", _first_terms_no_nullable),
    "doubled_as_affine": ("make_desc", _doubled_as_affine),
    "bracket_guard_any_lits": ("bracket_guard", _guard_any_lits),
    "mask_no_propagation": ("masking_states", _mask_direct_only),
}


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_planted_mistakes_are_caught(monkeypatch, mutant):
    fname, fn = MUTANTS[mutant]
    monkeypatch.setattr(C, fname, fn)
    changed = [n for n, (G, want) in _all_grammars().items() if C.make_scan_plan(G).to_bytes() != want]
    assert changed, f"{mutant} left every plan unchanged"


def test_arena_alt_dedup_is_needed(monkeypatch):
    """Without the alternation dedup the derivative DFA never closes (the state
    cap is hit) or changes shape."""
    def alt_no_dedup(self, a, b):
        tmp = []
        for x in (a, b):
            if self.k[x] == C.EMPTY:
                continue
            tmp.extend(self.kid[x] if self.k[x] == C.ALT else [x])
        tmp = tmp[:C.kMaxKids]
        if not tmp:
            return self.empty_id
        return tmp[0] if len(tmp) == 1 else self.mk(C.ALT, kids=tmp)
    monkeypatch.setattr(C.Arena, "alt", alt_no_dedup)
    changed = [n for n, (G, want) in _all_grammars().items() if C.make_scan_plan(G).to_bytes() != want]
    assert changed


# ---------------------------------------------------------------------------
SCANS = json.loads(_read(os.path.join(GOLD, "scans.json")))


@pytest.mark.parametrize("case", sorted(SCANS["cases"]))
def test_oracle_on_probe_plans(case):
    """The oracle reproduces the reference CPU backend on every probe plan."""
    rec = SCANS["cases"][case]
    d = W.alpha(rec["seed"], rec["n"], rec["alphabet"].encode("latin-1"))
    assert O.fnv(d) == rec["input_hash"]
    r = O.full(_fixture(rec["plan"]), d)
    assert r["positions"].size == rec["n_struct"] and r["bracket_pos"].size == rec["n_brackets"]
    for k in ("unclosed_mode", "balanced", "well_formed", "max_depth"):
        assert r[k] == rec[k], k
    for k, h in rec["hash"].items():
        assert O.fnv(r[k]) == h, k
    if "arrays" in rec:
        for k, v in rec["arrays"].items():
            assert np.array_equal(r[k], np.asarray(v, dtype=r[k].dtype)), k
