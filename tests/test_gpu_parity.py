"""GPU parity: the HIP backend (through the C ABI) against the CPU oracle,
bit-exact on every output — structural positions and bytes, the structural
bitmap, bracket positions/bytes, bracket pairs in both conventions, and the
balanced / well_formed / max_depth / unclosed_mode flags.

Inputs: every golden case (the reference's KATs, traps, generated documents,
adversarial alphabets), size sweeps across the tile and lane boundaries,
unaligned inputs, repeated calls on one scanner, the host and chunked
boundaries, and the full-size workloads through size-independent properties.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from pars_amd import Plan, Scanner
from pars_amd import _native as N
from pars_amd import workloads as W
from tests.gpu_util import d2h, gpu_full, say, to_device

pytestmark = pytest.mark.gpu

PLAN_NAMES = ["json", "csv", "clike", "clike6", "toml", "ini", "xml"]
PLANS = {g: Plan.load(g) for g in PLAN_NAMES}
_scanners: dict = {}


def scanner(torch, g: str) -> Scanner:
    if g not in _scanners:
        sc = Scanner(PLANS[g])
        sc.set_option(N.PARS_OPT_WANT_CHARS, 1)
        _scanners[g] = sc
    return _scanners[g]


def compare(g: str, data: np.ndarray, got: dict, label: str = ""):
    exp = O.full(PLANS[g].blob, data)
    tag = f"{g}:{label}"
    assert got["n_struct"] == exp["positions"].size, tag
    assert got["n_brackets"] == exp["bracket_pos"].size, tag
    assert np.array_equal(got["positions"], exp["positions"]), tag
    if "chars" in got:
        assert np.array_equal(got["chars"], exp["chars"]), tag
    assert np.array_equal(got["bracket_pos"], exp["bracket_pos"]), tag
    assert np.array_equal(got["bracket_char"], exp["bracket_char"]), tag
    for k in ("balanced", "well_formed", "max_depth", "unclosed_mode"):
        assert got[k] == exp[k], f"{tag} {k}: gpu {got[k]} cpu {exp[k]}"
    if "pair" in got:
        assert np.array_equal(got["pair"], exp["pair"]), tag
    if "pair_pos" in got:
        assert np.array_equal(got["pair_pos"], O.pair_pos_device(exp["positions"], exp["bracket_pos"], exp["pair"])), tag
    if "struct_bitmap" in got:
        sb = O.scan(PLANS[g].blob, data, bitmaps=True)["struct_bitmap"]
        assert np.array_equal(got["struct_bitmap"], sb), tag


# ---------------------------------------------------------------------------
def test_golden_cases(torch_cuda, golden):
    from tests.test_oracle import golden_input

    for name, case in sorted(golden["cases"].items()):
        d = golden_input(case)
        say(f"[gpu] golden {name} plan={case['plan']} n={d.size}")
        got = gpu_full(torch_cuda, scanner(torch_cuda, case["plan"]), d)
        compare(case["plan"], d, got, name)
        assert got["n_struct"] == case["n_struct"]
        if d.size:
            assert O.fnv(got["positions"]) == case["hash"]["positions"]
            assert O.fnv(got["pair"]) == case["hash"]["pair"]


SIZES = [1, 2, 3, 31, 63, 64, 65, 127, 128, 1000, 16383, 16384, 16385, 32768 + 7, 5 * 16384 + 63,
         64 * 16384 + 1, 300_001]


@pytest.mark.parametrize("g", PLAN_NAMES)
def test_size_sweep_random_alphabet(torch_cuda, g):
    for n in SIZES:
        d = W.alpha(1000 + n, n, W.ALPHABETS[g])
        say(f"[gpu] sweep {g} n={n}")
        compare(g, d, gpu_full(torch_cuda, scanner(torch_cuda, g), d), f"n={n}")


@pytest.mark.parametrize("g,gen", [("json", "json"), ("csv", "csv"), ("clike6", "clike"), ("clike", "clike")])
def test_generated_documents_multi_tile(torch_cuda, g, gen):
    for n in (70_001, 1 << 20, (5 << 20) + 3):
        d = {"json": lambda: W.json(21, n, 16, 0.3), "csv": lambda: W.csv(21, n),
             "clike": lambda: W.clike(21, n)}[gen]()
        say(f"[gpu] generated {g} n={n}")
        compare(g, d, gpu_full(torch_cuda, scanner(torch_cuda, g), d), f"n={n}")


@pytest.mark.parametrize("path", [0, 1])
@pytest.mark.parametrize("rows", [1, 2, 3, 4])
@pytest.mark.parametrize("g", ["json", "csv", "clike6"])
def test_rows_per_tile_and_paths(torch_cuda, g, rows, path):
    """Every tile shape and both evaluators (fast path, general walk) agree with the oracle."""
    sc = Scanner(PLANS[g])
    sc.set_option(N.PARS_OPT_WANT_CHARS, 1)
    sc.set_option(N.PARS_OPT_ROWS_PER_TILE, rows)
    sc.set_option(N.PARS_OPT_PATH, path)
    gen = {"json": lambda s, n: W.json(s, n, 12, 0.3), "csv": W.csv, "clike6": W.clike}[g]
    for n in (100, 3 * 16384 + 5, (1 << 20) + 77, 3 << 20):
        for d in (gen(n, n), W.alpha(n + 1, n, W.ALPHABETS[g])):
            say(f"[gpu] rows={rows} path={path} {g} n={n}")
            compare(g, d, gpu_full(torch_cuda, sc, d), f"rows={rows} path={path} n={n}")
    sc.close()


@pytest.mark.parametrize("waves", [6, 8])
@pytest.mark.parametrize("rows", [2, 4])
@pytest.mark.parametrize("g", ["json", "csv"])
def test_waves_per_block(torch_cuda, g, rows, waves):
    """Six- and eight-wave blocks (24 and 32 KiB rows) of the bit-parallel
    evaluators agree with the oracle, also with the chunked boundary."""
    sc = Scanner(PLANS[g])
    sc.set_option(N.PARS_OPT_WANT_CHARS, 1)
    sc.set_option(N.PARS_OPT_ROWS_PER_TILE, rows)
    sc.set_option(N.PARS_OPT_WAVES_PER_BLOCK, waves)
    gen = {"json": lambda s, n: W.json(s, n, 12, 0.3), "csv": W.csv}[g]
    for n in (100, 5 * 24576 + 7, (1 << 20) + 77, 6 << 20):
        for d in (gen(n, n), W.alpha(n + 1, n, W.ALPHABETS[g])):
            say(f"[gpu] waves={waves} rows={rows} {g} n={n}")
            compare(g, d, gpu_full(torch_cuda, sc, d, want_pairs=True), f"waves={waves} rows={rows} n={n}")
    assert sc.get_option(N.PARS_OPT_WAVES_PER_BLOCK) == waves
    sc.close()
    # chunks of one document with the carry, cut inside rows of the wider blocks
    n = 900_001
    d = gen(n, n)
    ck = Scanner(PLANS[g])
    ck.set_option(N.PARS_OPT_ROWS_PER_TILE, rows)
    ck.set_option(N.PARS_OPT_WAVES_PER_BLOCK, waves)
    ck.reserve_struct(n)
    t, base = to_device(torch_cuda, d)
    carry = N.ChunkCarry()
    bounds = [0, 24576, 24577, 32768 * 9, 300_003, n]
    for a, b in zip(bounds[:-1], bounds[1:]):
        carry = ck.scan_chunk(base + a, b - a, carry, None, a)
    exp = O.full(PLANS[g].blob, d)
    assert np.array_equal(d2h(ck.struct_ptr(), carry.out_base, np.uint32), exp["positions"])
    r = ck.pair_chunks(carry, None, want_pairs=True)
    assert np.array_equal(d2h(r.bracket_pair, r.n_brackets, np.int32), exp["pair"])
    ck.close()


def test_unaligned_inputs(torch_cuda):
    d = W.json(31, 200_000)
    for off in (1, 3, 5, 8, 13):
        say(f"[gpu] unaligned offset={off}")
        compare("json", d, gpu_full(torch_cuda, scanner(torch_cuda, "json"), d, offset=off), f"off={off}")


def test_repeated_calls_shrinking_and_growing(torch_cuda):
    sc = scanner(torch_cuda, "json")
    for n in (3 << 20, 10, 1 << 20, 0, 77, 4 << 20, 16384):
        d = W.alpha(n, n, W.ALPHABETS["json"]) if n % 2 else W.json(n, n)
        say(f"[gpu] repeat n={n}")
        if n == 0:
            t, ptr = to_device(torch_cuda, d)
            r = sc.scan_device(ptr, 0)
            assert r.n_struct == 0 and r.balanced and r.well_formed and r.max_depth == 0
            continue
        compare("json", d, gpu_full(torch_cuda, sc, d), f"n={n}")


def test_without_pairs_and_flags_only(torch_cuda):
    d = W.json(41, 1 << 20)
    got = gpu_full(torch_cuda, scanner(torch_cuda, "json"), d, want_pairs=False, run_pairs=False)
    exp = O.full(PLANS["json"].blob, d)
    assert np.array_equal(got["positions"], exp["positions"])
    assert "pair" not in got and "pair_pos" not in got


def test_scan_host(torch_cuda):
    for g, d in (("json", W.json(51, 300_000)), ("clike6", W.clike(51, 300_000)),
                 ("json", W.alpha(52, 99_999, W.ALPHABETS["json"]))):
        say(f"[gpu] scan_host {g}")
        sc = scanner(torch_cuda, g)
        r = sc.scan_host(bytes(d), want_pairs=True)
        exp = O.full(PLANS[g].blob, d)
        assert np.array_equal(r.structural, exp["positions"])
        assert np.array_equal(r.pair_pos, O.pair_pos_device(exp["positions"], exp["bracket_pos"], exp["pair"]))
        assert (r.balanced, r.well_formed, r.max_depth, r.unclosed_mode) == (
            exp["balanced"], exp["well_formed"], exp["max_depth"], exp["unclosed_mode"])


@pytest.mark.parametrize("g", ["json", "csv", "clike6", "clike"])
def test_scan_chunk_any_boundaries(torch_cuda, g):
    """Chunked stage 1 with the carry, at boundaries inside strings, escapes
    and comments, then R6 across all chunks; equals the whole-document oracle."""
    n = 700_000
    d = {"json": lambda: W.json(61, n), "csv": lambda: W.csv(61, n), "clike6": lambda: W.clike(61, n),
         "clike": lambda: W.clike(61, n)}[g]()
    rng = np.random.default_rng(7)
    cuts = sorted(set(rng.integers(1, n, 9).tolist()) | {16384, 16385, 16384 * 3})
    bounds = [0] + cuts + [n]
    sc = Scanner(PLANS[g])
    sc.reserve_struct(n)
    t, base = to_device(torch_cuda, d)
    carry = N.ChunkCarry()
    for a, b in zip(bounds[:-1], bounds[1:]):
        say(f"[gpu] chunk {g} [{a},{b})")
        carry = sc.scan_chunk(base + a, b - a, carry, None, a)
    # a length or end past 2^32-1 is refused before anything is launched (a
    # negative length, wrapped to size_t, once slipped through)
    for bad_n, bad_base in ((2**64 - 5091, 0), (100, 2**32 - 50)):
        with pytest.raises(Exception):
            sc.scan_chunk(base, bad_n, carry, None, bad_base)
    exp = O.full(PLANS[g].blob, d)
    assert carry.out_base == exp["positions"].size
    got = d2h(sc.struct_ptr(), carry.out_base, np.uint32)
    assert np.array_equal(got, exp["positions"])
    fs = O.final_states(PLANS[g].blob, d)
    nd = PLANS[g].n_descs
    assert [(carry.scan >> (3 * i)) & 7 for i in range(nd)] == fs[:nd].tolist()
    r = sc.pair_chunks(carry, None, want_pairs=True)
    assert (r.balanced, r.well_formed, r.max_depth, r.unclosed_mode) == (
        exp["balanced"], exp["well_formed"], exp["max_depth"], exp["unclosed_mode"])
    assert np.array_equal(d2h(r.bracket_pair, r.n_brackets, np.int32), exp["pair"])
    assert np.array_equal(d2h(r.pair_pos, r.n_struct, np.int32),
                          O.pair_pos_device(exp["positions"], exp["bracket_pos"], exp["pair"]))
    sc.close()


def test_bracket_capacity_regrow(torch_cuda):
    """All-bracket input exceeds the optimistic bracket capacity: the call regrows and reruns."""
    d = np.frombuffer(b"[" * 300_000 + b"]" * 300_000 + b"{}" * 50_000, np.uint8)
    sc = Scanner(PLANS["json"])
    say("[gpu] regrow")
    got = gpu_full(torch_cuda, sc, d)
    compare("json", d, got, "regrow")
    assert got["max_depth"] == 300_000
    sc.close()


# ---------------------------------------------------------------------------
# full-size workloads: exact against the oracle plus size-independent properties
# ---------------------------------------------------------------------------
def _properties(g: str, d: np.ndarray, got: dict):
    pos = got["positions"]
    assert np.all(np.diff(pos.astype(np.int64)) > 0)
    cls = O.classifier(PLANS[g].blob)
    assert np.all(cls[d[pos]] & 1)
    assert int(np.unpackbits(got["struct_bitmap"].view(np.uint8)).sum()) == pos.size
    p = got["pair"]
    idx = np.nonzero(p >= 0)[0]
    assert np.array_equal(p[p[idx]], idx)


def test_json_2gib_full_size(torch_cuda):
    n = 1 << 31
    say("[gpu] 2GiB json: generate")
    d = W.json(2, n, 32, 0.30)
    sc = Scanner(PLANS["json"])
    say("[gpu] 2GiB json: scan")
    got = gpu_full(torch_cuda, sc, d)
    _properties("json", d, got)
    assert got["balanced"] and got["well_formed"] and got["max_depth"] == 32
    say("[gpu] 2GiB json: oracle")
    exp = O.full(PLANS["json"].blob, d)
    assert np.array_equal(got["positions"], exp["positions"])
    assert np.array_equal(got["pair"], exp["pair"])
    assert got["unclosed_mode"] == exp["unclosed_mode"]
    sc.close()


def test_clike6_1gib_full_size(torch_cuda):
    n = 1 << 30
    say("[gpu] 1GiB clike6: generate")
    d = W.clike(3, n, 24)
    sc = Scanner(PLANS["clike6"])
    say("[gpu] 1GiB clike6: scan")
    got = gpu_full(torch_cuda, sc, d)
    _properties("clike6", d, got)
    say("[gpu] 1GiB clike6: oracle")
    exp = O.full(PLANS["clike6"].blob, d)
    assert np.array_equal(got["positions"], exp["positions"])
    assert np.array_equal(got["pair"], exp["pair"])
    assert (got["balanced"], got["max_depth"]) == (exp["balanced"], exp["max_depth"])
    sc.close()


def test_cpp_facade_end_to_end(torch_cuda, tmp_path):
    """The C++ facade, driven like gpu::scanner<S>, equals the oracle."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(__file__), "cpp", "facade_check")
    d = W.json(71, 50_000)
    inp = tmp_path / "in.json"
    d.tofile(inp)
    say("[gpu] C++ facade")
    r = subprocess.run([exe, os.path.join(os.path.dirname(os.path.dirname(__file__)), "pars_amd", "plans",
                                          "json.plan"), str(inp)], stdout=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0
    lines = r.stdout.split()
    k, bal, wf, md = (int(v) for v in lines[:4])
    exp = O.full(PLANS["json"].blob, d)
    assert k == exp["positions"].size and (bool(bal), bool(wf), md) == (exp["balanced"], exp["well_formed"],
                                                                       exp["max_depth"])
    assert np.array_equal(np.array(lines[4:], dtype=np.int64), exp["positions"].astype(np.int64))
