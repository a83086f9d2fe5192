"""The `cutrialdive` binary on the GPU: output files byte-identical to the
oracle's factor lists printed in the reference format
(factoring_results.hpp:183-201: "<n> - p1, p2\\n", ascending, "<n> - \\n" when
empty), and --resume from a checkpoint in the reference's v1 layout giving
the same file as an uninterrupted run."""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

from oracle import oracle

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_host import CLI, ckpt_bytes, ckpt_parse  # noqa: E402

pytestmark = pytest.mark.gpu


def text_of(lists, n0):
    return "".join(f"{n0 + i} - {', '.join(str(p) for p in fs)}\n" for i, fs in enumerate(lists))


def same_text(got: str, expect: str):
    g, e = got.splitlines(), expect.splitlines()
    bad = [(a, b) for a, b in zip(g, e) if a != b]
    assert len(g) == len(e) and not bad, (len(g), len(e), bad[:3])
    assert got == expect


def cli(*args, timeout=120):
    print("cli:", *args, flush=True)
    r = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert "Error" not in r.stdout, r.stdout
    return r


@pytest.mark.parametrize("spec,seq,param,n0,n1,f0,f1", [
    ("proth,1279", 1, 1279, 0, 80, 2, 1 << 16),
    ("mersenne", 0, 0, 2, 400, 2, 1 << 24),
    ("smarandache", 3, 10, 10, 40, 2, 1 << 28),
    ("riesel,5", 2, 5, 1, 300, 1 << 20, (1 << 20) + (1 << 22)),
    ("fibonacci", 4, 0, 1, 500, 3, 1 << 22),
    ("lucas", 5, 0, 1, 500, 2, 1 << 22),
    ("tribonacci", 8, 0, 2, 400, 2, 1 << 22),
    ("proth_fixed_n,100", 6, 100, 1, 300, 2, 1 << 22),
    ("riesel_fixed_n,33", 7, 33, 2, 300, 2, 1 << 22),
    ("smarandache,7", 3, 7, 1, 200, (1 << 33) - (1 << 20), 1 << 33),
])
def test_output_file_matches_oracle(tmp_path, spec, seq, param, n0, n1, f0, f1):
    out = tmp_path / "out.txt"
    cli("--num-seq", spec, "-s", n0, "-e", n1, "--tf-start", f0, "--tf-end", f1, "--no-progress",
        "--max-factors-per-number", 1024, "--output", out)
    expect = text_of(oracle.trial_factor(seq, param, n0, n1, f0, f1, 1024), n0)
    same_text(out.read_text(), expect)
    assert not os.path.exists(str(out) + ".chkpnt")


def test_reference_line_on_the_console():
    r = cli("--num-seq", "proth,1279", "-s", "79", "--tf-bits", "16", "--no-progress")
    assert "79 - 3, 2063\n" in r.stdout
    assert "Trial factoring took" in r.stdout


def test_readme_lines_with_tf_bits():
    # README.md:64-70 of the reference: --tf-bits 37 includes the prime 2
    r = cli("--num-seq", "smarandache", "-s", "10", "-e", "13", "--tf-bits", "37", "--no-progress")
    assert "10 - 2, 5, 1234567891\n11 - 3, 7, 13, 67, 107, 630803\n12 - 2, 3, 2437, 2110805449\n" in r.stdout


def test_resume_equals_uninterrupted_run(tmp_path):
    seq, param, n0, n1, f0, f1, mx = 1, 3, 1, 400, 2, 1 << 23, 64
    primes = oracle.sieve(2, 1 << 22)
    last = int(primes[-1])
    counts, factors = oracle.trial_factor_raw(seq, param, n0, n1, f0, last + 1, mx)
    assert int(counts.max()) <= mx
    rows = []
    for i, c in enumerate(counts.tolist()):
        rows.append((c, sorted(int(x) for x in factors[i * mx:i * mx + c])))
    out = tmp_path / "o.txt"
    ck = tmp_path / "o.txt.chkpnt"
    ck.write_bytes(ckpt_bytes(seq, str(param), n0, n1, f0, f1, str(out), mx, last, rows))
    assert ckpt_parse(ck.read_bytes())["last"] == last
    r = cli("--resume", ck, "--no-progress", "--checkpoint-period", "1")
    assert f"Resuming at the smallest prime > {last}." in r.stdout
    expect = text_of(oracle.trial_factor(seq, param, n0, n1, f0, f1, mx), n0)
    same_text(out.read_text(), expect)
    assert not ck.exists()


def test_resume_keeps_counts_beyond_max(tmp_path):
    # a checkpoint whose counts exceed max: the stored subset stays, counts keep adding up
    seq, param, n0, n1, f0, f1, mx = 0, 0, 2, 40, 2, 1 << 20, 2
    last = 1 << 10
    counts, factors = oracle.trial_factor_raw(seq, param, n0, n1, f0, last, 1024)
    rows = [(int(c), sorted(int(x) for x in factors[i * 1024:i * 1024 + c])[:mx]) for i, c in enumerate(counts)]
    out = tmp_path / "m.txt"
    ck = tmp_path / "m.txt.chkpnt"
    ck.write_bytes(ckpt_bytes(seq, "", n0, n1, f0, f1, str(out), mx, last - 1, rows))
    cli("--resume", ck, "--no-progress")
    full = oracle.trial_factor(seq, param, n0, n1, f0, f1, 1024)
    lines = out.read_text().splitlines()
    for i, line in enumerate(lines):
        n, _, rest = line.partition(" - ")
        assert int(n) == n0 + i
        got = [int(x) for x in rest.split(", ")] if rest else []
        assert len(got) == min(len(full[i]), mx)
        assert set(got) <= set(full[i])
        # the factors stored in the checkpoint are kept
        assert set(rows[i][1]) <= set(got)


def test_autotest_passes():
    r = cli("--autotest")
    assert "All self tests passed." in r.stdout
