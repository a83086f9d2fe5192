"""CPU tests of the window arithmetic shared by the device sieve and the host
schedulers (cutrialdive_amd/csrc/ctd_ranges.hpp).

Prime windows reach f1 = 2^64 - 1 (--tf-bits 64, command_line.cpp:317-325 of
the reference), so every segment, batch and tile bound near the top of the
range must saturate instead of wrapping.  The header is compiled with g++ into
tests/native/ranges_tool.cpp and every answer is checked here against Python's
exact integers.
"""
from __future__ import annotations

import math
import os
import random
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cutrialdive_amd", "csrc")
M64 = (1 << 64) - 1
SEG_SPAN = 1 << 20  # integers per sieve segment (ctd_sieve.hpp)


@pytest.fixture(scope="module")
def tool(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("native") / "ranges_tool")
    subprocess.run(["g++", "-O1", "-std=c++20", "-Wall", "-Werror", f"-I{CSRC}", "-o", out,
                    os.path.join(ROOT, "tests", "native", "ranges_tool.cpp")], check=True)
    return out


def ask(tool, lines):
    r = subprocess.run([tool], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True, timeout=60)
    return r.stdout.split("\n")[:len(lines)]


def test_isqrt_every_width_and_near_2_64(tool):
    rng = random.Random(5)
    xs = [0, 1, 2, 3, 4, 15, 16, 17, M64, M64 - 1, (1 << 64) - (1 << 33), ((1 << 32) - 1) ** 2,
          ((1 << 32) - 1) ** 2 - 1, ((1 << 32) - 1) ** 2 + 1, (1 << 62), (1 << 62) - 1]
    xs += [rng.getrandbits(rng.randint(1, 64)) for _ in range(400)]
    xs += [(r * r + d) & M64 for r in (rng.getrandbits(32) for _ in range(100)) for d in (-1, 0, 1) if 0 <= r * r + d <= M64]
    got = ask(tool, [f"isqrt {x}" for x in xs])
    for x, g in zip(xs, got):
        assert int(g) == math.isqrt(x), x
    # the sieving base of f1 = 2^64 - 1 ends at 2^32 - 1, not 2^32
    assert int(ask(tool, [f"isqrt {M64 - 1}"])[0]) == (1 << 32) - 1


def test_sat_add_and_block_end(tool):
    cases = [(0, 0), (M64, 0), (M64, 1), (M64 - SEG_SPAN + 1, SEG_SPAN - 1), (M64 - SEG_SPAN + 2, SEG_SPAN - 1),
             (1 << 63, 1 << 63), ((1 << 63) - 1, 1 << 63)]
    for (a, b), g in zip(cases, ask(tool, [f"sat_add {a} {b}" for a, b in cases])):
        assert int(g) == min(a + b, M64)
    ends = [((1 << 64) - (1 << 20) + 2, SEG_SPAN, M64), (M64 - 1, SEG_SPAN, M64), (0, SEG_SPAN, M64),
            (100, 1 << 40, 200), (M64 - 10, M64, M64), (5, 3, 6)]
    for (lo, ln, end), g in zip(ends, ask(tool, [f"block_end {lo} {ln} {end}" for lo, ln, end in ends])):
        assert int(g) == min(lo + ln, end)


def test_seg_last_saturates_at_the_top_segment(tool):
    # --tf-bits 64: f0 = 2, so segments start at even lo = 2 + k 2^20; the last
    # starts above 2^64 - 2^20 and its lo + span - 1 would wrap to a small number
    lo_top = 2 + ((M64 - 2) // SEG_SPAN) * SEG_SPAN
    assert lo_top > (1 << 64) - SEG_SPAN
    cases = [(lo_top, SEG_SPAN, M64), ((1 << 64) - 1000 & ~1, SEG_SPAN, M64), (0, SEG_SPAN, M64),
             (1 << 40, SEG_SPAN, (1 << 40) + 12345), (1 << 40, SEG_SPAN, 1 << 41)]
    for (lo, sp, end), g in zip(cases, ask(tool, [f"seg_last {a} {b} {c}" for a, b, c in cases])):
        want = min(lo + sp, end) - 1
        assert int(g) == want
        # every sieving number d with d^2 <= the segment's last candidate marks
        # in it: near 2^64 that includes every d < 2^32
        if lo == lo_top:
            assert int(g) >= ((1 << 32) - 1) ** 2


def test_whole_spans_never_rounds_to_zero(tool):
    cases = [(0, SEG_SPAN, 1 << 20), (1, SEG_SPAN, 1 << 20), (SEG_SPAN, SEG_SPAN, 1 << 20),
             (SEG_SPAN + 1, SEG_SPAN, 1 << 20), (M64, SEG_SPAN, 1 << 20), (1 << 63, SEG_SPAN, 1 << 20),
             (1 << 31, SEG_SPAN, 1 << 20)]
    for (ln, sp, mx), g in zip(cases, ask(tool, [f"whole_spans {a} {b} {c}" for a, b, c in cases])):
        want = min(max(-(-ln // sp), 1), mx) * sp
        assert int(g) == want and int(g) > 0


@pytest.mark.parametrize("lo0,length,end", [
    (2, 1 << 40, M64),                                   # whole 64-bit range, 2^40-integer blocks
    ((1 << 64) - (1 << 20) + 12344, SEG_SPAN, M64),      # a window inside the last 2^20
    ((1 << 64) - (1 << 22), 1 << 21, M64),               # tiles that straddle 2^64 - 2^20
    (M64 - 1, SEG_SPAN, M64),                            # one candidate
    (1 << 40, 1 << 28, (1 << 40) + (1 << 30) + 7),
])
def test_block_loops_cover_the_window_exactly(tool, lo0, length, end):
    n, covered = map(int, ask(tool, [f"blocks {lo0} {length} {end}"])[0].split())
    assert covered == end - lo0
    assert n == -(-(end - lo0) // length)
