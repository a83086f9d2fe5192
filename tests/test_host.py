"""CPU tests of the drop-in boundary: the C ABI exports, the C++ host
(checkpoint v1 bytes, CLI parsing, --print / --print-expression).

The checkpoint layout is restated here from checkpoint.cpp:63-143,172-198 of
the reference (native endianness, std::string::size_type = u64 lengths):
  u32 version (1), u32 seq_id, u64 len + seq_params bytes,
  u64 n0, n1, f0, f1, u64 len + output path bytes, u32 max_factors_per_number,
  u64 last_processed_prime,
  u64 numbers_count, u32 max, u64 n0, then per index: u32 count followed by
  min(count, max) u64 factors.
"""
from __future__ import annotations

import ctypes
import json
import os
import re
import struct
import subprocess
import sys
from fractions import Fraction

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cutrialdive_amd")
LIB = os.path.join(PKG, "libctd_hip.so")
CLI = os.path.join(PKG, "cutrialdive")
HEADER = os.path.join(ROOT, "include", "ctd_hip.h")
KATS = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_kats.json")))

sys.set_int_max_str_digits(1 << 20)


@pytest.fixture(scope="session")
def built():
    if not (os.path.exists(LIB) and os.path.exists(CLI)):
        from cutrialdive_amd import build

        build.build_all()
    return CLI


@pytest.fixture(scope="session")
def ckpt_tool(built, tmp_path_factory):
    out = str(tmp_path_factory.mktemp("native") / "ckpt_tool")
    subprocess.run(
        ["g++", "-O1", "-std=c++20", f"-I{ROOT}/include", f"-I{PKG}/host", "-o", out,
         os.path.join(ROOT, "tests", "native", "ckpt_tool.cpp"), os.path.join(PKG, "host", "ctd_host.cpp"),
         f"-L{PKG}", "-lctd_hip", f"-Wl,-rpath,{PKG}"],
        check=True)
    return out


def run_cli(*args, check_rc=None):
    r = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=60)
    if check_rc is not None:
        assert r.returncode == check_rc, (r.returncode, r.stdout, r.stderr)
    return r


# ---------------------------------------------------------------- C ABI exports

def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(ctd_hip_\w+)\s*\(", text)))


def test_header_declares_the_engine_entry_points():
    fns = header_functions()
    for f in ["ctd_hip_open", "ctd_hip_close", "ctd_hip_last_error", "ctd_hip_trial_factor",
              "ctd_hip_trial_factor_device", "ctd_hip_sieve", "ctd_hip_abi_version"]:
        assert f in fns


def test_library_exports_every_declared_symbol(built):
    lib = ctypes.CDLL(LIB)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing


def test_python_binding_covers_the_header(built):
    from cutrialdive_amd import _native

    assert set(header_functions()) <= set(_native.SIGNATURES)


def test_missing_library_fails_loudly(tmp_path):
    # the product path has no CPU fallback: without the HIP library it raises
    from cutrialdive_amd import _native

    with pytest.raises(RuntimeError, match="build the HIP engine"):
        _native.load(str(tmp_path / "libctd_hip.so"))


def test_abi_version_without_a_device(built):
    lib = ctypes.CDLL(LIB)
    lib.ctd_hip_abi_version.restype = ctypes.c_uint32
    assert lib.ctd_hip_abi_version() >= 1


# ---------------------------------------------------------------- checkpoint v1 bytes

def ckpt_parse(data: bytes):
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from("<" + fmt, data, off)
        off += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    def take_str():
        nonlocal off
        n = take("Q")
        s = data[off:off + n].decode()
        off += n
        return s

    d = {"version": take("I"), "seq_id": take("I"), "params": take_str()}
    d["n0"], d["n1"], d["f0"], d["f1"] = take("QQQQ")
    d["output"] = take_str()
    d["max"] = take("I")
    d["last"] = take("Q")
    cnt, mx, n0 = take("Q"), take("I"), take("Q")
    d["buffer"] = (cnt, mx, n0)
    rows = []
    for _ in range(cnt):
        c = take("I")
        k = min(c, mx)
        rows.append((c, list(struct.unpack_from(f"<{k}Q", data, off))))
        off += 8 * k
    d["rows"] = rows
    assert off == len(data), "trailing bytes"
    return d


def ckpt_bytes(seq_id, params, n0, n1, f0, f1, output, mx, last, rows, version=1):
    b = struct.pack("<II", version, seq_id)
    b += struct.pack("<Q", len(params)) + params.encode()
    b += struct.pack("<QQQQ", n0, n1, f0, f1)
    b += struct.pack("<Q", len(output)) + output.encode()
    b += struct.pack("<IQ", mx, last)
    b += struct.pack("<QIQ", len(rows), mx, n0)
    for c, fs in rows:
        assert len(fs) == min(c, mx)
        b += struct.pack("<I", c) + struct.pack(f"<{len(fs)}Q", *fs)
    return b


def tool_rows(n, mx):
    rows = []
    for i in range(n):
        c = (5 * i + 3) % (mx + 3)
        rows.append((c, [1000003 * i + 17 * j for j in range(min(c, mx))]))
    return rows


def test_checkpoint_write_layout(ckpt_tool, tmp_path):
    p = tmp_path / "o.txt.chkpnt"
    subprocess.run([ckpt_tool, "write", str(p), "3", "16", "5", "40", "1099511627776", "1099511627776000",
                    "/data/o.txt", "4", "1099511628211"], check=True)
    d = ckpt_parse(p.read_bytes())
    assert d["version"] == 1 and d["seq_id"] == 3 and d["params"] == "16"
    assert (d["n0"], d["n1"], d["f0"], d["f1"]) == (5, 40, 1 << 40, 1099511627776000)
    assert d["output"] == "/data/o.txt" and d["max"] == 4 and d["last"] == 1099511628211
    assert d["buffer"] == (35, 4, 5)
    assert d["rows"] == tool_rows(35, 4)
    # the temporary file is renamed over the checkpoint
    assert [f.name for f in tmp_path.iterdir()] == ["o.txt.chkpnt"]


def test_checkpoint_write_without_output_path(ckpt_tool, tmp_path):
    p = tmp_path / "c"
    subprocess.run([ckpt_tool, "write", str(p), "0", "", "2", "3", "2", "1024", "-", "1", "1021"], check=True)
    d = ckpt_parse(p.read_bytes())
    assert d["params"] == "" and d["output"] == "" and d["rows"] == tool_rows(1, 1)


def test_checkpoint_load_reference_bytes(ckpt_tool, tmp_path):
    rows = [(0, []), (2, [3, 2063]), (5, [7, 11, 13]), (3, [5, 17, 257])]
    p = tmp_path / "r.chkpnt"
    p.write_bytes(ckpt_bytes(1, "1279", 76, 80, 2, 1 << 16, "out.txt", 3, 65521, rows))
    r = subprocess.run([ckpt_tool, "dump", str(p)], capture_output=True, text=True, check=True)
    lines = r.stdout.splitlines()
    assert lines[:10] == ["seq_id 1", "params 1279", "n0 76", "n1 80", "f0 2", "f1 65536", "output out.txt",
                          "max 3", "last 65521", "buffer 4 3 76"]
    assert lines[10:] == ["0:", "2: 3 2063", "5: 7 11 13", "3: 5 17 257"]


def test_checkpoint_round_trip_is_byte_identical(ckpt_tool, tmp_path):
    p = tmp_path / "a"
    subprocess.run([ckpt_tool, "write", str(p), "8", "", "100", "164", "3", "99999", "x/y.txt", "2", "99991"],
                   check=True)
    d = ckpt_parse(p.read_bytes())
    again = ckpt_bytes(d["seq_id"], d["params"], d["n0"], d["n1"], d["f0"], d["f1"], d["output"], d["max"],
                       d["last"], d["rows"])
    assert again == p.read_bytes()


def test_checkpoint_rejects_other_versions(ckpt_tool, tmp_path):
    p = tmp_path / "v2"
    p.write_bytes(ckpt_bytes(0, "", 2, 3, 2, 5, "o", 1, 3, [(1, [3])], version=2))
    r = subprocess.run([ckpt_tool, "dump", str(p)], capture_output=True, text=True)
    assert r.returncode == 1
    assert "Unexpected version number `2' found in checkpoint file" in r.stderr


def test_checkpoint_missing_file(ckpt_tool, tmp_path):
    r = subprocess.run([ckpt_tool, "dump", str(tmp_path / "nope")], capture_output=True, text=True)
    assert r.returncode == 3 and r.stdout.strip() == "missing"


def test_cli_resume_of_missing_checkpoint(built, tmp_path):
    r = run_cli("--resume", tmp_path / "nope.chkpnt", check_rc=1)
    assert "Failed to load checkpoint data from file" in r.stderr


# ---------------------------------------------------------------- CLI parsing (command_line.cpp:223-373)

def test_cli_usage_help_version(built):
    r = run_cli(check_rc=1)
    assert "Type 'cutrialdive --help' for more." in r.stderr
    assert "--num-seq" in run_cli("--help", check_rc=0).stdout
    assert "--resume" in run_cli("-h", check_rc=0).stdout
    assert run_cli("--version", check_rc=0).stdout.startswith("cutrialdive version ")


@pytest.mark.parametrize("args,msg", [
    (["-s", "3", "--num-seq", "mersenne"], "Invalid command line, option --num-seq must come first"),
    (["--num-seq", "foo", "-s", "1"], "Invalid number sequence id `foo'"),
    (["--num-seq", "proth,", "-s", "1"], "Invalid value for option --num-seq"),
    (["--num-seq", "mersenne", "-s", "2", "-e", "9"], "Missing mandatory option"),
    (["--num-seq", "mersenne", "-s", "2", "--tf-start", "3"], "Missing mandatory option"),
    (["--num-seq", "mersenne", "-s"], "Missing value for option -s"),
    (["--num-seq", "mersenne", "-s", "2", "--tf-bits", "65"],
     "The value of option --tf-bits must be an integer between 0 and 64"),
    # the check runs on the option after the last range option (command_line.cpp:334-337)
    (["--num-seq", "mersenne", "-s", "2", "--tf-bits", "20", "--tf-start", "3", "--no-progress"],
     "Mixing --tf-bits with --tf-start and --tf-end is not allowed"),
    (["--num-seq", "mersenne", "-s", "2", "--tf-bits", "20", "--segment-bits", "64"],
     "value of option --segment-bits must be an integer in range [0, 63]"),
    (["--num-seq", "mersenne", "-s", "2", "--tf-bits", "20", "--segment-bits", "20", "--segment-length", "5"],
     "Invalid command line, --segment-length is not allowed when --segment-bits is present"),
    (["--num-seq", "mersenne", "--verbosity", "5"], "Invalid value for option --verbosity"),
    (["--checkpoint-period", "5", "--num-seq", "mersenne"], "option --checkpoint-period is unexpected"),
    (["--num-seq", "proth,x", "-s", "1", "--tf-bits", "10"], "Missing or invalid k value `x' in Proth number spec"),
    (["--num-seq", "riesel", "-s", "1", "--tf-bits", "10"], "Missing or invalid k value `' in Riesel sequence spec"),
    (["--num-seq", "smarandache,65", "-s", "1", "--tf-bits", "10"],
     "Base `65' is not supported for Smarandache numbers"),
    (["--num-seq", "proth,0", "-s", "1", "--tf-bits", "10"], "Proth parameter k must be non-zero"),
    (["--num-seq", "proth_fixed_n,0", "-s", "1", "--tf-bits", "10"], "Invalid n value zero"),
])
def test_cli_parse_errors(built, args, msg):
    r = run_cli(*args, check_rc=1)
    assert msg in r.stderr, r.stderr


@pytest.mark.parametrize("args", [
    ["--num-seq", "mersenne", "-s", "5", "-e", "5", "--tf-bits", "20"],
    ["--num-seq", "mersenne", "-s", "5", "--tf-start", "100", "--tf-end", "100"],
])
def test_cli_empty_ranges_need_no_device(built, args):
    r = run_cli(*args, check_rc=0)
    assert "is empty" in r.stdout


# ---------------------------------------------------------------- --print / --print-expression

def concat(b, n):
    v = 0
    for i in range(1, n + 1):
        d, t = 1, i
        while t >= b:
            t //= b
            d += 1
        v = v * b ** d + i
    return v


def evaluate(expr: str):
    return eval(re.sub(r"(\d+)", r"Fraction(\1)", expr.replace("^", "**")), {"Fraction": Fraction})


def test_smarandache_expression_reference_strings(built):
    # smarandache_tests.cpp:311-323
    for n, s in KATS["smarandache_expression_base10"].items():
        assert run_cli("--num-seq", "smarandache", "--print-expression", n, check_rc=0).stdout.strip() == s


# n lists of smarandache_gp_tests.cpp:153-212 (checked there against PARI)
GP_CASES = {
    10: [0, 1, 2, 3, 4, 5, 6, 9, 10, 11, 16, 29, 77, 100, 999, 1000, 1001],
    2: [0, 1, 2, 3, 4, 5, 7, 8, 9, 15, 16, 17, 31, 32, 33, 1647],
    3: [0, 1, 2, 3, 4, 39, 81, 82, 145],
    4: [0, 1, 2, 3, 4, 5, 6, 15, 16, 17],
    16: [0, 1, 2, 3, 8, 15, 16, 17, 19, 31, 255, 256, 257],
    17: [0, 1, 2, 16, 17, 18, 31, 288, 289, 290],
}


@pytest.mark.parametrize("base", sorted(GP_CASES))
def test_smarandache_expression_and_value_evaluate_to_concatenation(built, base):
    for n in GP_CASES[base]:
        e = run_cli("--num-seq", f"smarandache,{base}", "--print-expression", n, check_rc=0).stdout.strip()
        v = run_cli("--num-seq", f"smarandache,{base}", "--print", n, check_rc=0).stdout.strip()
        assert evaluate(e) == concat(base, n), (base, n)
        assert int(v) == concat(base, n), (base, n)


def test_small_gp_values(built):
    # smarandache_gp_tests.cpp:155-196 literal values
    for b, n, v in [(10, 11, 1234567891011), (2, 2, 6), (2, 3, 27), (2, 4, 220), (3, 2, 5), (3, 3, 48),
                    (3, 4, 436), (4, 2, 6), (4, 3, 27), (4, 4, 436), (4, 5, 6981)]:
        assert int(run_cli("--num-seq", f"smarandache,{b}", "--print", n, check_rc=0).stdout) == v


def fib_pair(n):
    a, b = 0, 1
    for _ in range(n):
        a, b = b, a + b
    return a


def trib(n):
    t = [0, 0, 1]
    while len(t) <= n:
        t.append(t[-1] + t[-2] + t[-3])
    return t[n]


@pytest.mark.parametrize("spec,n,value", [
    ("mersenne", 0, 0), ("mersenne", 1, 1), ("mersenne", 127, 2 ** 127 - 1),
    ("proth,3", 70, 3 * 2 ** 70 + 1), ("riesel,12345678910111213", 200, 12345678910111213 * 2 ** 200 - 1),
    ("proth_fixed_n,5", 7, 7 * 32 + 1), ("riesel_fixed_n,100", 9, 9 * 2 ** 100 - 1),
    ("fibonacci", 0, 0), ("fibonacci", 1, 1), ("fibonacci", 1000, fib_pair(1000)),
    ("lucas", 0, 2), ("lucas", 1, 1), ("lucas", 500, fib_pair(499) + fib_pair(501)),
    ("tribonacci", 0, 0), ("tribonacci", 2, 1), ("tribonacci", 700, trib(700)),
])
def test_print_value(built, spec, n, value):
    assert int(run_cli("--num-seq", spec, "--print", n, check_rc=0).stdout) == value


@pytest.mark.parametrize("spec,n,text", [
    ("mersenne", 7, "2^7-1"),
    ("proth,3", 5, "3*2^5 + 1"),
    ("riesel,5", 9, "5*2^9 - 1"),
    ("proth_fixed_n,12", 3, "3*2^12+1"),
    ("riesel_fixed_n,12", 3, "3*2^12-1"),
    ("fibonacci", 3, "Expression not available"),
    ("tribonacci", 3, "Expression not available"),
])
def test_print_expression(built, spec, n, text):
    assert run_cli("--num-seq", spec, "--print-expression", n, check_rc=0).stdout.strip() == text
