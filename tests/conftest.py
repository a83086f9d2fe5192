import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libctd_hip.so on cuda:0)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def engine():
    import cutrialdive_amd as ctd

    eng = ctd.Engine(0)
    yield eng
    eng.close()


@pytest.fixture(scope="session")
def mersenne_lists(tmp_path_factory):
    """Factor lists of the Mersenne numbers M_n, n in [n0, D), over the odd
    primes below 2^B, from the multiplicative orders of 2 that
    tests/native/mersenne_orders.cpp finds without stepping any index:
    p | 2^n - 1 <=> ord_p(2) | n.  Returns a function (B, D, n0) -> one
    ascending list per index."""
    import subprocess

    import numpy as np

    d = tmp_path_factory.mktemp("orders")
    exe = str(d / "mersenne_orders")
    subprocess.run(["g++", "-O2", "-fopenmp", "-o", exe, os.path.join(ROOT, "tests", "native", "mersenne_orders.cpp")],
                   check=True)

    def lists(B, D, n0=2):
        out = str(d / f"orders_{B}_{D}.bin")
        subprocess.run([exe, str(B), str(D), out], check=True, capture_output=True, timeout=600)
        pairs = np.fromfile(out, dtype="<u4").reshape(-1, 2)
        res = [[] for _ in range(n0, D)]
        for p, order in pairs.tolist():  # ascending p: every list comes out sorted
            for n in range(max(order, (n0 + order - 1) // order * order), D, order):
                res[n - n0].append(p)
        return res

    return lists
