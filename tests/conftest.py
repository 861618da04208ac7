import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REF_TESTS = "/root/reference/tests/methods"
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


def seeded_problem(n, na, nb, seed, e_core=0.0):
    """Chemist integrals from the seeded generator (oracle copy) in Forte's layout."""
    from oracle import oracle as O
    h, eri = O.seeded_integrals(n, seed)
    aa, ab = O.chem_to_forte(eri)
    return dict(n=n, na=na, nb=nb, h=h, eri=eri, aa=aa, ab=ab, e_core=e_core)


def fcidump_path(k):
    """INTDUMP of the reference's integrals-fcidump-k test, copied under tests/golden."""
    return os.path.join(GOLDEN, f"integrals-fcidump-{k}.INTDUMP")
