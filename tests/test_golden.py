"""Committed golden vectors (tests/golden/cas_fixtures.npz, written by make_golden.py) against the
oracle (CPU) and against the GPU solver (gpu)."""
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN

FIX = os.path.join(GOLDEN, "cas_fixtures.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(FIX, allow_pickle=False)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_generator_digests(oracle, gold):
    h, eri = oracle.seeded_integrals(8, 1)
    assert sha(h) == str(gold["cfg1_h_sha256"]) and sha(eri) == str(gold["cfg1_eri_sha256"])
    for n, seed in ((14, 2), (16, 3), (12, 2000), (10, 1000)):
        h, eri = oracle.seeded_integrals(n, seed)
        assert sha(h) == str(gold[f"seeded_{n}_{seed}_h_sha256"])
        assert sha(eri) == str(gold[f"seeded_{n}_{seed}_eri_sha256"])


def test_oracle_against_golden(oracle, gold):
    h, eri = gold["cfg1_h"], gold["cfg1_eri"]
    o = oracle.Oracle.from_chem(h, eri, 4, 4)
    r = o.solve(nroot=3, e_conv=1e-12)
    assert np.abs(r["energies"] - gold["cas88_E3"]).max() < 1e-9
    assert abs(r["energies"][0] - float(gold["cfg1_E0"])) < 1e-9
    assert np.array_equal(o.strings(0), gold["strings_8_4"])
    for p in range(8):
        for q in range(8):
            assert np.array_equal(np.stack([np.asarray(x, dtype=np.float64) for x in o.vo(0, p, q)]),
                                  gold[f"vo_8_4_{p}_{q}"])
    rd = o.rdms(r["evecs"][0])
    for name, g in zip(["g1a", "g1b", "g2aa", "g2ab", "g2bb"], rd):
        assert np.abs(g - gold["cfg1_" + name]).max() < 1e-8
    for n, seed in ((6, 66), (8, 88)):
        h2, eri2 = oracle.seeded_integrals(n, seed)
        o2 = oracle.Oracle.from_chem(h2, eri2, n // 2, n // 2, e_core=0.5)
        assert np.abs(o2.sigma(gold[f"sigma{n}{n}_C"]) - gold[f"sigma{n}{n}_S"]).max() < 1e-12


@pytest.mark.gpu
def test_gpu_against_golden(gold):
    import forte_amd
    h, eri = gold["cfg1_h"], gold["cfg1_eri"]
    ints = forte_amd.make_active_space_integrals(h, eri)
    s = forte_amd.FCISolver(4, 4, 1, 3, ints)
    s.set_e_convergence(1e-12)
    s.compute_energy()
    assert np.abs(np.array(s.energies()) - gold["cas88_E3"]).max() < 1e-9
    assert np.array_equal(s.strings(0), gold["strings_8_4"].astype(np.uint64))
    for p in range(8):
        for q in range(8):
            got = np.stack([np.asarray(x, dtype=np.float64) for x in s.vo_list(0, p, q)])
            assert np.array_equal(got, gold[f"vo_8_4_{p}_{q}"])
    for pq in range(28):
        got = np.stack([np.asarray(x, dtype=np.float64) for x in s.oo_list(0, pq)])
        assert np.array_equal(got, gold[f"oo_8_4_{pq}"])
    vv = []
    for p in range(8):
        for q in range(p):
            for r in range(8):
                for t in range(r):
                    if (p, q) != (r, t):
                        vv.extend(np.asarray(x, dtype=np.float64) for x in s.vvoo_list(0, p, q, r, t))
    assert sha(np.concatenate(vv)) == str(gold["vvoo_8_4_sha256"])
    rd = s.rdms([(0, 0)], 2)[0]
    for name in ["g1a", "g1b", "g2aa", "g2ab", "g2bb"]:
        assert np.abs(getattr(rd, name)() - gold["cfg1_" + name]).max() < 1e-8
    for n, seed in ((6, 66), (8, 88)):
        h2, eri2 = forte_amd.seeded_integrals(n, seed)
        s2 = forte_amd.FCISolver(n // 2, n // 2, 1, 1, forte_amd.make_active_space_integrals(h2, eri2, 0.5))
        ref = gold[f"sigma{n}{n}_S"]
        assert np.abs(s2.sigma(gold[f"sigma{n}{n}_C"]) - ref).max() < 1e-12 * np.abs(ref).max()
