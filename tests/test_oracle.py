"""The CPU oracle, pinned against the reference's own known answers and against independent
brute force. Nothing here needs a GPU."""
import itertools
from math import comb

import numpy as np
import pytest

from conftest import fcidump_path, seeded_problem

# Known-answer energies of the reference's FCIDUMP tests (tests/methods/integrals-fcidump-k/input.dat)
KAT = {2: -1.163012998021681, 3: -1.926739016209154, 4: -1.252893688642302, 5: -1.134690658798176}


def kat_oracle(O, k):
    from forte_amd.fcidump import chem_to_forte, read_fcidump
    d = read_fcidump(fcidump_path(k))
    nel, ms2 = d["nelec"], d["ms2"]
    na = (nel + ms2) // 2
    aa, ab = chem_to_forte(d["eri"])
    return O.Oracle(d["norb"], na, nel - na, d["hcore"], d["hcore"], aa, ab, aa, d["enuc"]), ms2 + 1


@pytest.mark.parametrize("k", sorted(KAT))
def test_kat_energy(oracle, k):
    o, mult = kat_oracle(oracle, k)
    r = o.solve(nroot=1, multiplicity=mult, e_conv=1e-12)
    assert r["converged"]
    assert abs(r["energies"][0] - KAT[k]) < 1e-9
    # and against a dense diagonalization of the Slater-rules Hamiltonian
    ev = np.linalg.eigvalsh(o.dense_hamiltonian())
    assert abs(ev[0] - KAT[k]) < 1e-9


def test_fcidump_reader_matches_chem_to_forte(oracle):
    from forte_amd.fcidump import chem_to_forte, read_fcidump
    d = read_fcidump(fcidump_path(3))
    assert d["norb"] == 4 and d["nelec"] == 4 and d["ms2"] == 0
    assert abs(d["enuc"] - 3.89442719099991574438) < 1e-15
    eri = d["eri"]
    # 8-fold symmetry
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        assert np.array_equal(eri, eri.transpose(perm))
    aa, ab = chem_to_forte(eri)
    aa2, ab2 = oracle.chem_to_forte(eri)
    assert np.array_equal(aa, aa2) and np.array_equal(ab, ab2)


@pytest.mark.parametrize("n", range(1, 13))
def test_strings_colex_closed_form(oracle, n):
    """Forte's next_permutation order == colex rank of the bit-reversed occupation."""
    for k in range(n + 1):
        s = oracle.make_strings(n, k)
        assert len(s) == comb(n, k)
        for idx, x in enumerate(s):
            occ = sorted(n - 1 - i for i in range(n) if (int(x) >> i) & 1)
            assert sum(comb(c, t + 1) for t, c in enumerate(occ)) == idx


def test_strings_first_last(oracle):
    s = oracle.make_strings(6, 2)
    assert int(s[0]) == 0b110000 and int(s[-1]) == 0b000011


def _apply(ops, det):
    """Apply second-quantized ops (list of (orbital, create)) right to left on a bit string."""
    sign = 1
    for orb, cr in reversed(ops):
        occ = (det >> orb) & 1
        if occ == cr:
            return 0, 0
        if bin(det & ((1 << orb) - 1)).count("1") % 2:
            sign = -sign
        det ^= 1 << orb
    return sign, det


def test_vo_oo_vvoo_lists_are_the_operators(oracle):
    """Every list entry is J = sign a+p a_q I (VO), a+p a+q a_q a_p (OO), a+p a+q a_s a_r (VVOO),
    checked by explicit operator algebra; list sizes are the reference's counts."""
    p_ = seeded_problem(6, 3, 2, 5)
    o = oracle.Oracle(6, 3, 2, p_["h"], p_["h"], p_["aa"], p_["ab"], p_["aa"])
    for spin in (0, 1):
        S = o.strings(spin)
        k = 3 if spin == 0 else 2
        for p in range(6):
            for q in range(6):
                sg, I, J = o.vo(spin, p, q)
                nred = 6 - 1 - (p != q)
                assert len(sg) == (comb(nred, k - 1) if 0 <= k - 1 <= nred else 0)
                for s_, i, j in zip(sg, I, J):
                    sgn, d = _apply([(p, 1), (q, 0)], int(S[i]))
                    assert d == int(S[j]) and sgn == s_
        for pq in range(o.npair()):
            p, q = o.pair(pq)
            sg, I, J = o.oo(spin, pq)
            for s_, i, j in zip(sg, I, J):
                sgn, d = _apply([(p, 1), (q, 1), (q, 0), (p, 0)], int(S[i]))
                assert d == int(S[j]) and i == j and sgn == s_ == 1
        for p, q, r, s in itertools.product(range(6), repeat=4):
            if not (p > q and r > s) or (p == r and q == s):
                continue
            sg, I, J = o.vvoo(spin, p, q, r, s)
            for s_, i, j in zip(sg, I, J):
                sgn, d = _apply([(p, 1), (q, 1), (s, 0), (r, 0)], int(S[i]))
                assert d == int(S[j]) and sgn == s_


def test_sign_table_of_reference_determinant_test():
    """tests/code/test_determinant.cc:160-190: alpha occupation {0,3,4,59}."""
    det = (1 << 0) | (1 << 3) | (1 << 4) | (1 << 59)
    want = {0: 1, 1: -1, 2: -1, 3: -1, 4: 1, 5: -1, 58: -1, 59: -1, 60: 1, 63: 1}
    for n, v in want.items():
        assert (-1) ** bin(det & ((1 << n) - 1)).count("1") == v
    # the same convention drives the oracle's VO signs (checked above by operator algebra)


@pytest.mark.parametrize("n,na,nb,seed", [(4, 2, 2, 1), (5, 3, 2, 2), (6, 3, 3, 3), (6, 2, 4, 4),
                                          (5, 1, 1, 6), (6, 5, 1, 7), (4, 4, 0, 8)])
def test_sigma_against_brute_force(oracle, n, na, nb, seed):
    p = seeded_problem(n, na, nb, seed, e_core=0.37)
    o = oracle.Oracle(n, na, nb, p["h"], p["h"], p["aa"], p["ab"], p["aa"], 0.37)
    rng = np.random.default_rng(seed)
    C = rng.standard_normal(o.ndet)
    s = o.sigma(C)
    sb = o.brute_sigma(C)
    assert np.abs(s - sb).max() < 1e-11 * max(1.0, np.abs(sb).max())
    H = o.dense_hamiltonian()
    assert np.abs(H @ C - sb).max() < 1e-11 * max(1.0, np.abs(sb).max())


@pytest.mark.parametrize("n,na,nb,seed", [(4, 2, 2, 1), (5, 3, 2, 2), (6, 3, 3, 3), (6, 2, 4, 4),
                                          (5, 1, 1, 6), (6, 5, 1, 7), (4, 4, 0, 8), (7, 3, 3, 9)])
def test_sigma_rows_against_brute_force(oracle, n, na, nb, seed):
    """The Slater-Condon row sigma (the checker at sizes beyond a full oracle sigma) against the
    explicit second-quantized sigma, on a context with lists and on a strings-only one."""
    p = seeded_problem(n, na, nb, seed, e_core=0.37)
    C = np.random.default_rng(seed).standard_normal(oracle.binom(n, na) * oracle.binom(n, nb))
    for lists in (True, False):
        o = oracle.Oracle(n, na, nb, p["h"], p["h"], p["aa"], p["ab"], p["aa"], 0.37, lists=lists)
        sb = o.brute_sigma(C)
        rows = np.arange(o.ndet)[::-1].copy()
        got = o.sigma_rows(C, rows)
        assert np.abs(got - sb[rows]).max() < 1e-11 * max(1.0, np.abs(sb).max())


@pytest.mark.parametrize("n,na,nb,seed", [(4, 2, 2, 1), (6, 3, 3, 3), (6, 2, 4, 4), (6, 5, 1, 7),
                                          (4, 4, 0, 8), (8, 4, 4, 5)])
def test_hdiag_is_the_hamiltonian_diagonal(oracle, n, na, nb, seed):
    """orc_hdiag (the checker of the GPU diagonal, bit-exact there) equals the diagonal of the
    Slater-rules Hamiltonian and the brute-force sigma of unit vectors."""
    p = seeded_problem(n, na, nb, seed, e_core=0.7)
    o = oracle.Oracle(n, na, nb, p["h"], p["h"], p["aa"], p["ab"], p["aa"], 0.7)
    d = o.hdiag()
    H = o.dense_hamiltonian()
    assert np.abs(d - np.diag(H)).max() < 1e-12 * max(1.0, np.abs(d).max())
    for I in np.random.default_rng(seed).choice(o.ndet, size=min(5, o.ndet), replace=False):
        e = np.zeros(o.ndet)
        e[I] = 1.0
        assert abs(o.brute_sigma(e)[I] - d[I]) < 1e-12 * max(1.0, abs(d[I]))
    lite = oracle.Oracle(n, na, nb, p["h"], p["h"], p["aa"], p["ab"], p["aa"], 0.7, lists=False)
    assert np.array_equal(lite.hdiag(), d)


def test_strings_only_context_refuses_list_functions(oracle):
    p = seeded_problem(6, 3, 3, 1)
    o = oracle.Oracle(6, 3, 3, p["h"], p["h"], p["aa"], p["ab"], p["aa"], lists=False)
    for call in (lambda: o.sigma(np.zeros(o.ndet)), lambda: o.vo(0, 1, 2), lambda: o.rdms(np.zeros(o.ndet)),
                 lambda: o.solve()):
        with pytest.raises(RuntimeError):
            call()


def test_sigma_terms_sum(oracle):
    p = seeded_problem(6, 3, 3, 11, e_core=1.5)
    o = oracle.Oracle(6, 3, 3, p["h"], p["h"], p["aa"], p["ab"], p["aa"], 1.5)
    C = np.random.default_rng(1).standard_normal(o.ndet)
    total = o.sigma(C)
    parts = sum(o.sigma(C, mask) for mask in (1, 2, 4, 8, 16, 32))
    assert np.abs(total - parts).max() < 1e-12
    assert np.allclose(o.sigma(C, 1), 1.5 * C)


@pytest.mark.parametrize("n,na,nb", [(4, 2, 2), (5, 3, 2), (6, 3, 3)])
def test_rdms_against_brute_force(oracle, n, na, nb):
    p = seeded_problem(n, na, nb, 21)
    o = oracle.Oracle(n, na, nb, p["h"], p["h"], p["aa"], p["ab"], p["aa"])
    C = np.random.default_rng(2).standard_normal(o.ndet)
    for a, b in zip(o.rdms(C), o.brute_rdms(C)):
        assert np.abs(a - b).max() < 1e-12 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("n,na,nb", [(4, 2, 2), (5, 3, 2), (5, 1, 3)])
def test_transition_rdms_against_brute_force(oracle, n, na, nb):
    """<L| ... |R> with two different vectors (the transition RDMs' checker)."""
    p = seeded_problem(n, na, nb, 23)
    o = oracle.Oracle(n, na, nb, p["h"], p["h"], p["aa"], p["ab"], p["aa"])
    rng = np.random.default_rng(5)
    L, R = rng.standard_normal(o.ndet), rng.standard_normal(o.ndet)
    got, ref = o.rdms(L, R), o.brute_rdms(L, R)
    for k in (0, 1, 3):  # g1a, g1b, g2ab
        assert np.abs(got[k] - ref[k]).max() < 1e-12 * max(1.0, np.abs(ref[k]).max())
    # same spin: Forte computes the elements with pair(p,q) >= pair(r,s) and copies each to its
    # transpose (fci_vector_rdm.cc:240-285), so only that half equals <L|a+p a+q a_s a_r|R>
    pidx = lambda a, b: a * (a - 1) // 2 + b  # noqa: E731
    for k in (2, 4):
        for pp, q, r, t in itertools.product(range(n), repeat=4):
            if pp > q and r > t:
                A, B = pidx(pp, q), pidx(r, t)
                want = ref[k][pp, q, r, t] if A >= B else ref[k][r, t, pp, q]
                assert abs(got[k][pp, q, r, t] - want) < 1e-12 * max(1.0, np.abs(ref[k]).max())


def test_energy_from_rdms_and_traces(oracle):
    p = seeded_problem(6, 3, 3, 31)
    o = oracle.Oracle(6, 3, 3, p["h"], p["h"], p["aa"], p["ab"], p["aa"], 0.25)
    r = o.solve(nroot=1, e_conv=1e-12)
    c = r["evecs"][0]
    g1a, g1b, g2aa, g2ab, g2bb = o.rdms(c)
    assert abs(np.trace(g1a) - 3) < 1e-10 and abs(np.trace(g1b) - 3) < 1e-10
    assert abs(np.einsum("pqpq", g2aa) - 6) < 1e-10 and abs(np.einsum("pqpq", g2ab) - 9) < 1e-10
    e = 0.25 + np.einsum("pq,pq", p["h"], g1a + g1b) + 0.25 * np.einsum("pqrs,pqrs", p["aa"], g2aa + g2bb) \
        + np.einsum("pqrs,pqrs", p["ab"], g2ab)
    assert abs(e - r["energies"][0]) < 1e-9


def test_fci_solve_matches_dense(oracle):
    p = seeded_problem(6, 3, 3, 41)
    o = oracle.Oracle(6, 3, 3, p["h"], p["h"], p["aa"], p["ab"], p["aa"])
    r = o.solve(nroot=3, e_conv=1e-12)
    H = o.dense_hamiltonian()
    ev, V = np.linalg.eigh(H)
    singlets = [ev[i] for i in range(len(ev)) if abs(o.spin2(V[:, i])) < 1e-6][:3]
    assert np.allclose(r["energies"], singlets, atol=1e-9)
    assert np.allclose(r["spin2"], 0.0, atol=1e-8)


def test_spin2_triplet(oracle):
    p = seeded_problem(5, 3, 2, 7)
    o = oracle.Oracle(5, 3, 2, p["h"], p["h"], p["aa"], p["ab"], p["aa"])
    r = o.solve(nroot=1, multiplicity=2, e_conv=1e-12)
    assert abs(r["spin2"][0] - 0.75) < 1e-8


def test_seeded_generator(oracle):
    h, eri = oracle.seeded_integrals(8, 1)
    assert np.array_equal(h, h.T)
    assert h[0, 0] == -2.0 and h[7, 7] == -0.5
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        assert np.abs(eri - eri.transpose(perm)).max() < 1e-14
    M = eri.reshape(64, 64)
    assert np.linalg.eigvalsh(M).min() > -1e-10  # PSD
    h2, eri2 = oracle.seeded_integrals(8, 1)
    assert np.array_equal(h, h2) and np.array_equal(eri, eri2)


# ---- Davidson-Liu on dense matrices, after tests/pytest/dl_solver/test_dl.py ------------------
def synthetic(size):
    M = np.zeros((size, size))
    for i in range(size):
        M[i, i] = -1.0 + i * 0.1
        for j in range(i):
            M[i, j] = M[j, i] = 0.05 / (1.0 + abs(i - j))
    return M


M4 = np.array([[-1, 1, 1, 1], [1, 0, 1, 1], [1, 1, 0, 1], [1, 1, 1, 0]], dtype=float)


@pytest.mark.parametrize("size,nroot", [(10, 1), (10, 5), (10, 10), (100, 3), (100, 10), (1000, 4)])
def test_dl_synthetic(oracle, size, nroot):
    M = synthetic(size)
    ok, ev, _, _ = oracle.dl_dense(M, nroot, hdiag=np.diag(M), guesses=[[(i, 1.0)] for i in range(nroot)])
    assert np.allclose(ev, np.linalg.eigvalsh(M)[:nroot])


def test_dl_small_sizes(oracle):
    for size in range(1, 25):
        for nroot in range(1, size + 1):
            M = synthetic(size)
            ok, ev, _, _ = oracle.dl_dense(M, nroot, hdiag=np.diag(M), guesses=[[(i, 1.0)] for i in range(nroot)])
            assert np.allclose(ev, np.linalg.eigvalsh(M)[:nroot])


def test_dl_4x4_variants(oracle):
    ev = np.linalg.eigvalsh(M4)
    _, e, _, _ = oracle.dl_dense(M4, 1, hdiag=np.diag(M4), guesses=[[(i, 1.0)] for i in range(4)])
    assert np.isclose(e[0], ev[0])
    _, e, _, _ = oracle.dl_dense(M4, 1, hdiag=np.diag(M4), guesses=[[(0, 0.1)]])
    assert np.isclose(e[0], ev[0])
    # project out the lowest root -> the second one
    w, V = np.linalg.eigh(M4)
    _, e, _, _ = oracle.dl_dense(M4, 1, hdiag=np.diag(M4), project_out=[[(i, v) for i, v in enumerate(V[:, 0])]])
    assert np.isclose(e[0], ev[1])


def test_dl_errors(oracle):
    """tests/pytest/dl_solver/test_dl_errors.py"""
    h = np.array([-1.0, 0, 0, 0])
    with pytest.raises(RuntimeError):
        oracle.dl_dense(M4, 1, hdiag=h, guesses=[[(0, 0.0)]])
    with pytest.raises(RuntimeError):
        oracle.dl_dense(M4, 2, hdiag=np.diag(M4), guesses=[[(0, 0.1)]])
    with pytest.raises(RuntimeError):
        oracle.dl_dense(M4, 1, hdiag=np.diag(M4), guesses=[[(0, 0.1)]], no_sigma=True)
    with pytest.raises(RuntimeError):
        oracle.dl_dense(M4, 1, hdiag=None)


# ---- point-group symmetry and doubly occupied cores (the reference's D2h FCIDUMP tests) -----

KAT_SYM = {  # tests/methods/integrals-fcidump-1 and -6 (input.dat; the same INTDUMP file)
    1: (dict(frozen_docc=[2, 0, 0, 0], restricted_docc=[2, 0, 0, 0], active=[2, 2, 2, 2]), -112.74446815362198, 784),
    6: (dict(frozen_docc=[2, 0, 0, 0], restricted_docc=[2, 0, 0, 0], active=[2, 2, 2, 2], nel=13, multiplicity=2),
        -112.181397694724, 392),
}


@pytest.mark.parametrize("k", sorted(KAT_SYM))
def test_kat_symmetry_dense(oracle, k):
    """Core folding, MO spaces and the irrep sector against the reference's known answers, by a
    dense diagonalization of the Slater-rules Hamiltonian restricted to the sector."""
    import os
    from forte_amd.active_space import active_space_from_fcidump, forte_layout
    from conftest import GOLDEN
    kw, ref, ndet = KAT_SYM[k]
    a = active_space_from_fcidump(os.path.join(GOLDEN, "integrals-fcidump-1.INTDUMP.gz"), **kw)
    h, aa, ab = forte_layout(a["h"], a["eri"])
    o = oracle.Oracle(h.shape[0], a["na"], a["nb"], h, h, aa, ab, aa, a["e_core"])
    mask = oracle.sector_mask(o, a["orbsym"], a["irrep"])
    assert mask.sum() == ndet  # "Number of determinants" of the reference output
    ev = np.linalg.eigvalsh(o.dense_hamiltonian()[np.ix_(mask, mask)])
    assert abs(ev[0] - ref) < 1e-9


@pytest.mark.parametrize("n,k,orbsym", [(6, 3, [0, 1, 2, 3, 0, 1]), (8, 4, [0, 0, 1, 1, 2, 2, 3, 3]),
                                        (7, 2, [0, 4, 1, 5, 2, 6, 3]), (5, 5, [1, 1, 0, 2, 3])])
def test_string_classes(oracle, n, k, orbsym):
    """Per-irrep string addresses: consecutive within each irrep, in enumeration order, with
    class sizes equal to an independent count over all k-subsets."""
    irrep, addr = oracle.string_classes(n, k, orbsym)
    for h in set(irrep.tolist()):
        assert np.array_equal(addr[irrep == h], np.arange(np.sum(irrep == h)))
    count = {}
    for c in itertools.combinations(range(n), k):
        h = 0
        for p in c:
            h ^= orbsym[p]
        count[h] = count.get(h, 0) + 1
    assert {h: int(np.sum(irrep == h)) for h in count} == count

