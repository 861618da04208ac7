"""The CPU restatement of Forte's CSF spin adaptation (oracle/spin_adapt.py), pinned by

  - the reference's own printed CSF / coupling / determinant counts for the spin-adapted test
    inputs (tests/methods/*/output.ref: "Number of CSFs", "Number of couplings", "Number of
    determinants"), with the active orbitals per irrep those outputs list;
  - what the mathematics fixes: the CSFs are orthonormal and span exactly the eigenspace of S^2
    with eigenvalue S(S+1) inside the (Ms, irrep) determinant space, with S^2 built here from
    explicit spin-orbital ladder operators (independent of the oracle's <S^2> code);
  - eigenvalues of H in the CSF basis equal those of H restricted to that eigenspace.
"""
import numpy as np
import pytest

from conftest import seeded_problem
from oracle.spin_adapt import SpinAdapter, configurations, count_csfs, det_occupations, spin_couplings

# (active orbitals per irrep, na, nb, 2S, irrep) -> (determinants, CSFs, couplings), as printed by
# the reference for its spin-adapted tests
GOLDEN_COUNTS = [
    ("sa-fci-1", [3, 0, 1, 1, 0, 3, 1, 1], 3, 3, 0, 0, (1824, 714, 3396)),
    ("sa-fci-2", [3, 0, 1, 1, 0, 3, 1, 1], 3, 3, 4, 4, (1792, 330, 3780)),
    ("sa-fci-6", [9, 1, 4, 4], 4, 4, 0, 0, (2342800, 596304, 11119508)),
    ("sa-fci-7 (4a3b)", [7, 0, 2, 2], 4, 3, 1, 2, (13596, 7240, 54616)),
    ("sa-fci-7 (3a4b)", [7, 0, 2, 2], 3, 4, 1, 2, (13596, 7240, 54616)),
    ("sa-fci-rdms-2", [4, 0, 1, 1], 3, 2, 1, 0, (76, 64, 100)),
]


@pytest.mark.parametrize("case", GOLDEN_COUNTS, ids=[c[0] for c in GOLDEN_COUNTS])
def test_counts_match_the_reference_outputs(case):
    _, per_irrep, na, nb, twoS, irrep, expect = case
    assert count_csfs(per_irrep, na, nb, twoS, irrep) == expect


def orbsym_of(per_irrep):
    return [h for h, m in enumerate(per_irrep) for _ in range(m)]


@pytest.mark.parametrize("per_irrep,na,nb,twoS,irrep", [
    ([3, 0, 1, 1, 0, 3, 1, 1], 3, 3, 0, 0), ([3, 0, 1, 1, 0, 3, 1, 1], 3, 3, 4, 4), ([4, 0, 1, 1], 3, 2, 1, 0),
    ([2, 1, 2, 1], 3, 3, 2, 3), ([6], 3, 3, 0, 0), ([6], 4, 2, 2, 0), ([7], 4, 3, 3, 0)])
def test_explicit_adapter_counts(per_irrep, na, nb, twoS, irrep):
    """The explicit CSF list (what the GPU builds) has the combinatorial counts."""
    sa = SpinAdapter(sum(per_irrep), na, nb, twoS, orbsym_of(per_irrep), irrep)
    nd, ncsf, ncoup = count_csfs(per_irrep, na, nb, twoS, irrep)
    assert (sa.ncsf, sa.ncoupling) == (ncsf, ncoup)
    assert len(set(sa.idx.tolist())) <= nd


def test_couplings_and_occupations_orders():
    """N = 4 singlets: the two genealogical couplings (uudd, udud) in next_permutation order; the
    six Ms = 0 spin patterns in lexicographic order."""
    assert spin_couplings(4, 0) == [(0, 0, 1, 1), (0, 1, 0, 1)]
    assert spin_couplings(3, 1) == [(0, 0, 1), (0, 1, 0)]
    assert spin_couplings(0, 0) == [()]
    assert det_occupations(4, 0) == [(0, 0, 1, 1), (0, 1, 0, 1), (0, 1, 1, 0), (1, 0, 0, 1), (1, 0, 1, 0),
                                     (1, 1, 0, 0)]
    # configurations ordered by (open shells, closed shells) as integers
    confs = configurations(4, 2, 2)
    keys = [(s, d) for d, s in confs]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)


# ---- S^2 from explicit ladder operators ------------------------------------------------------

def _ladder_apply(vec, n, up):
    """S+ (up) or S- on {(alpha, beta): c}; spin orbitals ordered alpha 0..n-1, then beta."""
    out = {}
    for (a, b), c in vec.items():
        for p in range(n):
            # S+ = a+_{p alpha} a_{p beta}; S- = a+_{p beta} a_{p alpha}
            if up:
                if not (b >> p) & 1 or (a >> p) & 1:
                    continue
                occ = a | (b << n)
                s1 = (-1) ** bin(occ & ((1 << (n + p)) - 1)).count("1")  # annihilate beta p
                occ ^= 1 << (n + p)
                s2 = (-1) ** bin(occ & ((1 << p) - 1)).count("1")  # create alpha p
                key = (a | (1 << p), b & ~(1 << p))
            else:
                if not (a >> p) & 1 or (b >> p) & 1:
                    continue
                occ = a | (b << n)
                s1 = (-1) ** bin(occ & ((1 << p) - 1)).count("1")
                occ ^= 1 << p
                s2 = (-1) ** bin(occ & ((1 << (n + p)) - 1)).count("1")
                key = (a & ~(1 << p), b | (1 << p))
            out[key] = out.get(key, 0.0) + s1 * s2 * c
    return out


def s2_matrix(n, na, nb):
    from oracle.oracle import make_strings
    sa, sb = make_strings(n, na), make_strings(n, nb)
    ra = {int(s): i for i, s in enumerate(sa)}
    rb = {int(s): i for i, s in enumerate(sb)}
    nd = len(sa) * len(sb)
    S2 = np.zeros((nd, nd))
    sz = 0.5 * (na - nb)
    for i, a in enumerate(sa):
        for j, b in enumerate(sb):
            col = i * len(sb) + j
            v = _ladder_apply(_ladder_apply({(int(a), int(b)): 1.0}, n, True), n, False)  # S- S+
            for (x, y), c in v.items():
                S2[ra[x] * len(sb) + rb[y], col] += c
            S2[col, col] += sz * (sz + 1)
    return S2


@pytest.mark.parametrize("n,na,nb,twoS", [(4, 2, 2, 0), (4, 2, 2, 2), (4, 2, 2, 4), (5, 3, 2, 1), (5, 3, 2, 3),
                                          (6, 3, 3, 0), (6, 3, 3, 2), (5, 4, 2, 2), (6, 2, 2, 0)])
def test_csfs_are_an_orthonormal_basis_of_the_spin_eigenspace(n, na, nb, twoS):
    sa = SpinAdapter(n, na, nb, twoS)
    T = sa.matrix()
    assert np.abs(T @ T.T - np.eye(sa.ncsf)).max() < 1e-13
    S2 = s2_matrix(n, na, nb)
    assert np.abs(S2 - S2.T).max() < 1e-13
    S = twoS / 2
    # every CSF is an eigenfunction with S(S+1) ...
    assert np.abs(S2 @ T.T - S * (S + 1) * T.T).max() < 1e-12
    # ... and together they span the whole eigenspace
    w, U = np.linalg.eigh(S2)
    Us = U[:, np.abs(w - S * (S + 1)) < 1e-8]
    assert Us.shape[1] == sa.ncsf
    P = T.T @ T
    assert np.abs(P @ Us - Us).max() < 1e-12


@pytest.mark.parametrize("n,na,nb,twoS,seed", [(6, 3, 3, 0, 1), (6, 3, 3, 2, 2), (5, 3, 2, 1, 3), (6, 4, 2, 2, 4)])
def test_csf_hamiltonian_spectrum(oracle, n, na, nb, twoS, seed):
    p = seeded_problem(n, na, nb, seed)
    o = oracle.Oracle(n, na, nb, p["h"], p["h"], p["aa"], p["ab"], p["aa"], 0.0)
    H = o.dense_hamiltonian()
    sa = SpinAdapter(n, na, nb, twoS)
    T = sa.matrix()
    e_csf = np.linalg.eigvalsh(T @ H @ T.T)
    S2 = s2_matrix(n, na, nb)
    w, U = np.linalg.eigh(S2)
    S = twoS / 2
    Us = U[:, np.abs(w - S * (S + 1)) < 1e-8]
    e_ref = np.linalg.eigvalsh(Us.T @ H @ Us)
    assert np.abs(e_csf - e_ref).max() < 1e-10
    # the CSF diagonal without the full preconditioner: sum c^2 H_II
    assert np.abs(sa.hdiag(np.diag(H)) - np.diag(T @ np.diag(np.diag(H)) @ T.T)).max() < 1e-12


def test_transforms_are_transposes():
    sa = SpinAdapter(6, 3, 3, 0, [0, 1, 0, 1, 2, 3], 1)
    rng = np.random.default_rng(0)
    d = rng.standard_normal(sa.ndet)
    c = rng.standard_normal(sa.ncsf)
    assert abs(sa.det_to_csf(d) @ c - d @ sa.csf_to_det(c)) < 1e-12
    T = sa.matrix()
    assert np.abs(sa.det_to_csf(d) - T @ d).max() < 1e-13
