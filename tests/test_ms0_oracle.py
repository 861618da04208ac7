"""The Ms = 0 spin-flip symmetry the GPU solver uses (fci_sigma_ms0_device, FCI_MS0_SYMMETRY), pinned
on the oracle's sigma terms and dense Hamiltonian:

  - in this determinant order (C[Ia][Ib], alpha-major) every eigenvector of spin S with na == nb
    satisfies C^T = (-1)^S C;
  - with equal alpha and beta integrals the beta same-spin term is the alpha one on the transposed
    vector, H_bb C = (H_aa C^T)^T;
  - hence for C^T = eta C: H C = P + eta P^T with P = e_core C / 2 + H_aa C + H_ab C / 2, the form the
    device builds (the alpha-beta term of an eta-symmetric C is itself eta-symmetric).
"""
import numpy as np
import pytest

from conftest import seeded_problem

H0, H1A, H1B, H2AB, H2AA, H2BB = 1, 2, 4, 8, 16, 32


def spin_of(o, c):
    s2 = o.spin2(c)
    return int(round((-1 + np.sqrt(1 + 4 * s2)) / 2))


@pytest.mark.parametrize("n,k", [(4, 2), (5, 2), (6, 3), (6, 2), (7, 3), (8, 4), (5, 1), (6, 1)])
def test_eigenvectors_have_spin_flip_sign(oracle, n, k):
    p = seeded_problem(n, k, k, 3)
    o = oracle.Oracle(n, k, k, p["h"], p["h"], p["aa"], p["ab"], p["aa"], 0.0)
    w, V = np.linalg.eigh(o.dense_hamiltonian())
    N = int(round(np.sqrt(o.ndet)))
    seen = set()
    for i in range(len(w)):
        # degenerate levels may mix spins; take vectors of a definite S only
        s2 = o.spin2(V[:, i])
        S = spin_of(o, V[:, i])
        if abs(s2 - S * (S + 1)) > 1e-8:
            continue
        M = V[:, i].reshape(N, N)
        assert np.abs(M.T - (-1) ** S * M).max() < 1e-9, (i, S)
        seen.add(S)
    assert {0, 1} <= seen


@pytest.mark.parametrize("n,k,seed", [(5, 2, 1), (6, 3, 2), (7, 3, 3), (6, 1, 4)])
def test_beta_same_spin_is_the_transposed_alpha_term(oracle, n, k, seed):
    p = seeded_problem(n, k, k, seed)
    o = oracle.Oracle(n, k, k, p["h"], p["h"], p["aa"], p["ab"], p["aa"], 0.0)
    N = int(round(np.sqrt(o.ndet)))
    C = np.random.default_rng(seed).standard_normal((N, N))
    bb = o.sigma(C.reshape(-1), H1B | H2BB).reshape(N, N)
    aa_t = o.sigma(np.ascontiguousarray(C.T).reshape(-1), H1A | H2AA).reshape(N, N).T
    assert np.abs(bb - aa_t).max() < 1e-12 * max(1.0, np.abs(bb).max())


@pytest.mark.parametrize("eta", [1, -1])
@pytest.mark.parametrize("n,k,seed", [(6, 3, 5), (7, 2, 6), (5, 1, 7)])
def test_symmetric_sigma_form(oracle, n, k, seed, eta):
    p = seeded_problem(n, k, k, seed, e_core=0.7)
    o = oracle.Oracle(n, k, k, p["h"], p["h"], p["aa"], p["ab"], p["aa"], 0.7)
    N = int(round(np.sqrt(o.ndet)))
    X = np.random.default_rng(seed).standard_normal((N, N))
    C = X + eta * X.T
    full = o.sigma(C.reshape(-1)).reshape(N, N)
    P = (0.5 * o.sigma(C.reshape(-1), H0) + o.sigma(C.reshape(-1), H1A | H2AA)
         + 0.5 * o.sigma(C.reshape(-1), H2AB)).reshape(N, N)
    assert np.abs(full - (P + eta * P.T)).max() < 1e-11 * np.abs(full).max()
    # the alpha-beta term of an eta-symmetric vector is eta-symmetric
    B = o.sigma(C.reshape(-1), H2AB).reshape(N, N)
    assert np.abs(B - eta * B.T).max() < 1e-12 * np.abs(B).max()


def test_unequal_spin_integrals_break_the_identity(oracle):
    """A planted difference between the alpha and beta same-spin integrals makes the transposed-alpha
    form wrong: the identity above is not vacuous."""
    n, k = 5, 2
    p = seeded_problem(n, k, k, 8)
    bb = p["aa"].copy()
    bb[0, 1, 0, 1] += 0.3
    bb[1, 0, 1, 0] += 0.3
    o = oracle.Oracle(n, k, k, p["h"], p["h"], p["aa"], p["ab"], bb, 0.0)
    N = int(round(np.sqrt(o.ndet)))
    C = np.random.default_rng(0).standard_normal((N, N))
    b = o.sigma(C.reshape(-1), H1B | H2BB).reshape(N, N)
    a_t = o.sigma(np.ascontiguousarray(C.T).reshape(-1), H1A | H2AA).reshape(N, N).T
    assert np.abs(b - a_t).max() > 1e-3
