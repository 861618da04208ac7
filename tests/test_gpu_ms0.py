"""The Ms = 0 spin-flip-symmetric sigma (fci_sigma_ms0_device) and solves kept in that subspace
(FCI_MS0_SYMMETRY) against the oracle's sigma, the full GPU sigma, the oracle's dense spectrum and
the same solves without the symmetry."""
import sys

import numpy as np
import pytest

from conftest import seeded_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_solver(p, nroot=1, mult=1, bb=None):
    import forte_amd
    ints = forte_amd.ActiveSpaceIntegrals(p["n"], p["h"], p["h"], p["aa"], p["ab"], p["aa"] if bb is None else bb,
                                          p["e_core"])
    return forte_amd.FCISolver(p["na"], p["nb"], mult, nroot, ints)


def sym_vector(N, eta, seed):
    X = np.random.default_rng(seed).standard_normal((N, N))
    C = X + eta * X.T
    return (C / np.linalg.norm(C)).reshape(-1)


def run_sigma(g, C, dtype, eta=None):
    import forte_amd
    c = forte_amd.DeviceArray(g.ndet(), dtype)
    s = forte_amd.DeviceArray(g.ndet(), dtype)
    c.upload(C.astype(dtype))
    if eta is None:
        g.sigma_arrays(c, s)
    else:
        g.sigma_ms0_arrays(c, s, eta)
    return s.download().astype(np.float64)


SHAPES = [(4, 2), (5, 1), (6, 3), (6, 1), (8, 4), (9, 4), (10, 5), (12, 6), (16, 2), (13, 3)]


@pytest.mark.parametrize("eta", [1, -1])
@pytest.mark.parametrize("n,k", SHAPES)
def test_sigma_ms0_vs_oracle(oracle, n, k, eta):
    say(f"sigma ms0 n={n} k={k} eta={eta}")
    p = seeded_problem(n, k, k, 31 + n, e_core=0.37)
    g = gpu_solver(p)
    o = oracle.Oracle(n, k, k, p["h"], p["h"], p["aa"], p["ab"], p["aa"], p["e_core"])
    N = int(round(np.sqrt(o.ndet)))
    C = sym_vector(N, eta, n + k)
    ref = o.sigma(C)
    got = run_sigma(g, C, "float64", eta)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    M = got.reshape(N, N)
    say(f"  FP64 max error {err:.2e}")
    assert err < 1e-12
    assert np.array_equal(M.T, eta * M)  # exactly eta-symmetric
    got32 = run_sigma(g, C, "float32", eta)
    err32 = np.linalg.norm(got32 - ref) / np.linalg.norm(ref)
    full32 = run_sigma(g, C, "float32")
    errf = np.linalg.norm(full32 - ref) / np.linalg.norm(ref)
    M32 = got32.reshape(N, N)
    say(f"  mixed relative error {err32:.2e} (full mixed sigma {errf:.2e})")
    # as accurate as the full mixed sigma (one more FP32 rounding in the transpose-add)
    assert err32 < 1.5 * errf + 1e-7
    assert err32 < 3e-6
    assert np.array_equal(M32.T, eta * M32)


@pytest.mark.parametrize("n,k", [(14, 7), (16, 8)])
def test_sigma_ms0_production_shapes(n, k):
    """CAS(14,14) and CAS(16,16) (the benchmark's shape): the symmetric form against the full GPU
    sigma, which the oracle pins at these shapes on sampled rows (test_gpu_configs.py)."""
    say(f"sigma ms0 production shape n={n} k={k}")
    p = seeded_problem(n, k, k, 3)
    g = gpu_solver(p)
    N = int(round(np.sqrt(g.ndet())))
    for eta in (1, -1):
        C = sym_vector(N, eta, 7 + eta)
        full = run_sigma(g, C, "float64")
        sym = run_sigma(g, C, "float64", eta)
        err = np.abs(sym - full).max() / np.abs(full).max()
        say(f"  eta {eta}: FP64 symmetric vs full max error {err:.2e}")
        assert err < 1e-12
        full32 = run_sigma(g, C, "float32")
        sym32 = run_sigma(g, C, "float32", eta)
        e32 = np.linalg.norm(sym32 - full) / np.linalg.norm(full)
        f32 = np.linalg.norm(full32 - full) / np.linalg.norm(full)
        say(f"  eta {eta}: mixed symmetric {e32:.2e}, mixed full {f32:.2e} (vs FP64 full)")
        assert e32 < 1e-6
        del full, sym, full32, sym32


# (Quintet targets are left out: on these seeded Hamiltonians the reference's Davidson-Liu, with
# the other multiplicities projected out of the guess space only, lands on singlets, which share
# the quintets' spin-flip sign, with or without the symmetry.)
SOLVES = [
    # n, k, nroot, mult, precision
    (8, 4, 3, 1, "fp64"),
    (8, 4, 2, 3, "fp64"),
    (9, 4, 2, 1, "fp64"),
    (10, 5, 2, 1, "mixed"),
    (10, 5, 1, 3, "mixed"),
    (6, 2, 4, 1, "fp64"),
]


@pytest.mark.parametrize("n,k,nroot,mult,prec", SOLVES)
def test_solve_ms0_vs_oracle_and_full(oracle, n, k, nroot, mult, prec):
    say(f"solve ms0 n={n} k={k} nroot={nroot} mult={mult} {prec}")
    p = seeded_problem(n, k, k, 41 + n, e_core=0.2)
    res = {}
    for on in (True, False):
        g = gpu_solver(p, nroot, mult)
        g.set_ms0_symmetry(on)
        # the reference run is the FP64 solve in CSFs of the multiplicity. Determinant solves without
        # the symmetry are not a reliable reference here: the other multiplicities are projected out
        # of the guess space only, and rounding lets them back in. With FP32 sigma noise the second
        # of two (10, 5) singlet roots drifts to a triplet lying between them (<S^2> = 2, no
        # convergence in 400 iterations); in FP64 the (10, 5) triplet does not converge in 400
        # iterations either. The symmetric subspace excludes the states of the other spin parity.
        g.set_precision(prec if on else "fp64")
        g.set_spin_adapt(not on)
        g.set_e_convergence(1e-11)
        g.set_r_convergence(1e-7)
        g.set_maxiter(400)
        g.compute_energy()
        st = g.stats()
        say(f"  symmetry {on}: {g.energies()} {st}")
        assert st["converged"]
        eta = (-1) ** ((mult - 1) // 2)
        assert st["ms0_symmetry"] == (eta if on else 0)  # the spin-adapted reference: option off
        res[on] = (np.array(g.energies()), np.array(g.spin2()), [g.evec(r) for r in range(nroot)])
        if on:
            N = int(round(np.sqrt(g.ndet())))
            for v in res[on][2]:
                M = v.reshape(N, N)
                assert np.array_equal(M.T, eta * M)
    assert np.abs(res[True][0] - res[False][0]).max() < 1e-9
    S = (mult - 1) / 2
    assert np.abs(res[True][1] - S * (S + 1)).max() < 1e-8
    if n <= 9:
        o = oracle.Oracle(n, k, k, p["h"], p["h"], p["aa"], p["ab"], p["aa"], p["e_core"])
        w, V = np.linalg.eigh(o.dense_hamiltonian())
        ref = [w[i] for i in range(len(w)) if abs(o.spin2(V[:, i]) - S * (S + 1)) < 1e-6][:nroot]
        assert np.abs(res[True][0] - np.array(ref)).max() < 1e-9


def test_ms0_with_point_group_gas_olsen_and_learned():
    """The symmetric subspace combined with a point-group sector, GAS restrictions, Olsen's correction
    and the learned preconditioner (whose corrections are projected onto the subspace)."""
    import forte_amd
    import forte_amd.learned
    spaces = [[2, 0, 1, 1], [1, 1, 0, 2]]
    orbsym, gas_of = forte_amd.gas_layout(spaces)
    n = len(orbsym)
    p = seeded_problem(n, 4, 4, 5)
    out = {}
    for case in ("sector", "gas", "olsen", "learned"):
        e = {}
        for on in (True, False):
            ints = forte_amd.ActiveSpaceIntegrals(n, p["h"], p["h"], p["aa"], p["ab"], p["aa"], 0.0)
            nroot = 1 if case == "learned" else 2  # the learned preconditioner targets one root
            if case == "sector":
                ints.set_active_mo_symmetry(orbsym)
                g = forte_amd.FCISolver(4, 4, 1, nroot, ints, 1)
            else:
                g = forte_amd.FCISolver(4, 4, 1, nroot, ints)
            if case == "gas":
                g.set_gas(gas_of, [5, 0], [8, 4])
            if case == "olsen":
                g.set_preconditioner("olsen")
            if case == "learned":
                g.set_lp_weights(forte_amd.learned.load_weights())
                g.set_preconditioner("learned")
            g.set_ms0_symmetry(on)
            g.set_e_convergence(1e-11)
            g.set_r_convergence(1e-7)
            g.set_maxiter(400)
            g.set_die_if_not_converged(False)
            g.compute_energy()
            st = g.stats()
            say(f"  {case} symmetry {on}: {g.energies()} {st}")
            assert st["converged"], (case, on)
            assert st["ms0_symmetry"] == (1 if on else 0)
            e[on] = np.array(g.energies())
        say(f"  {case}: {e[True]} vs {e[False]}")
        assert np.abs(e[True] - e[False]).max() < 1e-9
        out[case] = e[True]


def test_ms0_not_used_where_it_does_not_apply():
    import forte_amd
    p = seeded_problem(6, 3, 3, 2)
    # unequal alpha and beta same-spin integrals
    bb = p["aa"].copy()
    bb[0, 1, 0, 1] += 0.1  # antisymmetrized <pq||rs>: keep its symmetries
    bb[1, 0, 1, 0] += 0.1
    bb[0, 1, 1, 0] -= 0.1
    bb[1, 0, 0, 1] -= 0.1
    g = gpu_solver(p, bb=bb)
    g.compute_energy()
    assert g.stats()["ms0_symmetry"] == 0
    c = forte_amd.DeviceArray(g.ndet(), "float64")
    s = forte_amd.DeviceArray(g.ndet(), "float64")
    with pytest.raises(RuntimeError, match="equal alpha and beta integrals"):
        g.sigma_ms0_arrays(c, s, 1)
    # na != nb
    q = seeded_problem(6, 4, 2, 2)
    g = gpu_solver(q, 1, 3)
    g.compute_energy()
    assert g.stats()["ms0_symmetry"] == 0
    # spin adaptation keeps CSF vectors; its sigma takes the symmetric form unless switched off
    e = {}
    for on in (True, False):
        g = gpu_solver(p)
        g.set_spin_adapt(True)
        g.set_ms0_symmetry(on)
        g.set_e_convergence(1e-11)
        e[on] = g.compute_energy()
        assert g.stats()["ms0_symmetry"] == (1 if on else 0)
    assert abs(e[True] - e[False]) < 1e-9
    g = gpu_solver(p)
    with pytest.raises(RuntimeError, match="eta"):
        g.sigma_ms0_arrays(c, s, 2)


def test_ms0_warm_restarts():
    """Switching the symmetry on between two solves of one solver (a fresh guess: the kept basis
    lives in determinant coordinates), then new (rotated) integrals, which restart from the kept
    packed basis with the roots to project out re-packed."""
    import forte_amd
    n, k = 8, 4
    h, eri = forte_amd.seeded_integrals(n, 12)
    g = forte_amd.FCISolver(k, k, 1, 2, forte_amd.make_active_space_integrals(h, eri))
    g.set_e_convergence(1e-11)
    g.set_r_convergence(1e-7)
    g.set_maxiter(400)
    g.set_ms0_symmetry(False)
    g.compute_energy()
    e0 = np.array(g.energies())
    assert g.stats()["ms0_symmetry"] == 0
    g.set_ms0_symmetry(True)
    g.compute_energy()
    e1 = np.array(g.energies())
    say(f"  without {e0}, then with the symmetry {e1} {g.stats()}")
    assert g.stats()["ms0_symmetry"] == 1
    assert np.abs(e1 - e0).max() < 1e-9
    U, _ = np.linalg.qr(np.eye(n) + 0.05 * np.random.default_rng(1).standard_normal((n, n)))
    h2 = U.T @ h @ U
    eri2 = np.einsum("pqrs,pa,qb,rc,sd->abcd", eri, U, U, U, U, optimize=True)
    g.set_active_space_integrals(forte_amd.make_active_space_integrals(h2, eri2))
    g.compute_energy()
    e2 = np.array(g.energies())
    say(f"  rotated orbitals {e2} {g.stats()}")
    assert np.abs(e2 - e0).max() < 1e-8
    N = int(round(np.sqrt(g.ndet())))
    for r in range(2):
        M = g.evec(r).reshape(N, N)
        assert np.array_equal(M.T, M)
