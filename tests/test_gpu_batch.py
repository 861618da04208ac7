"""The batched solver (fci_batch_solve: problems of one shape advance together through
Davidson-Liu iterations captured in a graph) against the oracle and the single-problem solver."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sample_rows(c, count, seed):
    rng = np.random.default_rng(seed)
    rand = rng.choice(c.size, size=min(count, c.size), replace=False)
    top = np.argsort(-np.abs(c))[:64]
    return np.unique(np.concatenate([rand, top]))


def test_batch_small_problems_match_oracle(oracle):
    """n = 6 and 8 members of the seeded batch, two roots: energies against full oracle solves."""
    from forte_amd.batch import seeded_batch, solve_batch
    probs = [p for p in seeded_batch(40) if p[0] <= 8][:12]
    say("batch n<=8:", probs)
    opts = {"E_CONVERGENCE": 1e-11, "R_CONVERGENCE": 1e-7, "DL_MAXITER": 300}
    res, wall, st = solve_batch(probs, nroot=2, options=opts)
    say(f"  {wall:.2f}s stats {st}")
    for (n, seed), r in zip(probs, res):
        assert r["converged"], (n, seed)
        h, eri = oracle.seeded_integrals(n, seed)
        ref = oracle.Oracle.from_chem(h, eri, n // 2, n // 2).solve(nroot=2, e_conv=1e-11, r_conv=1e-7,
                                                                     maxiter=300, want_evecs=False)["energies"]
        assert np.abs(np.array(r["energies"]) - ref).max() < 1e-9, (n, seed, r["energies"], ref)


def test_batch_matches_single_solver():
    """A mixed batch (n = 6..12) against one FCISolver per problem with the same options."""
    import forte_amd
    from forte_amd.batch import seeded_batch, solve_batch
    probs = seeded_batch(64)
    sel = []
    for n in (6, 8, 10, 12):
        sel += [p for p in probs if p[0] == n][:2]
    opts = {"E_CONVERGENCE": 1e-10, "R_CONVERGENCE": 1e-6, "DL_MAXITER": 400}
    res, wall, st = solve_batch(sel, nroot=2, options=opts)
    say(f"batch of {len(sel)}: {wall:.2f}s stats {st}")
    for (n, seed), r in zip(sel, res):
        h, eri = forte_amd.seeded_integrals(n, seed)
        s = forte_amd.FCISolver(n // 2, n // 2, 1, 2, forte_amd.make_active_space_integrals(h, eri))
        s.set_options(opts)
        s.compute_energy()
        say(f"  n={n} seed={seed}: batch {r['energies']} ({r['iterations']} it) single {s.energies()} "
            f"({s.stats()['iterations']} it)")
        assert r["converged"]
        assert np.abs(np.array(r["energies"]) - np.array(s.energies())).max() < 1e-9


def test_batch_n10_n12_members(oracle):
    """n = 10 members against full oracle solves; n = 12 (853,776 determinants, beyond a quick
    oracle solve) through oracle Slater-rules sigma rows of the batched eigenvectors (residuals),
    the energy from oracle RDMs and <S^2>."""
    from forte_amd.batch import seeded_batch, solve_batch
    probs = seeded_batch(64)
    sel = [p for p in probs if p[0] == 10][:2] + [p for p in probs if p[0] == 12][:2]
    opts = {"E_CONVERGENCE": 1e-10, "R_CONVERGENCE": 1e-6, "DL_MAXITER": 400}
    res, wall, st = solve_batch(sel, nroot=2, options=opts, want_evecs=True)
    say(f"batch members {sel}: {wall:.2f}s")
    for (n, seed), r in zip(sel, res):
        assert r["converged"], (n, seed)
        h, eri = oracle.seeded_integrals(n, seed)
        aa, ab = oracle.chem_to_forte(eri)
        V = r["evecs"]
        assert np.abs(V @ V.T - np.eye(2)).max() < 1e-10
        if n == 10:
            ref = oracle.Oracle.from_chem(h, eri, 5, 5).solve(nroot=2, e_conv=1e-10, r_conv=1e-6, maxiter=400,
                                                              want_evecs=False)["energies"]
            say(f"  n=10 seed {seed}: batch {r['energies']} oracle {ref}")
            assert np.abs(np.array(r["energies"]) - ref).max() < 1e-9
            continue
        o = oracle.Oracle.from_chem(h, eri, 6, 6)
        for root in range(2):
            c = V[root]
            rows = sample_rows(c, 2000, seed + root)
            hc = o.sigma_rows(c, rows)
            res_ = np.abs(hc - r["energies"][root] * c[rows]).max()
            g1a, g1b, g2aa, g2ab, g2bb = o.rdms(c)
            e_rdm = (np.einsum("pq,pq", h, g1a + g1b) + 0.25 * np.einsum("pqrs,pqrs", aa, g2aa + g2bb)
                     + np.einsum("pqrs,pqrs", ab, g2ab))
            say(f"  n=12 seed {seed} root {root}: E {r['energies'][root]:.12f} <H>_RDM {e_rdm:.12f} "
                f"max sampled |Hc - Ec| {res_:.2e} <S^2> {o.spin2(c):.2e}")
            assert res_ < 2e-6
            assert abs(e_rdm - r["energies"][root]) < 1e-9
            assert abs(o.spin2(c)) < 1e-6


def test_batch_rejects_unsupported_options():
    import forte_amd
    h, eri = forte_amd.seeded_integrals(6, 1)
    ints = forte_amd.make_active_space_integrals(h, eri)
    with pytest.raises(RuntimeError):
        forte_amd._forte_amd.batch_solve([ints], [3], [3], 2, 1, {"FCI_PRECISION": "mixed"}, False)
    with pytest.raises(RuntimeError):
        forte_amd._forte_amd.batch_solve([ints], [3], [3], 5, 1, {}, False)


def test_batch_guess_errors_reach_the_caller():
    """The guess construction runs over host threads for the problems of a class; an error there
    (no determinant of 3 + 3 electrons in 6 orbitals has S = 4) comes back as the call's error."""
    import forte_amd
    ints = [forte_amd.make_active_space_integrals(*forte_amd.seeded_integrals(6, s)) for s in (1, 2, 3, 4)]
    with pytest.raises(RuntimeError, match="multiplicity"):
        forte_amd._forte_amd.batch_solve(ints, [3] * 4, [3] * 4, 1, 9, {}, False)


@pytest.mark.parametrize("mult", [1, 3])
def test_batch_spin_flip_symmetry(mult):
    """Symmetric classes (packed Davidson-Liu vectors, sigma without the beta same-spin term) against
    spin-adapted single solves, singlets and triplets, and against the same batch with the symmetry
    off where that converges (its n = 10 triplet does not in 400 iterations: the determinant solve
    without the symmetry lets singlet components back in); the eigenvectors come back as exactly
    (-1)^S-symmetric determinant vectors."""
    import forte_amd
    from forte_amd.batch import solve_many
    sel = [(6, 3), (6, 5), (8, 2), (8, 7), (10, 1)]
    ints = []
    for n, seed in sel:
        h, eri = forte_amd.seeded_integrals(n, seed)
        ints.append(forte_amd.make_active_space_integrals(h, eri))
    na = [n // 2 for n, _ in sel]
    out = {}
    for on in (True, False):
        opts = {"E_CONVERGENCE": 1e-11, "R_CONVERGENCE": 1e-7, "DL_MAXITER": 400, "FCI_MS0_SYMMETRY": on}
        out[on] = solve_many(ints, na, na, nroot=2, multiplicity=mult, options=opts, want_evecs=True)
    eta = 1 if mult == 1 else -1
    a, b = out[True], out[False]
    for i, (n, seed) in enumerate(sel):
        ref = forte_amd.FCISolver(n // 2, n // 2, mult, 2, ints[i])
        ref.set_spin_adapt(True)
        ref.set_e_convergence(1e-11)
        ref.set_r_convergence(1e-7)
        ref.set_maxiter(400)
        ref.compute_energy()
        say(f"  n={n} seed={seed} mult={mult}: symmetric {a['energies'][i]} ({a['iterations'][i]} it), "
            f"full {b['energies'][i]} ({b['iterations'][i]} it, converged {b['converged'][i]}), "
            f"CSF solve {ref.energies()}")
        assert a["converged"][i]
        assert np.abs(np.array(a["energies"][i]) - np.array(ref.energies())).max() < 1e-9
        if b["converged"][i]:
            assert np.abs(np.array(a["energies"][i]) - np.array(b["energies"][i])).max() < 1e-9
        V = np.asarray(a["evecs"][i])
        N = int(round(np.sqrt(V.shape[1])))
        for v in V:
            M = v.reshape(N, N)
            assert np.array_equal(M.T, eta * M)
        assert np.abs(V @ V.T - np.eye(2)).max() < 1e-10
