"""GenCI's GAS spaces on the GPU (fci_set_gas) against the oracle's restricted Hamiltonian and the
reference's printed determinant counts, and the MCSCF repeated-solve pattern (new integrals into a
solved solver, average RDMs)."""
import sys

import numpy as np
import pytest

from conftest import seeded_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gas_solver(p, spaces, gmin, gmax, nroot=1, mult=1, irrep=None):
    import forte_amd
    orbsym, gas_of = forte_amd.gas_layout(spaces)
    ints = forte_amd.ActiveSpaceIntegrals(p["n"], p["h"], p["h"], p["aa"], p["ab"], p["aa"], p["e_core"])
    if irrep is not None:
        ints.set_active_mo_symmetry(orbsym)
        s = forte_amd.FCISolver(p["na"], p["nb"], mult, nroot, ints, irrep)
    else:
        s = forte_amd.FCISolver(p["na"], p["nb"], mult, nroot, ints)
    s.set_gas(gas_of, gmin, gmax)
    return s, orbsym, gas_of


GAS_COUNTS = [
    ("gasci-2", [[2, 0, 1, 1], [1, 0, 0, 1]], 4, 4, 0, [6, 0], [8, 4], 31),
    ("gasci-3", [[1, 0, 0, 0], [3, 0, 1, 2]], 5, 5, 0, [0, 0], [2, 10], 133),
    ("gasci-3 GAS1MAX 1", [[1, 0, 0, 0], [3, 0, 1, 2]], 5, 5, 0, [0, 0], [1, 10], 68),
    ("gasci-4 B1", [[1, 0, 0, 0], [3, 0, 1, 2]], 5, 5, 2, [0, 0], [1, 10], 40),
]


@pytest.mark.parametrize("case", GAS_COUNTS, ids=[c[0] for c in GAS_COUNTS])
def test_gas_space_size_matches_the_reference_outputs(case):
    name, spaces, na, nb, irrep, gmin, gmax, expect = case
    n = sum(sum(c) for c in spaces)
    say(f"{name}: n={n} na={na} nb={nb} irrep={irrep} gas_min={gmin} gas_max={gmax}")
    s, _, _ = gas_solver(seeded_problem(n, na, nb, 1), spaces, gmin, gmax, irrep=irrep)
    assert s.space_size() == expect


SOLVES = [
    # spaces, na, nb, gmin, gmax, nroot, mult, irrep, spin_adapt
    ([[4], [4]], 3, 3, [2, 0], [5, 4], 3, 1, None, False),
    ([[4], [4]], 3, 3, [2, 0], [5, 4], 2, 1, None, True),
    ([[3], [2], [3]], 4, 3, [3, 1, 0], [6, 4, 2], 2, 2, None, False),
    ([[2, 0, 1, 1], [1, 1, 0, 2]], 4, 4, [5, 0], [8, 4], 2, 1, 0, False),
    ([[2, 0, 1, 1], [1, 1, 0, 2]], 4, 4, [5, 0], [8, 4], 1, 3, 1, True),
]


@pytest.mark.parametrize("spaces,na,nb,gmin,gmax,nroot,mult,irrep,sa", SOLVES)
def test_gas_solve_vs_restricted_dense_hamiltonian(oracle, spaces, na, nb, gmin, gmax, nroot, mult, irrep, sa):
    n = sum(sum(c) for c in spaces)
    say(f"GAS solve n={n} na={na} nb={nb} gas {spaces} min {gmin} max {gmax} nroot {nroot} mult {mult} "
        f"irrep {irrep} spin_adapt {sa}")
    p = seeded_problem(n, na, nb, 17, e_core=0.1)
    s, orbsym, gas_of = gas_solver(p, spaces, gmin, gmax, nroot, mult, irrep)
    s.set_spin_adapt(sa)
    s.set_e_convergence(1e-12)
    s.set_r_convergence(1e-8)
    s.set_maxiter(300)
    s.compute_energy()
    assert s.stats()["converged"]
    o = oracle.Oracle(n, na, nb, p["h"], p["h"], p["aa"], p["ab"], p["aa"], p["e_core"])
    mask = oracle.gas_mask(n, na, nb, gas_of, gmin, gmax)
    if irrep is not None:
        mask &= oracle.sector_mask(o, orbsym, irrep)
    assert s.space_size() == int(mask.sum())
    idx = np.nonzero(mask)[0]
    H = o.dense_hamiltonian()[np.ix_(idx, idx)]
    w, V = np.linalg.eigh(H)
    # keep the eigenvalues of the requested multiplicity: <S^2> of each restricted eigenvector
    S = (mult - 1) / 2
    ref = []
    for k in range(len(w)):
        c = np.zeros(o.ndet)
        c[idx] = V[:, k]
        if abs(o.spin2(c) - S * (S + 1)) < 1e-6:
            ref.append(w[k])
        if len(ref) == nroot:
            break
    E = np.array(s.energies())
    say(f"  gpu {E} oracle {ref} ({s.space_size()} determinants)")
    assert np.abs(E - np.array(ref)).max() < 1e-9
    for r in range(nroot):
        c = s.evec(r)
        assert np.abs(c[~mask]).max() == 0.0
        assert abs(o.spin2(c) - S * (S + 1)) < 1e-8


def test_unrestricted_gas_equals_fci():
    import forte_amd
    p = seeded_problem(8, 4, 4, 2)
    s, _, _ = gas_solver(p, [[5], [3]], [0, 0], [8, 6], 2)
    s.set_e_convergence(1e-11)
    e_gas = s.compute_energy()
    ints = forte_amd.ActiveSpaceIntegrals(8, p["h"], p["h"], p["aa"], p["ab"], p["aa"], 0.0)
    f = forte_amd.FCISolver(4, 4, 1, 2, ints)
    f.set_e_convergence(1e-11)
    assert s.space_size() == f.ndet()
    assert abs(e_gas - f.compute_energy()) < 1e-9


def test_genci_factory_and_mixed_precision(oracle):
    """make_active_space_method("GENCI") with GASn / GASnMIN / GASnMAX, FP64 and mixed precision."""
    import forte_amd
    spaces = [[3, 0, 1, 1], [2, 0, 1, 2]]
    orbsym, gas_of = forte_amd.gas_layout(spaces)
    n = len(orbsym)
    p = seeded_problem(n, 5, 5, 9)
    e = {}
    for prec in ("fp64", "mixed"):
        ints = forte_amd.ActiveSpaceIntegrals(n, p["h"], p["h"], p["aa"], p["ab"], p["aa"], 0.0)
        s = forte_amd.make_active_space_method("GENCI", 5, 5, 1, 1, ints,
                                               {"GAS1": spaces[0], "GAS2": spaces[1], "GAS1MIN": [7],
                                                "E_CONVERGENCE": 1e-11, "R_CONVERGENCE": 1e-7,
                                                "DL_MAXITER": 400, "FCI_PRECISION": prec})
        e[prec] = s.compute_energy()
        say(f"  GENCI {prec}: {e[prec]:.12f} ({s.space_size()} of {s.ndet()} determinants) {s.stats()}")
    o = oracle.Oracle(n, 5, 5, p["h"], p["h"], p["aa"], p["ab"], p["aa"], 0.0)
    mask = oracle.gas_mask(n, 5, 5, gas_of, [7, 0], [10, 10])
    idx = np.nonzero(mask)[0]
    w = np.linalg.eigvalsh(o.dense_hamiltonian()[np.ix_(idx, idx)])
    assert abs(e["fp64"] - w[0]) < 1e-9
    assert abs(e["mixed"] - w[0]) < 1e-8


def test_mcscf_repeated_solves_orbital_rotation():
    """The CASSCF micro-iteration pattern: new (rotated) integrals into a solved solver, solve again
    from the previous basis, average RDMs. FCI is invariant under orbital rotations, so the energies
    repeat and the 1-RDM rotates with the orbitals: g1' = U^T g1 U."""
    import forte_amd
    n, na, nb = 10, 5, 5
    h, eri = forte_amd.seeded_integrals(n, 4)
    ints = forte_amd.make_active_space_integrals(h, eri)
    s = forte_amd.FCISolver(na, nb, 1, 2, ints)
    s.set_e_convergence(1e-11)
    s.set_r_convergence(1e-7)
    s.set_maxiter(400)
    s.compute_energy()
    E0 = np.array(s.energies())
    avg0 = forte_amd.compute_average_rdms(s, [0.5, 0.5], 2)
    it0 = s.stats()["iterations"]
    rng = np.random.default_rng(0)
    for step in range(3):
        # a small rotation, as one macro-iteration would apply
        K = rng.standard_normal((n, n)) * 0.05
        U, _ = np.linalg.qr(np.eye(n) + K - K.T)
        h = U.T @ h @ U
        eri = np.einsum("pqrs,pa,qb,rc,sd->abcd", eri, U, U, U, U, optimize=True)
        s.set_active_space_integrals(forte_amd.make_active_space_integrals(h, eri))
        s.compute_energy()
        E = np.array(s.energies())
        avg = forte_amd.compute_average_rdms(s, [0.5, 0.5], 2)
        say(f"  step {step}: energies {E} ({s.stats()['iterations']} iterations, first solve {it0})")
        assert np.abs(E - E0).max() < 1e-8
        g1 = avg0["g1a"] + avg0["g1b"]
        assert abs(np.trace(avg["g1a"] + avg["g1b"]) - (na + nb)) < 1e-10
        # rotate the reference RDM along
        g1 = U.T @ g1 @ U
        avg0 = {"g1a": U.T @ avg0["g1a"] @ U, "g1b": U.T @ avg0["g1b"] @ U}
        assert np.abs(avg["g1a"] + avg["g1b"] - g1).max() < 1e-6
