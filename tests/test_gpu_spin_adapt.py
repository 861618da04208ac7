"""CSF spin adaptation on the GPU (CI_SPIN_ADAPT) against the oracle's SpinAdapter restatement, the
reference's printed CSF counts and the determinant-basis solver."""
import sys

import numpy as np
import pytest

from conftest import seeded_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def solver(p, nroot=1, mult=1, orbsym=None, irrep=0):
    import forte_amd
    ints = forte_amd.ActiveSpaceIntegrals(p["n"], p["h"], p["h"], p["aa"], p["ab"], p["aa"], p["e_core"])
    if orbsym is not None:
        ints.set_active_mo_symmetry(list(orbsym))
        return forte_amd.FCISolver(p["na"], p["nb"], mult, nroot, ints, irrep)
    return forte_amd.FCISolver(p["na"], p["nb"], mult, nroot, ints)


GOLDEN_COUNTS = [
    ("sa-fci-1", [3, 0, 1, 1, 0, 3, 1, 1], 3, 3, 0, 0, (714, 3396)),
    ("sa-fci-2", [3, 0, 1, 1, 0, 3, 1, 1], 3, 3, 4, 4, (330, 3780)),
    ("sa-fci-6", [9, 1, 4, 4], 4, 4, 0, 0, (596304, 11119508)),
    ("sa-fci-7", [7, 0, 2, 2], 4, 3, 1, 2, (7240, 54616)),
    ("sa-fci-rdms-2", [4, 0, 1, 1], 3, 2, 1, 0, (64, 100)),
]


@pytest.mark.parametrize("case", GOLDEN_COUNTS, ids=[c[0] for c in GOLDEN_COUNTS])
def test_csf_counts_match_the_reference_outputs(case):
    name, per_irrep, na, nb, twoS, irrep, expect = case
    orbsym = [h for h, m in enumerate(per_irrep) for _ in range(m)]
    n = len(orbsym)
    say(f"{name}: n={n} na={na} nb={nb} 2S={twoS} irrep={irrep}")
    p = seeded_problem(n, na, nb, 3)
    s = solver(p, 1, twoS + 1, orbsym, irrep)
    assert tuple(s.csf_size()) == expect


CASES = [(6, 3, 3, 0, None, 0), (6, 3, 3, 2, None, 0), (7, 4, 3, 1, None, 0), (7, 4, 3, 3, None, 0),
         (8, 4, 4, 0, [0, 1, 0, 1, 2, 3, 2, 3], 1), (8, 5, 3, 2, [0, 0, 1, 1, 2, 2, 3, 3], 0), (5, 1, 1, 0, None, 0),
         (6, 6, 0, 6, None, 0)]


@pytest.mark.parametrize("n,na,nb,twoS,orbsym,irrep", CASES)
def test_transforms_and_diagonal_vs_oracle(oracle, n, na, nb, twoS, orbsym, irrep):
    from oracle.spin_adapt import SpinAdapter
    say(f"transforms n={n} na={na} nb={nb} 2S={twoS} orbsym={orbsym} irrep={irrep}")
    p = seeded_problem(n, na, nb, 7, e_core=0.25)
    s = solver(p, 1, twoS + 1, orbsym, irrep)
    ref = SpinAdapter(n, na, nb, twoS, orbsym, irrep)
    assert tuple(s.csf_size()) == (ref.ncsf, ref.ncoupling)
    rng = np.random.default_rng(n + na)
    d = rng.standard_normal(s.ndet())
    c = rng.standard_normal(ref.ncsf)
    assert np.abs(s.det_to_csf(d) - ref.det_to_csf(d)).max() < 1e-13
    assert np.abs(s.csf_to_det(c) - ref.csf_to_det(c)).max() < 1e-13
    o = oracle.Oracle(n, na, nb, p["h"], p["h"], p["aa"], p["ab"], p["aa"], p["e_core"])
    hd = o.hdiag()
    assert np.abs(s.hdiag_csf() - ref.hdiag(hd)).max() < 1e-11
    # the full preconditioner: <CSF|H|CSF> from the dense Slater-rules H (scalar energy included)
    s.set_spin_adapt_full_preconditioner(True)
    T = ref.matrix()
    H = o.dense_hamiltonian()
    assert np.abs(s.hdiag_csf() - np.diag(T @ H @ T.T)).max() < 1e-11


@pytest.mark.parametrize("n,na,nb,twoS", [(18, 2, 2, 0), (17, 3, 1, 2), (18, 3, 2, 1)])
def test_transforms_beyond_16_orbitals_vs_oracle(n, na, nb, twoS):
    """Above 16 orbitals the string ranks are summed electron by electron (the byte tables cover
    16-bit strings): both transforms against the oracle's SpinAdapter."""
    from oracle.spin_adapt import SpinAdapter
    say(f"transforms n={n} na={na} nb={nb} 2S={twoS}")
    p = seeded_problem(n, na, nb, 9)
    s = solver(p, 1, twoS + 1)
    ref = SpinAdapter(n, na, nb, twoS)
    assert tuple(s.csf_size()) == (ref.ncsf, ref.ncoupling)
    rng = np.random.default_rng(n + na + twoS)
    d = rng.standard_normal(s.ndet())
    c = rng.standard_normal(ref.ncsf)
    assert np.abs(s.det_to_csf(d) - ref.det_to_csf(d)).max() < 1e-13
    assert np.abs(s.csf_to_det(c) - ref.csf_to_det(c)).max() < 1e-13


@pytest.mark.parametrize("n,na,nb,twoS", [(13, 7, 6, 1), (14, 7, 7, 0), (14, 7, 7, 2), (12, 8, 4, 4)])
def test_transforms_of_large_open_shell_classes_sampled(n, na, nb, twoS):
    """Every open-shell class, including those of 10 to 14 open shells that the dense kernels do not
    take (row-per-wave gathers, pattern-per-wave scatters, scatters with lanes over configurations),
    at sizes beyond a full oracle transform: sampled configurations of each class, their CSF offsets
    from the oracle's configuration order, rows and columns of T built from the oracle's couplings,
    overlaps and determinant signs, against det_to_csf of a random determinant vector and
    csf_to_det of a random CSF vector; and the round trip det_to_csf(csf_to_det(c)) = c."""
    from oracle.oracle import make_strings
    from oracle.spin_adapt import (_popc, configurations, csf_determinant, det_occupations, overlap,
                                   spin_couplings)
    say(f"large classes n={n} na={na} nb={nb} 2S={twoS}")
    p = seeded_problem(n, na, nb, 7)
    s = solver(p, 1, twoS + 1)
    confs = [c for c in configurations(n, na, nb) if _popc(c[1]) >= twoS]
    cpl, occ = {}, {}
    offs, by_class = [], {}
    off = 0
    for k, (docc, socc) in enumerate(confs):
        N = _popc(socc)
        if N not in cpl:
            cpl[N], occ[N] = spin_couplings(N, twoS), det_occupations(N, na - nb)
        offs.append(off)
        off += len(cpl[N])
        by_class.setdefault(N, []).append(k)
    assert s.csf_size()[0] == off
    ra = {int(x): i for i, x in enumerate(make_strings(n, na))}
    sb = make_strings(n, nb)
    rb = {int(x): i for i, x in enumerate(sb)}
    rng = np.random.default_rng(n + twoS)
    d = rng.standard_normal(s.ndet())
    c = rng.standard_normal(off)
    got_csf, got_det = s.det_to_csf(d), s.csf_to_det(c)
    assert np.abs(s.det_to_csf(got_det) - c).max() < 1e-12
    for N, members in sorted(by_class.items()):
        pick = sorted({members[0], members[-1], *rng.choice(members, size=min(3, len(members)), replace=False)})
        rows = sorted({0, len(cpl[N]) - 1, *rng.choice(len(cpl[N]), size=min(4, len(cpl[N])), replace=False)})
        cols = sorted({0, len(occ[N]) - 1, *rng.choice(len(occ[N]), size=min(4, len(occ[N])), replace=False)})
        err_c = err_d = 0.0
        for k in pick:
            docc, socc = confs[k]
            dets = [csf_determinant(docc, socc, o) for o in occ[N]]
            addr = [ra[a] * len(sb) + rb[b] for a, b, _ in dets]
            for i in rows:
                want = sum(sg * ov * d[ad] for (_, _, sg), ad, o in zip(dets, addr, occ[N])
                           if (ov := overlap(N, cpl[N][i], o)) != 0.0)
                err_c = max(err_c, abs(got_csf[offs[k] + i] - want))
            for j in cols:
                _, _, sg = dets[j]
                want = sum(sg * overlap(N, cp, occ[N][j]) * c[offs[k] + i] for i, cp in enumerate(cpl[N]))
                err_d = max(err_d, abs(got_det[addr[j]] - want))
        say(f"  N={N}: {len(members)} configurations, {len(cpl[N])} CSFs x {len(occ[N])} determinants each; "
            f"max errors det->csf {err_c:.1e} csf->det {err_d:.1e}")
        assert err_c < 1e-12 and err_d < 1e-12


SOLVES = [(6, 3, 3, 0, 3, None, 0), (6, 3, 3, 2, 2, None, 0), (7, 4, 3, 1, 2, None, 0), (7, 4, 3, 3, 1, None, 0),
          (8, 4, 4, 0, 2, [0, 1, 0, 1, 2, 3, 2, 3], 1), (8, 4, 4, 4, 1, None, 0)]


@pytest.mark.parametrize("n,na,nb,twoS,nroot,orbsym,irrep", SOLVES)
def test_spin_adapted_solve_vs_oracle(oracle, n, na, nb, twoS, nroot, orbsym, irrep):
    """Energies, <S^2> and eigenvectors of the CSF-basis Davidson-Liu against the spectrum of H in
    the oracle's CSF basis; the determinant-basis eigenvector is the CSF one transformed."""
    from oracle.spin_adapt import SpinAdapter
    say(f"CSF solve n={n} na={na} nb={nb} 2S={twoS} nroot={nroot} orbsym={orbsym} irrep={irrep}")
    p = seeded_problem(n, na, nb, 11, e_core=0.5)
    s = solver(p, nroot, twoS + 1, orbsym, irrep)
    s.set_spin_adapt(True)
    s.set_e_convergence(1e-12)
    s.set_r_convergence(1e-8)
    s.compute_energy()
    assert s.stats()["converged"]
    ref = SpinAdapter(n, na, nb, twoS, orbsym, irrep)
    o = oracle.Oracle(n, na, nb, p["h"], p["h"], p["aa"], p["ab"], p["aa"], p["e_core"])
    T = ref.matrix()
    w, V = np.linalg.eigh(T @ o.dense_hamiltonian() @ T.T)
    E = np.array(s.energies())
    say(f"  gpu {E} oracle {w[:nroot]}")
    assert np.abs(E - w[:nroot]).max() < 1e-9  # the dense H holds the scalar energy
    S = twoS / 2
    assert np.abs(np.array(s.spin2()) - S * (S + 1)).max() < 1e-9
    for r in range(nroot):
        cc = s.evec_csf(r)
        cd = s.evec(r)
        assert np.abs(ref.csf_to_det(cc) - cd).max() < 1e-12
        assert abs(abs(cc @ V[:, r]) - 1.0) < 1e-8


def test_spin_adapted_matches_determinant_solve_and_rdms(oracle):
    """CAS(10,10) singlet and triplet: the CSF-basis solve lands on the determinant solve's energies
    and RDMs."""
    p = seeded_problem(10, 5, 5, 21)
    for mult in (1, 3):
        a = solver(p, 2, mult)
        a.set_e_convergence(1e-11)
        a.set_r_convergence(1e-7)
        a.set_maxiter(500)
        a.compute_energy()
        b = solver(p, 2, mult)
        b.set_spin_adapt(True)
        b.set_e_convergence(1e-11)
        b.set_r_convergence(1e-7)
        b.set_maxiter(500)
        b.compute_energy()
        say(f"  mult {mult}: det {a.energies()} ({a.stats()['iterations']} it), "
            f"csf {b.energies()} ({b.stats()['iterations']} it), ncsf {b.ncsf()} of {b.ndet()} dets")
        assert np.abs(np.array(a.energies()) - np.array(b.energies())).max() < 1e-9
        ra, rb = a.rdms([(0, 0)], 2)[0], b.rdms([(0, 0)], 2)[0]
        for x, y in zip([ra.g1a(), ra.g2aa(), ra.g2ab()], [rb.g1a(), rb.g2aa(), rb.g2ab()]):
            assert np.abs(x - y).max() < 1e-7


@pytest.mark.parametrize("n,na,nb,mult", [(12, 6, 6, 1), (13, 7, 6, 2)])
def test_spin_adapted_mixed_precision(n, na, nb, mult):
    """Mixed precision in the CSF basis (FP32 determinant sigma, FP64 refinement) reaches the FP64
    CSF energy; the 13-orbital doublet runs the FP32 transforms of classes with 11 open shells
    (156 configurations: lanes over configurations) and 13 (one configuration)."""
    p = seeded_problem(n, na, nb, 5)
    e = {}
    for prec in ("fp64", "mixed"):
        s = solver(p, 1, mult)
        s.set_spin_adapt(True)
        s.set_precision(prec)
        s.set_e_convergence(1e-10)
        s.set_r_convergence(1e-6)
        s.set_maxiter(500)
        e[prec] = s.compute_energy()
        say(f"  {prec}: {e[prec]:.12f} {s.stats()}")
        assert s.stats()["converged"]
    assert abs(e["fp64"] - e["mixed"]) < 1e-8


@pytest.mark.parametrize("k,kw,ref", [
    (1, dict(frozen_docc=[2, 0, 0, 0], restricted_docc=[2, 0, 0, 0], active=[2, 2, 2, 2]), -112.74446815362198),
    (6, dict(frozen_docc=[2, 0, 0, 0], restricted_docc=[2, 0, 0, 0], active=[2, 2, 2, 2], nel=13, multiplicity=2),
     -112.181397694724)])
def test_spin_adapted_d2h_kats(k, kw, ref):
    """The reference's D2h known answers integrals-fcidump-1 / -6 with CI_SPIN_ADAPT."""
    import os
    import forte_amd
    from conftest import GOLDEN
    prob = forte_amd.fci_problem_from_fcidump(os.path.join(GOLDEN, "integrals-fcidump-1.INTDUMP.gz"), **kw)
    s = forte_amd.FCISolver(prob["na"], prob["nb"], prob["multiplicity"], 1, prob["ints"], prob["irrep"])
    s.set_spin_adapt(True)
    s.set_e_convergence(1e-12)
    e = s.compute_energy()
    say(f"  D2h KAT {k} with spin adaptation: {e:.12f} ({s.ncsf()} CSFs)")
    assert abs(e - ref) < 1e-9


def test_spin_adapt_rejects_learned_preconditioner():
    import forte_amd.learned
    p = seeded_problem(6, 3, 3, 1)
    s = solver(p)
    s.set_spin_adapt(True)
    s.set_lp_weights(forte_amd.learned.load_weights())
    s.set_preconditioner("learned")
    with pytest.raises(RuntimeError, match="CI_SPIN_ADAPT"):
        s.compute_energy()
