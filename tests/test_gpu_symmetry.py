"""Point-group symmetry on the GPU: the reference's D2h FCIDUMP known answers (frozen and
restricted docc folded into the core), per-irrep string classes bit-exact against the oracle, and
the lowest root of every irrep of a symmetric problem against a dense diagonalization of its sector."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, seeded_problem

pytestmark = pytest.mark.gpu

KAT_SYM = {
    1: (dict(frozen_docc=[2, 0, 0, 0], restricted_docc=[2, 0, 0, 0], active=[2, 2, 2, 2]), -112.74446815362198, 784),
    6: (dict(frozen_docc=[2, 0, 0, 0], restricted_docc=[2, 0, 0, 0], active=[2, 2, 2, 2], nel=13, multiplicity=2),
        -112.181397694724, 392),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("k", sorted(KAT_SYM))
def test_kat_d2h_fcidump(k):
    import forte_amd
    kw, ref, ndet = KAT_SYM[k]
    say("D2h KAT", k)
    p = forte_amd.fci_problem_from_fcidump(os.path.join(GOLDEN, "integrals-fcidump-1.INTDUMP.gz"), **kw)
    s = forte_amd.FCISolver(p["na"], p["nb"], p["multiplicity"], 1, p["ints"], p["irrep"])
    s.set_e_convergence(1e-12)
    e = s.compute_energy()
    say(f"  E {e:.12f} reference {ref:.12f} sector {s.sector_size()}")
    assert s.sector_size() == ndet
    assert abs(e - ref) < 1e-9
    mult = p["multiplicity"]
    assert abs(s.spin2()[0] - (mult * mult - 1) / 4.0) < 1e-8
    v = s.evec_sector(0)
    assert v.size == ndet and abs(np.linalg.norm(v) - 1.0) < 1e-10
    c = s.evec(0)
    assert abs(np.linalg.norm(c) - 1.0) < 1e-10  # nothing outside the sector


@pytest.mark.parametrize("n,na,nb,orbsym", [(6, 3, 3, [0, 1, 2, 3, 0, 1]), (8, 4, 4, [0, 0, 1, 1, 2, 2, 3, 3]),
                                            (7, 2, 4, [0, 4, 1, 5, 2, 6, 3]), (10, 5, 4, [0, 1, 2, 3, 4, 5, 6, 7, 0, 1])])
def test_string_classes_bit_exact(oracle, n, na, nb, orbsym):
    import forte_amd
    p = seeded_problem(n, na, nb, 3)
    ints = forte_amd.ActiveSpaceIntegrals(n, p["h"], p["h"], p["aa"], p["ab"], p["aa"])
    ints.set_active_mo_symmetry(orbsym)
    s = forte_amd.FCISolver(na, nb, 1, 1, ints, 0)
    for spin, k in ((0, na), (1, nb)):
        ir, ad = s.string_classes(spin)
        ri, ra = oracle.string_classes(n, k, orbsym)
        assert np.array_equal(ir, ri) and np.array_equal(ad, ra), spin


def symmetric_problem(n, na, nb, orbsym, seed):
    """Seeded integrals with every symmetry-forbidden element set to zero."""
    p = seeded_problem(n, na, nb, seed)
    s = np.array(orbsym)
    h = p["h"] * (s[:, None] == s[None, :])
    allowed = (s[:, None, None, None] ^ s[None, :, None, None] ^ s[None, None, :, None] ^ s[None, None, None, :]) == 0
    eri = None
    from oracle import oracle as O
    _, eri = O.seeded_integrals(n, seed)
    eri = eri * allowed
    aa, ab = O.chem_to_forte(eri)
    return h, aa, ab


@pytest.mark.parametrize("n,na,nb,orbsym,mult", [(6, 3, 3, [0, 0, 1, 1, 2, 3], 1), (7, 3, 2, [0, 1, 0, 2, 1, 3, 0], 2)])
def test_lowest_root_of_every_irrep(oracle, n, na, nb, orbsym, mult):
    import forte_amd
    h, aa, ab = symmetric_problem(n, na, nb, orbsym, 70 + n)
    o = oracle.Oracle(n, na, nb, h, h, aa, ab, aa)
    H = o.dense_hamiltonian()
    for irrep in range(4):
        mask = oracle.sector_mask(o, orbsym, irrep)
        if mask.sum() == 0:
            continue
        ints = forte_amd.ActiveSpaceIntegrals(n, h, h, aa, ab, aa)
        ints.set_active_mo_symmetry(orbsym)
        s = forte_amd.FCISolver(na, nb, mult, 1, ints, irrep)
        s.set_e_convergence(1e-12)
        s.set_r_convergence(1e-8)
        s.set_die_if_not_converged(False)
        s.set_maxiter(200)
        try:
            e = s.compute_energy()
        except RuntimeError as err:  # no guess of this multiplicity in a small sector
            say(f"  irrep {irrep}: {err}")
            continue
        ev, vecs = np.linalg.eigh(H[np.ix_(mask, mask)])
        # the lowest root of the requested multiplicity in the sector
        S2 = []
        for j in range(min(len(ev), 12)):
            full = np.zeros(o.ndet)
            full[mask] = vecs[:, j]
            S2.append(o.spin2(full))
        want = [ev[j] for j in range(len(S2)) if abs(S2[j] - (mult * mult - 1) / 4.0) < 1e-6][0]
        say(f"  irrep {irrep}: sector {mask.sum()} E {e:.12f} dense {want:.12f}")
        assert s.sector_size() == mask.sum()
        assert abs(e - want) < 1e-9
        assert np.abs(s.evec(0)[~mask]).max() == 0.0
