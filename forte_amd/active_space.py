"""Active spaces with point-group symmetry and doubly occupied core orbitals, from an FCIDUMP file.

The reference's FCIDUMP path (forte/proc/fcidump.py:262-364, forte/modules/objects_factory_fcidump.py:
28-160) converts the file's orbital irreps to Cotton order, assumes the orbitals are stored irrep by
irrep (Pitzer order), and takes, within each irrep, FROZEN_DOCC orbitals first, then RESTRICTED_DOCC,
then ACTIVE (forte/base_classes/mo_space_info.cc). Frozen and restricted doubly occupied orbitals
are folded into a scalar and a one-body operator (ForteIntegrals' frozen-core operator and
ActiveSpaceIntegrals' restricted-docc operator, forte/integrals/active_space_integrals.cc:44-165):

    E_core = E_file + sum_c 2 h_cc + sum_{c,d} [2 (cc|dd) - (cd|dc)]
    h'_pq  = h_pq + sum_c [2 (pq|cc) - (pc|cq)]            (c, d over all docc orbitals)

The FCI problem is then the active orbitals (in Pitzer order) with their irreps, and the state's
irrep selects the determinants (the product of the occupied orbitals' irreps).
"""
import numpy as np

from .fcidump import chem_to_forte, read_fcidump

# FCIDUMP irrep label (1-based) -> Cotton-order irrep index, per point group (the conversion of
# forte/proc/fcidump.py:245-259; index 0 is unused)
DUMP_TO_COTTON = {
    "c1": [-1, 0], "ci": [-1, 0, 1], "c2": [-1, 0, 1], "cs": [-1, 0, 1],
    "d2": [-1, 0, 3, 2, 1], "c2v": [-1, 0, 2, 3, 1], "c2h": [-1, 0, 2, 3, 1],
    "d2h": [-1, 0, 7, 6, 1, 5, 2, 3, 4],
}
NIRREP = {"c1": 1, "ci": 2, "c2": 2, "cs": 2, "d2": 4, "c2v": 4, "c2h": 4, "d2h": 8}


def cotton_orbsym(d):
    """(orbsym, isym, nirrep) of a read_fcidump dict in Cotton order (C1 without a point group)."""
    pg = (d.get("pntgrp") or "").lower()
    if not pg:
        return [0] * d["norb"], 0, 1
    m = DUMP_TO_COTTON[pg]
    return [m[x] for x in d["orbsym"]], m[d["isym"]], NIRREP[pg]


def mo_spaces(orbsym, nirrep, frozen_docc=None, restricted_docc=None, active=None):
    """Absolute orbital indices of the doubly occupied core (frozen then restricted) and of the
    active orbitals, and the active orbitals' irreps, for Pitzer-ordered orbitals. Missing
    spaces: no docc; all orbitals active."""
    nmopi = [orbsym.count(h) for h in range(nirrep)]
    if list(orbsym) != sorted(orbsym):
        raise ValueError("orbitals must be stored irrep by irrep (Pitzer order)")
    zero = [0] * nirrep
    fd = list(frozen_docc) if frozen_docc is not None else zero
    rd = list(restricted_docc) if restricted_docc is not None else zero
    act = list(active) if active is not None else [nmopi[h] - fd[h] - rd[h] for h in range(nirrep)]
    for name, v in (("FROZEN_DOCC", fd), ("RESTRICTED_DOCC", rd), ("ACTIVE", act)):
        if len(v) != nirrep:
            raise ValueError(f"{name} must have {nirrep} entries")
    core, actv, asym = [], [], []
    off = 0
    for h in range(nirrep):
        if fd[h] + rd[h] + act[h] > nmopi[h]:
            raise ValueError(f"irrep {h}: more orbitals requested than the {nmopi[h]} available")
        core += list(range(off, off + fd[h] + rd[h]))
        actv += list(range(off + fd[h] + rd[h], off + fd[h] + rd[h] + act[h]))
        asym += [h] * act[h]
        off += nmopi[h]
    return core, actv, asym


def fold_core(hcore, eri, e_file, core, active):
    """(h_active, eri_active, E_core): the doubly occupied core folded into a scalar and a
    one-body operator over the active orbitals (chemist-notation ERIs)."""
    c = np.asarray(core, dtype=int)
    a = np.asarray(active, dtype=int)
    e = float(e_file)
    if c.size:
        J = np.einsum("pqcc->pq", eri[:, :, c][:, :, :, c])
        K = np.einsum("pccq->pq", eri[:, c][:, :, c])
        e += 2.0 * np.trace(hcore[np.ix_(c, c)]) + np.trace(2.0 * J[np.ix_(c, c)] - K[np.ix_(c, c)])
        h = hcore + 2.0 * J - K
    else:
        h = hcore
    return (np.ascontiguousarray(h[np.ix_(a, a)]), np.ascontiguousarray(eri[np.ix_(a, a, a, a)]), e)


def active_space_from_fcidump(fname, frozen_docc=None, restricted_docc=None, active=None, nel=None,
                              multiplicity=None, ms=None, root_sym=None):
    """The FCI problem of an FCIDUMP file and Forte's MO-space options:
    dict(h, eri (chemist, active orbitals), e_core, orbsym (active), irrep, na, nb, multiplicity,
    ndocc). Electron counts follow objects_factory_fcidump.py:40-76 (NEL, MULTIPLICITY, MS
    overrides; twice_ms = (multiplicity + 1) % 2 by default) minus the docc electrons."""
    d = read_fcidump(fname)
    orbsym, isym, nirrep = cotton_orbsym(d)
    nelec = d["nelec"] if nel is None else int(nel)
    if multiplicity is None:
        multiplicity = d["ms2"] + 1
    twice_ms = (multiplicity + 1) % 2 if ms is None else int(round(2.0 * ms))
    if (nelec - twice_ms) % 2:
        raise ValueError(f"the value of MS ({twice_ms}/2) is incompatible with the number of electrons ({nelec})")
    na = (nelec + twice_ms) // 2
    nb = nelec - na
    irrep = isym if root_sym is None else int(root_sym)
    core, act, asym = mo_spaces(orbsym, nirrep, frozen_docc, restricted_docc, active)
    h, eri, ecore = fold_core(d["hcore"], d["eri"], d["enuc"], core, act)
    nc = len(core)
    if na - nc < 0 or nb - nc < 0:
        raise ValueError("more doubly occupied orbitals than electrons")
    return dict(h=h, eri=eri, e_core=ecore, orbsym=asym, irrep=irrep, na=na - nc, nb=nb - nc,
                multiplicity=multiplicity, ndocc=nc, nirrep=nirrep)


def forte_layout(h, eri):
    """(oei, tei_aa, tei_ab) in Forte's active-space layout."""
    aa, ab = chem_to_forte(eri)
    return np.ascontiguousarray(h), aa, ab
