"""CPU restatement of Forte's CSF spin adaptation (forte/sparse_ci/ci_spin_adaptation.cc).

Test infrastructure only: imported by tests/, never by the product package. Pure Python and
numpy, sized for the small cases the tests build explicitly; `count_csfs` counts configurations
combinatorially so that the reference's printed CSF / coupling counts of larger active spaces can
be checked as well.

Conventions (those of the reference):
  - a configuration is (docc, socc) as orbital bit masks; configurations are ordered by
    (socc, docc) compared as integers (ConfigurationImpl is a BitArray whose high word holds the
    singly occupied orbitals, bitarray.hpp:301-326), and only those with at least 2S open shells
    are kept (ci_spin_adaptation.cc:151-156);
  - the CSFs of a configuration with N open shells are the genealogical spin couplings of
    make_spin_couplings (ci_spin_adaptation.cc:232-258): N steps, 0 = up / 1 = down, the first step
    up, never a negative partial spin, in std::next_permutation order of steps 1..N-1;
  - the determinants of a configuration are the det occupations of make_determinant_occupations
    (:260-279): 0 = alpha / 1 = beta on the open shells in ascending orbital order, in
    std::next_permutation order;
  - a coefficient is the Clebsch-Gordan product of overlap() (:60-81) times the sign of creating
    the open-shell electrons, highest open shell first, on top of the closed shells (:197-230).
"""
import math
from itertools import combinations

import numpy as np


def clebsch_gordan(twoS, twoM, dtwoS, dtwoM):
    """ci_spin_adaptation.cc:46-58 (twoS, twoM: the coupled spin after the step)."""
    if dtwoS == 1:
        return math.sqrt(0.5 * (twoS + dtwoM * twoM) / twoS)
    if dtwoS == -1:
        return -dtwoM * math.sqrt(0.5 * (twoS + 2.0 - dtwoM * twoM) / (twoS + 2.0))
    return 0.0


def overlap(N, coupling, occ):
    """<CSF path | determinant path> (ci_spin_adaptation.cc:60-81)."""
    ov = 1.0
    p = q = 0
    for i in range(N):
        dS = 1 - 2 * coupling[i]
        dM = 1 - 2 * occ[i]
        p += dS
        q += dM
        if abs(q) > p:
            return 0.0
        ov *= clebsch_gordan(p, q, dS, dM)
    return ov


def _ordered_01(N, nones, fixed_first_zero):
    """All 0/1 sequences of length N with `nones` ones, in lexicographic order (what
    std::next_permutation enumerates from the sorted start); optionally position 0 held at 0."""
    start = 1 if fixed_first_zero else 0
    out = []
    for ones in combinations(range(start, N), nones):
        s = [0] * N
        for k in ones:
            s[k] = 1
        out.append(tuple(s))
    out.sort()
    return out


def spin_couplings(N, twoS):
    """make_spin_couplings (ci_spin_adaptation.cc:232-258)."""
    if N == 0:
        return [()]
    nup = (N + twoS) // 2
    res = []
    for c in _ordered_01(N, N - nup, True):
        p, ok = 0, True
        for x in c:
            p += 1 - 2 * x
            if p < 0:
                ok = False
        if ok:
            res.append(c)
    return res


def det_occupations(N, twoMs):
    """make_determinant_occupations (ci_spin_adaptation.cc:260-279)."""
    if N == 0:
        return [()]
    nup = (N + twoMs) // 2
    return _ordered_01(N, N - nup, False)


def _bits(x):
    return [i for i in range(x.bit_length()) if (x >> i) & 1]


def _popc(x):
    return bin(x).count("1")


def configurations(n, na, nb, orbsym=None, irrep=0):
    """Sorted configurations (docc, socc) of the determinant space with na / nb electrons (and,
    with orbsym, determinant irrep `irrep`)."""
    twoMs = na - nb
    confs = []
    for d in range(0, min(na, nb) + 1):
        N = na + nb - 2 * d
        if N > n - d or N < abs(twoMs):
            continue
        for docc in combinations(range(n), d):
            dm = sum(1 << i for i in docc)
            rest = [i for i in range(n) if not (dm >> i) & 1]
            for socc in combinations(rest, N):
                sm = sum(1 << i for i in socc)
                if orbsym is not None:
                    h = 0
                    for i in socc:
                        h ^= orbsym[i]
                    if h != irrep:
                        continue
                confs.append((sm, dm))
    confs.sort()
    return [(dm, sm) for sm, dm in confs]


def csf_determinant(docc, socc, occ):
    """(alpha, beta, sign) of the determinant with closed shells `docc` and the open shells of
    `socc` (ascending) occupied by spins `occ` (0 alpha / 1 beta), created highest first on top of
    the closed shells (ci_spin_adaptation.cc:205-218, determinant.hpp:305-362)."""
    so = _bits(socc)
    a, b, sign = docc, docc, 1.0
    for k in range(len(so) - 1, -1, -1):
        p = so[k]
        below = (1 << p) - 1
        if occ[k]:
            sign *= (-1.0) ** (_popc(b & below) + _popc(a))
            b |= 1 << p
        else:
            sign *= (-1.0) ** _popc(a & below)
            a |= 1 << p
    return a, b, sign


class SpinAdapter:
    """CSF <-> determinant map for determinants in the C1 layout C[Ia * Nb + Ib] (Forte string
    addresses). csf[i] = sum_j coeff * det[idx] over the entries of CSF i (det_C_to_csf_C,
    ci_spin_adaptation.cc:89-104); csf_C_to_det_C is the transpose (:106-121)."""

    def __init__(self, n, na, nb, twoS, orbsym=None, irrep=0):
        from oracle.oracle import make_strings
        self.n, self.na, self.nb, self.twoS = n, na, nb, twoS
        twoMs = na - nb
        sa, sb = make_strings(n, na), make_strings(n, nb)
        ra = {int(s): i for i, s in enumerate(sa)}
        rb = {int(s): i for i, s in enumerate(sb)}
        self.nb_str = len(sb)
        self.ndet = len(sa) * len(sb)
        self.confs = [c for c in configurations(n, na, nb, orbsym, irrep) if _popc(c[1]) >= twoS]
        self.bounds = [0]
        self.idx, self.coeff = [], []
        cache = {}
        for docc, socc in self.confs:
            N = _popc(socc)
            if N not in cache:
                cp, oc = spin_couplings(N, twoS), det_occupations(N, twoMs)
                cache[N] = [[(j, overlap(N, c, o)) for j, o in enumerate(oc)] for c in cp], oc
            rows, oc = cache[N]
            for row in rows:
                for j, ov in row:
                    if ov != 0.0:
                        a, b, s = csf_determinant(docc, socc, oc[j])
                        self.idx.append(ra[a] * len(sb) + rb[b])
                        self.coeff.append(s * ov)
                self.bounds.append(len(self.idx))
        self.ncsf = len(self.bounds) - 1
        self.ncoupling = len(self.idx)
        self.idx = np.array(self.idx, dtype=np.int64)
        self.coeff = np.array(self.coeff)

    def matrix(self):
        """Dense T (ncsf x ndet): CSF i = sum_j T[i, j] |det j>."""
        T = np.zeros((self.ncsf, self.ndet))
        for i in range(self.ncsf):
            b0, b1 = self.bounds[i], self.bounds[i + 1]
            T[i, self.idx[b0:b1]] = self.coeff[b0:b1]
        return T

    def det_to_csf(self, det):
        out = np.zeros(self.ncsf)
        for i in range(self.ncsf):
            b0, b1 = self.bounds[i], self.bounds[i + 1]
            out[i] = self.coeff[b0:b1] @ det[self.idx[b0:b1]]
        return out

    def csf_to_det(self, csf):
        out = np.zeros(self.ndet)
        for i in range(self.ncsf):
            b0, b1 = self.bounds[i], self.bounds[i + 1]
            out[self.idx[b0:b1]] += self.coeff[b0:b1] * csf[i]
        return out

    def hdiag(self, hdiag_det):
        """FCISolver::form_Hdiag_csf without the full preconditioner
        (fci_solver_initial_guess.cc:134-141): sum of c^2 times the determinant energies."""
        out = np.zeros(self.ncsf)
        for i in range(self.ncsf):
            b0, b1 = self.bounds[i], self.bounds[i + 1]
            out[i] = np.sum(self.coeff[b0:b1] ** 2 * hdiag_det[self.idx[b0:b1]])
        return out


def count_csfs(norb_per_irrep, na, nb, twoS, irrep=0):
    """(ndet, ncsf, ncoupling) of the sector without building it: configurations are counted
    per (open shells, irrep) over the irreps' orbital counts, and the CSFs / nonzero couplings of
    N open shells are those of make_spin_couplings / overlap. The determinant count is that of
    the FCI sector."""
    twoMs = na - nb
    nh = len(norb_per_irrep)
    # ways[h] = dict[(ndocc, nsocc, irrep of socc)] -> count over the orbitals of irreps < h
    ways = {(0, 0, 0): 1}
    for h, m in enumerate(norb_per_irrep):
        new = {}
        for (d, s, g), w in ways.items():
            for dd in range(m + 1):
                for ss in range(m - dd + 1):
                    c = math.comb(m, dd) * math.comb(m - dd, ss)
                    key = (d + dd, s + ss, g ^ (h if ss % 2 else 0))
                    new[key] = new.get(key, 0) + w * c
        ways = new
    ncsf = ncoup = ndet = 0
    per_N = {}
    for (d, N, g), w in ways.items():
        if g != irrep or 2 * d + N != na + nb or N < abs(twoMs) or (N + twoMs) % 2:
            continue
        if d + (N + twoMs) // 2 != na:
            continue
        ndet += w * math.comb(N, (N - twoMs) // 2)
        if N < twoS:
            continue
        if N not in per_N:
            cp, oc = spin_couplings(N, twoS), det_occupations(N, twoMs)
            nz = sum(1 for c in cp for o in oc if overlap(N, c, o) != 0.0)
            per_N[N] = (len(cp), nz)
        ncsf += w * per_N[N][0]
        ncoup += w * per_N[N][1]
    return ndet, ncsf, ncoup
