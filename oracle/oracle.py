"""ctypes wrapper of the CPU oracle (oracle/fci_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package. See fci_oracle.h for the reference files each
function restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libfci_oracle.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_u32 = ctypes.POINTER(ctypes.c_uint32)
_u64 = ctypes.POINTER(ctypes.c_uint64)
_i64 = ctypes.POINTER(ctypes.c_int64)
_i32 = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_binom.restype = ctypes.c_uint64
        L.orc_binom.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_make_strings.restype = ctypes.c_int64
        L.orc_make_strings.argtypes = [ctypes.c_int, ctypes.c_int, _u64]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int] * 3 + [_d] * 5 + [ctypes.c_double]
        L.orc_create_strings.restype = ctypes.c_void_p
        L.orc_create_strings.argtypes = [ctypes.c_int] * 3 + [_d] * 5 + [ctypes.c_double]
        L.orc_has_lists.restype = ctypes.c_int
        L.orc_has_lists.argtypes = [ctypes.c_void_p]
        L.orc_sigma_rows.argtypes = [ctypes.c_void_p, _d, ctypes.c_int64, _i64, _d]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_last_error.restype = ctypes.c_char_p
        for f in ("orc_nstr",):
            getattr(L, f).restype = ctypes.c_int64
            getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_ndet.restype = ctypes.c_int64
        L.orc_ndet.argtypes = [ctypes.c_void_p]
        L.orc_strings.argtypes = [ctypes.c_void_p, ctypes.c_int, _u64]
        L.orc_address.restype = ctypes.c_int64
        L.orc_address.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64]
        L.orc_vo_size.restype = ctypes.c_int64
        L.orc_vo_size.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 3
        L.orc_vo_get.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 3 + [_d, _u32, _u32]
        L.orc_npair.restype = ctypes.c_int64
        L.orc_npair.argtypes = [ctypes.c_void_p]
        L.orc_pair.argtypes = [ctypes.c_void_p, ctypes.c_int64, _i32, _i32]
        L.orc_oo_size.restype = ctypes.c_int64
        L.orc_oo_size.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64]
        L.orc_oo_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, _d, _u32, _u32]
        L.orc_vvoo_size.restype = ctypes.c_int64
        L.orc_vvoo_size.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5
        L.orc_vvoo_get.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5 + [_d, _u32, _u32]
        L.orc_sigma.argtypes = [ctypes.c_void_p, _d, _d, ctypes.c_int]
        L.orc_sigma_timers.argtypes = [ctypes.c_void_p, _d]
        L.orc_sigma_sample.argtypes = [ctypes.c_void_p, _d, _d, ctypes.c_int, _d, _d]
        L.orc_sigma_aabb_range.argtypes = [ctypes.c_void_p, _d, _d, ctypes.c_int, ctypes.c_int]
        L.orc_hdiag.argtypes = [ctypes.c_void_p, _d]
        L.orc_spin2.restype = ctypes.c_double
        L.orc_spin2.argtypes = [ctypes.c_void_p, _d]
        L.orc_rdms.argtypes = [ctypes.c_void_p, _d, _d, ctypes.c_int] + [_d] * 5
        L.orc_dense_hamiltonian.argtypes = [ctypes.c_void_p, _d]
        L.orc_brute_sigma.argtypes = [ctypes.c_void_p, _d, _d]
        L.orc_brute_rdms.argtypes = [ctypes.c_void_p, _d, _d] + [_d] * 5
        L.orc_brute_rdm3.argtypes = [ctypes.c_void_p, _d, _d] + [_d] * 4
        L.orc_jacobi.argtypes = [ctypes.c_int, _d, _d, _d]
        L.orc_dl_dense.restype = ctypes.c_int
        L.orc_dl_dense.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, _d, _d,
                                   ctypes.c_int, _i64, _i64, _d, ctypes.c_int, _i64, _i64, _d,
                                   _d, _d, _i32]
        L.orc_fci_solve.restype = ctypes.c_int
        L.orc_fci_solve.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _d, _d, _d, _i32]
        L.orc_seeded_integrals.argtypes = [ctypes.c_int, ctypes.c_uint64, _d, _d]
        L.orc_chem_to_forte.argtypes = [ctypes.c_int, _d, _d, _d]
        _lib = L
    return _lib


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


def string_classes(n, k, orbsym):
    """FCIStringLists::make_fci_strings with symmetry (fci_string_lists.cc:258-293,
    fci_string_address.cc:36-59): the strings in next_permutation order, each filed under its irrep
    (XOR of its orbitals' irreps) at the next address of that irrep. Returns (irrep, address) per
    string in enumeration order (the C1 order)."""
    strs = make_strings(n, k)
    irrep = np.zeros(len(strs), dtype=np.int64)
    addr = np.zeros(len(strs), dtype=np.int64)
    seen = {}
    for i, x in enumerate(strs):
        h = 0
        for p in range(n):
            if (int(x) >> p) & 1:
                h ^= orbsym[p]
        irrep[i] = h
        addr[i] = seen.get(h, 0)
        seen[h] = addr[i] + 1
    return irrep, addr


def sector_mask(o, orbsym, irrep):
    """Determinants (C1 order) whose irrep is `irrep`."""
    def sym(s):
        h = 0
        for p in range(o.n):
            if (int(s) >> p) & 1:
                h ^= orbsym[p]
        return h
    sa = np.array([sym(x) for x in o.strings(0)], dtype=np.int64)
    sb = np.array([sym(x) for x in o.strings(1)], dtype=np.int64)
    return ((sa[:, None] ^ sb[None, :]) == irrep).ravel()


def gas_mask(n, na, nb, gas_of_orbital, gas_min, gas_max):
    """Determinants (C1 order) of GenCI's GAS space: every GAS space g holds between gas_min[g] and
    gas_max[g] electrons of both spins (the occupation patterns of genci/ci_occupation.cc:64-123
    are exactly those with these per-space totals)."""
    ng = len(gas_min)
    orbs = [sum(1 << p for p in range(n) if gas_of_orbital[p] == g) for g in range(ng)]
    def counts(strs):
        return np.array([[bin(int(x) & m).count("1") for m in orbs] for x in strs], dtype=np.int64).reshape(-1, ng)
    ca, cb = counts(make_strings(n, na)), counts(make_strings(n, nb))
    tot = ca[:, None, :] + cb[None, :, :]
    ok = np.all((tot >= np.array(gas_min)) & (tot <= np.array(gas_max)), axis=2)
    return ok.ravel()


class DLOpts(ctypes.Structure):
    _fields_ = [("collapse_per_root", ctypes.c_int), ("subspace_per_root", ctypes.c_int),
                ("maxiter", ctypes.c_int), ("e_convergence", ctypes.c_double),
                ("r_convergence", ctypes.c_double)]


class FCIOpts(ctypes.Structure):
    _fields_ = [("nroot", ctypes.c_int), ("multiplicity", ctypes.c_int),
                ("guess_per_root", ctypes.c_int), ("ndets_per_guess", ctypes.c_int),
                ("collapse_per_root", ctypes.c_int), ("subspace_per_root", ctypes.c_int),
                ("maxiter", ctypes.c_int), ("e_convergence", ctypes.c_double),
                ("r_convergence", ctypes.c_double)]


def binom(n, k):
    return int(lib().orc_binom(n, k))


def make_strings(n, k):
    cnt = binom(n, k)
    out = np.zeros(max(cnt, 1), dtype=np.uint64)
    lib().orc_make_strings(n, k, _p(out, ctypes.c_uint64))
    return out[:cnt]


def seeded_integrals(n, seed):
    h = np.zeros((n, n))
    eri = np.zeros((n, n, n, n))
    lib().orc_seeded_integrals(n, seed, _p(h), _p(eri))
    return h, eri


def chem_to_forte(eri):
    n = eri.shape[0]
    aa = np.zeros_like(eri)
    ab = np.zeros_like(eri)
    lib().orc_chem_to_forte(n, _p(np.ascontiguousarray(eri)), _p(aa), _p(ab))
    return aa, ab


def jacobi(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    ev = np.zeros(n)
    V = np.zeros((n, n))
    lib().orc_jacobi(n, _p(A), _p(ev), _p(V))
    return ev, V  # V[k] is eigenvector k


def dl_dense(M, nroot, hdiag=None, guesses=None, project_out=None, collapse_per_root=1,
             subspace_per_root=5, maxiter=50, e_conv=1e-12, r_conv=1e-6, no_sigma=False):
    """The reference's DavidsonLiuSolver driven by add_test_sigma_builder(M)."""
    M = np.ascontiguousarray(M, dtype=np.float64)
    size = M.shape[0]

    def sparse(vs):
        vs = vs or []
        off = [0]
        idx, val = [], []
        for v in vs:
            for i, c in v:
                idx.append(i)
                val.append(c)
            off.append(len(idx))
        return (np.array(off, dtype=np.int64), np.array(idx or [0], dtype=np.int64),
                np.array(val or [0.0]), len(vs))

    goff, gidx, gval, ng = sparse(guesses)
    poff, pidx, pval, np_ = sparse(project_out)
    o = DLOpts(collapse_per_root, subspace_per_root, maxiter, e_conv, r_conv)
    evals = np.zeros(nroot)
    evecs = np.zeros((nroot, size))
    it = ctypes.c_int(0)
    hd = None if hdiag is None else _p(np.ascontiguousarray(hdiag, dtype=np.float64))
    rc = lib().orc_dl_dense(size, nroot, ctypes.byref(o), None if no_sigma else _p(M), hd, ng,
                            _p(goff, ctypes.c_int64), _p(gidx, ctypes.c_int64), _p(gval), np_,
                            _p(poff, ctypes.c_int64), _p(pidx, ctypes.c_int64), _p(pval),
                            _p(evals), _p(evecs), ctypes.byref(it))
    if rc < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return rc == 1, evals, evecs, it.value


class Oracle:
    """One FCI problem (strings, lists, sigma, RDMs) in Forte's layout."""

    def __init__(self, n, na, nb, oei_a, oei_b, tei_aa, tei_ab, tei_bb, e_core=0.0, lists=True):
        """lists=False builds strings and addresses only: sigma_rows, hdiag and
        dense_hamiltonian work, the string-list functions do not."""
        self.n, self.na, self.nb = n, na, nb
        self._keep = [np.ascontiguousarray(x, dtype=np.float64)
                      for x in (oei_a, oei_b, tei_aa, tei_ab, tei_bb)]
        create = lib().orc_create if lists else lib().orc_create_strings
        self.h = create(n, na, nb, *[_p(x) for x in self._keep], float(e_core))
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())
        self.lists = bool(lists)
        self.nas = int(lib().orc_nstr(self.h, 0))
        self.nbs = int(lib().orc_nstr(self.h, 1))
        self.ndet = self.nas * self.nbs

    @classmethod
    def from_chem(cls, h, eri, na, nb, e_core=0.0, lists=True):
        aa, ab = chem_to_forte(eri)
        return cls(h.shape[0], na, nb, h, h, aa, ab, aa, e_core, lists)

    def sigma_rows(self, C, rows):
        """(H C)[rows] by the Slater-Condon rules over the determinants each row connects to."""
        C = np.ascontiguousarray(C, dtype=np.float64).reshape(-1)
        if C.size != self.ndet:
            raise ValueError("C has the wrong length")
        rows = np.ascontiguousarray(rows, dtype=np.int64).reshape(-1)
        if rows.size and (rows.min() < 0 or rows.max() >= self.ndet):
            raise ValueError("row index out of range")
        out = np.zeros(rows.size)
        lib().orc_sigma_rows(self.h, _p(C), rows.size, _p(rows, ctypes.c_int64), _p(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def strings(self, spin):
        out = np.zeros(max(self.nas if spin == 0 else self.nbs, 1), dtype=np.uint64)
        lib().orc_strings(self.h, spin, _p(out, ctypes.c_uint64))
        return out[: (self.nas if spin == 0 else self.nbs)]

    def address(self, spin, s):
        return int(lib().orc_address(self.h, spin, int(s)))

    def _need_lists(self):
        if not self.lists:
            raise RuntimeError("this oracle context was built without excitation lists")

    def _list(self, size, getter, *args):
        self._need_lists()
        sign = np.zeros(max(size, 1))
        I = np.zeros(max(size, 1), dtype=np.uint32)
        J = np.zeros(max(size, 1), dtype=np.uint32)
        getter(self.h, *args, _p(sign), _p(I, ctypes.c_uint32), _p(J, ctypes.c_uint32))
        return sign[:size], I[:size], J[:size]

    def vo(self, spin, p, q):
        self._need_lists()
        return self._list(int(lib().orc_vo_size(self.h, spin, p, q)), lib().orc_vo_get, spin, p, q)

    def npair(self):
        return int(lib().orc_npair(self.h))

    def pair(self, pq):
        p, q = ctypes.c_int(), ctypes.c_int()
        lib().orc_pair(self.h, pq, ctypes.byref(p), ctypes.byref(q))
        return p.value, q.value

    def oo(self, spin, pq):
        self._need_lists()
        return self._list(int(lib().orc_oo_size(self.h, spin, pq)), lib().orc_oo_get, spin,
                          ctypes.c_int64(pq))

    def vvoo(self, spin, p, q, r, s):
        self._need_lists()
        return self._list(int(lib().orc_vvoo_size(self.h, spin, p, q, r, s)), lib().orc_vvoo_get,
                          spin, p, q, r, s)

    def sigma(self, C, mask=63):
        self._need_lists()
        C = np.ascontiguousarray(C, dtype=np.float64).reshape(-1)
        out = np.zeros(self.ndet)
        lib().orc_sigma(self.h, _p(C), _p(out), mask)
        return out

    def sigma_timers(self):
        t = np.zeros(6)
        lib().orc_sigma_timers(self.h, _p(t))
        return t

    def sigma_sample(self, C, stride):
        """Strided timing sample of the Forte-structured sigma: (seconds[6], fraction[6])."""
        self._need_lists()
        C = np.ascontiguousarray(C, dtype=np.float64).reshape(-1)
        out = np.zeros(self.ndet)
        t, w = np.zeros(6), np.zeros(6)
        lib().orc_sigma_sample(self.h, _p(C), _p(out), int(stride), _p(t), _p(w))
        return t, w

    def sigma_aabb_range(self, C, rs_begin, rs_end, out=None):
        self._need_lists()
        C = np.ascontiguousarray(C, dtype=np.float64).reshape(-1)
        if out is None:
            out = np.zeros(self.ndet)
        lib().orc_sigma_aabb_range(self.h, _p(C), _p(out), rs_begin, rs_end)
        return out

    def hdiag(self):
        out = np.zeros(self.ndet)
        lib().orc_hdiag(self.h, _p(out))
        return out

    def spin2(self, C):
        self._need_lists()
        C = np.ascontiguousarray(C, dtype=np.float64).reshape(-1)
        return float(lib().orc_spin2(self.h, _p(C)))

    def rdms(self, CL, CR=None, level=2):
        self._need_lists()
        CL = np.ascontiguousarray(CL, dtype=np.float64).reshape(-1)
        CR = CL if CR is None else np.ascontiguousarray(CR, dtype=np.float64).reshape(-1)
        n = self.n
        g1a, g1b = np.zeros((n, n)), np.zeros((n, n))
        g2 = [np.zeros((n, n, n, n)) for _ in range(3)]
        lib().orc_rdms(self.h, _p(CL), _p(CR), level, _p(g1a), _p(g1b), *[_p(g) for g in g2])
        return g1a, g1b, g2[0], g2[1], g2[2]

    def dense_hamiltonian(self):
        H = np.zeros((self.ndet, self.ndet))
        lib().orc_dense_hamiltonian(self.h, _p(H))
        return H

    def brute_sigma(self, C):
        C = np.ascontiguousarray(C, dtype=np.float64).reshape(-1)
        out = np.zeros(self.ndet)
        lib().orc_brute_sigma(self.h, _p(C), _p(out))
        return out

    def brute_rdms(self, CL, CR=None):
        CL = np.ascontiguousarray(CL, dtype=np.float64).reshape(-1)
        CR = CL if CR is None else np.ascontiguousarray(CR, dtype=np.float64).reshape(-1)
        n = self.n
        g1a, g1b = np.zeros((n, n)), np.zeros((n, n))
        g2 = [np.zeros((n, n, n, n)) for _ in range(3)]
        lib().orc_brute_rdms(self.h, _p(CL), _p(CR), _p(g1a), _p(g1b), *[_p(g) for g in g2])
        return g1a, g1b, g2[0], g2[1], g2[2]

    def brute_rdm3(self, CL, CR=None):
        """(g3aaa, g3aab, g3abb, g3bbb): <CL| a+p a+q a+r a_u a_t a_s |CR> at [p,q,r,s,t,u]."""
        CL = np.ascontiguousarray(CL, dtype=np.float64).reshape(-1)
        CR = CL if CR is None else np.ascontiguousarray(CR, dtype=np.float64).reshape(-1)
        n = self.n
        g = [np.zeros((n,) * 6) for _ in range(4)]
        lib().orc_brute_rdm3(self.h, _p(CL), _p(CR), *[_p(x) for x in g])
        return tuple(g)

    def solve(self, nroot=1, multiplicity=1, guess_per_root=1, ndets_per_guess=50,
              collapse_per_root=2, subspace_per_root=10, maxiter=100, e_conv=1e-10,
              r_conv=1e-6, want_evecs=True):
        self._need_lists()
        o = FCIOpts(nroot, multiplicity, guess_per_root, ndets_per_guess, collapse_per_root,
                    subspace_per_root, maxiter, e_conv, r_conv)
        e = np.zeros(nroot)
        s2 = np.zeros(nroot)
        ev = np.zeros((nroot, self.ndet)) if want_evecs else None
        it = ctypes.c_int(0)
        rc = lib().orc_fci_solve(self.h, ctypes.byref(o), _p(e), _p(s2),
                                 _p(ev) if want_evecs else None, ctypes.byref(it))
        if rc < 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return dict(converged=rc == 1, energies=e, spin2=s2, evecs=ev, niter=it.value)
