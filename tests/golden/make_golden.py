"""Writes tests/golden/cas_fixtures.npz: small golden vectors for the FCI path.

Produced by the CPU oracle (oracle/fci_oracle.c) and numpy.linalg.eigh on the dense
Slater-rules Hamiltonian; run from the repository root:  python tests/golden/make_golden.py
Contents (seeded integrals of the benchmark generator):
  cfg1_h, cfg1_eri         CAS(8,8) seed 1 integrals, with their SHA-256 digests
  cfg1_E0                  lowest singlet energy (dense eigh)
  cas88_E3                 three lowest singlet energies of CAS(8,8) seed 1 (dense eigh)
  cfg1_g1a..g2bb           root-0 RDMs of CAS(8,8) seed 1 (oracle, eigh eigenvector)
  strings_8_4              the 70 strings of 4-in-8 in Forte order
  vo_8_4_<p>_<q>           (sign, I, J) of every VO list for 4 electrons in 8 orbitals
  oo_8_4_<pq>              every OO list
  vvoo_8_4_sha256          SHA-256 of all VVOO lists concatenated in (p,q,r,s) order
  seeded_<n>_<seed>_*_sha256  digests of the larger seeded integral sets (CAS(14,14) seed 2,
                           CAS(16,16) seed 3, CAS(12,12) seed 2000, CAS(10,10) seed 1000)
  sigma66_C, sigma66_S     a seeded vector and H C for CAS(6,6) seed 66 (e_core 0.5)
  sigma88_C, sigma88_S     the same for CAS(8,8) seed 88
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    out = {}
    h, eri = O.seeded_integrals(8, 1)
    out["cfg1_h"], out["cfg1_eri"] = h, eri
    out["cfg1_h_sha256"], out["cfg1_eri_sha256"] = np.array(sha(h)), np.array(sha(eri))
    o = O.Oracle.from_chem(h, eri, 4, 4)
    H = o.dense_hamiltonian()
    ev, V = np.linalg.eigh(H)
    singlet = [i for i in range(len(ev)) if abs(o.spin2(V[:, i])) < 1e-6]
    out["cfg1_E0"] = np.array(ev[singlet[0]])
    out["cas88_E3"] = ev[singlet[:3]]
    for name, g in zip(["g1a", "g1b", "g2aa", "g2ab", "g2bb"], o.rdms(V[:, singlet[0]])):
        out["cfg1_" + name] = g
    out["strings_8_4"] = o.strings(0)
    vv = []
    for p in range(8):
        for q in range(8):
            out[f"vo_8_4_{p}_{q}"] = np.stack([np.asarray(x, dtype=np.float64) for x in o.vo(0, p, q)])
    for pq in range(o.npair()):
        out[f"oo_8_4_{pq}"] = np.stack([np.asarray(x, dtype=np.float64) for x in o.oo(0, pq)])
    for p in range(8):
        for q in range(p):
            for r in range(8):
                for s in range(r):
                    if (p, q) != (r, s):
                        vv.extend(np.asarray(x, dtype=np.float64) for x in o.vvoo(0, p, q, r, s))
    out["vvoo_8_4_sha256"] = np.array(sha(np.concatenate(vv)))
    for n, seed in ((14, 2), (16, 3), (12, 2000), (10, 1000)):
        h2, eri2 = O.seeded_integrals(n, seed)
        out[f"seeded_{n}_{seed}_h_sha256"] = np.array(sha(h2))
        out[f"seeded_{n}_{seed}_eri_sha256"] = np.array(sha(eri2))
    for n, seed in ((6, 66), (8, 88)):
        h2, eri2 = O.seeded_integrals(n, seed)
        o2 = O.Oracle.from_chem(h2, eri2, n // 2, n // 2, e_core=0.5)
        C = np.random.default_rng(seed).standard_normal(o2.ndet)
        out[f"sigma{n}{n}_C"] = C
        out[f"sigma{n}{n}_S"] = o2.sigma(C)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "cas_fixtures.npz"), **out)


if __name__ == "__main__":
    main()
