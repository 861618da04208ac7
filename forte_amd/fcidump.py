"""FCIDUMP reader (the integral format of Forte's FCIDUMP path).

Mirrors forte/proc/fcidump.py:262-364 (fcidump_from_file) for the header keys and the 8-fold
permutational symmetry of the ERIs, and forte/modules/objects_factory_fcidump.py:68-96 for the
chemist -> physicist conversion. Lines are classified by their index pattern rather than by
their position in the file: four zero indices give the scalar (nuclear + frozen core) energy,
i>0 and j=k=l=0 an orbital energy, i,j>0 and k=l=0 a core Hamiltonian element, anything else an
ERI (ij|kl).
"""
import gzip

import numpy as np


def read_fcidump(fname):
    """Return dict(norb, nelec, ms2, isym, orbsym, pntgrp, uhf, enuc, hcore, eri[, epsilon]).
    orbsym and isym are as written in the file (the FCIDUMP convention, 1-based); a gzip-compressed
    file (name ending in .gz) is read the same way."""
    out = {}
    opener = gzip.open if str(fname).endswith(".gz") else open
    with opener(fname, "rt") as f:
        lines = f.read().splitlines()
    if lines[0].strip() != "&FCI":
        raise ValueError(f"{fname}: not an FCIDUMP file (first line is not &FCI)")
    i = 1
    header = []
    while True:
        line = lines[i]
        i += 1
        if "END" in line or line.strip() == "/":
            break
        header.append(line)
    text = " ".join(header)
    # split "KEY=value," pairs; ORBSYM holds a comma-separated list
    for part in text.replace(" ", "").split("="):
        pass
    keys = {}
    cur = None
    for tok in text.replace(" ", "").split(","):
        if "=" in tok:
            k, v = tok.split("=", 1)
            cur = k.upper()
            keys[cur] = [v] if v != "" else []
        elif tok != "" and cur is not None:
            keys[cur].append(tok)
    norb = int(keys["NORB"][0])
    out["norb"] = norb
    out["nelec"] = int(keys["NELEC"][0])
    out["ms2"] = int(keys.get("MS2", ["0"])[0])
    out["uhf"] = "TRUE" in "".join(keys.get("UHF", ["FALSE"])).upper()
    out["orbsym"] = [int(x) for x in keys.get("ORBSYM", ["1"] * norb)]
    out["isym"] = int(keys.get("ISYM", ["1"])[0])
    out["pntgrp"] = keys["PNTGRP"][0] if keys.get("PNTGRP") else None
    hcore = np.zeros((norb, norb))
    eri = np.zeros((norb, norb, norb, norb))
    eps = None
    enuc = 0.0
    for line in lines[i:]:
        t = line.split()
        if len(t) < 5:
            continue
        v = float(t[0].replace("D", "E").replace("d", "e"))
        p, q, r, s = (int(x) - 1 for x in t[1:5])
        if p < 0 and q < 0 and r < 0 and s < 0:
            enuc = v
        elif q < 0 and r < 0 and s < 0:
            if eps is None:
                eps = np.zeros(norb)
            eps[p] = v
        elif r < 0 and s < 0:
            hcore[p, q] = hcore[q, p] = v
        else:
            for a, b, c, d in ((p, q, r, s), (p, q, s, r), (q, p, r, s), (q, p, s, r),
                               (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)):
                eri[a, b, c, d] = v
    out["enuc"] = enuc
    out["hcore"] = hcore
    out["eri"] = eri
    if eps is not None:
        out["epsilon"] = eps
    return out


def chem_to_forte(eri):
    """tei_ab[p,q,r,s] = (pr|qs) = <pq|rs>; tei_aa = tei_bb = (pr|qs) - (ps|qr)."""
    ab = np.ascontiguousarray(eri.transpose(0, 2, 1, 3))
    aa = ab - np.ascontiguousarray(eri.transpose(0, 2, 3, 1))
    return aa, ab
