"""GPU parity at the benchmark configurations' own shapes.

Full oracle solves are affordable only up to about 63,504 determinants, so the checks at the large
shapes use what the oracle can compute exactly there:
  - rows of H c by the Slater-Condon rules over the determinants connected to each row
    (Oracle.sigma_rows), for ~2,000 sampled rows plus the largest |c| rows: they pin the GPU sigma
    element-wise and bound each root's residual H c - E c;
  - <S^2> and the 1-/2-RDMs of the GPU eigenvector from the oracle's Forte-structured loops, and
    the energy as the expectation value of H through those RDMs;
  - mixed precision against the FP64 GPU solve (energy and 2-RDM), as the configuration asks.
"""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sample_rows(c, count, seed):
    """`count` random determinant indices plus the 64 largest |c| (where errors would matter most)."""
    rng = np.random.default_rng(seed)
    rand = rng.choice(c.size, size=min(count, c.size), replace=False)
    top = np.argsort(-np.abs(c))[:64]
    return np.unique(np.concatenate([rand, top]))


def rdm_energy(h, aa, ab, rd, e_core=0.0):
    """<H> from the 1-/2-RDMs in Forte's layout (oei, antisymmetrized tei_aa, tei_ab)."""
    g1a, g1b, g2aa, g2ab, g2bb = rd
    return (e_core + np.einsum("pq,pq", h, g1a + g1b) + 0.25 * np.einsum("pqrs,pqrs", aa, g2aa + g2bb)
            + np.einsum("pqrs,pqrs", ab, g2ab))


def gpu_rdms(s, root=0):
    r = s.rdms([(root, root)], 2)[0]
    return r.g1a(), r.g1b(), r.g2aa(), r.g2ab(), r.g2bb()


# ---- configuration 2: CAS(14,14), 3 roots, FP64, subspace 8 per root, root-0 RDMs ------------

def test_cas14_three_roots_fp64(oracle):
    import forte_amd
    n, k, seed = 14, 7, 2
    h, eri = oracle.seeded_integrals(n, seed)
    aa, ab = oracle.chem_to_forte(eri)
    say(f"CAS(14,14) seed {seed}: 3 roots FP64, DL_SUBSPACE_PER_ROOT 8, e_conv 1e-9, r_conv 1e-6")
    s = forte_amd.FCISolver(k, k, 1, 3, forte_amd.ActiveSpaceIntegrals(n, h, h, aa, ab, aa, 0.0))
    s.set_e_convergence(1e-9)
    s.set_r_convergence(1e-6)
    s.set_subspace_per_root(8)
    s.set_maxiter(500)
    t0 = time.perf_counter()
    s.compute_energy()
    st = s.stats()
    E = np.array(s.energies())
    say(f"  solve {time.perf_counter() - t0:.2f}s energies {E} stats {st}")
    assert st["converged"]
    assert np.all(np.diff(E) >= 0)
    o = oracle.Oracle.from_chem(h, eri, k, k)
    evecs = [s.evec(r) for r in range(3)]
    G = np.array([[a @ b for b in evecs] for a in evecs])
    assert np.abs(G - np.eye(3)).max() < 1e-10
    for r, c in enumerate(evecs):
        rows = sample_rows(c, 2000, 100 + r)
        ref = o.sigma_rows(c, rows)
        sig = s.sigma(c)
        serr = np.abs(sig[rows] - ref).max() / np.abs(ref).max()
        res = np.abs(ref - E[r] * c[rows]).max()
        s2 = o.spin2(c)
        say(f"  root {r}: sigma rows rel err {serr:.2e}, max sampled |Hc - Ec| {res:.2e}, "
            f"<S^2> gpu {s.spin2()[r]:.3e} oracle {s2:.3e}")
        assert serr < 1e-12
        assert res < 2e-6  # r_conv 1e-6 bounds the residual's norm, so every element
        assert abs(s.spin2()[r] - s2) < 1e-8
        assert abs(s2) < 1e-6  # singlets
    t0 = time.perf_counter()
    got = gpu_rdms(s, 0)
    t_gpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    ref = o.rdms(evecs[0])
    say(f"  root-0 RDMs: gpu {t_gpu:.2f}s, oracle {time.perf_counter() - t0:.2f}s")
    for name, a, b in zip(("g1a", "g1b", "g2aa", "g2ab", "g2bb"), got, ref):
        err = np.abs(a - b).max()
        say(f"    {name} max |gpu - oracle| {err:.2e}")
        assert err < 1e-8, name
    e_rdm = rdm_energy(h, aa, ab, ref)
    say(f"  E0 {E[0]:.12f}, <H> from the oracle RDMs {e_rdm:.12f}")
    assert abs(e_rdm - E[0]) < 1e-9


# ---- configurations 3 and 5: CAS(16,16) mixed precision (and learned) vs the FP64 solve ----

CAS16 = dict(e_conv=1e-10, r_conv=1e-6, maxiter=500)


def _solve16(ints, precision, preconditioner="diagonal"):
    import forte_amd
    import forte_amd.learned
    s = forte_amd.FCISolver(8, 8, 1, 1, ints)
    s.set_e_convergence(CAS16["e_conv"])
    s.set_r_convergence(CAS16["r_conv"])
    s.set_maxiter(CAS16["maxiter"])
    s.set_precision(precision)
    if preconditioner == "learned":
        s.set_lp_weights(forte_amd.learned.load_weights())
        s.set_preconditioner("learned")
    t0 = time.perf_counter()
    e = s.compute_energy()
    say(f"  {precision}+{preconditioner}: E {e:.12f} in {time.perf_counter() - t0:.2f}s {s.stats()}")
    return s, e


@pytest.fixture(scope="module")
def cas16_fp64(oracle):
    import forte_amd
    n, seed = 16, 3
    h, eri = oracle.seeded_integrals(n, seed)
    aa, ab = oracle.chem_to_forte(eri)
    ints = forte_amd.ActiveSpaceIntegrals(n, h, h, aa, ab, aa, 0.0)
    say(f"CAS(16,16) seed {seed}, FP64 solve, e_conv {CAS16['e_conv']}, r_conv {CAS16['r_conv']}")
    s, e = _solve16(ints, "fp64")
    assert s.stats()["converged"]
    c = s.evec(0)
    t0 = time.perf_counter()
    rd = gpu_rdms(s)
    say(f"  FP64 RDMs {time.perf_counter() - t0:.2f}s")
    o = oracle.Oracle.from_chem(h, eri, 8, 8, lists=False)
    rows = sample_rows(c, 2000, 316)
    ref = o.sigma_rows(c, rows)
    out = dict(h=h, aa=aa, ab=ab, ints=ints, E=e, c=c, rdms=rd, rows=rows, ref=ref, sig64=s.sigma(c))
    del s
    return out


def test_cas16_fp64_sigma_rows_and_residual(cas16_fp64):
    d = cas16_fp64
    rows, ref, c = d["rows"], d["ref"], d["c"]
    serr = np.abs(d["sig64"][rows] - ref).max() / np.abs(ref).max()
    res = np.abs(ref - d["E"] * c[rows]).max()
    e_rdm = rdm_energy(d["h"], d["aa"], d["ab"], d["rdms"])
    say(f"CAS(16,16) FP64: sigma rows rel err {serr:.2e}, max sampled |Hc - Ec| {res:.2e}, "
        f"|E - <H>_RDM| {abs(e_rdm - d['E']):.2e}")
    assert serr < 1e-12
    assert res < 2e-6
    assert abs(e_rdm - d["E"]) < 1e-8
    g1a, g1b, g2aa, g2ab, g2bb = d["rdms"]
    assert abs(np.trace(g1a) - 8) < 1e-10 and abs(np.trace(g1b) - 8) < 1e-10
    assert abs(np.einsum("pqpq", g2ab) - 64) < 1e-9
    assert abs(np.einsum("pqpq", g2aa) - 56) < 1e-9


def test_cas16_mixed_sigma_rows(cas16_fp64):
    """FP32/BF16x3 sigma of the FP64 eigenvector on the sampled rows, against the oracle."""
    import forte_amd
    d = cas16_fp64
    s = forte_amd.FCISolver(8, 8, 1, 1, d["ints"])
    c32 = forte_amd.DeviceArray(s.ndet(), "float32")
    s32 = forte_amd.DeviceArray(s.ndet(), "float32")
    c32.upload(d["c"].astype(np.float32))
    s.sigma_arrays(c32, s32)
    got = s32.download().astype(np.float64)
    rows, ref = d["rows"], d["ref"]
    err = np.abs(got[rows] - ref).max() / np.abs(ref).max()
    nerr = np.linalg.norm(got - d["sig64"]) / np.linalg.norm(d["sig64"])
    say(f"CAS(16,16) mixed sigma: sampled rows rel err {err:.2e}, whole vector vs FP64 GPU {nerr:.2e}")
    assert err < 2e-6
    assert nerr < 1e-6


def test_cas16_mixed_solve_vs_fp64(cas16_fp64):
    """Configuration 3: mixed precision with FP64 refinement lands on the FP64 energy; the
    2-RDM error against the FP64 solve is reported."""
    d = cas16_fp64
    say("CAS(16,16) mixed + diagonal")
    s, e = _solve16(d["ints"], "mixed")
    st = s.stats()
    assert st["converged"] and st["sigma_builds_fp64"] < st["sigma_builds"]
    rd = gpu_rdms(s)
    errs = [np.abs(a - b).max() for a, b in zip(rd, d["rdms"])]
    say(f"  |dE| vs FP64 {abs(e - d['E']):.2e}; max RDM error vs FP64 g1a {errs[0]:.2e} g1b {errs[1]:.2e} "
        f"g2aa {errs[2]:.2e} g2ab {errs[3]:.2e} g2bb {errs[4]:.2e}")
    assert abs(e - d["E"]) < 1e-8
    assert max(errs) < 1e-5
    c = s.evec(0)
    res = np.abs(d["ref"] - d["E"] * c[d["rows"]] * np.sign(c @ d["c"])).max()
    assert res < 1e-5


def test_cas16_mixed_learned_solve(cas16_fp64):
    """Configuration 5: the learned preconditioner changes only the path; the answer is the FP64 one."""
    d = cas16_fp64
    say("CAS(16,16) mixed + learned")
    s, e = _solve16(d["ints"], "mixed", "learned")
    assert s.stats()["converged"]
    say(f"  |dE| vs FP64 {abs(e - d['E']):.2e}")
    assert abs(e - d["E"]) < 1e-8
