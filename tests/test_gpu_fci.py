"""GPU parity: the gfx950 solver against the CPU oracle and the reference's known answers."""
import itertools
import sys

import numpy as np
import pytest

from conftest import fcidump_path, seeded_problem

pytestmark = pytest.mark.gpu

KAT = {2: -1.163012998021681, 3: -1.926739016209154, 4: -1.252893688642302, 5: -1.134690658798176}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_solver(p, nroot=1, mult=1):
    import forte_amd
    ints = forte_amd.ActiveSpaceIntegrals(p["n"], p["h"], p["h"], p["aa"], p["ab"], p["aa"], p["e_core"])
    return forte_amd.FCISolver(p["na"], p["nb"], mult, nroot, ints)


def cpu_oracle(O, p):
    return O.Oracle(p["n"], p["na"], p["nb"], p["h"], p["h"], p["aa"], p["ab"], p["aa"], p["e_core"])


TABLE_CASES = [(4, 2, 2), (5, 3, 2), (6, 3, 3), (6, 1, 5), (7, 3, 4), (8, 4, 4), (8, 2, 6), (3, 3, 0)]


@pytest.mark.parametrize("n,na,nb", TABLE_CASES)
def test_tables_bit_exact(oracle, n, na, nb):
    say("tables", n, na, nb)
    p = seeded_problem(n, na, nb, 3)
    g, o = gpu_solver(p), cpu_oracle(oracle, p)
    for spin in (0, 1):
        assert np.array_equal(g.strings(spin), o.strings(spin))
        for a, b in itertools.product(range(n), repeat=2):
            for x, y in zip(g.vo_list(spin, a, b), o.vo(spin, a, b)):
                assert np.array_equal(x, y), ("vo", spin, a, b)
        for pq in range(n * (n - 1) // 2):
            for x, y in zip(g.oo_list(spin, pq), o.oo(spin, pq)):
                assert np.array_equal(x, y), ("oo", spin, pq)
        for a, b, c, d in itertools.product(range(n), repeat=4):
            if a > b and c > d and not (a == c and b == d):
                for x, y in zip(g.vvoo_list(spin, a, b, c, d), o.vvoo(spin, a, b, c, d)):
                    assert np.array_equal(x, y), ("vvoo", spin, a, b, c, d)


@pytest.mark.parametrize("n,na,nb", [(6, 3, 3), (8, 4, 4), (10, 5, 4), (12, 6, 6)])
def test_hdiag_bit_exact(oracle, n, na, nb):
    say("hdiag", n, na, nb)
    p = seeded_problem(n, na, nb, 4, e_core=0.7)
    assert np.array_equal(gpu_solver(p).hdiag(), cpu_oracle(oracle, p).hdiag())


SIGMA_CASES = [(2, 1, 1), (3, 2, 1), (4, 2, 2), (4, 3, 2), (5, 1, 1), (6, 3, 3), (6, 2, 4), (6, 5, 1),
               (4, 4, 0), (4, 0, 4), (6, 6, 6), (7, 1, 6), (8, 4, 4), (8, 3, 5), (9, 2, 2), (10, 5, 5),
               (10, 2, 8), (12, 6, 6), (12, 4, 7),
               # 14 to 18 orbitals: 7, 9, 10 and 11 row tiles of pairs, the alpha-beta layouts of CAS(14,14)
               # to CAS(18,18) (few electrons keep the oracle fast)
               (14, 3, 3), (16, 2, 2), (16, 3, 3), (17, 2, 2), (18, 3, 3),
               # the alpha-beta instances the CAS(16,16) and CAS(14,14) solves run (55 and 49 alpha
               # single excitations: 16 k-steps, 9 and 7 row tiles, the shared last row tile at 9)
               (16, 5, 1), (14, 7, 1), (16, 1, 5)]


@pytest.mark.parametrize("n,na,nb", SIGMA_CASES)
def test_sigma_fp64(oracle, n, na, nb):
    say("sigma fp64", n, na, nb)
    p = seeded_problem(n, na, nb, 7 + n, e_core=0.31)
    g, o = gpu_solver(p), cpu_oracle(oracle, p)
    C = np.random.default_rng(n * 100 + na).standard_normal(o.ndet)
    ref = o.sigma(C)
    got = g.sigma(C)
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(got - ref).max() < 1e-12 * scale


@pytest.mark.parametrize("n,na,nb", [(6, 3, 3), (8, 4, 4), (10, 5, 5), (12, 6, 6), (9, 3, 5), (14, 3, 3), (16, 2, 2),
                                     (16, 3, 3), (17, 2, 2), (18, 3, 3),
                                     # same-spin cliques of 36, 45 and 55 members (the alpha side,
                                     # even row length): the static-size pass-1 kernels
                                     (9, 2, 2), (10, 2, 3), (12, 4, 5), (12, 3, 2),
                                     # 6 row tiles (n = 13) and the production k-step counts
                                     (13, 3, 3), (16, 5, 1), (14, 7, 1)])
def test_sigma_mixed(oracle, n, na, nb):
    import forte_amd
    say("sigma mixed", n, na, nb)
    p = seeded_problem(n, na, nb, 17 + n)
    g, o = gpu_solver(p), cpu_oracle(oracle, p)
    C = np.random.default_rng(n).standard_normal(o.ndet)
    C /= np.linalg.norm(C)
    ref = o.sigma(C)
    c32 = forte_amd.DeviceArray(o.ndet, "float32")
    s32 = forte_amd.DeviceArray(o.ndet, "float32")
    c32.upload(C.astype(np.float32))
    g.sigma_arrays(c32, s32)
    got = s32.download().astype(np.float64)
    # BF16x3 products with FP32 accumulation and the operator diagonal kept in FP32: measured
    # 1.4e-7 to 3.0e-7 of the vector norm up to 14 orbitals, 1.0e-6 for the few-electron
    # 16-18-orbital shapes. (Production shape: 1.5e-7 at CAS(16,16), bounded at 1e-6 in
    # test_gpu_configs.py, which rejects the 1.1e-6 measured when the operator diagonal went
    # through the BF16 products.)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    say("  relative error", err)
    assert err < (3e-6 if n >= 16 else 6e-7)


def test_sigma_device_fp64_matches_host(oracle):
    import forte_amd
    p = seeded_problem(8, 4, 4, 5)
    g = gpu_solver(p)
    C = np.random.default_rng(0).standard_normal(g.ndet())
    c = forte_amd.DeviceArray(g.ndet(), "float64")
    s = forte_amd.DeviceArray(g.ndet(), "float64")
    c.upload(C)
    g.sigma_arrays(c, s)
    ref = g.sigma(C)
    assert np.abs(s.download() - ref).max() < 1e-12 * np.abs(ref).max()
    t = g.sigma_timings()
    assert len(t) == 3 and all(x >= 0 for x in t)
    ms, t3 = g.sigma_bench(c, s, 3)
    assert ms > 0 and len(t3) == 3
    assert np.abs(s.download() - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("k", sorted(KAT))
def test_kat_energy_gpu(k):
    import forte_amd
    say("kat", k)
    ints, na, nb, mult = forte_amd.integrals_from_fcidump(fcidump_path(k))
    s = forte_amd.FCISolver(na, nb, mult, 1, ints)
    s.set_e_convergence(1e-12)
    e = s.compute_energy()
    assert abs(e - KAT[k]) < 1e-9
    assert abs(s.spin2()[0] - (mult * mult - 1) / 4.0) < 1e-8


@pytest.mark.parametrize("n,na,nb,nroot,mult", [(6, 3, 3, 3, 1), (8, 4, 4, 1, 1), (7, 4, 3, 2, 2),
                                                (8, 4, 4, 2, 3), (10, 5, 5, 1, 1)])
def test_energies_vs_oracle(oracle, n, na, nb, nroot, mult):
    say("energies", n, na, nb, nroot, mult)
    p = seeded_problem(n, na, nb, 1 if n == 8 else 50 + n)
    g, o = gpu_solver(p, nroot, mult), cpu_oracle(oracle, p)
    g.set_e_convergence(1e-10)
    eg = g.compute_energy()
    r = o.solve(nroot=nroot, multiplicity=mult, e_conv=1e-10)
    assert np.abs(np.array(g.energies()) - r["energies"]).max() < 1e-9
    assert np.abs(np.array(g.spin2()) - r["spin2"]).max() < 1e-7
    assert eg == g.energies()[0]


def test_rdms_vs_oracle(oracle):
    import forte_amd
    say("rdms")
    p = seeded_problem(8, 4, 4, 1)
    g, o = gpu_solver(p, 2), cpu_oracle(oracle, p)
    g.set_e_convergence(1e-12)
    g.set_r_convergence(1e-8)
    g.compute_energy()
    for (l, r) in [(0, 0), (1, 1), (0, 1)]:
        rd = g.rdms([(l, r)], 2)[0]
        cl, cr = g.evec(l), g.evec(r)
        ref = o.rdms(cl, cr)
        for a, b in zip([rd.g1a(), rd.g1b(), rd.g2aa(), rd.g2ab(), rd.g2bb()], ref):
            assert np.abs(a - b).max() < 1e-12, (l, r)
    # spin-free combination (fci_vector_rdm.cc:121-126) and traces
    rd = g.rdms([(0, 0)], 2, forte_amd.RDMsType.spin_free)[0]
    G1, G2 = rd.SF_G1(), rd.SF_G2()
    assert abs(np.trace(G1) - 8) < 1e-10
    assert abs(np.einsum("pqpq", G2) - 8 * 7) < 1e-9
    # energy from RDMs equals the eigenvalue
    r0 = g.rdms([(0, 0)], 2)[0]
    e = np.einsum("pq,pq", p["h"], r0.g1a() + r0.g1b()) + 0.25 * np.einsum("pqrs,pqrs", p["aa"], r0.g2aa() + r0.g2bb()) \
        + np.einsum("pqrs,pqrs", p["ab"], r0.g2ab())
    assert abs(e - g.energies()[0]) < 1e-9


def test_rdms_against_oracle_eigenvector(oracle):
    """Root 0 RDMs of the GPU solve agree with the oracle's own solve (sign-invariant)."""
    p = seeded_problem(8, 4, 4, 1)
    g, o = gpu_solver(p), cpu_oracle(oracle, p)
    g.set_e_convergence(1e-12)
    g.set_r_convergence(1e-8)
    g.compute_energy()
    r = o.solve(nroot=1, e_conv=1e-12, r_conv=1e-8)
    rd = g.rdms([(0, 0)], 2)[0]
    ref = o.rdms(r["evecs"][0])
    for a, b in zip([rd.g1a(), rd.g1b(), rd.g2aa(), rd.g2ab(), rd.g2bb()], ref):
        assert np.abs(a - b).max() < 1e-8


@pytest.mark.parametrize("n,na,nb,mult", [(6, 3, 3, 1), (6, 4, 2, 3), (7, 3, 4, 2), (5, 3, 2, 2), (6, 2, 2, 1),
                                           (7, 4, 4, 1)])
def test_rdms3_vs_oracle(oracle, n, na, nb, mult):
    """3-RDMs (aaa, aab, abb, bbb) against brute-force operator strings; for root_left != root_right
    the reference's convention <right| ... |left>."""
    import forte_amd
    say("rdms3", n, na, nb)
    p = seeded_problem(n, na, nb, 33 + n)
    g, o = gpu_solver(p, 2, mult), cpu_oracle(oracle, p)
    g.set_e_convergence(1e-12)
    g.set_r_convergence(1e-8)
    g.compute_energy()
    for (l, r) in [(0, 0), (1, 1), (0, 1)]:
        rd = g.rdms([(l, r)], 3)[0]
        ref = o.brute_rdm3(g.evec(r), g.evec(l))
        for name, a, b in zip(("aaa", "aab", "abb", "bbb"), [rd.g3aaa(), rd.g3aab(), rd.g3abb(), rd.g3bbb()], ref):
            assert np.abs(a - b).max() < 1e-12, (name, l, r)
        lower = o.rdms(g.evec(l), g.evec(r))
        assert np.abs(rd.g2ab() - lower[3]).max() < 1e-12 and np.abs(rd.g1a() - lower[0]).max() < 1e-12
    # spin-free G3 (fci_vector_rdm.cc:127-131) and its partial trace: sum_r G3[p,q,r,s,t,r] = (N - 2) G2[p,q,s,t]
    sf = g.rdms([(0, 0)], 3, forte_amd.RDMsType.spin_free)[0]
    G2, G3 = sf.SF_G2(), sf.SF_G3()
    assert np.abs(np.einsum("pqrstr->pqst", G3) - (na + nb - 2) * G2).max() < 1e-10


def test_rdms3_partial_traces_cas8():
    """CAS(8,8) ground state: every 3-RDM block contracts to the 2-RDMs."""
    say("rdms3 traces CAS(8,8)")
    p = seeded_problem(8, 4, 4, 1)
    g = gpu_solver(p)
    g.set_e_convergence(1e-12)
    g.compute_energy()
    rd = g.rdms([(0, 0)], 3)[0]
    g2aa, g2ab, g2bb = rd.g2aa(), rd.g2ab(), rd.g2bb()
    aaa, aab, abb, bbb = rd.g3aaa(), rd.g3aab(), rd.g3abb(), rd.g3bbb()
    assert np.abs(np.einsum("pqrstr->pqst", aaa) - 2 * g2aa).max() < 1e-11
    assert np.abs(np.einsum("pqrstr->pqst", bbb) - 2 * g2bb).max() < 1e-11
    assert np.abs(np.einsum("pqrstr->pqst", aab) - 4 * g2aa).max() < 1e-11
    assert np.abs(np.einsum("pqrsqu->prsu", aab) - 3 * g2ab).max() < 1e-11
    assert np.abs(np.einsum("pqrsqu->prsu", abb) - 3 * g2ab).max() < 1e-11


def test_transition_rdms(oracle):
    """ActiveSpaceMethod::transition_rdms between a singlet and a triplet solver (Ms = 0, same
    integrals) and between two roots of one solver, against the oracle's <L|a+p a_q|R>."""
    import forte_amd
    say("transition rdms")
    p = seeded_problem(6, 3, 3, 29)
    s1, t1 = gpu_solver(p, 2, 1), gpu_solver(p, 1, 3)
    for g in (s1, t1):
        g.set_e_convergence(1e-12)
        g.set_r_convergence(1e-8)
        g.compute_energy()
    o = cpu_oracle(oracle, p)
    for (l, r, a, b) in [(0, 0, s1, t1), (1, 0, s1, t1), (0, 1, s1, s1), (0, 0, t1, s1)]:
        rd = a.transition_rdms([(l, r)], b, 1)[0]
        ref = o.rdms(a.evec(l), b.evec(r), level=1)
        assert np.abs(rd.g1a() - ref[0]).max() < 1e-12, (l, r)
        assert np.abs(rd.g1b() - ref[1]).max() < 1e-12, (l, r)
    # singlet -> triplet (Ms = 0): the alpha and beta blocks are opposite, the spin-free sum vanishes
    rd = s1.transition_rdms([(0, 0)], t1, 1, forte_amd.RDMsType.spin_free)[0]
    assert np.abs(rd.SF_G1()).max() < 1e-10
    # the same roots through rdms() agree with transition_rdms() of a solver with itself
    assert np.abs(s1.rdms([(0, 1)], 1)[0].g1a() - s1.transition_rdms([(0, 1)], s1, 1)[0].g1a()).max() < 1e-14
    with pytest.raises(RuntimeError, match="order 2 or higher"):
        s1.transition_rdms([(0, 0)], t1, 2)
    with pytest.raises(RuntimeError, match="nroot_right"):
        s1.transition_rdms([(0, 1)], t1, 1)
    q = seeded_problem(6, 4, 2, 29)
    d = gpu_solver(q, 1, 3)
    d.compute_energy()
    with pytest.raises(RuntimeError, match="alfa and beta electrons"):
        s1.transition_rdms([(0, 0)], d, 1)


def test_warm_restart_and_new_integrals(oracle):
    import forte_amd
    p = seeded_problem(6, 3, 3, 61)
    g = gpu_solver(p)
    e1 = g.compute_energy()
    e2 = g.compute_energy()
    assert abs(e1 - e2) < 1e-9
    q = seeded_problem(6, 3, 3, 62)
    g.set_active_space_integrals(forte_amd.ActiveSpaceIntegrals(6, q["h"], q["h"], q["aa"], q["ab"], q["aa"]))
    e3 = g.compute_energy()
    r = cpu_oracle(oracle, q).solve(nroot=1)
    assert abs(e3 - r["energies"][0]) < 1e-9


def test_mixed_precision_solve(oracle):
    say("mixed solve")
    p = seeded_problem(10, 5, 5, 9)
    g64, gmx = gpu_solver(p), gpu_solver(p)
    g64.set_e_convergence(1e-10)
    gmx.set_e_convergence(1e-10)
    gmx.set_precision("mixed")
    e64 = g64.compute_energy()
    emx = gmx.compute_energy()
    st = gmx.stats()
    say("  fp64", e64, g64.stats(), "mixed", emx, st)
    assert abs(e64 - emx) < 1e-9
    assert st["sigma_builds_fp64"] < st["sigma_builds"]


def test_mixed_precision_solve_cas14():
    """CAS(14,14), 11.8 million determinants: the production kernels at scale (alpha-beta with the
    A operand in LDS, same-spin pass 1 with a compile-time clique size of 36 and staged stores).
    The mixed solve and its FP64 refinement must land on the FP64 energy."""
    import forte_amd
    say("mixed solve CAS(14,14)")
    h, eri = forte_amd.seeded_integrals(14, 5)
    ints = forte_amd.make_active_space_integrals(h, eri)
    out = {}
    for prec in ("fp64", "mixed"):
        s = forte_amd.FCISolver(7, 7, 1, 1, ints)
        s.set_e_convergence(1e-10)
        s.set_r_convergence(1e-5)
        s.set_maxiter(400)
        s.set_precision(prec)
        out[prec] = (s.compute_energy(), s.stats())
        say(" ", prec, out[prec])
    assert out["mixed"][1]["converged"] and out["fp64"][1]["converged"]
    assert abs(out["mixed"][0] - out["fp64"][0]) < 1e-8
    assert out["mixed"][1]["sigma_builds_fp64"] < out["mixed"][1]["sigma_builds"]


def test_errors_are_runtime_errors():
    import forte_amd
    p = seeded_problem(4, 2, 2, 1)
    g = gpu_solver(p)
    with pytest.raises(RuntimeError):
        g.rdms([(0, 0)], 2)  # before compute_energy
    with pytest.raises(RuntimeError):
        g.set_precision("fp16")
    g.compute_energy()
    with pytest.raises(RuntimeError):
        g.rdms([(0, 5)], 1)
    with pytest.raises(RuntimeError):
        g.rdms([(0, 0)], 4)
    with pytest.raises(RuntimeError):
        forte_amd.FCISolver(5, 2, 1, 1, forte_amd.ActiveSpaceIntegrals(4, p["h"], p["h"], p["aa"], p["ab"], p["aa"]))


def test_threaded_solves_match_oracle(oracle):
    """Concurrent solvers on per-thread streams give the same energies as the oracle."""
    from forte_amd.batch import seeded_batch, solve_batch_threads
    say("threaded batch")
    probs = [p for p in seeded_batch(24) if p[0] <= 8][:8]
    res, _ = solve_batch_threads(probs, nroot=2, threads=4, options={"E_CONVERGENCE": 1e-11, "R_CONVERGENCE": 1e-7,
                                                                      "DL_MAXITER": 300})
    for (n, seed), r in zip(probs, res):
        assert r["converged"]
        h, eri = oracle.seeded_integrals(n, seed)
        o = oracle.Oracle.from_chem(h, eri, n // 2, n // 2)
        ref = o.solve(nroot=2, e_conv=1e-11, r_conv=1e-7, maxiter=300)["energies"]
        assert np.abs(np.array(r["energies"]) - ref).max() < 1e-9, (n, seed)


def test_fp64_sigma_run_to_run():
    """The FP64 alpha-beta row sums and the alpha-beta 2-RDM add in scheduling order (LDS and
    global FP64 atomics), so repeated evaluations may differ in the last bits; bound that
    difference (Forte's sums run in a fixed order)."""
    import forte_amd
    p = seeded_problem(12, 6, 6, 11)
    g = gpu_solver(p)
    C = np.random.default_rng(5).standard_normal(g.ndet())
    C /= np.linalg.norm(C)
    c = forte_amd.DeviceArray(g.ndet(), "float64")
    s = forte_amd.DeviceArray(g.ndet(), "float64")
    c.upload(C)
    outs = []
    for _ in range(3):
        g.sigma_arrays(c, s)
        outs.append(s.download())
    norm = np.linalg.norm(outs[0])
    d = max(np.abs(o - outs[0]).max() for o in outs[1:])
    say(f"  FP64 sigma run-to-run max difference {d:.2e} (|sigma| {norm:.3e})")
    assert d <= 1e-14 * norm
    g.set_e_convergence(1e-10)
    g.compute_energy()
    r1 = g.rdms([(0, 0)], 2)[0]
    r2 = g.rdms([(0, 0)], 2)[0]
    dr = max(np.abs(a - b).max() for a, b in zip([r1.g1a(), r1.g2aa(), r1.g2ab(), r1.g2bb()],
                                                 [r2.g1a(), r2.g2aa(), r2.g2ab(), r2.g2bb()]))
    say(f"  RDM run-to-run max difference {dr:.2e}")
    assert dr <= 1e-13


@pytest.mark.parametrize("n,na,nb,nroot,mult", [(8, 4, 4, 2, 1), (7, 4, 3, 2, 2), (10, 5, 5, 1, 1)])
def test_olsen_preconditioner_same_answer(oracle, n, na, nb, nroot, mult):
    """Olsen's correction changes the Davidson-Liu path, not the eigenpairs."""
    p = seeded_problem(n, na, nb, 31)
    res = {}
    for pc in ("diagonal", "olsen"):
        g = gpu_solver(p, nroot, mult)
        g.set_preconditioner(pc)
        g.set_e_convergence(1e-12)
        g.set_r_convergence(1e-8)
        g.set_maxiter(500)
        g.compute_energy()
        assert g.stats()["converged"]
        res[pc] = (np.array(g.energies()), g.stats()["iterations"])
    say(f"  diagonal {res['diagonal'][1]} iterations, olsen {res['olsen'][1]}")
    assert np.abs(res["diagonal"][0] - res["olsen"][0]).max() < 1e-10
    if n <= 8:
        ref = cpu_oracle(oracle, p).solve(nroot=nroot, multiplicity=mult, e_conv=1e-12, r_conv=1e-8)
        assert np.abs(res["olsen"][0] - np.array(ref["energies"])).max() < 1e-9


@pytest.mark.parametrize("n,na,nb", [(16, 2, 2), (15, 2, 3), (14, 1, 2)])
def test_rdms_large_cliques(oracle, n, na, nb):
    """Same-spin 2-RDM cliques of more than 96 members (C(n - k + 2, 2) = 120, 105, 91/78 here):
    the member products are split into tiles."""
    say("rdms large cliques", n, na, nb)
    p = seeded_problem(n, na, nb, 41)
    g, o = gpu_solver(p, 1, abs(na - nb) + 1), cpu_oracle(oracle, p)
    g.set_e_convergence(1e-12)
    g.set_r_convergence(1e-8)
    g.set_maxiter(300)
    g.compute_energy()
    rd = g.rdms([(0, 0)], 2)[0]
    ref = o.rdms(g.evec(0))
    for name, a, b in zip(("g1a", "g1b", "g2aa", "g2ab", "g2bb"), [rd.g1a(), rd.g1b(), rd.g2aa(), rd.g2ab(), rd.g2bb()],
                          ref):
        assert np.abs(a - b).max() < 1e-12, name


@pytest.mark.parametrize("n,na,nb", [(19, 2, 2), (19, 3, 2)])
def test_sigma_nineteen_orbitals(oracle, n, na, nb):
    """The largest orbital count of the alpha-beta kernels (190 packed pairs, 12 row tiles)."""
    import forte_amd
    say("sigma n=19", n, na, nb)
    p = seeded_problem(n, na, nb, 5 + na)
    g, o = gpu_solver(p), cpu_oracle(oracle, p)
    C = np.random.default_rng(n + na).standard_normal(o.ndet)
    C /= np.linalg.norm(C)
    ref = o.sigma(C)
    assert np.abs(g.sigma(C) - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
    c32 = forte_amd.DeviceArray(o.ndet, "float32")
    s32 = forte_amd.DeviceArray(o.ndet, "float32")
    c32.upload(C.astype(np.float32))
    g.sigma_arrays(c32, s32)
    err = np.linalg.norm(s32.download().astype(np.float64) - ref) / np.linalg.norm(ref)
    say("  mixed relative error", err)
    assert err < 3e-6


def test_twenty_orbitals_sigma_is_refused():
    """Beyond 19 orbitals the alpha-beta kernels raise before any launch."""
    p = seeded_problem(20, 2, 2, 1)
    g = gpu_solver(p)
    with pytest.raises(RuntimeError, match="too many orbitals"):
        g.sigma(np.ones(g.ndet()))
