#!/usr/bin/env python3
"""CAS(16,16) FCI on one MI355X: sigma throughput and time to solution.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--solve 0|1] [--cpu 0|1] [--n 16]
                    [--legs all|name,name] [--rconvs 1e-4,1e-6]

A step is one sigma build H C over all 165,636,900 determinants of CAS(16,16) (Ms = 0, C1,
seeded integrals, seed 3) with the mixed-precision kernels (FP32 C and sigma, BF16x3 MFMA
contractions accumulated in FP32). `value` is the sigma throughput in GB/s, defined as
9 * N_det * 4 bytes / t_sigma (C read by the two same-spin and the alpha-beta passes, sigma
written and re-read), inputs resident in HBM. The line also carries:
  roofline      MFMA roofline of the dominant kernel (k_ab3_mixed, the alpha-beta contraction),
                timed with HIP events on the stream it runs on; flops per launch = the BF16 MFMA
                flops the kernel executes, 3 * 2 * N_det * P * E (P = n(n+1)/2 packed beta pairs,
                E = k(n-k) alpha single excitations, BF16x3 = 3 BF16 products; the diagonal alpha
                term is a rank-1 VALU update), against the 2.5 PFLOP/s dense BF16 peak. The
                packed-pair GEMM count 3 * 2 * N_det * P^2 is reported beside it (packed_pair_*),
                not as the fraction: the kernel never runs those flops.
                traffic = HBM bytes per launch from the committed rocprofv3 counter summary;
  kernel_rooflines  every sigma kernel group, mixed and FP64, against its bound (same-spin:
                    2 N_det (s + oo + vvoo) flops on the FP32 / FP64 VALU; alpha-beta: executed
                    MFMA flops);
  fp64_sigma    the FP64 sigma of the plain solver on the same vector;
  ms0_symmetric_sigma  the sigma of a spin-flip-symmetric vector (C^T = C) without its beta
                same-spin term (fci_sigma_ms0_device), mixed and FP64;
  time_to_solution  full Davidson-Liu solves to e_conv 1e-9 at r_conv 1e-4 and at Forte's
                    default 1e-6: FP64 with the diagonal preconditioner, mixed precision
                    (FP32/BF16x3 sigma + FP64 refinement) with the learned preconditioner (shipped
                    weights), the diagonal one and Olsen's correction, all in the Ms = 0
                    spin-flip-symmetric subspace (the default); FP64 and mixed+diagonal also with
                    the symmetry off ("/full_sigma", every sigma built in full as in Forte);
  configs       CAS(8,8) FP64 and CAS(14,14) FP64 3-root solves (with root-0 RDM time) and their
                CPU baselines;
  spin_adaptation  CAS(14,14) (or CAS(n,n) with --spin-adapt-n) solved in determinants and in
                   CSFs (CI_SPIN_ADAPT);
  batched       512 small problems in one batched call against 16 concurrent solvers;
  cpu_baseline  the CPU restatement of Forte's sigma (oracle/fci_oracle.c, one core) on a
                strided sample of its loops, extrapolated to one full sigma.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BF16_PEAK_TFLOPS = 2500.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBPS = 8000.0


def host_cpu():
    """Model name and core counts of the host the CPU baseline runs on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return dict(model=model, logical_cpus=os.cpu_count(), usable_cpus=usable)


def traffic_from_profiles(n):
    """HBM bytes per launch of the alpha-beta kernel from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", f"k_ab3_mixed_cas{n}_{n}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_baseline(n, h, eri, sample_stride):
    """Forte-structured CPU sigma (oracle, FP64, 1 thread) on a strided sample, extrapolated."""
    import numpy as np
    from oracle import oracle as O

    t0 = time.time()
    o = O.Oracle.from_chem(h, eri, n // 2, n // 2)
    setup = time.time() - t0
    C = np.random.default_rng(7).standard_normal(o.ndet)
    C /= np.linalg.norm(C)
    t, w = o.sigma_sample(C, sample_stride)
    est = float(sum(ti / wi for ti, wi in zip(t, w) if wi > 0))
    return dict(value=round(9 * o.ndet * 8 / est / 1e9, 4), unit="GB/s", cores=1, kind="port",
                host=host_cpu(),
                seconds_per_sigma_extrapolated=round(est, 2), sampled_seconds=round(float(t.sum()), 2),
                list_setup_seconds=round(setup, 2),
                sample=(f"FP64 Forte-structured sigma (oracle/fci_oracle.c) of the same CAS({n},{n}) "
                        f"vector: every term's outer loop (H1: pq, H2_aaaa: pair pq, H2_aabb: rs) run "
                        f"on indices i % {sample_stride} == 0, each term extrapolated by the fraction "
                        f"of its list entries covered"))


def batch_cpu_baseline(probs, gpu_results):
    """Oracle (one core) on the batched workload: full FP64 solves of two problems per size for
    n = 6, 8, 10; for n = 12 one sampled sigma (stride 8) times 2 x the GPU iteration count.
    Extrapolated to the whole batch by the problem counts per size."""
    import numpy as np
    from oracle import oracle as O
    per_n_time = {}
    for nn in (6, 8, 10):
        sel = [p for p in probs if p[0] == nn][:2]
        t0 = time.time()
        for n_, seed in sel:
            h_, eri_ = O.seeded_integrals(n_, seed)
            O.Oracle.from_chem(h_, eri_, n_ // 2, n_ // 2).solve(nroot=2, maxiter=600, e_conv=1e-9, r_conv=1e-5)
        per_n_time[nn] = (time.time() - t0) / len(sel)
    sel = [p for p in probs if p[0] == 12][:1]
    if sel:
        n_, seed = sel[0]
        h_, eri_ = O.seeded_integrals(n_, seed)
        o = O.Oracle.from_chem(h_, eri_, 6, 6)
        C = np.random.default_rng(1).standard_normal(o.ndet)
        t, w = o.sigma_sample(C, 8)
        sig = float(sum(ti / wi for ti, wi in zip(t, w) if wi > 0))
        its = [r["iterations"] for r in gpu_results if r["n"] == 12]
        per_n_time[12] = sig * 2 * (sum(its) / len(its))
    counts = {}
    for n_, _ in probs:
        counts[n_] = counts.get(n_, 0) + 1
    total = sum(counts[k] * per_n_time[k] for k in counts if k in per_n_time)
    return dict(value=round(len(probs) / total, 4), unit="problems/s", cores=1, kind="port",
                host=host_cpu(),
                seconds_extrapolated=round(total, 1),
                sample="oracle FP64 solves of 2 problems each for n = 6, 8, 10; n = 12 from one strided "
                       "sigma sample x 2 x the GPU iteration count; scaled by the problem counts")


def other_configs(forte_amd, with_cpu):
    """CAS(8,8) FP64 (seed 1, e_conv 1e-10) and CAS(14,14) FP64 3 roots (seed 2, subspace 8 per
    root, e_conv 1e-9) solved on the GPU, with root-0 RDM time; CPU baselines from the oracle: a
    full CAS(8,8) solve, and for CAS(14,14) a strided sigma sample extrapolated to one sigma times
    the GPU's sigma count."""
    import numpy as np
    out = {}
    for key, n, seed, nroot, e_conv, spr in (("cas8_fp64_1root", 8, 1, 1, 1e-10, 10),
                                             ("cas14_fp64_3roots", 14, 2, 3, 1e-9, 8)):
        h, eri = forte_amd.seeded_integrals(n, seed)
        ints = forte_amd.make_active_space_integrals(h, eri)
        s = forte_amd.FCISolver(n // 2, n // 2, 1, nroot, ints)
        s.set_e_convergence(e_conv)
        s.set_r_convergence(1e-6)
        s.set_subspace_per_root(spr)
        s.set_maxiter(500)
        forte_amd.device_synchronize()
        t0 = time.perf_counter()
        s.compute_energy()
        forte_amd.device_synchronize()
        dt = time.perf_counter() - t0
        t0 = time.perf_counter()
        s.rdms([(0, 0)], 2)
        forte_amd.device_synchronize()
        t_rdm = time.perf_counter() - t0
        st = s.stats()
        d = dict(workload=f"CAS({n},{n}) seed {seed}, {nroot} root(s), FP64, e_conv {e_conv:g}, r_conv 1e-6, "
                          f"DL_SUBSPACE_PER_ROOT {spr}",
                 seconds=round(dt, 4), rdm_seconds=round(t_rdm, 4), energies=list(s.energies()),
                 **{k: st[k] for k in ("iterations", "sigma_builds", "converged")})
        print(f"# {key}: {d}", file=sys.stderr, flush=True)
        if with_cpu:
            from oracle import oracle as O
            o = O.Oracle.from_chem(h, eri, n // 2, n // 2)
            if n <= 10:
                t0 = time.perf_counter()
                r = o.solve(nroot=nroot, e_conv=e_conv, r_conv=1e-6, subspace_per_root=spr, maxiter=500,
                            want_evecs=False)
                tc = time.perf_counter() - t0
                d["cpu_baseline"] = dict(seconds=round(tc, 3), cores=1, kind="port", host=host_cpu(),
                                         sample="full oracle FCISolver solve (Forte-structured FP64 sigma, "
                                                "Davidson-Liu), same thresholds",
                                         abs_dE=float(np.abs(np.array(r["energies"]) - np.array(s.energies())).max()),
                                         iterations=r["niter"])
            else:
                C = np.random.default_rng(7).standard_normal(o.ndet)
                t, w = o.sigma_sample(C, 16)
                est = float(sum(ti / wi for ti, wi in zip(t, w) if wi > 0))
                d["cpu_baseline"] = dict(seconds=round(est * st["sigma_builds"], 1), cores=1, kind="port",
                                         host=host_cpu(), sigma_seconds=round(est, 2),
                                         sample="oracle Forte-structured FP64 sigma on outer-loop indices "
                                                "i % 16 == 0, extrapolated to one sigma, times the GPU "
                                                "solve's sigma count (Davidson-Liu and RDM time not included)")
            print(f"# {key} cpu: {d['cpu_baseline']}", file=sys.stderr, flush=True)
        out[key] = d
        del s
    return out


def spin_adapt_leg(forte_amd, n=14):
    """CAS(n,n) singlet (n = 14 by default), seed 3, FP64, e_conv 1e-9 / r_conv 1e-6: the
    determinant-basis solve and the CSF-basis one (CI_SPIN_ADAPT), with the CSF and coupling counts."""
    h, eri = forte_amd.seeded_integrals(n, 3)
    ints = forte_amd.make_active_space_integrals(h, eri)
    k = n // 2
    out = {"workload": f"CAS({n},{n}) seed 3 singlet, FP64, e_conv 1e-9, r_conv 1e-6"}
    for label, sa in (("determinants", False), ("csf", True)):
        s = forte_amd.FCISolver(k, k, 1, 1, ints)
        s.set_spin_adapt(sa)
        s.set_e_convergence(1e-9)
        s.set_r_convergence(1e-6)
        s.set_maxiter(400)
        if sa:
            out["ncsf"], out["ncoupling"] = s.csf_size()
        forte_amd.device_synchronize()
        t0 = time.perf_counter()
        e = s.compute_energy()
        forte_amd.device_synchronize()
        st = s.stats()
        out[label] = dict(seconds=round(time.perf_counter() - t0, 3), energy=e, iterations=st["iterations"],
                          converged=st["converged"])
        print(f"# spin adaptation {label}: {out[label]}", file=sys.stderr, flush=True)
        del s
    out["ndet"] = int(forte_amd.FCISolver(k, k, 1, 1, ints).ndet())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16)
    ap.add_argument("--solve", type=int, default=1)
    ap.add_argument("--cpu", type=int, default=1)
    ap.add_argument("--cpu-stride", type=int, default=64)
    ap.add_argument("--maxiter", type=int, default=400)
    ap.add_argument("--pcie", type=int, default=1, help="time one FP64 sigma on host arrays (PCIe both ways)")
    ap.add_argument("--batch", type=int, default=512,
                    help="also solve this many seeded small problems (n in {6,8,10,12}, 2 roots, FP64) "
                         "with concurrent solvers; 0 to skip")
    ap.add_argument("--batch-threads", type=int, default=16)
    ap.add_argument("--configs", type=int, default=1,
                    help="also solve CAS(8,8) (1 root) and CAS(14,14) (3 roots, RDMs) in FP64")
    ap.add_argument("--legs", default="all",
                    help="time-to-solution legs to run, comma-separated names as in the output "
                         "(e.g. mixed+diagonal,fp64+diagonal/full_sigma), or all")
    ap.add_argument("--rconvs", default="1e-4,1e-6", help="residual thresholds of the time-to-solution legs")
    ap.add_argument("--spin-adapt-n", type=int, default=14, choices=(8, 10, 12, 14, 16),
                    help="active orbitals of the spin adaptation leg (CAS(n,n) singlet)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a plain copy kernel whose "
                         "FETCH_SIZE/WRITE_SIZE counters calibrate the traffic figure)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this FCI solver runs on one GPU; multi-GPU runs are not part of this build "
              "(the sigma build would shard by alpha-row blocks with one all-gather of C per build)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import forte_amd
    import forte_amd.learned

    n = args.n
    na = nb = n // 2
    h, eri = forte_amd.seeded_integrals(n, 3)
    ints = forte_amd.make_active_space_integrals(h, eri)
    t0 = time.time()
    solver = forte_amd.FCISolver(na, nb, 1, 1, ints)
    setup_s = time.time() - t0
    nd = solver.ndet()
    x = np.random.default_rng(0).standard_normal(nd)
    x /= np.linalg.norm(x)

    # ---- mixed-precision sigma (the timed steps) ----------------------------------------------
    C = forte_amd.DeviceArray(nd, "float32")
    S = forte_amd.DeviceArray(nd, "float32")
    C.upload(x.astype(np.float32))
    print(f"# CAS({n},{n}) ndet={nd} setup {setup_s:.2f}s; warmup {args.warmup} steps", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        solver.sigma_device(C.ptr(), S.ptr(), 1)
    forte_amd.device_synchronize()
    # timed region: exactly `steps` sigma builds; each build records HIP events around its kernel
    # groups on the stream they run on (sigma_timings), summed here per group
    t3 = [0.0, 0.0, 0.0]
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.sigma_device(C.ptr(), S.ptr(), 1)
        for i, v in enumerate(solver.sigma_timings()):
            t3[i] += v
    forte_amd.device_synchronize()
    wall = time.perf_counter() - t0
    t3 = [v / args.steps for v in t3]
    ms_per_step = 1e3 * wall / args.steps
    ms_event = sum(t3)
    value = 9 * nd * 4 / (ms_per_step * 1e-3) / 1e9
    print(f"# mixed sigma {ms_per_step:.2f} ms/step (events {ms_event:.2f}); beta-same {t3[0]:.2f} "
          f"alpha-same {t3[1]:.2f} alpha-beta {t3[2]:.2f} ms", file=sys.stderr, flush=True)
    np_ = n * (n + 1) // 2
    E = na * (n - na)
    s_ab = na * (n - na + 1)
    # alpha-beta flops: the packed-GEMM formulation 2*N_det*P^2 fixed by the benchmark definition
    # (x3 BF16 products for BF16x3); this kernel executes the cheaper 2*N_det*P*E (E = k(n-k) < P)
    ab_flops = 3 * 2 * nd * np_ * np_
    ab_exec = 3 * 2 * nd * np_ * E
    ab_useful = 2 * nd * s_ab * s_ab
    achieved = ab_exec / (t3[2] * 1e-3) / 1e12
    # same-spin term of one spin (H1 folded in): F = 2 N_det (s + oo + vvoo), VALU-bound
    kk = na
    oo = kk * (kk - 1) // 2
    same_flops = 2 * nd * (kk * (n - kk + 1) + oo + oo * ((n - kk + 2) * (n - kk + 1) // 2 - 1))
    del C, S

    if args.calib_copy_mb > 0:
        nc = args.calib_copy_mb * (1 << 20) // 4
        A = forte_amd.DeviceArray(nc, "float32")
        B = forte_amd.DeviceArray(nc, "float32")
        B.copy_from(A)
        forte_amd.device_synchronize()
        print(f"# calibration copy of {nc * 4} bytes", file=sys.stderr, flush=True)
        del A, B

    # ---- PCIe-inclusive FP64 sigma (host arrays in and out; reported, never the value) ---------
    pcie_ms = None
    if args.pcie:
        forte_amd.device_synchronize()
        t0 = time.perf_counter()
        solver.sigma(x)
        pcie_ms = 1e3 * (time.perf_counter() - t0)
        print(f"# fp64 sigma with host buffers (PCIe both ways) {pcie_ms:.1f} ms", file=sys.stderr, flush=True)

    # ---- FP64 sigma on the same vector -----------------------------------------------------------
    C64 = forte_amd.DeviceArray(nd, "float64")
    S64 = forte_amd.DeviceArray(nd, "float64")
    C64.upload(x)
    solver.sigma_bench(C64, S64, 1)
    ms64, t3_64 = solver.sigma_bench(C64, S64, max(1, args.steps // 4))
    print(f"# fp64 sigma {ms64:.2f} ms; beta-same {t3_64[0]:.2f} alpha-same {t3_64[1]:.2f} "
          f"alpha-beta {t3_64[2]:.2f} ms", file=sys.stderr, flush=True)
    del C64, S64

    # ---- the spin-flip-symmetric sigma on an Ms = 0 singlet-symmetric vector (C^T = C) -----------
    # the solves below keep their vectors in that subspace, where the beta same-spin term is the
    # transposed alpha one (fci_sigma_ms0_device)
    N = int(round(np.sqrt(nd)))
    xs = x.reshape(N, N)
    xs = xs + xs.T
    xs = (xs / np.linalg.norm(xs)).reshape(-1)
    ms0 = {}
    for dt_name, dt in (("mixed", "float32"), ("fp64", "float64")):
        Cs = forte_amd.DeviceArray(nd, dt)
        Ss = forte_amd.DeviceArray(nd, dt)
        Cs.upload(xs.astype(dt))
        solver.sigma_ms0_arrays(Cs, Ss, 1)
        reps = max(1, args.steps // (2 if dt == "float32" else 4))
        forte_amd.device_synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            solver.sigma_ms0_arrays(Cs, Ss, 1)
        ms0[dt_name] = 1e3 * (time.perf_counter() - t0) / reps
        del Cs, Ss
    del xs
    print(f"# spin-flip-symmetric sigma: mixed {ms0['mixed']:.2f} ms, fp64 {ms0['fp64']:.2f} ms", file=sys.stderr,
          flush=True)

    result = {
        "metric": "CAS(16,16) FCI time to |dE|<1e-8 Eh: mixed+learned vs FP64 plain; sigma GB/s",
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+bf16x3",
        "data": "synthetic (seeded integrals, seed 3; random normalized C)",
        "config": {"workload": f"FCI sigma CAS({n},{n}) Ms=0 C1", "ndet": nd, "n_orbitals": n,
                   "n_alpha": na, "n_beta": nb, "precision": "mixed"},
        "roofline": {"bound": "mfma", "kernel": "k_ab3_mixed (alpha-beta BF16x3 contraction)",
                     "achieved": round(achieved, 2), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / BF16_PEAK_TFLOPS, 4), "traffic": traffic_from_profiles(n),
                     "kernel_ms": round(t3[2], 3),
                     "algorithmic_flops": ab_exec,
                     "formulation": "3 x 2 N_det P E BF16 MFMA flops the kernel executes (P = n(n+1)/2 packed "
                                    "beta pairs, E = k(n-k) alpha single excitations, BF16x3 = 3 products)",
                     "packed_pair_flops": ab_flops,
                     "packed_pair_tflops": round(ab_flops / (t3[2] * 1e-3) / 1e12, 2),
                     "packed_pair_formulation": "3 x 2 N_det P^2 (dense packed-pair GEMM; counts flops the "
                                                "kernel never runs, so its fraction is not a roofline)",
                     "useful_flops": ab_useful,
                     "algorithmic_bytes": 3 * nd * 4},
        "sigma_split_ms": {"beta_same_spin_with_transposes": round(t3[0], 3),
                           "alpha_same_spin": round(t3[1], 3), "alpha_beta": round(t3[2], 3)},
        "fp64_sigma": {"ms": round(ms64, 3), "GBps": round(9 * nd * 8 / (ms64 * 1e-3) / 1e9, 3),
                       "split_ms": [round(t, 3) for t in t3_64],
                       "host_buffers_pcie_inclusive_ms": None if pcie_ms is None else round(pcie_ms, 1)},
        "ms0_symmetric_sigma": {"vector": "C^T = C (Ms = 0, even S), the subspace the singlet solves keep",
                                "mixed_ms": round(ms0["mixed"], 3), "fp64_ms": round(ms0["fp64"], 3),
                                "note": "beta same-spin term = transposed alpha term, not built; "
                                        "wall time per call incl. one host synchronization"},
    }

    # every sigma kernel group against the roofline that bounds it
    ab64_exec = 2 * nd * np_ * E
    result["kernel_rooflines"] = {
        "mixed_alpha_same_spin": dict(ms=round(t3[1], 3), bound="valu-fp32", flops=same_flops,
                                      peak_tflops=FP32_PEAK_TFLOPS,
                                      frac=round(same_flops / FP32_PEAK_TFLOPS / 1e9 / t3[1], 4)),
        "mixed_beta_same_spin_with_transposes": dict(ms=round(t3[0], 3), bound="valu-fp32", flops=same_flops,
                                                     peak_tflops=FP32_PEAK_TFLOPS,
                                                     frac=round(same_flops / FP32_PEAK_TFLOPS / 1e9 / t3[0], 4)),
        "mixed_alpha_beta": dict(ms=round(t3[2], 3), bound="mfma-bf16", flops=ab_exec, peak_tflops=BF16_PEAK_TFLOPS,
                                 frac=round(ab_exec / BF16_PEAK_TFLOPS / 1e9 / t3[2], 4)),
        "fp64_alpha_same_spin": dict(ms=round(t3_64[1], 3), bound="valu-fp64", flops=same_flops,
                                     peak_tflops=FP64_PEAK_TFLOPS,
                                     frac=round(same_flops / FP64_PEAK_TFLOPS / 1e9 / t3_64[1], 4)),
        "fp64_beta_same_spin_with_transposes": dict(ms=round(t3_64[0], 3), bound="valu-fp64", flops=same_flops,
                                                    peak_tflops=FP64_PEAK_TFLOPS,
                                                    frac=round(same_flops / FP64_PEAK_TFLOPS / 1e9 / t3_64[0], 4)),
        "fp64_alpha_beta": dict(ms=round(t3_64[2], 3), bound="mfma-fp64", flops=ab64_exec,
                                peak_tflops=FP64_PEAK_TFLOPS,
                                frac=round(ab64_exec / FP64_PEAK_TFLOPS / 1e9 / t3_64[2], 4)),
    }

    # ---- time to solution ------------------------------------------------------------------------
    if args.solve:
        # (precision, preconditioner, Ms = 0 spin-flip symmetry); "full_sigma": the symmetry off, every
        # sigma with its beta same-spin term, as the reference builds it
        legs = [("fp64", "diagonal", True), ("fp64", "diagonal", False), ("mixed", "learned", True),
                ("mixed", "diagonal", True), ("mixed", "olsen", True), ("mixed", "diagonal", False)]
        lp_w = None
        if os.path.exists(forte_amd.learned.WEIGHTS):
            lp_w = forte_amd.learned.load_weights()
        else:
            legs = [leg for leg in legs if leg[1] != "learned"]

        def leg_name(prec, pc, sym):
            return f"{prec}+{pc}" + ("" if sym else "/full_sigma")

        if args.legs != "all":
            wanted = set(args.legs.split(","))
            legs = [leg for leg in legs if leg_name(*leg) in wanted]

        def solve_leg(prec, pc, sym, r_conv):
            s = forte_amd.FCISolver(na, nb, 1, 1, ints)
            s.set_e_convergence(1e-9)
            s.set_r_convergence(r_conv)
            s.set_maxiter(args.maxiter)
            s.set_die_if_not_converged(False)
            s.set_precision(prec)
            s.set_ms0_symmetry(sym)
            if pc == "learned":
                s.set_lp_weights(lp_w)
                s.set_preconditioner("learned")
            elif pc == "olsen":
                s.set_preconditioner("olsen")
            forte_amd.device_synchronize()
            t0 = time.perf_counter()
            e = s.compute_energy()
            forte_amd.device_synchronize()
            dt = time.perf_counter() - t0
            st = s.stats()
            print(f"# solve {leg_name(prec, pc, sym)} r_conv {r_conv:g}: E={e:.12f} {dt:.2f}s {st}", file=sys.stderr,
                  flush=True)
            # time outside the sigma builds (Davidson-Liu vector passes, guess, host steps) per iteration
            n32 = st["sigma_builds"] - st["sigma_builds_fp64"]
            m32, m64 = (ms0["mixed"], ms0["fp64"]) if st["ms0_symmetry"] else (ms_per_step, ms64)
            rest = dt - 1e-3 * (n32 * m32 + st["sigma_builds_fp64"] * m64)
            return dict(seconds=round(dt, 3), energy=e,
                        non_sigma_ms_per_iteration=round(1e3 * rest / max(1, st["iterations"]), 2),
                        **{k: st[k] for k in ("iterations", "sigma_builds", "sigma_builds_fp64",
                                              "converged", "fell_back", "ms0_symmetry")})

        def speedup(num, den):
            return round(num["seconds"] / den["seconds"], 3) if num and den else None

        result["time_to_solution"] = {}
        for r_conv in (float(x) for x in args.rconvs.split(",")):
            tts = {leg_name(*leg): solve_leg(*leg, r_conv) for leg in legs}
            ref = tts.get("fp64+diagonal")
            plain = tts.get("fp64+diagonal/full_sigma")
            mixed = tts.get("mixed+diagonal")
            best = tts.get("mixed+learned", mixed)
            result["time_to_solution"][f"r_conv_{r_conv:g}"] = {
                "runs": tts,
                "speedup_mixed_learned_vs_fp64": speedup(ref, best),
                "speedup_mixed_diagonal_vs_fp64": speedup(ref, mixed),
                "speedup_mixed_olsen_vs_fp64": speedup(ref, tts.get("mixed+olsen")),
                "speedup_mixed_diagonal_vs_fp64_full_sigma": speedup(plain, mixed),
                "speedup_ms0_symmetry_fp64": speedup(plain, ref),
                "speedup_ms0_symmetry_mixed": speedup(tts.get("mixed+diagonal/full_sigma"), mixed),
                "abs_dE_vs_fp64": {k: abs(v["energy"] - ref["energy"]) for k, v in tts.items()} if ref else None,
                "thresholds": {"e_conv": 1e-9, "r_conv": r_conv, "maxiter": args.maxiter},
                "learned_weights": os.path.relpath(forte_amd.learned.WEIGHTS, ROOT) if lp_w is not None else None}

    # ---- the other configurations: CAS(8,8) and CAS(14,14) FP64 solves ---------------------------
    if args.configs:
        result["configs"] = other_configs(forte_amd, args.cpu)
        result["spin_adaptation"] = spin_adapt_leg(forte_amd, args.spin_adapt_n)

    # ---- batched small problems -----------------------------------------------------------------
    if args.batch > 0:
        from forte_amd.batch import seeded_batch, solve_batch, solve_batch_threads
        probs = seeded_batch(args.batch)
        bopt = {"E_CONVERGENCE": 1e-9, "R_CONVERGENCE": 1e-5}
        solve_batch(probs[:8], nroot=2, options=bopt)  # warm-up (code objects, allocator)
        forte_amd.device_synchronize()
        res, wall, bst = solve_batch(probs, nroot=2, options=bopt)
        forte_amd.device_synchronize()
        by_n = {}
        for r in res:
            d = by_n.setdefault(r["n"], {"count": 0, "iterations": 0})
            d["count"] += 1
            d["iterations"] += r["iterations"]
        result["batched"] = {"workload": f"{args.batch} seeded CAS(n,n), n in {{6,8,10,12}}, 2 roots, FP64, "
                                         f"e_conv 1e-9, r_conv 1e-5",
                             "schedule": "one batched call: problems of one size advance together, one captured "
                                         "graph replay per Davidson-Liu iteration, autotuned chunk and sigma "
                                         "sub-batch per size",
                             "problems_per_s": round(len(res) / wall, 2), "seconds": round(wall, 3),
                             "all_converged": all(r["converged"] for r in res), "per_n": by_n,
                             "setup_seconds": round(bst["setup_seconds"], 3),
                             "iterate_seconds": round(bst["iterate_seconds"], 3), "classes": bst["classes"]}
        print(f"# batch: {len(res)} problems in {wall:.2f}s ({len(res) / wall:.1f}/s), "
              f"converged {result['batched']['all_converged']} {bst}", file=sys.stderr, flush=True)
        if args.batch_threads > 0:
            rt, wt = solve_batch_threads(probs, nroot=2, threads=args.batch_threads, options=bopt)
            forte_amd.device_synchronize()
            de = max(abs(a - b) for r1, r2 in zip(res, rt) for a, b in zip(r1["energies"], r2["energies"]))
            result["batched"]["threaded_schedule"] = {
                "problems_per_s": round(len(rt) / wt, 2), "seconds": round(wt, 3), "threads": args.batch_threads,
                "all_converged": all(r["converged"] for r in rt), "max_abs_dE_vs_batched": de}
            print(f"# threaded batch: {wt:.2f}s, max |dE| vs batched {de:.2e}", file=sys.stderr, flush=True)
        if args.cpu:
            result["batched"]["cpu_baseline"] = batch_cpu_baseline(probs, res)
            print(f"# batch cpu baseline {result['batched']['cpu_baseline']}", file=sys.stderr, flush=True)

    if args.cpu:
        result["cpu_baseline"] = cpu_baseline(n, h, eri, args.cpu_stride)
        # the same FP64 solves on the CPU, extrapolated: one sigma times the GPU FP64 solve's sigma
        # count (Davidson-Liu vector work not included)
        if "time_to_solution" in result:
            sps = result["cpu_baseline"]["seconds_per_sigma_extrapolated"]
            result["cpu_baseline"]["time_to_solution_extrapolated_seconds"] = {
                k: round(sps * v["runs"]["fp64+diagonal/full_sigma"]["sigma_builds"], 1)
                for k, v in result["time_to_solution"].items() if "fp64+diagonal/full_sigma" in v["runs"]}
        print(f"# cpu baseline {result['cpu_baseline']}", file=sys.stderr, flush=True)
    else:
        result["cpu_baseline"] = None

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
