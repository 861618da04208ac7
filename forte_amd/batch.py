"""Many independent small FCI problems on one GPU (the batched workload: 512 seeded problems with
n drawn from {6, 8, 10, 12}, two roots each, FP64).

`solve_batch` hands all problems to the library at once (fci_batch_solve): problems of one shape
share string tables and advance together through Davidson-Liu iterations that run entirely on the
device, one captured graph replay per iteration (Forte's repeated FCISolver::compute_energy,
forte/fci/fci_solver.cc:177-314, as in mcscf/mcscf_2step.cc:187-198).

`solve_batch_threads` is the other schedule: one full `FCISolver` per problem on concurrent host
threads (the library uses per-thread default streams, so their kernels overlap on the device).
Results come back in input order.
"""
import time
from concurrent.futures import ThreadPoolExecutor

from . import _forte_amd as _C

_MASK = (1 << 64) - 1


def _splitmix64(x):
    z = (x + 0x9E3779B97F4A7C15) & _MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
    return z ^ (z >> 31)


def seeded_batch(count=512, first_seed=1000, sizes=(6, 8, 10, 12)):
    """[(n, seed)] of the batched workload: problem i has seed first_seed + i and
    n = sizes[splitmix64(seed) % len(sizes)]."""
    return [(sizes[_splitmix64(first_seed + i) % len(sizes)], first_seed + i) for i in range(count)]


def _solve_one(n, seed, nroot, options):
    from . import make_active_space_integrals
    h, eri = _C.seeded_integrals(n, seed)
    s = _C.FCISolver(n // 2, n // 2, 1, nroot, make_active_space_integrals(h, eri))
    s.set_maxiter(600)
    s.set_die_if_not_converged(False)  # reported per problem instead
    if options:
        s.set_options(dict(options))
    e = s.compute_energy()
    st = s.stats()
    return dict(n=n, seed=seed, energies=list(s.energies()), energy=e, iterations=st["iterations"],
                converged=st["converged"])


def solve_many(ints, na, nb, nroot=1, multiplicity=1, options=None, want_evecs=False):
    """Solve independent problems in one batched call (fci_batch_solve): problem i has the
    ActiveSpaceIntegrals ints[i] and na[i] / nb[i] electrons; options take FCISolver's option
    names (DL_*, E_CONVERGENCE, R_CONVERGENCE). Returns dict(energies [P, nroot], iterations,
    converged, stats[, evecs])."""
    return _C.batch_solve(list(ints), [int(x) for x in na], [int(x) for x in nb], nroot, multiplicity,
                          dict(options or {}), want_evecs)


def solve_batch(problems, nroot=2, options=None, want_evecs=False):
    """Solve [(n, seed)] seeded CAS(n, n) problems (Ms = 0, singlets) in one batched call.
    Returns (results in input order, wall seconds, batch statistics)."""
    from . import make_active_space_integrals
    t0 = time.perf_counter()
    ints, na, nb = [], [], []
    for n, seed in problems:
        h, eri = _C.seeded_integrals(n, seed)
        ints.append(make_active_space_integrals(h, eri))
        na.append(n // 2)
        nb.append(n // 2)
    opts = {"DL_MAXITER": 600}
    opts.update(options or {})
    out = _C.batch_solve(ints, na, nb, nroot, 1, opts, want_evecs)
    wall = time.perf_counter() - t0
    res = []
    for i, (n, seed) in enumerate(problems):
        e = [float(x) for x in out["energies"][i]]
        r = dict(n=n, seed=seed, energies=e, energy=e[0], iterations=out["iterations"][i],
                 converged=out["converged"][i])
        if want_evecs:
            r["evecs"] = out["evecs"][i]
        res.append(r)
    return res, wall, out["stats"]


def solve_batch_threads(problems, nroot=2, threads=16, options=None):
    """Solve [(n, seed)] seeded CAS(n, n) problems (Ms = 0, singlets) with `threads` concurrent
    solvers. Returns (results in input order, wall seconds)."""
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        futs = [ex.submit(_solve_one, n, seed, nroot, options) for n, seed in problems]
        res = [f.result() for f in futs]
    return res, time.perf_counter() - t0


if __name__ == "__main__":
    # python -m forte_amd.batch [count]: time the seeded batch (a warm-up call on 8 problems first)
    import sys
    probs = seeded_batch(int(sys.argv[1]) if len(sys.argv) > 1 else 512)
    opts = {"E_CONVERGENCE": 1e-9, "R_CONVERGENCE": 1e-5}
    solve_batch(probs[:8], nroot=2, options=opts)
    res, wall, st = solve_batch(probs, nroot=2, options=opts)
    print(f"{len(probs)} problems in {wall:.3f} s, all converged: {all(r['converged'] for r in res)}", flush=True)
    print(st, flush=True)
