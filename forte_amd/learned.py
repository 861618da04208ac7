"""Learned Davidson preconditioner: training data, training on the GPU and the shipped weights.

The preconditioner scales Forte's diagonal Davidson correction r / (lambda - H_II)
(forte/helpers/davidson_liu_solver.cc:432-448) element-wise by exp(MLP(x_I)); an MLP whose output
layer is zero reproduces the reference exactly. Training pairs come from FP64 solves with the
diagonal preconditioner (FCISolver.lp_harvest): for sampled determinants of each iteration, the
features x_I, the diagonal correction d0_I and the exact correction e_I = c_I / (c.x) - x_I. The
loss of a group of samples (one iteration of one problem) is 1 - cos^2(d0 exp(y), e).

    python -m forte_amd.learned --n 12 --problems 200 --out forte_amd/data/lp_weights.npy

trains on seeded CAS(n,n) problems (seeds 2000 + j) and writes the weights. Everything runs on the
GPU through the library: harvest (device Davidson-Liu), forward/backward and Adam (device kernels).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import _forte_amd as _C

NIN, NH = 34, 128
NPARAM = _C.LP_NPARAM
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
WEIGHTS = os.path.join(DATA_DIR, "lp_weights.npy")


def init_params(seed=0):
    """He initialization of the two hidden layers; output layer zero (the diagonal preconditioner)."""
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((NH, NIN)) * np.sqrt(2.0 / NIN)
    W2 = rng.standard_normal((NH, NH)) * np.sqrt(2.0 / NH)
    return np.concatenate([W1.ravel(), np.zeros(NH), W2.ravel(), np.zeros(NH), np.zeros(NH),
                           np.zeros(1)]).astype(np.float32)


def load_weights(path=WEIGHTS):
    w = np.load(path, allow_pickle=False)
    if w.shape != (NPARAM,) or w.dtype != np.float32:
        raise ValueError(f"{path}: expected {NPARAM} float32 parameters")
    return w


def save_weights(params, path=WEIGHTS, meta=None):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, np.asarray(params, dtype=np.float32))
    if meta is not None:
        with open(os.path.splitext(path)[0] + ".json", "w") as f:
            json.dump(meta, f, indent=1)


def seeded_solver(n, seed, na=None, nb=None):
    """FCISolver for the seeded integrals of CAS(n, n) (Ms = 0 unless na/nb given)."""
    from . import make_active_space_integrals
    h, eri = _C.seeded_integrals(n, seed)
    na = n // 2 if na is None else na
    nb = n // 2 if nb is None else nb
    return _C.FCISolver(na, nb, 1, 1, make_active_space_integrals(h, eri))


def harvest(solvers, per_iter=1024, max_groups=64, seed=0, e_conv=1e-9, r_conv=1e-5, maxiter=400, log=None):
    """Training groups from FP64 diagonal-preconditioned solves of each solver."""
    Xs, ds, es = [], [], []
    for j, s in enumerate(solvers):
        s.set_e_convergence(e_conv)
        s.set_r_convergence(r_conv)
        s.set_maxiter(maxiter)
        t0 = time.time()
        try:
            X, d0, e = s.lp_harvest(per_iter, max_groups, seed + j)
        except RuntimeError as err:  # no converged eigenvector: no targets from this problem
            if log:
                log(f"harvest {j}: skipped ({str(err).splitlines()[0]})")
            continue
        ng = len(d0) // per_iter
        keep = []
        for g in range(ng):  # groups whose correction or target vanish carry no direction
            sl = slice(g * per_iter, (g + 1) * per_iter)
            fin = np.isfinite(X[sl]).all() and np.isfinite(d0[sl]).all() and np.isfinite(e[sl]).all()
            if fin and np.linalg.norm(e[sl]) > 1e-10 and np.linalg.norm(d0[sl]) > 0:
                keep.append(g)
        idx = (np.asarray(keep)[:, None] * per_iter + np.arange(per_iter)[None, :]).ravel()
        Xs.append(X[idx])
        ds.append(d0[idx])
        es.append(e[idx])
        if log:
            log(f"harvest {j}: {len(keep)} of {ng} groups in {time.time() - t0:.2f}s")
    return np.concatenate(Xs), np.concatenate(ds), np.concatenate(es)


def train(X, d0, e, group, epochs=10, lr=1e-3, reg=1e-4, seed=0, params=None):
    """Adam on the GPU, one group (one Davidson iteration of one problem) per step."""
    if params is None:
        params = init_params(seed)
    return _C.lp_train(np.asarray(params, np.float32), X, d0, e, group=group, epochs=epochs, lr=lr, reg=reg,
                       seed=seed)


def group_losses(params, X, d0, e, group):
    """Loss of every group (no regularization) for the given parameters."""
    out = []
    for g in range(len(d0) // group):
        sl = slice(g * group, (g + 1) * group)
        out.append(_C.lp_loss_grad(params, X[sl], d0[sl], e[sl], 0.0)[0])
    return np.array(out)


def compare_solves(params, seeds, n, e_conv=1e-9, r_conv=1e-5, maxiter=400, precision="fp64",
                   variants=("diagonal", "learned")):
    """Davidson-Liu iterations and wall time per preconditioner variant: "diagonal" (Forte's),
    "learned", "olsen" (the diagonal correction made orthogonal to the Ritz vector)."""
    out = []
    for sd in seeds:
        row = {"seed": sd}
        for pc in variants:
            s = seeded_solver(n, sd)
            s.set_e_convergence(e_conv)
            s.set_r_convergence(r_conv)
            s.set_maxiter(maxiter)
            s.set_die_if_not_converged(False)
            s.set_precision(precision)
            if pc == "learned":
                s.set_lp_weights(params)
                s.set_preconditioner("learned")
            elif pc == "olsen":
                s.set_preconditioner("olsen")
            t0 = time.time()
            e = s.compute_energy()
            st = s.stats()
            row[pc] = dict(energy=e, iterations=st["iterations"], seconds=round(time.time() - t0, 4),
                           converged=st["converged"], fell_back=st["fell_back"])
        out.append(row)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=12)
    ap.add_argument("--problems", type=int, default=200)
    ap.add_argument("--first-seed", type=int, default=2000)
    ap.add_argument("--per-iter", type=int, default=1024)
    ap.add_argument("--max-groups", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--reg", type=float, default=1e-4)
    ap.add_argument("--holdout", type=int, default=8, help="problems kept out of training")
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--compare", action="store_true",
                    help="only solve the held-out problems with the diagonal, learned (shipped weights) and "
                         "Olsen preconditioners and print the iteration counts")
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    seeds = list(range(a.first_seed, a.first_seed + a.problems))
    ntrain = a.problems - a.holdout
    if a.compare:
        cmp = compare_solves(load_weights(), seeds[ntrain:], a.n, variants=("diagonal", "learned", "olsen"))
        tot = {v: sum(r[v]["iterations"] for r in cmp) for v in ("diagonal", "learned", "olsen")}
        for row in cmp:
            log(f"seed {row['seed']}: " + ", ".join(f"{v} {row[v]['iterations']}" for v in tot))
        log("total iterations: " + ", ".join(f"{v} {t}" for v, t in tot.items()))
        print(json.dumps(dict(n=a.n, seeds=seeds[ntrain:], totals=tot, solves=cmp)))
        return
    t0 = time.time()
    X, d0, e = harvest((seeded_solver(a.n, s) for s in seeds[:ntrain]), a.per_iter, a.max_groups, log=log)
    Xh, dh, eh = harvest((seeded_solver(a.n, s) for s in seeds[ntrain:]), a.per_iter, a.max_groups, seed=10**6,
                         log=log)
    t_h = time.time() - t0
    log(f"training set {len(d0)} samples, held-out {len(dh)}; harvest {t_h:.1f}s")
    p0 = init_params(0)
    t0 = time.time()
    params, losses = train(X, d0, e, a.per_iter, a.epochs, a.lr, a.reg, seed=0, params=p0)
    t_t = time.time() - t0
    base = float(group_losses(p0, Xh, dh, eh, a.per_iter).mean())
    held = float(group_losses(params, Xh, dh, eh, a.per_iter).mean())
    log(f"epoch losses {np.round(losses, 5).tolist()}; held-out loss diagonal {base:.5f} learned {held:.5f}")
    cmp = compare_solves(params, seeds[ntrain:], a.n)
    for row in cmp:
        log(f"held-out seed {row['seed']}: diagonal {row['diagonal']['iterations']} iterations, "
            f"learned {row['learned']['iterations']} (fell back: {row['learned']['fell_back']})")
    meta = dict(n=a.n, problems=a.problems, first_seed=a.first_seed, holdout=a.holdout, per_iter=a.per_iter,
                max_groups=a.max_groups, epochs=a.epochs, lr=a.lr, reg=a.reg, train_samples=int(len(d0)),
                epoch_loss=[float(x) for x in losses], heldout_loss_diagonal=base, heldout_loss_learned=held,
                harvest_seconds=t_h, train_seconds=t_t, heldout_solves=cmp)
    save_weights(params, a.out, meta)
    print(json.dumps(meta))


if __name__ == "__main__":
    main()
