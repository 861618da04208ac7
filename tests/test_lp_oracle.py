"""CPU checks that pin the learned-preconditioner oracle (oracle/lp_oracle.py) independently of
its own formulas: finite-difference gradients, the y = 0 reduction to Forte's diagonal
preconditioner, feature bits against explicit occupations, and Adam's first step."""
import numpy as np

from oracle import lp_oracle as L


def _params(seed):
    rng = np.random.default_rng(seed)
    p = L.init_params(seed)
    p["w3"] = rng.standard_normal(L.NH) * 0.2
    p["b3"] = np.array([0.1])
    p["b1"] = rng.standard_normal(L.NH) * 0.1
    p["b2"] = rng.standard_normal(L.NH) * 0.1
    return p


def _data(m, seed):
    rng = np.random.default_rng(seed)
    X = np.concatenate([rng.uniform(0, 3, (m, 1)), rng.standard_normal((m, 1)),
                        rng.integers(0, 2, (m, 32)).astype(float)], axis=1)
    return X, rng.standard_normal(m), rng.uniform(0.5, 4.0, m), rng.standard_normal(m)


def test_gradients_against_finite_differences():
    p = _params(1)
    X, r, hd, e = _data(150, 2)
    lam = 0.1
    loss, g = L.loss_and_grads(p, X, r, hd, lam, e, reg=1e-3)
    v, gv = L.flatten(p), L.flatten(g)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(L.NH * L.NIN, 12, replace=False),          # W1
                          L.NH * L.NIN + rng.choice(L.NH, 4, replace=False),    # b1
                          L.NH * (L.NIN + 1) + rng.choice(L.NH * L.NH, 12, replace=False),  # W2
                          L.NPARAM - 1 - rng.choice(2 * L.NH + 1, 8, replace=False)])  # b2, w3, b3
    h = 1e-6
    for i in idx:
        vp, vm = v.copy(), v.copy()
        vp[i] += h
        vm[i] -= h
        fd = (L.loss_and_grads(L.unflatten(vp), X, r, hd, lam, e, reg=1e-3)[0] -
              L.loss_and_grads(L.unflatten(vm), X, r, hd, lam, e, reg=1e-3)[0]) / (2 * h)
        assert abs(fd - gv[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, gv[i])


def test_zero_output_layer_is_forte_diagonal():
    p = L.init_params(0)  # output layer zero-initialized: y = 0
    X, r, hd, _ = _data(64, 5)
    lam = 1.0
    hd[3] = lam + 1e-9  # |lambda - H| <= 1e-6 -> 0, davidson_liu_solver.cc:441-446
    d = L.precondition(p, X, r, hd, lam)
    ref = np.array([ri / (lam - hi) if abs(lam - hi) > 1e-6 else 0.0 for ri, hi in zip(r, hd)])
    assert np.array_equal(d, ref)


def test_scaled_correction_has_no_loss_change():
    # only the direction of d matters: a constant y (b3) leaves the cosine loss unchanged
    p0, p1 = L.init_params(0), L.init_params(0)
    p1["b3"] = np.array([1.7])
    X, r, hd, e = _data(80, 7)
    l0 = L.loss_and_grads(p0, X, r, hd, 0.2, e, reg=0.0)[0]
    l1 = L.loss_and_grads(p1, X, r, hd, 0.2, e, reg=0.0)[0]
    assert abs(l0 - l1) < 1e-12


def test_feature_bits_from_occupations():
    # n = 6, k = 3: aufbau occupation 0,1,2; holes count down from orbital 2, particles up from 3
    n, k = 6, 3
    strings = [0b000111, 0b001011, 0b110100, 0b111000]
    f = L.occupation_bits(strings, n, k)
    for s, row in zip(strings, f):
        occ = [(s >> i) & 1 for i in range(n)]
        holes = [1 - occ[k - 1 - j] if k - 1 - j >= 0 else 0 for j in range(8)]
        parts = [occ[k + j] if k + j < n else 0 for j in range(8)]
        assert list(row) == holes + parts
        assert row.sum() == 2 * sum(occ[k:])  # holes = particles for a fixed electron count


def test_adam_first_step_is_signed_lr():
    # after one step from zero moments, Adam moves every parameter by lr * g / (|g| + eps')
    p = _params(4)
    g = {k: np.random.default_rng(9).standard_normal(np.shape(v)) for k, v in p.items()}
    q = L.adam_step({k: v.copy() for k, v in p.items()}, g, {}, lr=1e-3, eps=1e-12)
    for k in p:
        assert np.allclose(p[k] - q[k], 1e-3 * np.sign(g[k]), atol=1e-9)
