"""CPU model of the learned Davidson preconditioner (numpy, FP64). Test infrastructure only.

Imported by tests/ (and nowhere in the product) to check the gfx950 kernels of the learned
preconditioner: feature construction, MLP forward pass, loss, back-propagated gradients and the
Adam update.

Preconditioner being learned. Forte's Davidson-Liu solver corrects the residual r of a root with
the diagonal preconditioner d_I = r_I / (lambda - H_II), zeroing elements with
|lambda - H_II| <= 1e-6 (forte/helpers/davidson_liu_solver.cc:432-448). The learned variant
scales that correction element-wise:

    d_I = r_I / (lambda - H_II) * exp(y_I),      y_I = MLP(x_I),

so y = 0 is exactly Forte's preconditioner. The MLP has three layers
(34 -> 128 -> 128 -> 1, tanh-form GELU after the first two) and reads the features

    x_I = [ (H_II - lambda) / 16,  r_I / rms(r),  16 alpha bits,  16 beta bits ]

where the bits describe the determinant relative to the aufbau (lowest-index-first) reference
of each spin: bit j < 8 is 1 when orbital k-1-j is empty (a hole), bit 8+j is 1 when orbital
k+j is occupied (a particle); positions outside 0..n-1 are 0.

Training target. For a Davidson iterate b (unit norm, current Ritz vector) with residual r and
the converged eigenvector c, the exact correction is e = c / (c.b) - b, orthogonalized to b.
Only the direction of d matters to the solver, so the loss over a set of determinants S is

    loss = 1 - (d.e)^2 / ((d.d)(e.e)) + reg * mean(y^2)

(all sums over S).
"""
import numpy as np

NIN = 34
NH = 128
HSCALE = 16.0
GELU_C = 0.7978845608028654  # sqrt(2/pi)


def occupation_bits(strings, n, k):
    """[nstr, 16] hole/particle bits of each string (bit i of a string = orbital i)."""
    s = np.asarray(strings, dtype=np.uint64)
    f = np.zeros((len(s), 16))
    for j in range(8):
        o = k - 1 - j
        if 0 <= o < n:
            f[:, j] = 1.0 - ((s >> np.uint64(o)) & np.uint64(1)).astype(np.float64)
        o = k + j
        if 0 <= o < n:
            f[:, 8 + j] = ((s >> np.uint64(o)) & np.uint64(1)).astype(np.float64)
    return f


def features(str_a, str_b, n, na, nb, hdiag, lam, r):
    """[ndet, 34] features in determinant order I = Ia * Nb + Ib."""
    fa = occupation_bits(str_a, n, na)
    fb = occupation_bits(str_b, n, nb)
    Na, Nb = len(fa), len(fb)
    rms = np.sqrt(np.mean(r * r))
    X = np.empty((Na * Nb, NIN))
    X[:, 0] = (hdiag - lam) / HSCALE
    X[:, 1] = r / rms if rms > 0 else 0.0
    X[:, 2:18] = np.repeat(fa, Nb, axis=0)
    X[:, 18:34] = np.tile(fb, (Na, 1))
    return X


def init_params(seed=0):
    """He-style initialization with the output layer at zero (y = 0: the diagonal preconditioner)."""
    rng = np.random.default_rng(seed)
    return {
        "W1": rng.standard_normal((NH, NIN)) * np.sqrt(2.0 / NIN),
        "b1": np.zeros(NH),
        "W2": rng.standard_normal((NH, NH)) * np.sqrt(2.0 / NH),
        "b2": np.zeros(NH),
        "w3": np.zeros(NH),
        "b3": np.zeros(1),
    }


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    u = GELU_C * (z + 0.044715 * z ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3 * 0.044715 * z * z)


def forward(p, X):
    z1 = X @ p["W1"].T + p["b1"]
    a1 = gelu(z1)
    z2 = a1 @ p["W2"].T + p["b2"]
    a2 = gelu(z2)
    y = a2 @ p["w3"] + p["b3"][0]
    return y, (X, z1, a1, z2, a2)


def precondition(p, X, r, hdiag, lam):
    """d = r / (lambda - H) * exp(MLP(x)), zero where |lambda - H| <= 1e-6."""
    y, _ = forward(p, X)
    den = lam - hdiag
    d = np.where(np.abs(den) > 1e-6, r / np.where(np.abs(den) > 1e-6, den, 1.0), 0.0)
    return d * np.exp(y)


def loss_and_grads(p, X, r, hdiag, lam, e, reg=1e-4):
    """Loss over the rows of X and its gradients by back-propagation."""
    y, (X, z1, a1, z2, a2) = forward(p, X)
    den = lam - hdiag
    d0 = np.where(np.abs(den) > 1e-6, r / np.where(np.abs(den) > 1e-6, den, 1.0), 0.0)
    d = d0 * np.exp(y)
    Sde, Sdd, See = d @ e, d @ d, e @ e
    m = len(y)
    loss = 1.0 - Sde * Sde / (Sdd * See) + reg * np.mean(y * y)
    gd = -2.0 * Sde * e / (Sdd * See) + 2.0 * Sde * Sde * d / (Sdd * Sdd * See)
    gy = gd * d + 2.0 * reg * y / m
    g = {"w3": a2.T @ gy, "b3": np.array([gy.sum()])}
    d2 = np.outer(gy, p["w3"]) * gelu_grad(z2)
    g["W2"] = d2.T @ a1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ p["W2"]) * gelu_grad(z1)
    g["W1"] = d1.T @ X
    g["b1"] = d1.sum(0)
    return loss, g


def adam_step(p, g, state, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma and Ba, 2015) with bias correction; state holds m, v and t."""
    t = state.get("t", 0) + 1
    state["t"] = t
    for k in p:
        m = state.setdefault("m_" + k, np.zeros_like(p[k]))
        v = state.setdefault("v_" + k, np.zeros_like(p[k]))
        m[:] = beta1 * m + (1 - beta1) * g[k]
        v[:] = beta2 * v + (1 - beta2) * g[k] * g[k]
        mh = m / (1 - beta1 ** t)
        vh = v / (1 - beta2 ** t)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)
    return p


PARAM_ORDER = ("W1", "b1", "W2", "b2", "w3", "b3")


def flatten(p):
    """Parameters as one vector in the order W1 (row-major [128][34]), b1, W2, b2, w3, b3."""
    return np.concatenate([np.asarray(p[k], dtype=np.float64).ravel() for k in PARAM_ORDER])


def unflatten(v):
    shapes = {"W1": (NH, NIN), "b1": (NH,), "W2": (NH, NH), "b2": (NH,), "w3": (NH,), "b3": (1,)}
    p, o = {}, 0
    for k in PARAM_ORDER:
        sz = int(np.prod(shapes[k]))
        p[k] = np.array(v[o:o + sz], dtype=np.float64).reshape(shapes[k])
        o += sz
    return p


NPARAM = NH * NIN + NH + NH * NH + NH + NH + 1
