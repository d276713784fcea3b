"""TEST INFRASTRUCTURE ONLY -- FP64 CPU restatement of the learned surrogate.

The surrogate has no counterpart in the reference: it replaces the
Gauss-Hermite averages of src/ff_dmd_I.cpp:180-242 (rho_bar, phi_bar and their
r- and alpha-derivatives for one pair) by one MLP per ordered type pair,
3 -> 64 -> 64 -> 64 -> 6 with SiLU, standardised inputs and outputs.  This
module restates the network's forward pass, its mean-squared loss and the
loss gradient by explicit back-propagation, in float64 numpy, so that tests can
check the GPU training kernel's gradient and the inference path's outputs.

Parameter layout of one network (flat float32 vector of 8966):
W1[3][64], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[64][6], b4[6];
W[k][o] multiplies input feature k into output o.
Normalisation vector (56 floats): r_lo, r_hi, a_lo, a_hi, r_c, r_s, a_c, a_s,
y_mean[4][6], y_std[4][6].

Pinned by: a central finite-difference check of the gradient
(tests/test_surrogate_oracle.py) and by the closed form of a network whose
hidden layers are zero (output = b4, gradient = 2 (b4 - t) / count).
"""
import numpy as np

HID, NIN, NOUT = 64, 3, 6
SHAPES = [("W1", (NIN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("W4", (HID, NOUT)), ("b4", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
NNORM = 8 + 2 * 4 * NOUT


def unpack(flat):
    """flat parameter vector of one network -> dict of float64 arrays"""
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def norm_parts(norm):
    norm = np.asarray(norm, dtype=np.float64)
    r_c, r_s, a_c, a_s = norm[4:8]
    y_mean = norm[8:8 + 4 * NOUT].reshape(4, NOUT)
    y_std = norm[8 + 4 * NOUT:8 + 8 * NOUT].reshape(4, NOUT)
    return r_c, r_s, a_c, a_s, y_mean, y_std


def inputs(norm, X):
    r_c, r_s, a_c, a_s, _, _ = norm_parts(norm)
    X = np.asarray(X, dtype=np.float64)
    return np.stack([(X[:, 0] - r_c) / r_s, (X[:, 1] - a_c) / a_s, (X[:, 2] - a_c) / a_s], axis=1)


def forward(flat, norm, tp, X):
    """de-standardised outputs [n][6] of network tp for samples X [n][3]"""
    p = unpack(flat)
    _, _, _, _, y_mean, y_std = norm_parts(norm)
    h = inputs(norm, X)
    for l in (1, 2, 3):
        h = silu(h @ p[f"W{l}"] + p[f"b{l}"])
    y = h @ p["W4"] + p["b4"]
    return y * y_std[tp] + y_mean[tp]


def loss_and_grad(flat, norm, tp, X, T, inv_count):
    """sum of squared standardised errors and the gradient of
    inv_count * sum (y - t)^2 with respect to the flat parameters"""
    p = unpack(flat)
    _, _, _, _, y_mean, y_std = norm_parts(norm)
    h0 = inputs(norm, X)
    zs, hs = [], [h0]
    h = h0
    for l in (1, 2, 3):
        z = h @ p[f"W{l}"] + p[f"b{l}"]
        zs.append(z)
        h = silu(z)
        hs.append(h)
    y = h @ p["W4"] + p["b4"]
    t = (np.asarray(T, dtype=np.float64) - y_mean[tp]) / y_std[tp]
    d = y - t
    dy = 2.0 * d * inv_count
    g = {"W4": hs[3].T @ dy, "b4": dy.sum(0)}
    dz = (dy @ p["W4"].T) * dsilu(zs[2])
    for l in (3, 2, 1):
        g[f"W{l}"] = hs[l - 1].T @ dz
        g[f"b{l}"] = dz.sum(0)
        if l > 1:
            dz = (dz @ p[f"W{l}"].T) * dsilu(zs[l - 2])
    return float((d * d).sum()), pack(g)


def glorot(rng):
    """random parameters of one network (for tests): Glorot-uniform weights, small biases"""
    d = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shp[0] + shp[1]))
            d[name] = rng.uniform(-lim, lim, size=shp)
        else:
            d[name] = rng.uniform(-0.1, 0.1, size=shp)
    return pack(d)
