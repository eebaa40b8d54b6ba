"""CPU restatement of the learned emission proposal (numpy, float64).

TEST INFRASTRUCTURE ONLY: the checker for the proposal kernels of the HIP
path; never imported by the product.  The proposal has no reference
counterpart: it replaces the cosine emission of sources/sampler.cc:17-32.

Function            restates                                  pinned by
------------------  ----------------------------------------  ------------------------------------------
encode              32-feature Fourier encoding of (u, v)     closed form at u = 0, 1/2
forward / softmax   32 -> 64 -> 64 -> 64 -> 1024 ReLU MLP     softmax rows sum to 1; zero output layer
                                                              gives the uniform 1/1024
loss_and_grad       contribution-weighted cross entropy and   central finite differences (float64)
                    its analytic backward
adam                Adam (Kingma & Ba 2015), bias-corrected   first step moves by lr * sign(g)
cdf_table           sequential FP32 inclusive CDF per anchor  monotone, ends at 1, differences = rows
used_deposits /     deposits some hit point's gather accepts, brute force over a k-d tree (scipy)
  contributions     per-photon FP32 sums of their mean flux
"""
import numpy as np

IN, HID, OUT = 32, 64, 1024
LAYOUT = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("W4", (OUT, HID)), ("b4", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in LAYOUT)


def unpack(w):
    out, o = {}, 0
    for name, shape in LAYOUT:
        k = int(np.prod(shape))
        out[name] = np.asarray(w[o:o + k], np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in LAYOUT])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def encode(uv):
    """Features 2q, 2q+1 = sin, cos(pi 2^q u) for q = 0..7, then the same of v."""
    uv = np.asarray(uv, np.float64).reshape(-1, 2)
    cols = []
    for c in range(2):
        for q in range(8):
            a = np.pi * (2.0 ** q) * uv[:, c]
            cols += [np.sin(a), np.cos(a)]
    return np.stack(cols, axis=1)


def _layers(w, uv, emulate):
    p = unpack(w)
    r = bf16 if emulate else (lambda x: np.asarray(x, np.float64))
    X = r(encode(uv))
    Ws = {k: r(p[k]) for k in ("W1", "W2", "W3", "W4")}
    A1 = X @ Ws["W1"].T + p["b1"]
    H1 = r(np.maximum(A1, 0.0))
    A2 = H1 @ Ws["W2"].T + p["b2"]
    H2 = r(np.maximum(A2, 0.0))
    A3 = H2 @ Ws["W3"].T + p["b3"]
    H3 = r(np.maximum(A3, 0.0))
    Z = H3 @ Ws["W4"].T + p["b4"]
    return X, Ws, (A1, H1), (A2, H2), (A3, H3), Z


def softmax(Z):
    m = Z.max(axis=1, keepdims=True)
    e = np.exp(Z - m)
    return e / e.sum(axis=1, keepdims=True)


def forward(w, uv, emulate=True):
    """Cell probabilities [n][1024] at light-sample coordinates uv [n][2]."""
    return softmax(_layers(w, uv, emulate)[-1])


def loss_and_grad(w, rec, emulate=True):
    """rec [n][4] = (u, v, cell, c).  Loss = -sum_s c_s log q(cell_s | uv_s) / sum c;
    returns (loss, gradient over the packed parameters)."""
    rec = np.asarray(rec, np.float64)
    uv, cell, c = rec[:, :2], rec[:, 2].astype(np.int64), rec[:, 3]
    wn = c / c.sum()
    r = bf16 if emulate else (lambda x: np.asarray(x, np.float64))
    X, Ws, (A1, H1), (A2, H2), (A3, H3), Z = _layers(w, uv, emulate)
    P = softmax(Z)
    n = len(rec)
    loss = -np.sum(wn * np.log(P[np.arange(n), cell]))
    dZ = P.copy()
    dZ[np.arange(n), cell] -= 1.0
    dZ = r(dZ * wn[:, None])
    g = {}
    g["W4"] = dZ.T @ H3
    g["b4"] = dZ.sum(axis=0)
    dA3 = r((dZ @ Ws["W4"]) * (H3 > 0))
    g["W3"] = dA3.T @ H2
    g["b3"] = dA3.sum(axis=0)
    dA2 = r((dA3 @ Ws["W3"]) * (H2 > 0))
    g["W2"] = dA2.T @ H1
    g["b2"] = dA2.sum(axis=0)
    dA1 = r((dA2 @ Ws["W2"]) * (H1 > 0))
    g["W1"] = dA1.T @ X
    g["b1"] = dA1.sum(axis=0)
    return loss, pack(g)


def loss_only(w, rec):
    rec = np.asarray(rec, np.float64)
    P = forward(w, rec[:, :2], emulate=False)
    wn = rec[:, 3] / rec[:, 3].sum()
    return -np.sum(wn * np.log(P[np.arange(len(rec)), rec[:, 2].astype(np.int64)]))


def adam(w, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in float32, as the device applies it."""
    f = np.float32
    w, m, v, g = (np.asarray(a, np.float32) for a in (w, m, v, g))
    m = f(b1) * m + f(1.0 - b1) * g
    v = f(b2) * v + f(1.0 - b2) * g * g
    c1 = f(1.0 / (1.0 - b1 ** t))
    c2 = f(1.0 / (1.0 - b2 ** t))
    w = w - f(lr) * (m * c1) / (np.sqrt(v * c2) + f(eps))
    return w, m, v


def cdf_table(prob):
    """Inclusive CDF rows summed sequentially in float32, divided by the total,
    last entry 1 (an all-zero row becomes uniform)."""
    prob = np.maximum(np.asarray(prob, np.float32), np.float32(0.0))
    out = np.empty_like(prob)
    for a in range(prob.shape[0]):
        s = np.cumsum(prob[a], dtype=np.float32)
        tot = s[-1]
        out[a] = s / tot if tot > 0 else np.arange(1, 1025, dtype=np.float32) / np.float32(1024)
        out[a, -1] = 1.0
    return out


def used_deposits(ph_pos, ph_nrm, hit_pos, hit_nrm, radius):
    """Deposits that some hit point's gather accepts: |x - p| <= r and
    n_p . n_x > 0 (the tests of sources/ppm_probabilistic.cc:300-334 in the
    fixed-radius form), by a k-d tree over the hit points."""
    from scipy.spatial import cKDTree
    hp = np.asarray(hit_pos, np.float64)
    hn = np.asarray(hit_nrm, np.float64)
    pp = np.asarray(ph_pos, np.float64)
    pn = np.asarray(ph_nrm, np.float64)
    used = np.zeros(len(pp), bool)
    if len(hp) == 0 or len(pp) == 0:
        return used
    lists = cKDTree(hp).query_ball_point(pp, radius)
    for i, lst in enumerate(lists):
        if lst:
            used[i] = bool(np.any(hn[lst] @ pn[i] > 0.0))
    return used


def contributions(ph_flux, ph_key, used):
    """Per emitted photon with a positive value: (photon index, c), c = FP32
    sum in bounce order of mean(flux) over its used deposits."""
    f = np.asarray(ph_flux, np.float32)
    val = ((f[:, 0] + f[:, 1]) + f[:, 2]) * np.float32(1.0 / 3.0)
    val = np.where(used, val, np.float32(0.0)).astype(np.float32)
    photon = np.asarray(ph_key, np.uint32) >> 6
    order = np.argsort(np.asarray(ph_key, np.uint32), kind="stable")
    out_p, out_c = [], []
    cur, acc = None, np.float32(0.0)
    for i in order:
        if photon[i] != cur:
            if cur is not None and acc > 0:
                out_p.append(cur)
                out_c.append(acc)
            cur, acc = photon[i], np.float32(0.0)
        acc = np.float32(acc + val[i])
    if cur is not None and acc > 0:
        out_p.append(cur)
        out_c.append(acc)
    return np.array(out_p, np.int64), np.array(out_c, np.float32)
