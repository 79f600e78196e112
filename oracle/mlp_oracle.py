"""CPU restatement of the Kastaun bracket predictor -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only to check the library's results.  It restates,
in numpy, what grace_amd/csrc/mlp.hh and mlp.hip compute:

  features()     mlp_features (per-state inputs, FP64 then rounded to FP32)
  init_params()  grmhd_mlp_init_params (SplitMix64, Glorot-uniform)
  forward()      BF16-operand forward pass with FP32-style accumulation
  grad()         loss gradient with the same rounding points as k_mlp_grad

The network has no counterpart in the reference.  Its pins: grad() against
central finite differences of forward() with rounding off (the calculus of
the backward pass), bf16() against torch's round-to-nearest-even bfloat16,
init_params() bit-exact against the library's host routine, and the
closed-form bracket of features() against the oracle's Newton solve for mu+
(x_lo <= mu+ h0 <= x_hi) -- see tests/test_mlp_oracle.py.
"""
import numpy as np

NF, KP, H, NO = 10, 16, 64, 2
W1, B1 = 0, KP * H
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H * NO
NP = B3 + NO
NSCALE = 24


def bf16(x):
    """Round-to-nearest-even to bfloat16, returned as float32 (finite x)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1].reshape(KP, H), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3].reshape(H, NO),
            p[B3:NP])


def _splitmix(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def init_params(seed):
    st = seed & 0xFFFFFFFFFFFFFFFF
    p = np.zeros(NP, dtype=np.float32)

    def uni():
        nonlocal st
        st, z = _splitmix(st)
        return (z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0

    a1, a2, a3 = np.sqrt(6.0 / (NF + H)), np.sqrt(6.0 / (2 * H)), np.sqrt(6.0 / (H + NO))
    for k in range(NF):
        for j in range(H):
            p[W1 + k * H + j] = np.float32(a1 * uni())
    for k in range(H * H):
        p[W2 + k] = np.float32(a2 * uni())
    for k in range(H * NO):
        p[W3 + k] = np.float32(a3 * uni())
    return p


def metric_parts(met):
    g = met[:6]
    gxx, gxy, gxz, gyy, gyz, gzz = g
    det = gxx * gyy * gzz + 2 * gxy * gxz * gyz - gxx * gyz ** 2 - gyy * gxz ** 2 - gzz * gxy ** 2
    inv = np.array([gyy * gzz - gyz ** 2, gxz * gyz - gxy * gzz, gxy * gyz - gxz * gyy,
                    gxx * gzz - gxz ** 2, gxy * gxz - gxx * gyz, gxx * gyy - gxy ** 2]) / det
    return g, inv, np.sqrt(det)


def features(cons, met, h0):
    """cons (10,n) densitised, met (10,n) -> (features float32 (10,n), log x_hi float64 (n))."""
    g, inv, sq = metric_parts(met)
    c = cons / sq
    D = np.maximum(c[0], 1e-300)
    q = c[4] / D
    r = c[1:4] / D
    b = c[7:10] / np.sqrt(D)
    r2 = (inv[0] * r[0] ** 2 + inv[3] * r[1] ** 2 + inv[5] * r[2] ** 2
          + 2 * (inv[1] * r[0] * r[1] + inv[2] * r[0] * r[2] + inv[4] * r[1] * r[2]))
    Bt2 = (g[0] * b[0] ** 2 + g[3] * b[1] ** 2 + g[5] * b[2] ** 2
           + 2 * (g[1] * b[0] * b[1] + g[2] * b[0] * b[2] + g[4] * b[1] * b[2]))
    rB2 = (r * b).sum(0) ** 2
    h02 = h0 * h0
    xlo = h0 / (h02 + r2)
    mu0 = 1.0 / h0
    x = 1.0 / (1.0 + mu0 * Bt2)
    rb = x * (r2 * x + mu0 * (x + 1.0) * rB2)
    xhi = h0 / np.sqrt(h02 + rb)
    qp = np.maximum(q, 0.0)

    def flog(v):   # log of the FP32-rounded, range-clamped argument
        return np.log(np.clip(v, 1e-37, 1e37).astype(np.float32).astype(np.float64))

    f = np.stack([flog(q), flog(r2 + 1e-30), flog(Bt2 + 1e-30),
                  np.minimum(1.0, rB2 / (r2 * Bt2 + 1e-300)), flog(xhi), flog(xlo),
                  np.log1p(qp.astype(np.float32).astype(np.float64)),
                  r2 / (1 + qp) ** 2, Bt2 / (1 + qp), flog(D)])
    f = f.astype(np.float32)
    return f, f[4].astype(np.float64), xlo, xhi


def normalise(feat, scale):
    mean = np.asarray(scale[:NF], dtype=np.float32)[:, None]
    inv = np.asarray(scale[NF:2 * NF], dtype=np.float32)[:, None]
    x = np.zeros((KP, feat.shape[1]), dtype=np.float32)
    x[:NF] = (feat.astype(np.float32) - mean) * inv
    return x


def forward(p, scale, feat, quant=True):
    """Normalised outputs (2,n) and the activations: x (16,n) layer-1 input,
    h1/h2 (64,n) FP32 post-activation values."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    x = q(normalise(feat, scale)).astype(np.float64)
    w1q, w2q, w3q = (q(w).astype(np.float64) for w in (w1, w2, w3))
    z1 = w1q.T @ x + b1[:, None]
    h1 = np.maximum(z1, 0.0)
    z2 = w2q.T @ q(h1).astype(np.float64) + b2[:, None]
    h2 = np.maximum(z2, 0.0)
    y = w3q.T @ q(h2).astype(np.float64) + b3[:, None]
    return y, dict(x=x, h1=h1, h2=h2, w2q=w2q, w3q=w3q)


def denormalise(y, scale):
    return y * np.asarray(scale[22:24], dtype=np.float64)[:, None] + np.asarray(scale[20:22], dtype=np.float64)[:, None]


def grad(p, scale, feat, target, quant=True):
    """Gradient (NP) and loss of (1/n) sum_s sum_j (y - t)^2 over finite targets."""
    n = feat.shape[1]
    y, a = forward(p, scale, feat, quant)
    tn = (target.astype(np.float64) - np.asarray(scale[20:22], dtype=np.float64)[:, None]) \
        / np.asarray(scale[22:24], dtype=np.float64)[:, None]
    ok = np.isfinite(tn)
    d = np.where(ok, y - np.where(ok, tn, 0.0), 0.0)
    loss = (d * d).sum() / n
    d3 = 2.0 * d / n
    q = bf16 if quant else (lambda v: np.asarray(v, dtype=np.float64))
    h1, h2, x = a["h1"], a["h2"], a["x"]
    h1q, h2q = q(h1).astype(np.float64), q(h2).astype(np.float64)
    dz2 = (a["w3q"] @ d3) * (h2 > 0)
    dz1 = (a["w2q"] @ dz2) * (h1 > 0)
    g = np.zeros(NP)
    g[W3:B3] = (h2q @ d3.T).ravel()
    g[B3:NP] = d3.sum(1)
    g[W2:B2] = (h1q @ dz2.T).ravel()
    g[B2:W3] = dz2.sum(1)
    gw1 = x @ dz1.T
    gw1[NF:] = 0.0
    g[W1:B1] = gw1.ravel()
    g[B1:W2] = dz1.sum(1)
    return g, loss
