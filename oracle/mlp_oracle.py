"""CPU oracle of the learned reinitialisation network (test infrastructure only).

The network has no counterpart in the reference; this file states what the
HIP kernels compute so the tests can check them:

* features  -- the 5x5x5 stencil (phi(nb) - phi(c)) / h, x offset fastest,
  indices clamped to the interior (the zero-gradient ghost rule of
  branches.f90:275-320), then phi(c) / h, then |grad phi| - 1 from
  fourth-order central differences, then (phi(c) / h)(1 / max(|grad phi|,
  0.1) - 1);
* forward   -- split-BF16 products: each FP32 operand v is hi = bf16(v) plus
  lo = bf16(v - hi) and a product is hi*hi + lo*hi + hi*lo accumulated in
  FP32; z1 = x W1 + b1, h1 = silu(z1), z2 = h1 W2 + b2, o = silu(z2) . w3 + b3;
* backward  -- the gradient of L = mean (o - y)^2 with the roundings the
  kernels apply: the weight gradients multiply the hi parts of x and h1 by
  dZ rounded to BF16, and dH1 = bf16(dZ2) hi(W2)^T;
* adam      -- beta 0.9 / 0.999, eps 1e-8, bias-corrected;
* hash32 / minibatch -- the sample draw of a training step;
* band      -- ordered |phi| < thr cell list.

`emulate=False` runs the same algebra in float64 with no rounding: that form
is pinned against central finite differences in the tests; the BF16 form is
pinned against torch's bfloat16 conversion.
"""
import numpy as np

NF, NH = 128, 64
OFF = {"W1": 0, "b1": NF * NH, "W2": NF * NH + NH, "b2": NF * NH + NH + NH * NH,
       "w3": NF * NH + 2 * NH + NH * NH, "b3": NF * NH + 3 * NH + NH * NH}
NPARAM = OFF["b3"] + 1


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out.reshape(np.shape(x))


def unpack(P):
    P = np.asarray(P)
    return (P[OFF["W1"]:OFF["b1"]].reshape(NF, NH), P[OFF["b1"]:OFF["W2"]],
            P[OFF["W2"]:OFF["b2"]].reshape(NH, NH), P[OFF["b2"]:OFF["w3"]],
            P[OFF["w3"]:OFF["b3"]], P[OFF["b3"]])


def silu(z):
    return z / (1 + np.exp(-z))


def dsilu(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def split(v):
    """(hi, lo) BF16 parts of float32 values, as float64 arrays."""
    v = np.asarray(v, dtype=np.float32)
    hi = bf16(v)
    lo = bf16(v - hi)
    return hi.astype(np.float64), lo.astype(np.float64)


def split_matmul(a, b):
    """hi(a) hi(b) + lo(a) hi(b) + hi(a) lo(b), summed exactly (float64)."""
    ah, al = split(a)
    bh, bl = split(b)
    return ah @ bh + al @ bh + ah @ bl


def forward(P, X, emulate=True):
    """Network output o (n,) and the cache for backward."""
    W1, b1, W2, b2, w3, b3 = unpack(P)
    if emulate:
        f32 = np.float32
        Xb = bf16(np.asarray(X, f32))
        W2b = bf16(np.asarray(W2, f32))
        z1 = split_matmul(X, W1).astype(f32) + b1.astype(f32)
        h1f = silu(z1).astype(f32)
        h1 = bf16(h1f)
        z2 = split_matmul(h1f, W2).astype(f32) + b2.astype(f32)
        h2 = silu(z2).astype(f32)
        o = (h2.astype(np.float64) @ w3.astype(np.float64)).astype(f32) + f32(b3)
    else:
        Xb, W1b, W2b = (np.asarray(a, dtype=np.float64) for a in (X, W1, W2))
        z1 = Xb @ W1b + b1
        h1 = silu(z1)
        z2 = h1 @ W2b + b2
        h2 = silu(z2)
        o = h2 @ np.asarray(w3, np.float64) + b3
    return o, (Xb, W2b, z1, h1, z2, h2, w3)


def loss_grad(P, X, Y, emulate=True, weight_c=0.0):
    """L = mean w (o - y)^2, w = c^2 / (y^2 + c^2) (1 for c = 0), and dL/dP
    (flat, float64)."""
    n = X.shape[0]
    o, (Xb, W2b, z1, h1, z2, h2, w3) = forward(P, X, emulate)
    y = np.asarray(Y, np.float64)
    e = o.astype(np.float64) - y
    w = weight_c ** 2 / (y * y + weight_c ** 2) if weight_c > 0 else np.ones_like(y)
    L = float(np.mean(w * e * e))
    dO = 2.0 * w * e / n
    g = np.zeros(NPARAM)
    g[OFF["w3"]:OFF["b3"]] = h2.astype(np.float64).T @ dO
    g[OFF["b3"]] = dO.sum()
    dz2 = dO[:, None] * np.asarray(w3, np.float64)[None, :] * dsilu(z2.astype(np.float64))
    g[OFF["b2"]:OFF["w3"]] = dz2.sum(0)
    dz2m = bf16(dz2.astype(np.float32)).astype(np.float64) if emulate else dz2
    g[OFF["W2"]:OFF["b2"]] = (h1.astype(np.float64).T @ dz2m).ravel()
    dh1 = dz2m @ W2b.astype(np.float64).T
    if emulate:
        dh1 = dh1.astype(np.float32).astype(np.float64)
    dz1 = dh1 * dsilu(z1.astype(np.float64))
    g[OFF["b1"]:OFF["W2"]] = dz1.sum(0)
    dz1m = bf16(dz1.astype(np.float32)).astype(np.float64) if emulate else dz1
    g[OFF["W1"]:OFF["b1"]] = (Xb.astype(np.float64).T @ dz1m).ravel()
    return L, g


def adam(P, m, v, g, lr, step):
    """One Adam step; returns new (P, m, v) in float32 like the kernel."""
    f32 = np.float32
    g = np.asarray(g, f32)
    m = f32(0.9) * m + f32(0.1) * g
    v = f32(0.999) * v + f32(0.001) * g * g
    bc1 = f32(1) - f32(0.9) ** step
    bc2 = f32(1) - f32(0.999) ** step
    P = P - f32(lr) * (m / bc1) / (np.sqrt(v / bc2) + f32(1e-8))
    return P.astype(f32), m.astype(f32), v.astype(f32)


def hash32(x):
    x = np.asarray(x, dtype=np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def minibatch(seed_step, nsamp, batch):
    """Sample indices of one training step (seed_step = the kernel's seed)."""
    pos = np.arange(batch, dtype=np.uint64)
    key = ((np.uint64(seed_step) * np.uint64(0x9E3779B9)) + pos) & 0xFFFFFFFF
    return (hash32(key) % np.uint64(nsamp)).astype(np.int64)


def step_seed(seed, step):
    return (seed * 2654435761 + step) & 0xFFFFFFFF


def band(phi, thr):
    """Ordered flat indices of the interior cells with |phi| < thr (phi: (nz, ny, nx))."""
    return np.flatnonzero(np.abs(phi.ravel()) < thr)


def features(phi, cells, inv_h, inv_12h=None):
    """Rows of the network input for the flat cell indices (phi: (nz, ny, nx)).
    inv_h = float32(1/h); inv_12h = float32(1/(12 h)) (default: from inv_h)."""
    nz, ny, nx = phi.shape
    if inv_12h is None:
        inv_12h = np.float32(1.0 / (12.0 / float(inv_h)))
    k, r = np.divmod(cells, nx * ny)
    j, i = np.divmod(r, nx)
    X = np.zeros((cells.size, NF), dtype=np.float32)
    pc = phi.ravel()[cells]

    def at(di, dj, dk):
        return phi[np.clip(k + dk, 0, nz - 1), np.clip(j + dj, 0, ny - 1), np.clip(i + di, 0, nx - 1)]

    q = 0
    for dk in range(-2, 3):
        for dj in range(-2, 3):
            for di in range(-2, 3):
                X[:, q] = (at(di, dj, dk) - pc).astype(np.float32) * np.float32(inv_h)
                q += 1
    X[:, 125] = pc.astype(np.float32) * np.float32(inv_h)
    g2 = np.zeros(cells.size, dtype=np.float32)
    for e in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
        m2, m1 = at(*(-2 * np.array(e))), at(*(-np.array(e)))
        p1, p2 = at(*e), at(*(2 * np.array(e)))
        d = ((m2 - p2) + phi.dtype.type(8) * (p1 - m1)).astype(np.float32) * np.float32(inv_12h)
        g2 += d * d
    gm = np.sqrt(g2)
    X[:, 126] = gm - np.float32(1)
    X[:, 127] = X[:, 125] * (np.float32(1) / np.maximum(gm, np.float32(0.1)) - np.float32(1))
    return X


def apply(P, phi, h, band_cells, ls_wid):
    """phi after the network pass on |phi| < band_cells h (h: the FP64 spacing;
    the threshold and h are rounded to phi's type, 1/h to FP32): the
    correction h o is scaled by clamp((|phi| - ls_wid) / ls_wid, 0, 1)."""
    h = float(h)
    cells = band(phi, np.asarray(band_cells * h, dtype=phi.dtype))
    X = features(phi, cells, np.float32(1.0 / h), np.float32(1.0 / (12.0 * h)))
    o, _ = forward(P, X)
    t = phi.dtype.type
    pc = phi.ravel()[cells]
    w = t(ls_wid)
    gate = np.minimum(np.maximum((np.abs(pc) - w) / w, t(0)), t(1))
    out = phi.copy().ravel()
    out[cells] = pc + t(h) * o.astype(phi.dtype) * gate
    return out.reshape(phi.shape)
