"""FP64 restatement of the learned Barnes-Hut correction's network.

TEST INFRASTRUCTURE ONLY: the checker for the GPU's BF16-MFMA forward and
backward kernels.  Only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg may import it; the product never does.

The network has no counterpart in the reference (grav_sim stops at the
monopole walk, src/acceleration_barnes_hut.c:78-248); what it computes is
fixed by its definition:

    16 -> 128 -> 128 -> 128 -> 3, GELU (tanh form) after each hidden layer,
    parameters in one flat vector: W1 (16x128) b1 W2 (128x128) b2
    W3 (128x128) b3 W4 (128x3) b4, each matrix row-major (inputs, outputs);
    loss = mean over rows of |Y - T|^2.

``emulate=True`` rounds both operands of every matrix product to bfloat16
(nearest even), as the GPU's matrix-core products do, and keeps everything
else in FP64.

The features are rotation invariant and the correction equivariant: the
network sees the walk's vectors in a per-particle orthonormal frame
(``local_frame``) and predicts the residual's components in that frame.

Pins (tests/test_mlp_oracle.py): gradients against central finite
differences; GELU against its closed form at known points and its derivative
against finite differences; ``bf16`` against torch's float32 -> bfloat16
conversion; ``local_frame`` orthonormal, right-handed, spanning (a, q);
``features16`` unchanged by a rotation of its inputs and reassembling |q|,
|v|; ``apply_correction`` rotating with its inputs.  The walk features it
consumes are pinned in the same file (quadrupole-walk identity, interaction
counts, a single accepted node in closed form).
"""
import numpy as np

F, H, O = 16, 128, 3
SHAPES = [("W1", (F, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
          ("W4", (H, O)), ("b4", (O,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
C_GELU = 0.7978845608028654  # sqrt(2 / pi)


def unpack(w):
    out, k = {}, 0
    for name, shp in SHAPES:
        sz = int(np.prod(shp))
        out[name] = np.asarray(w[k:k + sz], dtype=np.float64).reshape(shp)
        k += sz
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16(a):
    """float32 -> bfloat16 -> float64, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3 * 0.044715 * z * z)


def _mm(a, b, emulate):
    if emulate:
        return bf16(a) @ bf16(b)
    return a @ b


def forward(w, X, emulate=False):
    """Y (rows, 3) and the activations needed by the backward pass."""
    p = unpack(w)
    X = np.asarray(X, dtype=np.float64)
    Z1 = _mm(X, p["W1"], emulate) + p["b1"]
    H1 = gelu(Z1)
    Z2 = _mm(H1, p["W2"], emulate) + p["b2"]
    H2 = gelu(Z2)
    Z3 = _mm(H2, p["W3"], emulate) + p["b3"]
    H3 = gelu(Z3)
    Y = _mm(H3, p["W4"], emulate) + p["b4"]
    return Y, dict(X=X, Z1=Z1, H1=H1, Z2=Z2, H2=H2, Z3=Z3, H3=H3)


def loss(w, X, T, emulate=False):
    Y, _ = forward(w, X, emulate)
    return float(np.sum((Y - T) ** 2) / X.shape[0])


def gradients(w, X, T, emulate=False):
    """(flat gradient, loss) of mean over rows of |Y - T|^2."""
    p = unpack(w)
    Y, c = forward(w, X, emulate)
    B = X.shape[0]
    dY = 2.0 * (Y - T) / B
    g = {}
    g["W4"] = _mm(c["H3"].T, dY, emulate)
    g["b4"] = dY.sum(axis=0)
    dH3 = _mm(dY, p["W4"].T, emulate) * gelu_grad(c["Z3"])
    g["W3"] = _mm(c["H2"].T, dH3, emulate)
    g["b3"] = dH3.sum(axis=0)
    dH2 = _mm(dH3, p["W3"].T, emulate) * gelu_grad(c["Z2"])
    g["W2"] = _mm(c["H1"].T, dH2, emulate)
    g["b2"] = dH2.sum(axis=0)
    dH1 = _mm(dH2, p["W2"].T, emulate) * gelu_grad(c["Z1"])
    g["W1"] = _mm(c["X"].T, dH1, emulate)
    g["b1"] = dH1.sum(axis=0)
    return pack(g), float(np.sum((Y - T) ** 2) / B)


def random_params(seed, scale=1.0):
    """He-uniform weights and small random biases (test networks)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / shp[0]) * scale
            p[name] = rng.uniform(-lim, lim, size=shp)
        else:
            p[name] = rng.uniform(-0.1, 0.1, size=shp)
    return pack(p).astype(np.float32).astype(np.float64)


def local_frame(a, q, v):
    """Per-row orthonormal frames (e1, e2, e3), each (N, 3): e1 = a/|a|;
    e2 = the part of q orthogonal to e1, normalised (if that part is below
    1e-6 |q|: the same from v; then from the coordinate axis least aligned
    with e1); e3 = e1 x e2."""
    a, q, v = (np.asarray(t, dtype=np.float64) for t in (a, q, v))
    an = np.linalg.norm(a, axis=1)
    e1 = np.where(an[:, None] > 0, a / np.where(an > 0, an, 1.0)[:, None], np.array([1.0, 0.0, 0.0]))
    e2 = np.zeros_like(e1)
    done = np.zeros(len(a), dtype=bool)
    for w in (q, v):
        wn = np.linalg.norm(w, axis=1)
        u = w - np.sum(w * e1, axis=1)[:, None] * e1
        un = np.linalg.norm(u, axis=1)
        ok = ~done & (un > 1e-6 * wn) & (un > 0)
        e2[ok] = u[ok] / un[ok, None]
        done |= ok
    for i in np.flatnonzero(~done):
        k = int(np.argmin(np.abs(e1[i])))
        u = np.eye(3)[k] - e1[i, k] * e1[i]
        e2[i] = u / np.linalg.norm(u)
    e3 = np.cross(e1, e2)
    return e1, e2, e3


def features16(a, wf, G):
    """The 16 raw network inputs from coarse accelerations a (N, 3) and the
    walk features wf (N, 16) (oracle.bh_walk(..., features=True, oct=...)),
    all rotation invariant, in the frame of local_frame(a, q, v):
    0 G q.e1/|a|, 1 G q.e2/|a|, 2-4 G v.e1, v.e2, v.e3 /|a|, 5 G s2/|a|,
    6 G s3/|a|, 7 log10|a|, 8 log(1 + accepted nodes), 9 log(1 + particle
    interactions), 10 log10 max(G s3/|a|, 1e-12), 11 the same of s2,
    12 the same of |q|, 13-15 the octupole sum o.e1, o.e2, o.e3 over s3 (0 if
    s3 = 0)."""
    a = np.asarray(a, dtype=np.float64)
    wf = np.asarray(wf, dtype=np.float64)
    q, v = wf[:, 0:3], wf[:, 3:6]
    e1, e2, e3 = local_frame(a, q, v)
    an = np.linalg.norm(a, axis=1)
    g = np.where(an > 0, G / np.where(an > 0, an, 1.0), 0.0)
    f = np.zeros((len(a), 16))
    f[:, 0] = g * np.sum(q * e1, axis=1)
    f[:, 1] = g * np.sum(q * e2, axis=1)
    f[:, 2] = g * np.sum(v * e1, axis=1)
    f[:, 3] = g * np.sum(v * e2, axis=1)
    f[:, 4] = g * np.sum(v * e3, axis=1)
    s2, s3 = g * wf[:, 6], g * wf[:, 7]
    f[:, 5], f[:, 6] = s2, s3
    f[:, 7] = np.where(an > 0, np.log10(np.where(an > 0, an, 1.0)), -30.0)
    f[:, 8] = np.log1p(wf[:, 8])
    f[:, 9] = np.log1p(wf[:, 9])
    f[:, 10] = np.log10(np.maximum(s3, 1e-12))
    f[:, 11] = np.log10(np.maximum(s2, 1e-12))
    f[:, 12] = np.log10(np.maximum(g * np.linalg.norm(q, axis=1), 1e-12))
    o = wf[:, 10:13]
    is3 = np.where(wf[:, 7] > 0, 1.0 / np.where(wf[:, 7] > 0, wf[:, 7], 1.0), 0.0)
    f[:, 13] = np.sum(o * e1, axis=1) * is3
    f[:, 14] = np.sum(o * e2, axis=1) * is3
    f[:, 15] = np.sum(o * e3, axis=1) * is3
    return f


def apply_correction(a, wf, Y, G, tscale):
    """a + G (q + o) + G s3 tscale (Y0 e1 + Y1 e2 + Y2 e3): the quadrupole and
    octupole sums exactly, the network's output in the local frame."""
    a = np.asarray(a, dtype=np.float64)
    wf = np.asarray(wf, dtype=np.float64)
    e1, e2, e3 = local_frame(a, wf[:, 0:3], wf[:, 3:6])
    Y = np.asarray(Y, dtype=np.float64)
    s = (G * wf[:, 7] * tscale)[:, None]
    return a + G * (wf[:, 0:3] + wf[:, 10:13]) + s * (Y[:, 0:1] * e1 + Y[:, 1:2] * e2 + Y[:, 2:3] * e3)
