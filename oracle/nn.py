"""NumPy restatement of the learned shading surrogate -- TEST INFRASTRUCTURE ONLY.

Imported by tests/ and bench.py's checks, always as the reference of a comparison; the product
never imports it.  It restates, in float64 unless stated otherwise:

  features(n, V, L, r, m, albedo)   the 12 network inputs
  forward / backward                the 12-64-64-64-4 ReLU MLP and the
                                    training loss, gradients in the flat
                                    parameter layout the device kernels use
  adam                              one Adam update
  bf16 / infer_bf16                 round-to-nearest-even bfloat16 and the
                                    inference arithmetic (bfloat16 operands,
                                    float accumulation)

Pinned (tests/test_nn_oracle.py) by central finite differences of the loss,
the first Adam step's closed form (-lr * sign(g)), torch's bfloat16 rounding,
and the float64 forward.
"""
import numpy as np

SIZES = (12, 64, 64, 64, 4)
NOUT = 3            # colour channels; output 3 is the error estimate
HUBER_BETA = 0.01   # smooth-L1 transition of the colour loss
ERR_WEIGHT = 1.0    # weight of the error-head term


def layout():
    """[(name, shape, offset)] of the flat parameter vector: W1 [64][12], b1,
    W2 [64][64], b2, W3 [64][64], b3, W4 [4][64], b4 -- row-major, in order."""
    out, off = [], 0
    for i in range(4):
        o, k = SIZES[i + 1], SIZES[i]
        out.append((f'W{i + 1}', (o, k), off)); off += o * k
        out.append((f'b{i + 1}', (o,), off)); off += o
    return out


NPARAMS = sum(int(np.prod(s)) for _, s, _ in layout())


def unpack(p):
    return {name: p[off:off + int(np.prod(shape))].reshape(shape) for name, shape, off in layout()}


def init_params(seed):
    """He-uniform weights, zero biases (the device initialiser is given these)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS)
    for name, shape, off in layout():
        if name[0] == 'W':
            lim = np.sqrt(6.0 / shape[1])
            p[off:off + shape[0] * shape[1]] = rng.uniform(-lim, lim, shape[0] * shape[1])
    return p


def features(n, V, L, r, m, alb, dtype=np.float32):
    """The 12 inputs per sample: normal n (3), the view direction as its mirror
    image about n, R = reflect(-V, n) = 2 (n.V) n - V (3), the light direction as
    its cosine to R (1), log(roughness), metallic, albedo (3)."""
    T = dtype
    n, V, L = (np.asarray(a, T) for a in (n, V, L))
    NoV = ((n[:, 0] * V[:, 0]) + (n[:, 1] * V[:, 1])) + (n[:, 2] * V[:, 2])
    R = (T(2) * NoV)[:, None] * n - V
    L = np.broadcast_to(L, R.shape)
    RoL = ((R[:, 0] * L[:, 0]) + (R[:, 1] * L[:, 1])) + (R[:, 2] * L[:, 2])
    return np.concatenate([n, R, RoL[:, None], np.log(np.asarray(r, T))[:, None], np.asarray(m, T)[:, None],
                           np.asarray(alb, T)], axis=1).astype(T)


def forward(p, X):
    """(out [B][4], activations) of the float64 network."""
    P = unpack(np.asarray(p, np.float64))
    h = [np.asarray(X, np.float64)]
    for i in range(1, 4):
        h.append(np.maximum(0.0, h[-1] @ P[f'W{i}'].T + P[f'b{i}']))
    return h[-1] @ P['W4'].T + P['b4'], h


def huber_grad(d, beta=HUBER_BETA):
    return np.where(np.abs(d) < beta, d / beta, np.sign(d))


def loss(p, X, Y, part='all'):
    """Batch loss: mean over samples of mean_c smoothL1(o_c - y_c) ('colour')
    plus ERR_WEIGHT * 0.5 * (o_3 - max_c |o_c - y_c|)^2 ('error')."""
    o, _ = forward(p, X)
    d = o[:, :NOUT] - Y[:, :NOUT]
    a = np.abs(d)
    hub = np.where(a < HUBER_BETA, 0.5 * d * d / HUBER_BETA, a - 0.5 * HUBER_BETA).mean(axis=1)
    err = ERR_WEIGHT * 0.5 * (o[:, NOUT] - a.max(axis=1)) ** 2
    return float({'all': hub + err, 'colour': hub, 'error': err}[part].mean())


def backward(p, X, Y):
    """Gradient of loss() in the flat layout.  The error head's term reaches
    only W4[3] and b4[3]: the trunk is trained by the colour loss alone."""
    P = unpack(np.asarray(p, np.float64))
    o, h = forward(p, X)
    B = X.shape[0]
    d = o[:, :NOUT] - Y[:, :NOUT]
    e = np.abs(d).max(axis=1)
    go = np.zeros_like(o)
    go[:, :NOUT] = huber_grad(d) / (NOUT * B)
    go[:, NOUT] = ERR_WEIGHT * (o[:, NOUT] - e) / B
    g = {'W4': go.T @ h[3], 'b4': go.sum(axis=0)}
    delta = (go[:, :NOUT] @ P['W4'][:NOUT]) * (h[3] > 0)
    for i in (3, 2, 1):
        g[f'W{i}'] = delta.T @ h[i - 1]
        g[f'b{i}'] = delta.sum(axis=0)
        if i > 1:
            delta = (delta @ P[f'W{i}']) * (h[i - 1] > 0)
    out = np.zeros(NPARAMS)
    for name, shape, off in layout():
        out[off:off + int(np.prod(shape))] = g[name].reshape(-1)
    return out


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counts from 1); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(np.asarray(x, np.float32))
    return np.where(nan, np.float32(np.nan), out)


def infer_bf16(p, X):
    """Inference as the MFMA kernel computes it: bfloat16 inputs, weights and
    hidden activations, float32 biases, float64 accumulation here (the device
    accumulates in float32 in its own order: compare with a tolerance).
    Returns [B][4]: max(0, colour) and the error estimate."""
    P = unpack(np.asarray(p, np.float32))
    h = bf16(X).astype(np.float64)
    for i in range(1, 4):
        h = np.maximum(0.0, h @ bf16(P[f'W{i}']).astype(np.float64).T + P[f'b{i}'])
        h = bf16(h.astype(np.float32)).astype(np.float64)
    o = h @ bf16(P['W4']).astype(np.float64).T + P['b4']
    o[:, :NOUT] = np.maximum(0.0, o[:, :NOUT])
    return o
