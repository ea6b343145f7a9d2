"""CPU oracle of the surrogate ensemble (TEST INFRASTRUCTURE ONLY).

A numpy restatement of the network, loss, back-propagation and Adam that the
BF16 MFMA kernels implement (include/sph_surrogate.h), used only by tests/ to
check them.  The reference has no learned part; this oracle is pinned by
central finite differences of its own FP64 loss and by Adam's closed form
(tests/test_surrogate_oracle.py).

With bf16=True the oracle rounds where the kernels round - weights, inputs,
hidden activations, the output gradient and the hidden-layer gradients to
bfloat16 (round to nearest even) - and accumulates in float64, so the kernels
differ from it only by their float32 accumulation order.
"""
import numpy as np

N_IN, N_HID, N_OUT = 4, 256, 3
IN_PAD, OUT_PAD = 64, 64
ROWS = [N_HID, N_HID, N_HID, OUT_PAD]
COLS = [IN_PAD, N_HID, N_HID, N_HID]


def layout():
    """(w_offsets[4], b_offsets[4], per_member) of the flat parameter vector."""
    off, w = 0, []
    for r, c in zip(ROWS, COLS):
        w.append(off)
        off += r * c
    b = []
    for r in ROWS:
        b.append(off)
        off += r
    return w, b, off


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, members):
    """List over members of (W[4], b[4]) as float64 arrays (padded shapes)."""
    w_off, b_off, per = layout()
    p = np.asarray(params, np.float64).reshape(members, per)
    out = []
    for e in range(members):
        W = [p[e, w_off[l]:w_off[l] + ROWS[l] * COLS[l]].reshape(ROWS[l], COLS[l])
             for l in range(4)]
        b = [p[e, b_off[l]:b_off[l] + ROWS[l]] for l in range(4)]
        out.append((W, b))
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _r(a, on):
    return bf16(a) if on else np.asarray(a, np.float64)


def forward_member(W, b, x, bf=True):
    """x [B, 4] -> (y [B, 3], cache)."""
    B = x.shape[0]
    h = np.zeros((B, IN_PAD))
    h[:, :N_IN] = x
    hs, zs = [_r(h, bf)], []
    for l in range(4):
        z = hs[-1] @ _r(W[l], bf).T + b[l]
        zs.append(z)
        if l < 3:
            hs.append(_r(silu(z), bf))
    return zs[3][:, :N_OUT], (hs, zs)


def grads_member(W, b, x, t, mask, bf=True):
    """Loss and gradients (dW[4], db[4]) of one member on one batch."""
    y, (hs, zs) = forward_member(W, b, x, bf)
    n = max(1.0, float(mask.sum()))
    d = y - t
    loss = float((mask * d * d).sum() / n)
    dz = np.zeros((x.shape[0], OUT_PAD))
    dz[:, :N_OUT] = 2.0 * mask * d / n
    dz = _r(dz, bf)
    dW, db = [None] * 4, [None] * 4
    for l in range(3, -1, -1):
        dW[l] = dz.T @ hs[l]
        db[l] = dz.sum(0)
        if l == 0:
            break
        dz = _r((dz @ _r(W[l], bf)) * dsilu(zs[l - 1]), bf)
    return loss, dW, db


def gradients(params, members, x, t, mask, bf=True):
    """Flat gradient vector [E * P] and losses [E]; x [E, B, 4], t/mask [E, B, 3]."""
    w_off, b_off, per = layout()
    g = np.zeros((members, per))
    losses = np.zeros(members)
    for e, (W, b) in enumerate(unpack(params, members)):
        losses[e], dW, db = grads_member(W, b, x[e], t[e], mask[e], bf)
        for l in range(4):
            g[e, w_off[l]:w_off[l] + ROWS[l] * COLS[l]] = dW[l].ravel()
            g[e, b_off[l]:b_off[l] + ROWS[l]] = db[l]
    return g.ravel(), losses


def forward(params, members, x, bf=True):
    """y [E, B, 3] for x [E, B, 4]."""
    return np.stack([forward_member(W, b, x[e], bf)[0]
                     for e, (W, b) in enumerate(unpack(params, members))])


def adam(params, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias-corrected); returns (params, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
