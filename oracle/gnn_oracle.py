"""CPU oracle of the learned-compensation network (numpy, FP64).

TEST INFRASTRUCTURE ONLY.  Restates the network of the package's
csrc/gnn.hip -- features, 3 message-passing layers of width 32, sigmoid
head, mean-squared-error loss -- with hand-written backpropagation.  With
bf16=True the inputs of every dense product are rounded to bfloat16 (round
to nearest even) exactly where the GPU feeds its MFMA units, so GPU results
can be compared at FP32 accumulation accuracy.  The analytic gradients are
pinned against central finite differences in FP64 (bf16=False).  No
counterpart in the reference.
"""
from __future__ import annotations

import numpy as np

HID = 32
FEAT = 16
IN_DIM = (FEAT, HID, HID)
OFF_W = (0, 1024 + 32, 1024 + 32 + 2048 + 32)
OFF_B = (1024, 1024 + 32 + 2048, 1024 + 32 + 2048 + 32 + 2048)
OFF_WO = 1024 + 32 + 2 * (2048 + 32)
OFF_BO = OFF_WO + HID
N_PARAMS = OFF_BO + 1


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    W = [p[OFF_W[l]:OFF_W[l] + 2 * IN_DIM[l] * HID].reshape(2 * IN_DIM[l], HID) for l in range(3)]
    b = [p[OFF_B[l]:OFF_B[l] + HID] for l in range(3)]
    return W, b, p[OFF_WO:OFF_WO + HID], p[OFF_BO]


def features(rp, ci, va):
    """-> (feat[n, 16] float32, edge weights per CSR entry float32)."""
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    d = np.zeros(n)
    on = ci == rows
    d[rows[on]] = va[on]
    ad = np.where(np.abs(d) > 0, np.abs(d), 1.0)
    off = ~on
    w = np.zeros(ci.size)
    w[off] = va[off] / np.sqrt(ad[rows[off]] * ad[ci[off]])
    feat = np.zeros((n, FEAT))
    absv = np.where(off, np.abs(va), 0.0)
    sabs = np.bincount(rows, absv, n)
    ssum = np.bincount(rows, np.where(off, va, 0.0), n)
    mx = np.zeros(n)
    np.maximum.at(mx, rows[off], np.abs(va[off]))
    mn = np.full(n, np.inf)
    np.minimum.at(mn, rows[off], np.abs(va[off]))
    cnt = np.bincount(rows[off], minlength=n)
    mn[cnt == 0] = 0.0
    lower = np.bincount(rows, np.where(off & (ci < rows), np.abs(va), 0.0), n)
    feat[:, 0] = 1.0
    feat[:, 1] = sabs / ad
    feat[:, 2] = ssum / np.where(d != 0, d, 1.0)
    feat[:, 3] = mx / ad
    feat[:, 4] = mn / ad
    feat[:, 5] = cnt / 8.0
    feat[:, 6] = np.bincount(rows, w * w, n)
    feat[:, 7] = np.where(sabs > 0, lower / np.where(sabs > 0, sabs, 1.0), 0.0)
    return feat.astype(np.float32), w.astype(np.float32)


def aggregate(rp, ci, w, h):
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    out = np.zeros((n, h.shape[1]))
    np.add.at(out, rows, w[:, None].astype(np.float64) * h[ci])
    return out


def forward(p, rp, ci, feat, w, use_bf16=True):
    W, b, wo, bo = unpack(np.asarray(p, dtype=np.float64))
    rnd = bf16 if use_bf16 else (lambda x: np.asarray(x, dtype=np.float64))
    H = feat.astype(np.float64)
    cache = []
    for l in range(3):
        M = aggregate(rp, ci, w, H)
        X = np.concatenate([H, M], axis=1)
        Z = rnd(X) @ rnd(W[l]) + b[l]
        cache.append((H, M, X, Z))
        H = np.maximum(Z, 0.0)
    u = H @ wo + bo
    omega = 1.0 / (1.0 + np.exp(-u))
    return omega, H, cache


def loss_grad(p, rp, ci, feat, w, target, use_bf16=True):
    """-> (loss, grad[N_PARAMS]) of mean((omega - target)^2)."""
    n = rp.size - 1
    W, b, wo, bo = unpack(np.asarray(p, dtype=np.float64))
    rnd = bf16 if use_bf16 else (lambda x: np.asarray(x, dtype=np.float64))
    omega, H3, cache = forward(p, rp, ci, feat, w, use_bf16)
    diff = omega - target
    loss = float(np.sum(diff * diff) / n)
    du = 2.0 / n * diff * omega * (1.0 - omega)
    g = np.zeros(N_PARAMS)
    g[OFF_WO:OFF_WO + HID] = du @ H3
    g[OFF_BO] = du.sum()
    dH = np.outer(du, wo)
    # transpose graph for the aggregation backward
    rows = np.repeat(np.arange(n), np.diff(rp))
    for l in (2, 1, 0):
        H, M, X, Z = cache[l]
        D = IN_DIM[l]
        dZ = np.where(Z > 0, dH, 0.0)
        g[OFF_W[l]:OFF_W[l] + 2 * D * HID] = (rnd(X).T @ rnd(dZ)).ravel()
        g[OFF_B[l]:OFF_B[l] + HID] = dZ.sum(axis=0)
        if l == 0:
            break
        dX = rnd(dZ) @ rnd(W[l]).T
        dHd, dM = dX[:, :D], dX[:, D:]
        dHt = np.zeros_like(dHd)
        np.add.at(dHt, ci, w[:, None].astype(np.float64) * dM[rows])
        dH = dHd + dHt
    return loss, g
