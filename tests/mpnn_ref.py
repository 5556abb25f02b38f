"""CPU reference of the warm-start network (drape-2_amd/csrc/warmstart.hpp), forward and backward in numpy FP64.

With bf16=True the values that the GPU feeds to its BF16 MFMA products (features, weights, hidden states,
messages and the back-propagated GEMM operands) are rounded to BF16 (round to nearest even) at the same points,
so the GPU gradient can be compared tightly; with bf16=False it is the exact FP64 network, which the
finite-difference test pins.
"""
from __future__ import annotations

import numpy as np

from drape_amd import learn

IN, EDGE, H, R, OUT, OU = learn.IN, learn.EDGE, learn.HIDDEN, learn.ROUNDS, learn.OUT, learn.OUT_USED


def bf16(a):
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    lay = learn.param_layout()
    p = np.asarray(p, np.float64)
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in lay.items() if not k.startswith("_")}


def features(frame, n):
    f = np.zeros((n, IN))
    f[:, 0:3] = np.asarray(frame["v"], np.float64) / learn.VEL_SCALE
    f[:, 3:6] = np.asarray(frame["dv_prev"], np.float64) / learn.DV_SCALE
    pn = np.asarray(frame["phi_normal"], np.float64)
    f[:, 6] = pn[:, 3] / learn.PHI_SCALE
    f[:, 7:10] = pn[:, 0:3]
    return f


def edges(rs, nbr):
    src = np.repeat(np.arange(len(rs) - 1), np.diff(rs.astype(np.int64)))
    return src, nbr.astype(np.int64)


def edge_features(frame, rest2, src, dst):
    x = np.asarray(frame["x"], np.float64)
    v = np.asarray(frame["v"], np.float64)
    r = np.asarray(rest2, np.float64)
    e = np.zeros((len(src), EDGE))
    e[:, 0:2] = (r[dst] - r[src]) / learn.LEN_SCALE
    e[:, 2:5] = (x[dst] - x[src]) / learn.LEN_SCALE
    e[:, 5:8] = (v[dst] - v[src]) / learn.VEL_SCALE
    return e


def forward_backward(params, rs, nbr, rest2, frame, target=None, bf=True):
    """Returns (dv prediction [n, 3], loss, gradient dict or None)."""
    q = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    P = unpack(params)
    n = len(rs) - 1
    src, dst = edges(rs, nbr)
    deg = np.diff(rs.astype(np.int64)).astype(np.float64)
    e = edge_features(frame, rest2, src, dst)
    f = features(frame, n)
    fb = q(f)
    hEnc = np.maximum(fb @ q(P["We"]) + P["be"], 0.0)
    hF = hEnc.copy()
    hm_in, pq_, z_, U_ = [], [], [], []
    h_in = q(hF)
    for r in range(R):
        pq = h_in @ q(P[f"Wab{r}"])
        z = pq[src, :H] + pq[dst, H:] + e @ P[f"Wc{r}"] + P[f"bm{r}"]
        msum = np.zeros((n, H))
        np.add.at(msum, src, np.maximum(z, 0.0))
        m = np.where(deg[:, None] > 0, msum / np.maximum(deg, 1)[:, None], 0.0)
        hm = np.concatenate([h_in, q(m)], axis=1)
        U = np.maximum(hm @ q(P[f"Wu{r}"]) + P[f"bu{r}"], 0.0)
        hF = hF + U
        hm_in.append(hm)
        pq_.append(pq)
        z_.append(z)
        U_.append(U)
        h_in = q(hF)
    h3 = h_in
    out = h3 @ q(P["Wd"]) + P["bd"]
    dv = out[:, :OU] * learn.DV_SCALE
    if target is None:
        return dv, None, None
    t = np.asarray(target, np.float64) / learn.DV_SCALE
    diff = out[:, :OU] - t
    loss = float(np.sum(diff * diff) / n)
    G = {k: np.zeros_like(v) for k, v in P.items()}
    dout = np.zeros((n, OUT))
    dout[:, :OU] = 2.0 * diff / n
    doutb = q(dout)
    G["Wd"] = h3.T @ doutb
    G["bd"] = dout.sum(axis=0)
    dh = doutb @ q(P["Wd"]).T
    for r in reversed(range(R)):
        hm, pq, z, U = hm_in[r], pq_[r], z_[r], U_[r]
        dpre = dh * (U > 0)
        dpreb = q(dpre)
        G[f"Wu{r}"] = hm.T @ dpreb
        G[f"bu{r}"] = dpre.sum(axis=0)
        dHm = dpreb @ q(P[f"Wu{r}"]).T
        dm = dHm[:, H:]
        g = np.where(deg[src, None] > 0, dm[src] / np.maximum(deg[src], 1)[:, None], 0.0) * (z > 0)
        dP = np.zeros((n, H))
        dQ = np.zeros((n, H))
        np.add.at(dP, src, g)
        np.add.at(dQ, dst, g)
        G[f"Wc{r}"] = e.T @ g
        G[f"bm{r}"] = g.sum(axis=0)
        dpq = q(np.concatenate([dP, dQ], axis=1))
        G[f"Wab{r}"] = hm[:, :H].T @ dpq
        dh = dh + dHm[:, :H] + dpq @ q(P[f"Wab{r}"]).T
    dpre0 = dh * (hEnc > 0)
    G["We"] = fb.T @ q(dpre0)
    G["be"] = dpre0.sum(axis=0)
    return dv, loss, G


def flatten(G) -> np.ndarray:
    lay = learn.param_layout()
    out = np.zeros(lay["_total"][0])
    for k, (o, s) in lay.items():
        if not k.startswith("_"):
            out[o:o + int(np.prod(s))] = G[k].reshape(-1)
    return out
