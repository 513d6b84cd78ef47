"""CPU restatement of the learned warm start network (numpy, float64).

TEST INFRASTRUCTURE ONLY: the GPU tests compare the device's loss and
parameter gradient with this code; nothing in the product imports it.

The network has no counterpart in the reference.  This file states its
mathematics independently of the device code:

    z1_e = WP1 x_i + WQ1 x_j + G1 g_e + b1,   h1_i = mean_{e in adj(i)} relu(z1_e)
    z2_e = WP2 h1_i + WQ2 h1_j + G2 g_e + b2, h2_i = mean_{e in adj(i)} relu(z2_e)
    y_i  = Wo h2_i + bo,                       L = mean_{i,v} (y_iv - t_iv)^2

with the flat parameter layout of cfdg_gnn_get_params.  With quant=True the
operands that the device feeds to its BF16 MFMA products are rounded to
bfloat16 the same way (inputs x, weights [WP; WQ], h1, and the node gradients
dPQ entering the weight-gradient and dh1 products); everything else is exact
float64.  The analytic gradient is pinned by central finite differences in
tests/test_gnn_oracle.py.
"""
import numpy as np

F, FUSED, H, G, O = 32, 24, 64, 4, 5
PQ = 2 * H
SIZES = [("W1", (PQ, F)), ("G1", (H, G)), ("b1", (H,)), ("W2", (PQ, H)), ("G2", (H, G)), ("b2", (H,)),
         ("Wo", (O, H)), ("bo", (O,))]
NUM_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def bf16(a):
    """float32 -> bfloat16 (round to nearest, ties to even) -> float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(a32)
    out[nan] = np.nan
    return out


def _edges(adj_off):
    n = len(adj_off) - 1
    deg = np.diff(adj_off)
    src = np.repeat(np.arange(n), deg)
    return src, deg


def _seg_sum(vals, adj_off, n):
    out = np.zeros((n,) + vals.shape[1:])
    np.add.at(out, np.repeat(np.arange(n), np.diff(adj_off)), vals)
    return out


def forward(p, x, adj_off, adj, eg, quant=True):
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    w = unpack(p)
    n = len(adj_off) - 1
    src, deg = _edges(adj_off)
    degf = np.maximum(deg, 1).astype(np.float64)[:, None]
    pq1 = q(x) @ q(w["W1"]).T
    z1 = pq1[src, :H] + pq1[adj, H:] + eg @ w["G1"].T + w["b1"]
    h1 = _seg_sum(np.maximum(z1, 0.0), adj_off, n) / degf
    h1q = q(h1)
    pq2 = h1q @ q(w["W2"]).T
    z2 = pq2[src, :H] + pq2[adj, H:] + eg @ w["G2"].T + w["b2"]
    h2 = _seg_sum(np.maximum(z2, 0.0), adj_off, n) / degf
    y = h2 @ w["Wo"].T + w["bo"]
    return dict(w=w, src=src, degf=degf, pq1=pq1, z1=z1, h1q=h1q, pq2=pq2, z2=z2, h2=h2, y=y)


def loss_grad(p, x, t, adj_off, adj, rev, eg, quant=True):
    """Loss and flat gradient.  rev[e] = index of the reverse edge of e."""
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    f = forward(p, x, adj_off, adj, eg, quant)
    w, src, degf = f["w"], f["src"], f["degf"]
    n = len(adj_off) - 1
    d = f["y"] - t
    loss = float(np.mean(d * d))
    dy = 2.0 * d / d.size
    g = {"Wo": dy.T @ f["h2"], "bo": dy.sum(axis=0)}
    dh2 = dy @ w["Wo"]

    def edge_back(z, dh):
        dz = (dh[src] / degf[src]) * (z > 0.0)            # d/dz_e of the mean over adj(src)
        dP = _seg_sum(dz, adj_off, n)                      # z_e contains P_src
        dQ = np.zeros((n, H))
        np.add.at(dQ, adj, dz)                             # and Q_adj
        return dz, np.concatenate([dP, dQ], axis=1)

    dz2, dpq2 = edge_back(f["z2"], dh2)
    g["G2"] = dz2.T @ eg
    g["b2"] = dz2.sum(axis=0)
    g["W2"] = q(dpq2).T @ f["h1q"]
    dh1 = q(dpq2) @ q(w["W2"])
    dz1, dpq1 = edge_back(f["z1"], dh1)
    g["G1"] = dz1.T @ eg
    g["b1"] = dz1.sum(axis=0)
    g["W1"] = q(dpq1).T @ q(x)
    _ = rev  # the oracle scatters with np.add.at; rev is the device's gather order
    return loss, pack(g)
