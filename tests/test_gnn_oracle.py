"""The CPU restatement of the learned warm start network (oracle/gnn_ref.py)
pinned by central finite differences of its own loss: a dropped term, a
wrong index or a transposed operand in the analytic backward fails here.
Float64, no BF16 rounding; tolerance 1e-6 relative to the largest gradient
entry of the checked block."""
import numpy as np

import gnn_ref as G


def _graph(n, rng):
    """Random symmetric graph with a few self loops (boundary faces)."""
    nb = [set() for _ in range(n)]
    for i in range(n):
        for j in rng.choice(n, size=3, replace=False):
            if j != i:
                nb[i].add(int(j))
                nb[int(j)].add(i)
    lists = []
    for i in range(n):
        li = sorted(nb[i])
        if i % 5 == 0:
            li.append(i)  # self loop
        lists.append(li)
    off = np.zeros(n + 1, np.int32)
    off[1:] = np.cumsum([len(li) for li in lists])
    adj = np.array([j for li in lists for j in li], np.int32)
    rev = np.empty_like(adj)
    for i in range(n):
        for e in range(off[i], off[i + 1]):
            j = adj[e]
            rev[e] = next(k for k in range(off[j], off[j + 1]) if adj[k] == i)
    eg = rng.standard_normal((len(adj), G.G))
    return off, adj, rev, eg


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    n = 24
    off, adj, rev, eg = _graph(n, rng)
    x = np.zeros((n, G.F))
    x[:, :G.FUSED] = rng.standard_normal((n, G.FUSED))
    t = rng.standard_normal((n, G.O))
    p = rng.standard_normal(G.NUM_PARAMS) * 0.3
    loss, g = G.loss_grad(p, x, t, off, adj, rev, eg, quant=False)
    assert np.isfinite(loss)
    # check a sample of entries of every parameter block
    o = 0
    h = 1e-6
    for name, shape in G.SIZES:
        size = int(np.prod(shape))
        idx = o + rng.choice(size, size=min(size, 12), replace=False)
        if name == "W1":  # columns >= FUSED meet zero features: their gradient is exactly zero
            idx = np.array([o + r * G.F + c for r, c in zip(rng.integers(0, G.PQ, 12), rng.integers(0, G.FUSED, 12))])
        scale = np.abs(g[o:o + size]).max()
        for k in idx:
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            fd = (G.loss_grad(pp, x, t, off, adj, rev, eg, quant=False)[0] -
                  G.loss_grad(pm, x, t, off, adj, rev, eg, quant=False)[0]) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(scale, 1e-12) + 1e-9, (name, k - o, fd, g[k])
        o += size


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 3.0e38, 0.0])
    r = G.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2.0 ** -6 and r[3] == -2.5 and r[5] == 0.0
    assert np.isinf(G.bf16(np.array([3.4e38])))[0]
