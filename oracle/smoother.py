"""CPU oracle for the learned CMG smoother weights (numpy).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the device
kernels; the product never imports it.

The learned smoother has no counterpart in the reference engine: the
reference CMG cycle (rust/vendor/cmg/src/preconditioner/cycle.rs:136-311)
smooths with the fixed inverse diagonal 0.5 / d_i.  The learned variant keeps
that cycle and replaces 0.5 by a per-node weight omega_i in (0, OMEGA_MAX)
predicted by a small message-passing network from the level graph.  With
every omega_i < 1 the Jacobi smoother satisfies 2 D / omega - A > 0, so the
symmetric V-cycle stays positive definite.

Everything here is the definition the device kernels implement:
  features(level)       8 per-node inputs
  forward(params, ...)  the network (input layer, three message-passing
                        layers of width 64): omega per node of every
                        smoothed level
  cycle_forward/adjoint the training V-cycle (one visit per level) and its
                        exact adjoint, giving dL/domega
  loss_and_grad         L = mean_k |e_k - M A e_k|_A^2 / |e_k|_A^2 and dL/dparams

Pins (tests/test_smoother_oracle.py): the gradient against central finite
differences; omega = 0.5 everywhere reproduces the oracle's plain CMG cycle
(orc_cmg_apply) on single-visit hierarchies; the features against brute
force on small graphs.
"""
from __future__ import annotations

import numpy as np

F = 8          # input features per node
H = 64         # hidden width
OMEGA_MAX = 0.95
LAYERS = 3     # message-passing layers after the input layer

# flat parameter layout: W0, b0, then U_t, V_t, c_t per layer, then a, a0
OFF = {}
_o = 0
for _name, _size in ((("W0", H * F), ("b0", H))
                     + tuple(x for t in range(1, LAYERS + 1)
                             for x in ((f"U{t}", H * H), (f"V{t}", H * H), (f"c{t}", H)))
                     + (("a", H), ("a0", 1))):
    OFF[_name] = (_o, _size)
    _o += _size
P = _o
MATRICES = tuple(f"{k}{t}" for t in range(1, LAYERS + 1) for k in ("U", "V"))


def unpack(theta):
    out = {}
    for k, (o, s) in OFF.items():
        out[k] = theta[o:o + s]
    out["W0"] = out["W0"].reshape(H, F)
    for k in MATRICES:
        out[k] = out[k].reshape(H, H)
    return out


def initial_params(seed=1, scale=0.1):
    """Small random weights, a = 0 and a0 = logit(0.5 / OMEGA_MAX): omega = 0.5
    (the plain cycle) at the start of training."""
    rng = np.random.default_rng(seed)
    theta = np.zeros(P)
    for k in ("W0",) + MATRICES:
        o, s = OFF[k]
        fan_in = F if k == "W0" else H
        theta[o:o + s] = rng.standard_normal(s) * scale / np.sqrt(fan_in)
    q = 0.5 / OMEGA_MAX
    theta[OFF["a0"][0]] = np.log(q / (1.0 - q))
    return theta


# ----------------------------------------------------------------- hierarchy

def oracle_levels(cmg):
    """Level list from the oracle CMG: graph, labels, terminal kind, and the
    terminal operator (grounded dense inverse or the 0.5 / d Jacobi)."""
    out = []
    for lv in cmg.levels():
        n = lv["n"]
        W = np.zeros((0, 0))
        out.append(dict(n=n, u=lv["u"].astype(np.int64), v=lv["v"].astype(np.int64), w=lv["w"],
                        diag=lv["diag"], labels=None if lv["labels"] is None else lv["labels"].astype(np.int64),
                        coarse=lv["coarse"], terminal=lv["terminal"], repeat=lv["repeat"]))
        del W
    for lv in out:
        if lv["terminal"] == 1:       # direct (TR_DIRECT + 1)
            lv["G"] = grounded_inverse(lv)
    return out


def grounded_inverse(lv):
    """Inverse of the Laplacian with the last vertex of every component
    grounded, scattered to graph indices (zero rows/cols at the grounds)."""
    n = lv["n"]
    A = laplacian_dense(lv)
    comp = components(n, lv["u"], lv["v"])
    ground = np.zeros(n, bool)
    for c in np.unique(comp):
        ground[np.nonzero(comp == c)[0].max()] = True
    keep = np.nonzero(~ground)[0]
    G = np.zeros((n, n))
    if len(keep):
        G[np.ix_(keep, keep)] = np.linalg.inv(A[np.ix_(keep, keep)])
    return G


def components(n, u, v):
    parent = np.arange(n)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    for a, b in zip(u, v):
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    return np.array([find(i) for i in range(n)])


def laplacian_dense(lv):
    n = lv["n"]
    A = np.zeros((n, n))
    np.add.at(A, (lv["u"], lv["v"]), -lv["w"])
    np.add.at(A, (lv["v"], lv["u"]), -lv["w"])
    A[np.arange(n), np.arange(n)] += lv["diag"]
    return A


def adjacency(lv):
    import scipy.sparse as sp
    n = lv["n"]
    u, v, w = lv["u"], lv["v"], lv["w"]
    return sp.csr_matrix((np.concatenate([w, w]), (np.concatenate([u, v]), np.concatenate([v, u]))),
                         shape=(n, n))


def matvec(lv, x):
    Wm = lv.get("_W")
    if Wm is None:
        Wm = lv["_W"] = adjacency(lv)
    return lv["diag"][:, None] * x - Wm @ x


# ------------------------------------------------------------------ features

def features(lv, depth):
    """Per-node inputs of one smoothed level (original vertex order)."""
    n = lv["n"]
    u, v, w = lv["u"], lv["v"], lv["w"]
    diag = lv["diag"]
    lab = lv["labels"]
    s = np.bincount(u, weights=w, minlength=n) + np.bincount(v, weights=w, minlength=n)
    cnt = np.bincount(u, minlength=n) + np.bincount(v, minlength=n)
    wmax = np.zeros(n)
    np.maximum.at(wmax, u, w)
    np.maximum.at(wmax, v, w)
    same = lab[u] == lab[v]
    intra = np.bincount(u[same], weights=w[same], minlength=n) + np.bincount(v[same], weights=w[same], minlength=n)
    asz = np.bincount(lab, minlength=lv["coarse"])[lab]
    pos = diag > 0
    ld = np.where(pos, np.log(np.where(pos, diag, 1.0)), 0.0)
    mu = ld[pos].mean() if pos.any() else 0.0
    contrast = np.bincount(u, weights=w * (ld[v] - ld[u]), minlength=n) + \
        np.bincount(v, weights=w * (ld[u] - ld[v]), minlength=n)
    has = s > 0
    rs = np.where(has, 1.0 / np.where(has, s, 1.0), 0.0)
    X = np.zeros((n, F))
    X[:, 0] = 1.0
    X[:, 1] = np.where(pos, ld - mu, 0.0)
    X[:, 2] = np.log1p(cnt) / 4.0
    X[:, 3] = wmax * rs
    X[:, 4] = intra * rs
    X[:, 5] = np.log(asz) / 4.0
    X[:, 6] = depth / 8.0
    X[:, 7] = contrast * rs
    return X, rs


# ------------------------------------------------------------------- network

def forward(theta, X, lv, rs):
    """omega for one level and the activations the backward pass needs."""
    p = unpack(theta)
    Wm = lv.get("_W")
    if Wm is None:
        Wm = lv["_W"] = adjacency(lv)
    h0 = np.maximum(X @ p["W0"].T + p["b0"], 0.0)
    hs, ms = [h0], []
    h = h0
    for t in range(1, LAYERS + 1):
        m = rs[:, None] * (Wm @ h)
        h = np.maximum(h @ p[f"U{t}"].T + m @ p[f"V{t}"].T + p[f"c{t}"], 0.0)
        ms.append(m)
        hs.append(h)
    z = h @ p["a"] + p["a0"][0]
    sig = 1.0 / (1.0 + np.exp(-z))
    return OMEGA_MAX * sig, dict(hs=hs, ms=ms, sig=sig)


def bf16(x):
    """Round FP32 values to BF16 (nearest even), returned as FP32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return u.view(np.float32)


def forward_bf16(theta, X, lv, rs):
    """The solve path's inference of the same network: BF16 weights and
    activations (rounded where the device stores them), FP32 arithmetic;
    omega as FP64 of 0.95 sigmoid(z)."""
    p = unpack(theta)
    Wm = lv.get("_W")
    if Wm is None:
        Wm = lv["_W"] = adjacency(lv)
    W32 = Wm.astype(np.float32)
    X32 = np.asarray(X, np.float32)
    s = np.tile(p["b0"].astype(np.float32), (X.shape[0], 1))
    for f in range(F):
        s = s + p["W0"][:, f].astype(np.float32)[None, :] * X32[:, f:f + 1]
    h = bf16(np.maximum(s, np.float32(0)))
    for t in range(1, LAYERS + 1):
        m = bf16(np.asarray(rs, np.float32)[:, None] * (W32 @ h))
        U, V = bf16(p[f"U{t}"].astype(np.float32)), bf16(p[f"V{t}"].astype(np.float32))
        pre = h @ U.T + m @ V.T + p[f"c{t}"].astype(np.float32)
        h = np.maximum(pre, np.float32(0))
        if t < LAYERS:
            h = bf16(h)
    z = h @ p["a"].astype(np.float32) + np.float32(p["a0"][0])
    return OMEGA_MAX / (1.0 + np.exp(-z.astype(np.float64)))


def backward(theta, X, lv, rs, cache, g_omega):
    """dL/dtheta from dL/domega for one level."""
    p = unpack(theta)
    Wm = lv["_W"]
    grad = {k: np.zeros_like(v) for k, v in p.items()}
    hs, ms, sig = cache["hs"], cache["ms"], cache["sig"]
    dz = g_omega * OMEGA_MAX * sig * (1.0 - sig)
    grad["a"] = hs[LAYERS].T @ dz
    grad["a0"] = np.array([dz.sum()])
    dh = np.outer(dz, p["a"])
    for t in range(LAYERS, 0, -1):
        dpre = dh * (hs[t] > 0)
        grad[f"U{t}"] = dpre.T @ hs[t - 1]
        grad[f"V{t}"] = dpre.T @ ms[t - 1]
        grad[f"c{t}"] = dpre.sum(0)
        dm = dpre @ p[f"V{t}"]
        dh = dpre @ p[f"U{t}"] + Wm.T @ (rs[:, None] * dm)
    dpre = dh * (hs[0] > 0)
    grad["W0"] = dpre.T @ X
    grad["b0"] = dpre.sum(0)
    out = np.zeros(P)
    for k, (o, s) in OFF.items():
        out[o:o + s] = np.asarray(grad[k]).ravel()
    return out


# ---------------------------------------------------------- training V-cycle

def _restrict(lv, r):
    out = np.zeros((lv["coarse"], r.shape[1]))
    np.add.at(out, lv["labels"], r)
    return out


def cycle_forward(levels, ws, r):
    """y = M_w r with one visit per level (the training cycle).  ws[l] is the
    smoother inverse diagonal (omega / d) of every smoothed level.  Returns y
    and the per-level values the adjoint needs.  Level 0 is not centred; the
    coarse right-hand sides are centred by their lane means."""
    store = []

    def level(l, braw):
        lv = levels[l]
        b = braw - braw.mean(0) if l > 0 else braw
        if lv["terminal"]:
            store.append(dict(b=b))
            if lv["terminal"] == 1:
                return lv["G"] @ b
            return (np.where(lv["diag"] > 0, 0.5 / np.where(lv["diag"] > 0, lv["diag"], 1.0), 0.0))[:, None] * b
        w = ws[l][:, None]
        x1 = w * b
        r1 = b - matvec(lv, x1)
        ent = dict(b=b)
        store.append(ent)
        ec = level(l + 1, _restrict(lv, r1))
        pe = ec[lv["labels"]]
        s = r1 - matvec(lv, pe)
        ent["s"] = s
        return x1 + pe + w * s

    y = level(0, r)
    return y, store


def cycle_adjoint(levels, ws, store, lam):
    """Adjoint of cycle_forward: returns dL/dr and dL/dw per smoothed level."""
    gws = [None] * len(levels)

    def level(l, ly):
        lv = levels[l]
        ent = store[l]
        if lv["terminal"]:
            if lv["terminal"] == 1:
                lb = lv["G"].T @ ly
            else:
                lb = (np.where(lv["diag"] > 0, 0.5 / np.where(lv["diag"] > 0, lv["diag"], 1.0), 0.0))[:, None] * ly
        else:
            w = ws[l][:, None]
            b, s = ent["b"], ent["s"]
            gw = (ly * s).sum(1)
            ls = w * ly
            lpe = ly - matvec(lv, ls)
            lec = _restrict(lv, lpe)
            lc = level(l + 1, lec)
            lr1 = ls + lc[lv["labels"]]
            lx1 = ly - matvec(lv, lr1)
            gw = gw + (lx1 * b).sum(1)
            lb = lr1 + w * lx1
            gws[l] = gw
        return lb - lb.mean(0) if l > 0 else lb

    lr = level(0, lam)
    return lr, gws


def smoothed_levels(levels):
    return [l for l, lv in enumerate(levels) if not lv["terminal"]]


def prepare(levels):
    """Features and 1/s of every smoothed level."""
    out = {}
    for l in smoothed_levels(levels):
        out[l] = features(levels[l], l)
    return out


def omegas(theta, levels, feats):
    res = {}
    for l, (X, rs) in feats.items():
        res[l] = forward(theta, X, levels[l], rs)
    return res


def inverse_weights(levels, om):
    ws = [None] * len(levels)
    for l, (omega, _) in om.items():
        d = levels[l]["diag"]
        ws[l] = np.where(d > 0, omega / np.where(d > 0, d, 1.0), 0.0)
    return ws


def loss_and_grad(theta, levels, feats, E):
    """L = mean over columns of |e - M A e|_A^2 / |e|_A^2 and dL/dtheta."""
    om = omegas(theta, levels, feats)
    ws = inverse_weights(levels, om)
    lv0 = levels[0]
    Ae = matvec(lv0, E)
    q = (E * Ae).sum(0)
    y, store = cycle_forward(levels, ws, Ae)
    u = E - y
    Au = matvec(lv0, u)
    K = E.shape[1]
    L = float(np.mean((u * Au).sum(0) / q))
    lam = -2.0 * Au / (K * q)
    _, gws = cycle_adjoint(levels, ws, store, lam)
    grad = np.zeros(P)
    for l, (X, rs) in feats.items():
        d = levels[l]["diag"]
        g_omega = np.where(d > 0, gws[l] / np.where(d > 0, d, 1.0), 0.0)
        grad += backward(theta, X, levels[l], rs, om[l][1], g_omega)
    return L, grad


def loss(theta, levels, feats, E):
    om = omegas(theta, levels, feats)
    ws = inverse_weights(levels, om)
    lv0 = levels[0]
    Ae = matvec(lv0, E)
    q = (E * Ae).sum(0)
    y, _ = cycle_forward(levels, ws, Ae)
    u = E - y
    return float(np.mean((u * matvec(lv0, u)).sum(0) / q))
