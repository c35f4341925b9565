"""CPU oracle of the learned message-passing preconditioner.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing in the product).
The network has no counterpart in the reference -- it is the learned
preconditioner added beside the plain FP64 path -- so this file restates the
definition the HIP kernels implement, in float64 numpy, with optional emulation
of the kernels' BF16 rounding points:

  inputs per node i (per sample b), with d_i = a_ii, dbar = mean(d),
  delta_i = dbar / d_i, s_b = sqrt(mean_i r_i^2):
      f0 = (r_i / s_b) delta_i          (Jacobi-scaled residual)
      f1 = zs_i dbar / s_b              (zs = multicolor SGS applied to r)
      f2 = delta_i
  encoder:  h^0 = W_in f + b_in                                   (3 -> 32)
  layer l:  m_i = sum_{j != i} w_ij h_j,   w_ij = -a_ij / a_ii
            h^{l+1} = h^l + relu(W_l [h^l ; m] + b_l)             (64 -> 32)
  output:   out_i = w_out . h^L_i + c0 f0 + c1 f1,   z = out s_b / dbar

BF16 rounding points (round-to-nearest-even): every weight matrix/vector
(W_in, W_l, w_out, c), the message weights w_ij (rounded from float32), every
stored hidden state h^l, the aggregated message m as an MFMA operand, and the
backward signal delta = dL/du as an MFMA operand.
Biases and all accumulations are FP32 on the GPU (float64 here).

Loss (per sample; batch mean): the A-norm relative error of the optimally
scaled output direction with respect to the harvested error e (A e = r):
      L = 1 - <out, r>^2 / (<out, A out> <e, r>)
with A = D (I - W) built from the float32 weights (the operator the kernels use).

Pinning: backward() is pinned by central finite differences of loss() with
rounding off (tests/test_gnn_oracle.py); forward() by closed cases (zero
correction weights reproduce SGS exactly, a one-layer network with chosen
weights reproduces one Jacobi-weighted neighbour average).
"""
import numpy as np
import scipy.sparse as sp

H = 32
NIN = 3


def layout(num_layers):
    """Offsets of the parameter vector (FP32 master copy on the device)."""
    o = {"w_in": 0, "b_in": H * NIN}
    base = H * NIN + H
    for l in range(num_layers):
        o[f"w{l}"] = base + l * (H * 2 * H + H)
        o[f"b{l}"] = o[f"w{l}"] + H * 2 * H
    o["w_out"] = base + num_layers * (H * 2 * H + H)
    o["c"] = o["w_out"] + H
    o["total"] = o["c"] + 2
    return o


def bf16(x):
    """Round to the nearest BF16 (ties to even) via float32; returns float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


class Graph:
    """Message graph of a scalar CSR system (original row order)."""

    def __init__(self, row_ptr, col_idx, values):
        n = row_ptr.shape[0] - 1
        rows = np.repeat(np.arange(n), np.diff(row_ptr))
        cols = np.asarray(col_idx, dtype=np.int64)
        vals = np.asarray(values, dtype=np.float64)
        diag = rows == cols
        d = np.zeros(n)
        d[rows[diag]] = vals[diag]
        off = ~diag
        w = (-vals[off] / d[rows[off]]).astype(np.float32).astype(np.float64)
        self.n = n
        self.W = sp.csr_matrix((w, (rows[off], cols[off])), shape=(n, n))
        self.WT = self.W.T.tocsr()
        # message weights as the layers use them with rounding on (BF16)
        self.Wb = sp.csr_matrix((bf16(w), (rows[off], cols[off])), shape=(n, n))
        self.WbT = self.Wb.T.tocsr()
        self.d = d.astype(np.float32).astype(np.float64)
        self.dbar = float(np.mean(d))
        self.delta = (self.dbar / d).astype(np.float32).astype(np.float64)
        # every row sums to zero (pure-Neumann system): constants span A's null space
        rs = np.zeros(n)
        np.add.at(rs, rows, vals)
        self.neumann = bool(np.all(np.abs(rs) <= 1e-5 * np.abs(d)))  # FP32-assembled sums vanish to ~1e-7

    def aop(self, x):
        """A x = D (I - W) x for x [B, n]."""
        return self.d[None, :] * (x - (self.W @ x.T).T)

    def aop_sym(self, x):
        """(A + A^T) x = 2 D x - D W x - W^T D x."""
        dx = self.d[None, :] * x
        return 2.0 * dx - self.d[None, :] * (self.W @ x.T).T - (self.WT @ dx.T).T


def init_params(num_layers, seed=0):
    """A random network whose correction starts at zero (z = SGS(r))."""
    o = layout(num_layers)
    rng = np.random.default_rng(seed)
    th = np.zeros(o["total"])
    th[o["w_in"]:o["b_in"]] = rng.normal(0, 1 / np.sqrt(NIN), H * NIN)
    for l in range(num_layers):
        th[o[f"w{l}"]:o[f"b{l}"]] = rng.normal(0, 0.5 / np.sqrt(2 * H), H * 2 * H)
    th[o["c"] + 1] = 1.0
    return th


def _unpack(th, L, rnd):
    o = layout(L)
    R = bf16 if rnd else (lambda v: np.asarray(v, dtype=np.float64))
    P = {"w_in": R(th[o["w_in"]:o["b_in"]]).reshape(H, NIN), "b_in": th[o["b_in"]:o["b_in"] + H].astype(np.float64)}
    for l in range(L):
        P[f"w{l}"] = R(th[o[f"w{l}"]:o[f"b{l}"]]).reshape(H, 2 * H)
        P[f"b{l}"] = th[o[f"b{l}"]:o[f"b{l}"] + H].astype(np.float64)
    P["w_out"] = R(th[o["w_out"]:o["c"]])
    P["c"] = R(th[o["c"]:o["c"] + 2])
    return P


def forward(th, G, r, zs, num_layers=4, rnd=True):
    """Network output `out` [B, n] and the cache for backward()."""
    R = bf16 if rnd else (lambda v: v)
    P = _unpack(th, num_layers, rnd)
    r = np.atleast_2d(np.asarray(r, dtype=np.float64))
    zs = np.atleast_2d(np.asarray(zs, dtype=np.float64))
    s = np.sqrt(np.mean(r * r, axis=1))[:, None]
    f0 = r / s * G.delta[None, :]
    f1 = zs * G.dbar / s
    f2 = np.broadcast_to(G.delta[None, :], r.shape)
    F = np.stack([f0, f1, f2], axis=-1)
    h = R(F @ P["w_in"].T + P["b_in"])
    hs, As, masks = [h], [], []
    Wm = G.Wb if rnd else G.W
    for l in range(num_layers):
        m = np.stack([Wm @ h[b] for b in range(h.shape[0])])
        a = np.concatenate([h, R(m)], axis=-1)
        u = a @ P[f"w{l}"].T + P[f"b{l}"]
        h = R(h + np.maximum(u, 0.0))
        As.append(a)
        masks.append(u > 0)
        hs.append(h)
    out = h @ P["w_out"] + P["c"][0] * f0 + P["c"][1] * f1
    cache = dict(F=F, hs=hs, As=As, masks=masks, P=P, s=s)
    return out, cache


def apply(th, G, r, zs, num_layers=4, rnd=True):
    """Preconditioner output z = out s / dbar, projected off the constants when
    the system is pure Neumann (the loss does not determine that component)."""
    out, c = forward(th, G, r, zs, num_layers, rnd)
    z = out * c["s"] / G.dbar
    if G.neumann:
        z = z - z.mean(axis=-1, keepdims=True)
    return z


def loss(out, G, r, e):
    """Batch-mean A-norm loss and the per-sample (P, Q, S)."""
    r = np.atleast_2d(r)
    e = np.atleast_2d(e)
    P = np.sum(out * r, axis=1)
    Q = np.sum(out * G.aop(out), axis=1)
    S = np.sum(e * r, axis=1)
    return float(np.mean(1.0 - P * P / (Q * S))), (P, Q, S)


def backward(th, G, cache, out, r, e, num_layers=4, rnd=True):
    """Gradient of loss() with respect to the parameter vector."""
    R = bf16 if rnd else (lambda v: v)
    o = layout(num_layers)
    P_ = cache["P"]
    r = np.atleast_2d(r)
    B = r.shape[0]
    _, (P, Q, S) = loss(out, G, r, e)
    gout = (-2.0 * P / (Q * S))[:, None] * r + (P * P / (Q * Q * S))[:, None] * G.aop_sym(out)
    gout /= B
    grad = np.zeros(o["total"])
    F = cache["F"]
    hL = cache["hs"][-1]
    grad[o["w_out"]:o["c"]] = np.einsum("bn,bnf->f", gout, hL)
    grad[o["c"]] = np.sum(gout * F[..., 0])
    grad[o["c"] + 1] = np.sum(gout * F[..., 1])
    Gh = gout[..., None] * P_["w_out"][None, None, :]
    WmT = G.WbT if rnd else G.WT
    for l in reversed(range(num_layers)):
        dl = Gh * cache["masks"][l]
        dlr = R(dl)
        grad[o[f"w{l}"]:o[f"b{l}"]] = np.einsum("bnf,bng->fg", dlr, cache["As"][l]).ravel()
        grad[o[f"b{l}"]:o[f"b{l}"] + H] = dl.sum(axis=(0, 1))
        dA = dlr @ P_[f"w{l}"]
        dm = dA[..., H:]
        Gh = Gh + dA[..., :H] + np.stack([WmT @ dm[b] for b in range(B)])
    grad[o["w_in"]:o["b_in"]] = np.einsum("bnf,bnk->fk", Gh, F).ravel()
    grad[o["b_in"]:o["b_in"] + H] = Gh.sum(axis=(0, 1))
    return grad


def loss_and_grad(th, G, r, e, zs, num_layers=4, rnd=True):
    out, cache = forward(th, G, r, zs, num_layers, rnd)
    L, _ = loss(out, G, r, e)
    return L, backward(th, G, cache, out, r, e, num_layers, rnd)
