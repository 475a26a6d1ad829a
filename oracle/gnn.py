"""CPU reference of the learned preconditioner network (torch autograd).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py). No reference
counterpart: the reference's preconditioner slot is empty
(easier/numeric/preconditioner/__init__.py:4). This restatement of the
network defines the forward/backward the HIP kernels must reproduce; its own
gradients are pinned by central finite differences in float64
(tests/test_oracle_gnn.py).

Network (hidden width H = 32, L message-passing layers):
  s     = sqrt(mean(r^2))
  u_i   = (r_i / s, vfeat_i0, vfeat_i1)
  h0_i  = q(relu(u_i W_in + b_in))
  layer l:  P_i = h_i Wp,  Q_j = q(h_j Wq)
            pre_ij = P_i + Q_j + e_ij We + bm          (edge i<-j, e: 3 feats)
            a_i = sum_j relu(pre_ij)                    (row order)
            t_i = [h_i, q(a_i)] Wu + bu
            h_i <- q(h_i + relu(t_i))
  y_i   = h_i w_out + b_out
  z_i   = (r_i + s y_i) / d_i
q() rounds to bfloat16 (straight-through in the backward pass); matrices Wp,
Wq, Wu are used in bfloat16, everything else in float32.
"""
import math

import torch

H = 32


def init_params(L=3, seed=0):
    """Deterministic initial parameters (float32 masters): small random
    weights, zero decoder (so the untrained network is exactly Jacobi)."""
    g = torch.Generator().manual_seed(seed)

    def w(*shape, fan_in):
        return torch.randn(*shape, generator=g) * (1.0 / math.sqrt(fan_in))
    p = {'W_in': w(3, H, fan_in=3), 'b_in': torch.zeros(H)}
    for l in range(L):
        p[f'Wp{l}'] = w(H, H, fan_in=H)
        p[f'Wq{l}'] = w(H, H, fan_in=H)
        p[f'We{l}'] = w(3, H, fan_in=3)
        p[f'bm{l}'] = torch.zeros(H)
        p[f'Wu{l}'] = w(2 * H, H, fan_in=2 * H) * 0.5
        p[f'bu{l}'] = torch.zeros(H)
    p['w_out'] = torch.zeros(H)
    p['b_out'] = torch.zeros(())
    return p


def param_order(L):
    names = ['W_in', 'b_in']
    for l in range(L):
        names += [f'Wp{l}', f'Wq{l}', f'We{l}', f'bm{l}', f'Wu{l}', f'bu{l}']
    return names + ['w_out', 'b_out']


def _q(x, bf16):
    if not bf16:
        return x
    return x + (x.to(torch.bfloat16).to(x.dtype) - x).detach()


def forward(params, r, d, vfeat, efeat, row, col, L=3, bf16=True, dtype=torch.float32):
    """z = M(r). r, d: (n,) float64; vfeat (n, 2); efeat (nnz, 3) in the
    edge order of (row, col) (row-sorted); returns z (n,) float64."""
    P = {k: v.to(dtype) for k, v in params.items()}
    n = r.shape[0]
    s = torch.sqrt(torch.mean(r * r))
    if float(s) == 0.0:
        return torch.zeros_like(r)
    mat = lambda W: _q(W, bf16)
    u = torch.stack([(r / s).to(dtype), vfeat[:, 0].to(dtype), vfeat[:, 1].to(dtype)], 1)
    h = _q(torch.relu(u @ P['W_in'] + P['b_in']), bf16)
    ef = efeat.to(dtype)
    for l in range(L):
        Pm = h @ mat(P[f'Wp{l}'])
        Q = _q(h @ mat(P[f'Wq{l}']), bf16)
        pre = Pm[row] + Q[col] + ef @ P[f'We{l}'] + P[f'bm{l}']
        m = torch.relu(pre)
        a = torch.zeros((n, H), dtype=dtype).index_add_(0, row, m)
        t = torch.cat([h, _q(a, bf16)], 1) @ mat(P[f'Wu{l}']) + P[f'bu{l}']
        h = _q(h + torch.relu(t), bf16)
    y = h @ P['w_out'] + P['b_out']
    return (r + s * y.to(r.dtype)) / d


def loss(params, pairs, static, L=3, bf16=True, dtype=torch.float32):
    """mean over pairs of ||M(r) - e||^2 / ||e||^2."""
    d, vfeat, efeat, row, col = static
    tot = 0.0
    for r, e in pairs:
        z = forward(params, r, d, vfeat, efeat, row, col, L, bf16, dtype)
        tot = tot + ((z - e) ** 2).sum() / (e ** 2).sum()
    return tot / len(pairs)


def static_features(Av, Ae, row, col, points, logk_vertex, n_grid):
    """vfeat = (log(d/mean d), log kappa); efeat = (a_ij / sqrt(d_i d_j),
    (x_j - x_i) n, (y_j - y_i) n)."""
    Av = torch.as_tensor(Av, dtype=torch.float64)
    Ae = torch.as_tensor(Ae, dtype=torch.float64)
    row = torch.as_tensor(row)
    col = torch.as_tensor(col)
    pts = torch.as_tensor(points, dtype=torch.float64)
    vfeat = torch.stack([torch.log(Av / Av.mean()),
                         torch.as_tensor(logk_vertex, dtype=torch.float64).reshape(-1)], 1)
    efeat = torch.stack([Ae / torch.sqrt(Av[row] * Av[col]),
                         (pts[col, 0] - pts[row, 0]) * n_grid,
                         (pts[col, 1] - pts[row, 1]) * n_grid], 1)
    return vfeat.float(), efeat.float()
