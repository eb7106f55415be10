"""Torch (autograd) restatement of the differentiable layered min-sum forward, for
checking gradients.  Test infrastructure: float64 on the CPU.

Per iteration t and layer l (rows of the layer in parallel, codewords batched):
q_j = P[v_j] - C_j;  A_j = relu(|q_j| - beta[t,l]);  M_j = alpha[t,l] * min(cap, min_{k != j} A_k);
C_j <- (-1)^(signbit of the other q) * M_j;  P[v_j] <- q_j + C_j.
"""
import numpy as np
import torch


def layer_index(dec):
    """Per layer: (variable index [w, Z], C2V index [w, Z]) as long tensors."""
    Z = dec.z
    out, e0 = [], 0
    r = np.arange(Z)
    for l in range(dec.mrow):
        w = int(dec.row_weights[l])
        v = np.stack([dec.row_active_cols[e0 + j] * Z + (dec.row_offset[e0 + j] + r) % Z for j in range(w)])
        c = np.stack([(e0 + j) * Z + r for j in range(w)])
        out.append((torch.from_numpy(v.astype(np.int64)), torch.from_numpy(c.astype(np.int64))))
        e0 += w
    return out


def forward(dec, llr, alpha, beta, cap=1e30, idx=None, int8=False):
    """llr [B, N], alpha/beta [T, mrow] tensors (any may require grad) -> posteriors [B, N].

    int8: the int8 decoder's arithmetic on integer-valued inputs -- V2C values and
    posteriors clamped to [-128, 127] (torch.clamp passes the gradient inside the range,
    bounds included), offsets rounded with a straight-through gradient, and an offset of
    0 giving A = 0 at q = -128 (abs(-128) = -128 in the reference's int8 arithmetic)."""
    idx = idx or layer_index(dec)
    B = llr.shape[0]
    P = llr
    C = torch.zeros(B, dec.total_hc_ones * dec.z, dtype=llr.dtype)
    for t in range(alpha.shape[0]):
        for l, (v, c) in enumerate(idx):
            w, Z = v.shape
            q = P[:, v.reshape(-1)].reshape(B, w, Z) - C[:, c.reshape(-1)].reshape(B, w, Z)
            b = beta[t, l]
            if int8:
                q = torch.clamp(q, -128.0, 127.0)
                b = b + (torch.round(b) - b).detach()
            A = torch.relu(q.abs() - b)
            if int8:
                A = torch.where((b.detach() == 0) & (q.detach() == -128), torch.zeros_like(A), A)
            im = torch.argmin(A, dim=1, keepdim=True)                    # first minimum
            onehot = torch.zeros_like(A, dtype=torch.bool).scatter_(1, im, True)
            m1 = torch.gather(A, 1, im)
            m2 = torch.min(A.masked_fill(onehot, float("inf")), dim=1, keepdim=True).values
            x = torch.where(onehot, m2, m1)
            m = alpha[t, l] * torch.clamp(x, max=cap)
            sb = torch.signbit(q)
            neg = torch.logical_xor(sb, (sb.sum(dim=1, keepdim=True) % 2).bool())
            rn = torch.where(neg, -m, m)
            newp = q + rn
            if int8:
                newp = torch.clamp(newp, -128.0, 127.0)
            P = P.index_copy(1, v.reshape(-1), newp.reshape(B, -1))
            C = C.index_copy(1, c.reshape(-1), rn.reshape(B, -1))
    return P
