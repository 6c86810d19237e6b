"""PyTorch float64 restatement of the dense PairHMM, for reference gradients.

Test infrastructure only (like the rest of ``oracle/``): it yields
d(sum_p w_p ln L_p) / d(gap, ext, err[0..63]) by autograd for small pairs.

Recurrence as in batch_hmm.cpp:31-69 (modes 0/2: free start 1/(m+1) in
every column, terminal sum_j M+I of the last row; mode 1: anchored start with
D[0][j] = g x^(j-1), terminal M+I+D at (n, m)).  Rows are renormalised by
their maximum as in modes 0/1 (batch_hmm.cpp:170-178); that leaves ln L and
its gradient unchanged and keeps float64 in range.  The 1e-300 terminal clamp
is not applied (small test pairs never reach it).  Quality q uses error bin
min(q, 63) of the 64-entry table.
"""
from __future__ import annotations

import torch


def _deletion_matrix(m: int, g, x):
    """T with (T @ v)[j] = sum_{k<j} g x^(j-1-k) v[k] (the D recurrence along a row)."""
    j = torch.arange(m + 1, dtype=torch.float64)
    diff = j[:, None] - 1 - j[None, :]
    mask = diff >= 0
    powers = torch.where(mask, x ** torch.clamp(diff, min=0), torch.zeros((), dtype=torch.float64))
    return g * powers


def loglik(read: bytes, qual: bytes, hap: bytes, mode: int, gap, ext, err):
    """ln L of one pair as a differentiable float64 scalar."""
    n, m = len(read), len(hap)
    if n == 0 or m == 0:
        return torch.tensor(float("-inf"), dtype=torch.float64)
    g, x = gap, ext
    a, b = 1 - 2 * g, 1 - x
    T = _deletion_matrix(m, g, x)
    hapv = torch.tensor(list(hap), dtype=torch.int64)
    jj = torch.arange(m + 1, dtype=torch.float64)
    zero = torch.zeros(m + 1, dtype=torch.float64)
    if mode == 1:
        M = torch.where(jj == 0, torch.ones((), dtype=torch.float64), torch.zeros((), dtype=torch.float64))
        D = torch.where(jj > 0, g * x ** torch.clamp(jj - 1, min=0), torch.zeros((), dtype=torch.float64))
    else:
        M = torch.full((m + 1,), 1.0 / (m + 1), dtype=torch.float64)
        D = zero
    I = zero
    logs = torch.zeros((), dtype=torch.float64)
    for i in range(n):
        e = err[min(qual[i], 63)]
        emit = torch.where(hapv == read[i], 1 - e, e / 3)
        cM = torch.cat([torch.zeros(1, dtype=torch.float64), emit * (a * M[:-1] + b * (I[:-1] + D[:-1]))])
        cI = 0.25 * (g * M + x * I)
        cD = T @ cM
        s = torch.max(cM + cI + cD).detach()
        M, I, D = cM / s, cI / s, cD / s
        logs = logs + torch.log(s)
    if mode == 1:
        term = M[m] + I[m] + D[m]
    else:
        term = torch.sum(M + I)
    return torch.log(term) + logs


def value_and_grad(jobs, mode: int, gap: float, ext: float, err64, weights=None):
    """(ln L per job, d sum w ln L / d gap, / d ext, / d err[64]) in float64."""
    g = torch.tensor(float(gap), dtype=torch.float64, requires_grad=True)
    x = torch.tensor(float(ext), dtype=torch.float64, requires_grad=True)
    e = torch.tensor([float(v) for v in err64], dtype=torch.float64, requires_grad=True)
    vals = [loglik(s, q, h, mode, g, x, e) for s, q, h in jobs]
    w = [1.0] * len(jobs) if weights is None else [float(v) for v in weights]
    total = sum(wi * v for wi, v in zip(w, vals))
    total.backward()
    return (torch.stack(vals).detach().numpy(), float(g.grad), float(x.grad), e.grad.detach().numpy())
