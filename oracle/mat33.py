"""Oracle 3x3 matrix functions (CPU, float64).  Test infrastructure only.

The reference computes exp / log / the Frechet derivative of exp with its
own batched routines (torch_sim/math.py:30-224, 614-745) and checks them
against scipy.linalg (tests/test_math.py:18-110, 189+).  The oracle uses
scipy.linalg directly, which is the library both are pinned to, and restates
the reference's direction-by-direction cell-force contraction
(torch_sim/optimizers/cell_filters.py:142-178).
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import torch


def expm(a: torch.Tensor) -> torch.Tensor:
    """Batched matrix exponential via scipy.linalg.expm."""
    arr = a.detach().cpu().numpy()
    return torch.from_numpy(np.stack([sla.expm(m) for m in arr.reshape(-1, 3, 3)])).view_as(a)


def logm(a: torch.Tensor) -> torch.Tensor:
    """Batched principal matrix logarithm via scipy.linalg.logm (real part)."""
    arr = a.detach().cpu().numpy()
    out = np.stack([np.real(sla.logm(m)) for m in arr.reshape(-1, 3, 3)])
    return torch.from_numpy(out).view_as(a)


def expm_frechet(a: torch.Tensor, e: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """Batched (exp(A), L_exp(A, E)) via scipy.linalg.expm_frechet."""
    aa = a.detach().cpu().numpy().reshape(-1, 3, 3)
    ee = e.detach().cpu().numpy().reshape(-1, 3, 3)
    r, l = zip(*(sla.expm_frechet(x, y) for x, y in zip(aa, ee)))
    return (
        torch.from_numpy(np.stack(r)).view_as(a),
        torch.from_numpy(np.stack(l)).view_as(a),
    )


def frechet_cell_forces(log_f: torch.Tensor, grad: torch.Tensor) -> torch.Tensor:
    """sum_ij L_exp(logF, E_mn)_ij grad_ij over the 9 unit directions.

    Restates cell_filters.py:142-178.
    """
    n = log_f.shape[0]
    out = torch.zeros_like(grad)
    for m in range(3):
        for k in range(3):
            e = torch.zeros((n, 3, 3), dtype=log_f.dtype)
            e[:, m, k] = 1.0
            _, l = expm_frechet(log_f, e)
            out[:, m, k] = (l * grad).sum(dim=(1, 2))
    return out
