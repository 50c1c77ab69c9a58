"""Differentiable sigma_min(C - zI) over a recursion grid (no counterpart in the reference).

``SigmaMin`` is a torch.autograd.Function: the forward pass is psTriRecy on the GPU with the
certified right singular vector of every grid point kept in device memory; the backward
pass applies the adjoint kernels of librls_hip.so:

    d sigma / d Re z = -Re(u^H v),   d sigma / d Im z = Im(u^H v),
    d sigma / d Re C[i, j] = Re(conj(u_i) v_j),

with v the accepted right vector and u = (C - zI) v / ||(C - zI) v||.  C is parametrised
by real values either as a Toeplitz symbol (C[i, j] = t_{i-j}) or on a sparsity pattern
added to a fixed base matrix.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import PsTriRecyControl, Recursion, TriSolver, schur_factor


@dataclass
class MatrixSpec:
    """How real parameters map to C.

    kind "toeplitz": params t (2n - 1,), C[i, j] = t[i - j + n - 1].
    kind "pattern":  params c (nnz,), C = base + sum_e c[e] E(rows[e], cols[e]) (0-based)."""
    kind: str
    n: int
    rows: Optional[np.ndarray] = None
    cols: Optional[np.ndarray] = None
    base: Optional[np.ndarray] = None

    def matrix(self, params: np.ndarray) -> np.ndarray:
        n = self.n
        if self.kind == "toeplitz":
            if params.shape != (2 * n - 1,):
                raise ValueError(f"Toeplitz symbol must have 2n-1 = {2 * n - 1} entries")
            i = np.arange(n)
            return params[(i[:, None] - i[None, :]) + (n - 1)].astype(np.complex128)
        if self.kind == "pattern":
            C = (np.zeros((n, n), np.complex128) if self.base is None
                 else np.array(self.base, dtype=np.complex128))
            np.add.at(C, (self.rows, self.cols), params)
            return C
        raise ValueError(f"unknown matrix kind {self.kind!r}")


def solve_with_grad(spec: MatrixSpec, params: np.ndarray, rec: Recursion,
                    control: PsTriRecyControl, device: int = 0):
    """(sigma (nZ,), dsdz (nZ, 2), dsdparams (nZ, n_params)) over rec.Z."""
    C = spec.matrix(np.asarray(params, dtype=np.float64))
    T, Q = schur_factor(C)
    solver = TriSolver(T, device)
    solver.set_schur_vectors(Q)
    solver.keep_vectors(True)
    sigma, _ = solver.solve(rec.Z, control, rec)
    dsdz = solver.sigma_grad()
    if spec.kind == "toeplitz":
        dsdp = solver.grad_toeplitz()
    else:
        dsdp = solver.grad_pattern(spec.rows, spec.cols)
    return sigma, dsdz, dsdp


try:
    import torch

    class SigmaMin(torch.autograd.Function):
        """sigma = SigmaMin.apply(params, zxy, spec, rec, control).

        params: float64 tensor of the matrix parameters (see MatrixSpec); zxy: float64 tensor
        (nZ, 2) holding (Re z, Im z) of rec.Z in traversal order.  Returns sigma (nZ,)."""

        @staticmethod
        def forward(ctx, params, zxy, spec, rec, control):
            Z = rec.Z
            zc = zxy.detach().cpu().numpy()
            if zc.shape != (len(Z), 2) or not np.allclose(zc[:, 0] + 1j * zc[:, 1], Z, rtol=0, atol=1e-12):
                raise ValueError("zxy must hold (Re z, Im z) of rec.Z in traversal order")
            sigma, dsdz, dsdp = solve_with_grad(spec, params.detach().cpu().numpy(), rec, control)
            ctx.save_for_backward(torch.from_numpy(dsdz), torch.from_numpy(dsdp))
            return torch.from_numpy(sigma).to(params.device)

        @staticmethod
        def backward(ctx, g):
            dsdz, dsdp = ctx.saved_tensors
            g = g.detach().cpu().to(torch.float64)
            return (g @ dsdp).to(g.device), g[:, None] * dsdz, None, None, None

except ImportError:   # torch is optional for the adjoint entry points themselves
    SigmaMin = None
