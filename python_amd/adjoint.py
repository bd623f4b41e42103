"""Differentiable batched solve and d(populations)/dT_e for the ionization balance.

``SolveFn`` is a ``torch.autograd.Function`` over a batched FP64 solve (GSL operation order):
forward solves A x = b per cell and keeps the LU factors in the plan; backward calls the adjoint
kernel (hipla_adjoint), which reuses those factors to solve A^T lambda = dL/dx and returns
dL/dA = -lambda x^T and dL/db = lambda.  Gradients with respect to anything A and b are built
from (the electron temperature, below) then follow from torch's own chain rule.

``IonTe`` assembles the temperature-dependent ionization matrices of the adjoint workload in
torch: ionization i -> i+1 at C T^0.5 exp(-E / kT), recombination i+1 -> i at
alpha0 (T / 1e4)^-0.7, diagonals minus the column sums, and the abundance rows that
``populate_ion_rate_matrix`` places on each element's first ion (source/matrix_ion.c:633-655).
The reference obtains T_e from a root find around the solve (``calc_te``, source/ionization.c:486);
``newton_te`` is one Newton step on T_e per cell driven by the adjoint gradient.

A plan keeps the factors of its latest solve only: run backward before the next forward on the
same plan (backward raises if another solve ran in between).
"""
from __future__ import annotations

import numpy as np
import torch

from .generators import KB_EV
from ._lib import check, lib
from .solver import Plan, _ptr, _stream_handle


class SolveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, B, plan):
        # X starts at zero: a singular cell (code 4) is not written by the solve and reads 0
        Bd = B.detach().contiguous()
        X, codes, _ = plan.solve(A.detach().contiguous(), Bd, torch.zeros_like(Bd))
        ctx.plan = plan
        ctx.generation = plan.generation
        ctx.save_for_backward(X)
        ctx.mark_non_differentiable(codes)
        return X, codes

    @staticmethod
    def backward(ctx, gX, gcodes):
        (X,) = ctx.saved_tensors
        plan = ctx.plan
        if plan.generation != ctx.generation:
            raise RuntimeError("the plan was solved again after this forward pass: its kept LU factors belong to a "
                               "later solve, so this backward would be wrong (run backward first, or use one plan "
                               "per forward)")
        Lam = torch.empty_like(X)
        need_A = ctx.needs_input_grad[0]
        GradA = torch.empty(plan.totA, dtype=X.dtype, device=X.device) if need_A else None
        # singular cells (code 4) get lambda = 0 and dL/dA = 0 from the adjoint kernel
        plan.adjoint(X, gX.contiguous(), Lam, GradA)
        return GradA, Lam, None


def solve(A, B, plan: Plan):
    """x = A^-1 b per cell (flat ragged layout of ``plan``), differentiable in A and b."""
    X, codes = SolveFn.apply(A, B, plan)
    return X, codes


class IonTe:
    """Ionization matrices of order n whose rates depend on T_e (one parameter set per cell).

    ``assemble`` builds A(T) and b from torch operations (differentiable by torch's own chain
    rule, through the dense dL/dA of the adjoint); ``TeSolveFn`` (the default of
    ``loss_and_grad``) assembles on the GPU (hipla_te_assemble) and forms dL/dT_e directly from
    the adjoint's lambda (hipla_te_grad), with no dense dL/dA at all."""

    def __init__(self, te_batch: dict, device="cuda"):
        self.n = te_batch["A"].shape[1]
        self.ncells = te_batch["A"].shape[0]
        self.blocks = te_batch["blocks"]
        self.device = device
        self.terms = []
        for (s, m, ab, C, E, a0) in te_batch["params"]:
            if m > 1:
                t = lambda x: torch.as_tensor(x, dtype=torch.float64, device=device)
                self.terms.append((s, m, t(C), t(E), t(a0)))
        self.cons = [(s, m, ab) for (s, m, ab, *_r) in te_batch["params"]]
        self.offA = np.arange(self.ncells, dtype=np.int64) * self.n * self.n
        self.offV = np.arange(self.ncells, dtype=np.int64) * self.n
        # the GPU assembly's arrays: stage q joins rows row[q] and row[q] + 1
        rows, Cs, Es, A0s = [], [], [], []
        for (s, m, ab, C, E, a0) in te_batch["params"]:
            if m > 1:
                rows += [s + k for k in range(m - 1)]
                Cs.append(np.asarray(C).reshape(self.ncells, m - 1))
                Es.append(np.asarray(E).reshape(self.ncells, m - 1))
                A0s.append(np.asarray(a0).reshape(self.ncells, m - 1))
        if any(r < 0 or r + 1 >= self.n for r in rows) or self.n > 64:
            raise ValueError("stages must join rows inside the matrix, and n <= 64")
        if any(s < 0 or m < 1 or s + m > self.n for (s, m, _ab) in self.cons):
            raise ValueError("element blocks must lie inside the matrix")
        cat = (lambda xs: np.ascontiguousarray(np.concatenate(xs, axis=1))) if Cs else (
            lambda xs: np.zeros((self.ncells, 0)))
        dev = lambda x, dt: torch.as_tensor(np.ascontiguousarray(x), dtype=dt, device=device)
        self.nstage = len(rows)
        self.g_row = dev(np.asarray(rows, dtype=np.int32), torch.int32)
        self.g_C = dev(cat(Cs), torch.float64)
        self.g_E = dev(cat(Es), torch.float64)
        self.g_a0 = dev(cat(A0s), torch.float64)
        self.g_cs = dev(np.asarray([s for (s, m, ab) in self.cons], dtype=np.int32), torch.int32)
        self.g_cm = dev(np.asarray([m for (s, m, ab) in self.cons], dtype=np.int32), torch.int32)
        self.g_cab = dev(np.asarray([ab for (s, m, ab) in self.cons], dtype=np.float64), torch.float64)

    def assemble(self, T):
        """A(T) [ncells, n, n] and b [ncells, n] as differentiable torch tensors."""
        n, nc = self.n, self.ncells
        A = torch.zeros(nc, n, n, dtype=torch.float64, device=self.device)
        Tc = T[:, None]
        for (s, m, C, E, a0) in self.terms:
            k = torch.arange(m - 1, device=self.device)
            ion = C * torch.sqrt(Tc) * torch.exp(-E / (KB_EV * Tc))
            rec = a0 * (Tc / 1e4) ** -0.7
            A = A.index_put((torch.arange(nc, device=self.device)[:, None], s + k + 1, s + k), ion)
            A = A.index_put((torch.arange(nc, device=self.device)[:, None], s + k, s + k + 1), rec)
        d = torch.arange(n, device=self.device)
        A = A - torch.diag_embed(A.sum(dim=1))
        b = torch.zeros(nc, n, dtype=torch.float64, device=self.device)
        keep = torch.ones(n, dtype=torch.float64, device=self.device)
        ones = torch.zeros(n, n, dtype=torch.float64, device=self.device)
        for (s, m, ab) in self.cons:
            keep[s] = 0.0
            ones[s, s:s + m] = 1.0
            b[:, s] = ab
        A = A * keep[None, :, None] + ones[None]
        return A, b

    def assemble_gpu(self, T, A=None, b=None):
        """A(T) and b by hipla_te_assemble, in the flat layouts of ``plan()`` (not differentiable)."""
        T = T.detach().contiguous()
        if A is None:
            A = torch.empty(self.ncells * self.n * self.n, dtype=torch.float64, device=T.device)
        if b is None:
            b = torch.empty(self.ncells * self.n, dtype=torch.float64, device=T.device)
        check(lib().hipla_te_assemble(self.ncells, self.n, self.nstage, _ptr(self.g_row), _ptr(self.g_C),
                                      _ptr(self.g_E), _ptr(self.g_a0), len(self.cons), _ptr(self.g_cs),
                                      _ptr(self.g_cm), _ptr(self.g_cab), _ptr(T), _ptr(A), _ptr(b),
                                      _stream_handle(None)), "hipla_te_assemble")
        return A, b

    def grad_gpu(self, T, X, Lam, G=None):
        """dL/dT_e = -lambda^T (dA/dT) x per cell by hipla_te_grad."""
        T = T.detach().contiguous()
        if G is None:
            G = torch.empty(self.ncells, dtype=torch.float64, device=T.device)
        check(lib().hipla_te_grad(self.ncells, self.n, self.nstage, _ptr(self.g_row), _ptr(self.g_C),
                                  _ptr(self.g_E), _ptr(self.g_a0), len(self.cons), _ptr(self.g_cs), _ptr(T),
                                  _ptr(X), _ptr(Lam), _ptr(G), _stream_handle(None)), "hipla_te_grad")
        return G

    def plan(self):
        return Plan(np.full(self.ncells, self.n, dtype=np.int32), self.offA, self.offV, mode="fp64", graph=False,
                    keep_factors=True)


class TeSolveFn(torch.autograd.Function):
    """x(T_e) per cell: GPU assembly, the batched FP64 solve with kept factors; backward: the adjoint
    lambda (no dense dL/dA) and hipla_te_grad."""

    @staticmethod
    def forward(ctx, T, model, plan):
        A, b = model.assemble_gpu(T)
        X, codes, _ = plan.solve(A, b, torch.zeros_like(b))
        ctx.model, ctx.plan, ctx.generation = model, plan, plan.generation
        ctx.save_for_backward(T, X)
        ctx.mark_non_differentiable(codes)
        return X, codes

    @staticmethod
    def backward(ctx, gX, gcodes):
        T, X = ctx.saved_tensors
        plan = ctx.plan
        if plan.generation != ctx.generation:
            raise RuntimeError("the plan was solved again after this forward pass: its kept LU factors belong to a "
                               "later solve (run backward first, or use one plan per forward)")
        Lam, _ = plan.adjoint(X, gX.contiguous())
        return ctx.model.grad_gpu(T, X, Lam), None, None


def loss_and_grad(model: IonTe, plan: Plan, T, w, fused: bool = True):
    """Cooling proxy L_c = sum_i w_ci x_ci per cell, and dL_c/dT_c through the adjoint.

    fused=True: GPU assembly and dL/dT_e straight from lambda (TeSolveFn); fused=False: torch
    assembly, the dense dL/dA of SolveFn and torch's chain rule through the assembly."""
    T = T.detach().clone().requires_grad_(True)
    if fused:
        X, codes = TeSolveFn.apply(T, model, plan)
    else:
        A, b = model.assemble(T)
        X, codes = solve(A.reshape(-1), b.reshape(-1), plan)
    L = (X.view(model.ncells, model.n) * w).sum(dim=1)
    (g,) = torch.autograd.grad(L.sum(), T)
    return L.detach(), g.detach(), codes


def newton_te(model: IonTe, plan: Plan, T0, w, L_target):
    """One Newton step on T_e per cell toward L(T) = L_target."""
    L0, dL, codes = loss_and_grad(model, plan, T0, w)
    step = (L0 - L_target) / dL
    return T0 - step, L0, dL
