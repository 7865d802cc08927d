"""Gradients of a spectrum loss with respect to (K, J, Gamma, h).

SpectrumLoss evaluates L(theta) = sum_q (S(w_q) - target_q)^2 with
S(w) = Im <phi|(w + E0 - i eta - H(theta))^-1|phi>, phi = O|gs(theta)>, and
its gradient through the library's adjoint (esc_spectrum_grad).
spectrum_loss(theta, target, spec) wraps it as a torch.autograd.Function.
TableLoss does the same for a loss over a whole JsJs or JeJe table
(esc_table_grad).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._capi import DYN_KINDS, OPS, check, lib
from .engine import Honeycomb, SpinModel, _np_ptr, field111


class SpectrumLoss:
    """Fixed lattice, probe operator and spectral settings; theta varies.

    theta = (K, J, Gamma, h): K on all three bond families, Heisenberg J,
    Gamma, field strength h along [111]."""

    def __init__(self, lattice: Honeycomb, omega, eta=0.1, op="Jex", site=0, gs_steps=300, cf_steps=120,
                 checkpoint=10):
        self.lattice = lattice
        self.omega = np.ascontiguousarray(omega, np.float64)
        self.eta = float(eta)
        self.op = op
        self.site = int(site)
        self.gs_steps = int(gs_steps)
        self.cf_steps = int(cf_steps)
        self.checkpoint = int(checkpoint)

    def model(self, theta) -> SpinModel:
        K, J, G, h = (float(x) for x in theta)
        return SpinModel(self.lattice, K=(K, K, K), J=J, Gamma=G, B=field111(h))

    def evaluate(self, theta, target):
        """Returns dict(loss, grad[4], spectrum[nw], e0, de0[4])."""
        target = np.ascontiguousarray(target, np.float64)
        if target.shape != self.omega.shape:
            raise ValueError("target must have one value per frequency")
        m = self.model(theta)
        st = m.struct()
        loss, e0 = ctypes.c_double(), ctypes.c_double()
        spec = np.zeros(len(self.omega))
        grad = np.zeros(4)
        de0 = np.zeros(4)
        check(lib().esc_spectrum_grad(ctypes.byref(st), OPS[self.op], self.site, self.gs_steps, self.cf_steps,
                                      self.checkpoint, self.eta, _np_ptr(self.omega), len(self.omega),
                                      _np_ptr(target), ctypes.byref(loss), _np_ptr(spec), _np_ptr(grad),
                                      ctypes.byref(e0), _np_ptr(de0), None))
        return {"loss": loss.value, "grad": grad, "spectrum": spec, "e0": e0.value, "de0": de0}


class TableLoss(SpectrumLoss):
    """A loss over the whole JsJs or JeJe table (esc_table_grad):
    L = sum over frequencies and columns of (T - target)^2, T the table the
    CF-mode spectra write (JsJs: 14 columns, JeJe: 5, without the omega
    column); target has shape (nw, ncols)."""

    def __init__(self, lattice: Honeycomb, omega, eta=0.1, kind="JeJe", gs_steps=300, cf_steps=120, checkpoint=10):
        super().__init__(lattice, omega, eta=eta, gs_steps=gs_steps, cf_steps=cf_steps, checkpoint=checkpoint)
        if kind not in ("JsJs", "JeJe"):
            raise ValueError("kind must be JsJs or JeJe")
        self.kind = kind
        self.ncols = 14 if kind == "JsJs" else 5

    def evaluate(self, theta, target):
        """Returns dict(loss, grad[4], table (nw, 1 + ncols), e0, de0[4])."""
        target = np.ascontiguousarray(target, np.float64)
        if target.shape != (len(self.omega), self.ncols):
            raise ValueError(f"target must have shape ({len(self.omega)}, {self.ncols})")
        m = self.model(theta)   # holds the arrays the struct points into
        st = m.struct()
        loss, e0 = ctypes.c_double(), ctypes.c_double()
        table = np.zeros((len(self.omega), 1 + self.ncols))
        grad = np.zeros(4)
        de0 = np.zeros(4)
        check(lib().esc_table_grad(ctypes.byref(st), DYN_KINDS[self.kind], self.gs_steps, self.cf_steps,
                                   self.checkpoint, self.eta, _np_ptr(self.omega), len(self.omega), _np_ptr(target),
                                   ctypes.byref(loss), _np_ptr(table), _np_ptr(grad), ctypes.byref(e0), _np_ptr(de0),
                                   None))
        return {"loss": loss.value, "grad": grad, "table": table, "e0": e0.value, "de0": de0}


def _function():
    import torch

    class SpectrumLossFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, target, spec):
            r = spec.evaluate(theta.detach().cpu().numpy(), target.detach().cpu().numpy())
            s = torch.as_tensor(r["spectrum"], dtype=target.dtype, device=target.device)
            ctx.save_for_backward(torch.as_tensor(r["grad"], dtype=theta.dtype, device=theta.device),
                                  s - target.detach())
            return torch.as_tensor(r["loss"], dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, g):
            grad, resid = ctx.saved_tensors
            return g * grad, -2.0 * g * resid, None

    return SpectrumLossFunction


_FN = None


def spectrum_loss(theta, target, spec: SpectrumLoss):
    """Differentiable loss (torch): gradients flow to theta and target."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(theta, target, spec)
