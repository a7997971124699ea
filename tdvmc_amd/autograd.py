"""torch.autograd.Function over the library's ln psi and its adjoint.

    logpsi = LogPsi.apply(ctx, R, uR)     # R: [W, N, D] float64, uR: [P] float64

forward  : Re ln psi of every configuration (tdvmc_eval_configs);
backward : d/dR = gbar * Re grad ln psi, d/duR = sum_w gbar_w d ln psi_w / d alpha
           (tdvmc_adjoint).  Used as a test harness: torch.autograd.gradcheck
           compares it with finite differences of the forward.
Tensors are CPU float64; the library moves them to the GPU itself.
"""
from __future__ import annotations

import numpy as np
import torch


class LogPsi(torch.autograd.Function):
    @staticmethod
    def forward(fctx, tctx, R, uR, precision=0):
        uR_np = uR.detach().cpu().numpy().astype(np.float64)
        _, uI, phiR, phiI = tctx.get_params()
        tctx.set_params(uR_np, uI, phiR, phiI)
        out = tctx.eval_configs(R.detach().cpu().numpy(), precision=precision)
        fctx.tctx = tctx
        fctx.precision = precision
        fctx.save_for_backward(R, uR)
        return torch.from_numpy(out["expo"].copy())

    @staticmethod
    def backward(fctx, gout):
        R, uR = fctx.saved_tensors
        tctx = fctx.tctx
        _, uI, phiR, phiI = tctx.get_params()
        tctx.set_params(uR.detach().cpu().numpy(), uI, phiR, phiI)
        dR, da = tctx.adjoint(R.detach().cpu().numpy(), gout.detach().cpu().numpy().astype(np.float64),
                              precision=fctx.precision)
        return None, torch.from_numpy(dR).reshape(R.shape), torch.from_numpy(da), None
