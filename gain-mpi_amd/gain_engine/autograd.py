"""torch.autograd wrapper of the engine's adjoint (test harness).

mu = sum_p c_p M_p is formed on the device (gain_set_dipole_combination);
the forward returns the line strengths S of a set of (initial batch, final
block) tasks, the backward calls the hand-written adjoint
(gain_adjoint_dipole: dL/dmu as a rank-k update) and contracts it with the
basis matrices (gain_dipole_coeff_grad: dL/dc_p = <M_p, dL/dmu>).
"""
from __future__ import annotations

import numpy as np
import torch


class LineStrengths(torch.autograd.Function):
    """S(c) for tasks [(jindI, gammaI, levels, jindF, gammaF), ...]; output
    is the concatenation of the [nI, nF] S arrays (float64, device).  With
    reuse (the default) the forward keeps each task's pair line strengths
    and flags and the backward hands them to gain_adjoint_dipole_ls, which
    does not recompute the forward; reuse=False recomputes it
    (gain_adjoint_dipole)."""

    @staticmethod
    def forward(ctx, coef, eng, M, nc, tasks, reuse=True):
        # (the engine's calls wait for torch's stream: Engine.wait_for_torch)
        eng.set_dipole_combination(coef.detach().cpu().numpy(),
                                   M.data_ptr(), nc)
        outs, kept = [], []
        md2 = eng.maxdeg * eng.maxdeg
        for (jI, gI, levels, jF, gF) in tasks:
            nF = eng.blocks[(jF, gF)][0]
            S = torch.zeros((len(levels), nF), dtype=torch.float64,
                            device=M.device)
            ptrs = {"S": S.data_ptr()}
            if reuse:
                ls = torch.zeros((len(levels), nF, md2), dtype=torch.float64,
                                 device=M.device)
                fl = torch.zeros((len(levels), nF), dtype=torch.uint8,
                                 device=M.device)
                ptrs |= {"ls": ls.data_ptr(), "flags": fl.data_ptr()}
                kept.append((ls, fl))
            eng.lines_device(jI, gI, levels, jF, gF, ptrs)
            outs.append(S.reshape(-1))
        eng.synchronize()
        ctx.eng, ctx.M, ctx.nc, ctx.tasks = eng, M, nc, tasks
        ctx.kept = kept if reuse else None
        ctx.save_for_backward(coef)
        return torch.cat(outs)

    @staticmethod
    def backward(ctx, gS):
        (coef,) = ctx.saved_tensors
        eng, M, nc = ctx.eng, ctx.M, ctx.nc
        if ctx.kept is None:
            # the recomputed forward needs mu = sum_p c_p M_p again
            eng.set_dipole_combination(coef.detach().cpu().numpy(),
                                       M.data_ptr(), nc)
        dmu = torch.zeros(3 * nc * nc, dtype=torch.float64, device=M.device)
        gS = gS.contiguous().to(M.device)
        off = 0
        for t, (jI, gI, levels, jF, gF) in enumerate(ctx.tasks):
            nF = eng.blocks[(jF, gF)][0]
            n = len(levels) * nF
            part = gS[off:off + n].contiguous()
            if ctx.kept is None:
                eng.adjoint(jI, gI, levels, jF, gF, part.data_ptr(),
                            dmu.data_ptr(), on_device=True, accumulate=True)
            else:
                ls, fl = ctx.kept[t]
                eng.adjoint(jI, gI, levels, jF, gF, part.data_ptr(),
                            dmu.data_ptr(), on_device=True, accumulate=True,
                            ls=ls.data_ptr(), flags=fl.data_ptr())
            off += n
        eng.synchronize()
        dc = eng.dipole_coeff_grad(coef.shape[0], M.data_ptr(),
                                   dmu.data_ptr())
        return (torch.as_tensor(dc, dtype=coef.dtype), None, None, None,
                None, None)


def make_basis_matrices(P: int, nc: int, seed: int = 7) -> np.ndarray:
    """P dipole basis matrices M_p (3 x nc x nc, the dipole layout), each
    Gaussian and damped like the synthetic dipole."""
    from .synth import make_mu
    return np.stack([make_mu(nc, seed + p) for p in range(P)])
