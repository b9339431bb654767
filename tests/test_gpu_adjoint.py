"""Adjoint (dL/dmu, dL/dc for mu = sum_p c_p M_p) through the C ABI,
checked against torch.autograd.gradcheck (central finite differences of the
engine's own forward), explicit central differences with h = 1e-6, and an
FP64 reference gradient from a dense torch model assembled from the
textbook rotational factor (trove_wrappers.f90:329-381)."""
import numpy as np
import pytest

from gain_engine import synth
from gain_engine.engine import Engine
from oracle import oracle as orc

from helpers import SQRT2_REF_INV, textbook_rot

pytestmark = pytest.mark.gpu

NC, P = 16, 4


@pytest.fixture(scope="module")
def setup():
    import torch

    from gain_engine.autograd import make_basis_matrices
    prob = synth.make_problem([2, 3], NC, "pair2", blocks={(0, 0), (1, 1)})
    prob.params.erange_lower = (0.0, 1e9)
    M = make_basis_matrices(P, NC)
    rng = np.random.default_rng(3)
    coef = rng.standard_normal(P)
    prob.mu = np.tensordot(coef, M, 1)
    eng = Engine.from_problem(prob)
    Md = torch.as_tensor(M.reshape(P, -1), device="cuda:0").contiguous()
    levels = list(range(0, prob.blocks[(0, 0)].n, 3))
    tasks = [(0, 0, levels, 1, 1)]
    yield prob, eng, M, Md, coef, tasks
    eng.close()


def _loss_fn(eng, Md, tasks, target, w, mask):
    import torch

    from gain_engine.autograd import LineStrengths

    def loss(c):
        S = LineStrengths.apply(c, eng, Md, NC, tasks)
        Sm = S[mask]
        return (w * (torch.log(Sm) - target) ** 2).sum()
    return loss


def _prepare(setup):
    import torch

    from gain_engine.autograd import LineStrengths
    prob, eng, M, Md, coef, tasks = setup
    c = torch.tensor(coef, dtype=torch.float64, requires_grad=True)
    with torch.no_grad():
        S0 = LineStrengths.apply(c, eng, Md, NC, tasks)
        cstar = c + 0.05 * torch.tensor(
            np.random.default_rng(5).standard_normal(P))
        Sstar = LineStrengths.apply(cstar, eng, Md, NC, tasks)
    mask = (S0 > 0) & (Sstar > 0)
    w = torch.rand(int(mask.sum()), dtype=torch.float64,
                   generator=torch.Generator().manual_seed(0)).to(S0.device)
    target = torch.log(Sstar[mask])
    return c, _loss_fn(eng, Md, tasks, target, w, mask), mask, w, target


def test_adjoint_gradcheck(setup):
    import torch
    c, loss, _, _, _ = _prepare(setup)
    assert torch.autograd.gradcheck(loss, (c,), eps=1e-6, atol=1e-7,
                                    rtol=1e-6)


def test_adjoint_matches_central_differences(setup):
    import torch
    c, loss, _, _, _ = _prepare(setup)
    L = loss(c)
    (g,) = torch.autograd.grad(L, c)
    h = 1e-6
    for p in range(P):
        e = torch.zeros(P, dtype=torch.float64)
        e[p] = h
        fd = (loss(c.detach() + e) - loss(c.detach() - e)) / (2 * h)
        assert float(g[p]) == pytest.approx(float(fd), rel=1e-6, abs=1e-9)


def test_adjoint_matches_dense_fp64_reference(setup):
    """dL/dmu and dL/dc from a dense torch FP64 model of the same path."""
    import torch
    prob, eng, M, Md, coef, tasks = setup
    c, loss, mask, w, target = _prepare(setup)
    L = loss(c)
    (g_engine,) = torch.autograd.grad(L, c)

    bI, bF = prob.bases
    blkI, blkF = prob.blocks[(0, 0)], prob.blocks[(1, 1)]
    levels = tasks[0][2]
    tab = orc.threej_table(prob.jmax)
    kI, kF = bI.kval(), bF.kval()
    R = np.zeros((3, bF.nprim, bI.nprim))
    for rp in range(bF.nprim):
        for r in range(bI.nprim):
            if abs(int(kI[r]) - int(kF[rp])) > 1:
                continue
            f3j = tab[orc.threej_index(prob.jmax, bI.J, int(kI[r]), bF.J,
                                       int(kF[rp]))]
            a, v = textbook_rot(bI.J, bF.J, int(kI[r]), int(bI.tau[r]),
                                int(kF[rp]), int(bF.tau[rp]) & 1, f3j,
                                sqrt2=SQRT2_REF_INV)
            R[a, rp, r] = v
    XI = np.stack([orc.expand(bI, 0, 0, blkI.vector(l)) for l in levels])
    XF = np.stack([orc.expand(bF, 1, 0, blkF.vector(f))
                   for f in range(blkF.n)])
    dev = "cuda:0"
    Mt = torch.as_tensor(M, device=dev)
    ct = torch.tensor(coef, dtype=torch.float64, device=dev,
                      requires_grad=True)
    mu = torch.tensordot(ct, Mt, 1)  # [3][vI][vF]
    vI = torch.as_tensor(bI.vib, device=dev).long()
    vF = torch.as_tensor(bF.vib, device=dev).long()
    # D[a][r'][r] = R[a][r'][r] * mu[a][vib(r)][vib(r')]
    mug = mu[:, vI][:, :, vF].permute(0, 2, 1)
    D = (torch.as_tensor(R, device=dev) * mug).sum(0)
    H = torch.as_tensor(XI, device=dev) @ D.T          # [nI][nprimF]
    ls = H @ torch.as_tensor(XF, device=dev).T         # [nI][nF]
    S = (ls * ls).reshape(-1)
    Lref = (w * (torch.log(S[mask]) - target) ** 2).sum()
    (g_ref,) = torch.autograd.grad(Lref, ct)
    assert float(Lref) == pytest.approx(float(L), rel=1e-10)
    np.testing.assert_allclose(g_engine.numpy(), g_ref.cpu().numpy(),
                               rtol=1e-9, atol=1e-12 * float(
                                   g_ref.abs().max()))


def _dense_model(prob, levels, jI=0, gI=0, jF=1, gF=1):
    """Dense FP64 operators of the forward (reference layout of mu): the
    textbook rotational factor R[alpha][r'][r] and the expanded records."""
    bI, bF = prob.bases[jI], prob.bases[jF]
    blkI, blkF = prob.blocks[(jI, gI)], prob.blocks[(jF, gF)]
    tab = orc.threej_table(prob.jmax)
    kI, kF = bI.kval(), bF.kval()
    R = np.zeros((3, bF.nprim, bI.nprim))
    for rp in range(bF.nprim):
        for r in range(bI.nprim):
            if abs(int(kI[r]) - int(kF[rp])) > 1:
                continue
            f3j = tab[orc.threej_index(prob.jmax, bI.J, int(kI[r]), bF.J,
                                       int(kF[rp]))]
            a, v = textbook_rot(bI.J, bF.J, int(kI[r]), int(bI.tau[r]),
                                int(kF[rp]), int(bF.tau[rp]) & 1, f3j,
                                sqrt2=SQRT2_REF_INV)
            R[a, rp, r] = v
    XI = np.stack([orc.expand(bI, gI, 0, blkI.vector(l)) for l in levels])
    XF = np.stack([orc.expand(bF, gF, 0, blkF.vector(f))
                   for f in range(blkF.n)])
    return R, XI, XF


def _dense_S(R, XI, XF, mu, vI, vF):
    """S[i][F] = (x_F^T D x_I)^2 with D[r'][r] = sum_a R[a][r'][r]
    mu[a][vib(r)][vib(r')] (one-dimensional irreps)."""
    mug = mu[:, vI][:, :, vF].permute(0, 2, 1)
    D = (R * mug).sum(0)
    return (XI @ D.T @ XF.T) ** 2


def test_adjoint_dmu_elementwise_vs_dense_autograd():
    """dL/dmu of every dipole element (all 3 nc^2, not only its projection
    onto a few basis matrices) against torch autograd through the dense FP64
    model with mu itself as the leaf, for an unmasked loss linear in S and a
    frequency window that starts below zero, so pairs with
    fw0 <= nu < 1e-40 pass the filters but have S = 0 (Output.cpp:65): their
    dS must not reach dL/dmu."""
    import torch
    prob = synth.make_problem([2, 3], NC, "pair2", blocks={(0, 0), (1, 1)})
    prob.params.erange_lower = (0.0, 1e9)
    prob.params.freq_window = (-20000.0, 15000.0)
    # a final state degenerate with an initial one: nu = 0 passes the
    # window but forms no line
    blkI, blkF = prob.blocks[(0, 0)], prob.blocks[(1, 1)]
    blkF.energy[3] = blkI.energy[4]
    blkF.energy.sort()
    levels = list(range(0, blkI.n, 2))
    with Engine.from_problem(prob) as eng:
        nF = blkF.n
        dev = "cuda:0"
        res = eng.batched_lines(0, 0, levels, 1, 1)
        flags, S_eng = res["flags"], res["S"]
        nu = blkF.energy[None, :] - blkI.energy[levels][:, None]
        passed = (flags & 1) != 0
        formed = passed & (nu >= 1e-40)
        assert (passed & ~formed).any(), "no filtered pair with nu < 1e-40"
        assert (S_eng[passed & ~formed] == 0).all()
        w = torch.rand((len(levels), nF), dtype=torch.float64,
                       generator=torch.Generator().manual_seed(1))
        wd = w.to(dev).contiguous()
        dmu = torch.zeros(3 * NC * NC, dtype=torch.float64, device=dev)
        eng.adjoint(0, 0, levels, 1, 1, wd.data_ptr(), dmu.data_ptr())
        eng.synchronize()
        R, XI, XF = _dense_model(prob, levels)
        mu = torch.tensor(prob.mu, dtype=torch.float64, device=dev,
                          requires_grad=True)
        vI = torch.as_tensor(prob.bases[0].vib, device=dev).long()
        vF = torch.as_tensor(prob.bases[1].vib, device=dev).long()
        T = lambda a: torch.as_tensor(a, device=dev)
        S = _dense_S(T(R), T(XI), T(XF), mu, vI, vF)
        np.testing.assert_allclose(S_eng[formed],
                                   S.detach().cpu().numpy()[formed],
                                   rtol=1e-10, atol=1e-14 * S_eng.max())
        L = (wd * S * T(formed)).sum()
        (g,) = torch.autograd.grad(L, mu)
        g = g.reshape(-1)
        scale = float(g.abs().max())
        assert scale > 0
        np.testing.assert_allclose(dmu.cpu().numpy(), g.cpu().numpy(),
                                   rtol=1e-9, atol=1e-12 * scale)


def test_autograd_reuse_equals_recompute(setup):
    """dL/dc with the forward's line strengths handed to the backward
    (gain_adjoint_dipole_ls) equals dL/dc with the forward recomputed
    inside the adjoint (gain_adjoint_dipole)."""
    import torch

    from gain_engine.autograd import LineStrengths
    prob, eng, M, Md, coef, tasks = setup
    c, _, mask, w, target = _prepare(setup)
    grads = []
    for reuse in (True, False):
        S = LineStrengths.apply(c, eng, Md, NC, tasks, reuse)
        L = (w * (torch.log(S[mask]) - target) ** 2).sum()
        (g,) = torch.autograd.grad(L, c)
        grads.append(g.numpy())
    assert np.abs(grads[1]).max() > 0
    np.testing.assert_allclose(grads[0], grads[1], rtol=1e-12,
                               atol=1e-14 * np.abs(grads[1]).max())


@pytest.mark.parametrize("gI,gF", [(2, 2), (3, 4)])
def test_adjoint_given_ls_degenerate_irreps(gI, gF):
    """Td irreps E -> E and F1 -> F2 (two and three degenerate components,
    so each pair's line strengths sit at component slots): dL/dmu from the
    forward's ls and flags equals the recomputed one, for host and device
    memory, and <dL/dmu, D> equals the central difference of L = sum w S
    along a dipole direction D, which is exact up to rounding since S is
    quadratic in mu."""
    import torch
    nc = 20
    prob = synth.make_problem([2, 3], nc, "td5")
    prob.params.erange_lower = (0.0, 1e9)
    n0 = prob.blocks[(0, gI)].n
    levels = list(range(0, n0, 2))
    dev = "cuda:0"
    with Engine.from_problem(prob) as eng:
        nF = eng.blocks[(1, gF)][0]
        res = eng.batched_lines(0, gI, levels, 1, gF)
        assert (res["flags"] & 1).sum() > 0, "no pair passes the filters"
        w = np.random.default_rng(11).random((len(levels), nF))
        nmu = 3 * nc * nc
        d_re = np.zeros(nmu)
        d_ls = np.zeros(nmu)
        eng.adjoint(0, gI, levels, 1, gF, w, d_re, on_device=False)
        eng.adjoint(0, gI, levels, 1, gF, w, d_ls, on_device=False,
                    ls=res["ls"], flags=res["flags"])
        scale = np.abs(d_re).max()
        assert scale > 0
        np.testing.assert_allclose(d_ls, d_re, rtol=1e-12,
                                   atol=1e-14 * scale)
        # device memory
        wd = torch.as_tensor(w, device=dev).contiguous()
        lsd = torch.as_tensor(np.ascontiguousarray(res["ls"]),
                              device=dev).contiguous()
        fld = torch.as_tensor(np.ascontiguousarray(res["flags"]),
                              device=dev).contiguous()
        dd = torch.full((nmu,), 7.0, dtype=torch.float64, device=dev)
        eng.adjoint(0, gI, levels, 1, gF, wd.data_ptr(), dd.data_ptr(),
                    ls=lsd.data_ptr(), flags=fld.data_ptr())
        eng.synchronize()
        np.testing.assert_allclose(dd.cpu().numpy(), d_re, rtol=1e-12,
                                   atol=1e-14 * scale)
        # central difference along D through mu(t) = mu + t D
        D = np.random.default_rng(12).standard_normal(prob.mu.shape)
        Mt = torch.as_tensor(np.stack([prob.mu.reshape(-1), D.reshape(-1)]),
                             device=dev).contiguous()
        h = 0.5
        Lh = []
        for t in (h, -h):
            eng.set_dipole_combination(np.array([1.0, t]), Mt.data_ptr(), nc)
            Lh.append(float((w * eng.batched_lines(0, gI, levels, 1,
                                                   gF)["S"]).sum()))
        fd = (Lh[0] - Lh[1]) / (2 * h)
        ad = float(d_re @ D.reshape(-1))
        assert ad == pytest.approx(fd, rel=1e-10, abs=1e-12 * abs(Lh[0]))
