"""Adjoint at its full shape: mu = sum_{p<64} c_p M_p with M_p 3 x 2000 x
2000 (6.1 GB in HBM), J_I = 10 -> J_F = 11 in the two-irrep pair scheme
(N_I = 21000, N_F = 23000), 500 initial x 2000 final states with
orthonormal records from QR of a Gaussian.  Loss
L = sum_if w_if (log S_if - log S*_if)^2 with S* from a perturbed c, over
the strong lines (S > 1e-6 of the largest, where log S is well conditioned).

Checks of the hand-written backward (gain_adjoint_dipole + the contraction
with the M_p):
  * dL/dc against central differences (h = 1e-6) along four single
    coefficients and three random directions;
  * torch.autograd.gradcheck of L(c);
  * dL/dmu itself, along random dipole directions orthogonal to every M_p
    (invisible to dL/dc), against central differences of L(mu).
The forward is the batched FP64 path (dipole_GPU.cu:392-515,
GpuManager.cpp:733-746, Output.cpp:78-93), whose S the parity tests pin."""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

H = 1e-6


@pytest.fixture(scope="module")
def wl():
    from gain_engine import workload
    w = workload.build_adjoint()
    yield w
    w.close()


@pytest.fixture(scope="module")
def setup(wl):
    import torch

    from gain_engine.autograd import LineStrengths
    eng, M, nc = wl.eng, wl.M, wl.nc
    blkI = wl.prob.blocks[(0, 0)]
    tasks = [(0, 0, list(range(blkI.n)), 1, 1)]
    c0 = torch.tensor(wl.coef, dtype=torch.float64)
    with torch.no_grad():
        S0 = LineStrengths.apply(c0, eng, M, nc, tasks)
        cs = c0 + 0.05 * torch.tensor(
            np.random.default_rng(9).standard_normal(c0.shape[0]))
        Ss = LineStrengths.apply(cs, eng, M, nc, tasks)
    mask = (S0 > 1e-6 * S0.max()) & (Ss > 1e-6 * Ss.max())
    assert int(mask.sum()) > 100000
    w = torch.rand(int(mask.sum()), dtype=torch.float64,
                   generator=torch.Generator().manual_seed(2)).to(S0.device)
    target = torch.log(Ss[mask])

    def loss_of_S(S):
        return (w * (torch.log(S[mask]) - target) ** 2).sum()

    def loss(c):
        return loss_of_S(LineStrengths.apply(c, eng, M, nc, tasks))
    return c0, loss, loss_of_S, tasks


def _fd(f, x, d):
    import torch
    with torch.no_grad():
        return (float(f(x + H * d)) - float(f(x - H * d))) / (2 * H)


def test_adjoint_fullsize_dc_central_differences(setup):
    import torch
    c0, loss, _, _ = setup
    c = c0.clone().requires_grad_(True)
    (g,) = torch.autograd.grad(loss(c), c)
    gn = float(g.norm())
    dirs = []
    for p in (0, 17, 40, 63):
        e = torch.zeros_like(c0)
        e[p] = 1.0
        dirs.append(e)
    rng = np.random.default_rng(4)
    for _ in range(3):
        d = torch.tensor(rng.standard_normal(c0.shape[0]))
        dirs.append(d / d.norm())
    worst = 0.0
    for d in dirs:
        an = float(g @ d)
        fd = _fd(loss, c0, d)
        err = abs(fd - an) / (abs(an) + 1e-6 * gn)
        worst = max(worst, err)
        assert abs(fd - an) <= 1e-6 * abs(an) + 1e-8 * gn, (an, fd)
    print(f"dL/dc vs central differences: worst rel err {worst:.2e}")


def test_adjoint_fullsize_gradcheck(setup):
    import torch
    c0, loss, _, _ = setup
    c = c0.clone().requires_grad_(True)
    assert torch.autograd.gradcheck(loss, (c,), eps=H, atol=1e-8,
                                    rtol=1e-6)


def test_adjoint_fullsize_dmu_outside_span_central_differences(wl, setup):
    """Directional derivatives of L(mu) along dipole directions Delta with
    <Delta, M_p> = 0 for every p: the part of dL/dmu that dL/dc never
    sees."""
    import torch
    c0, _, loss_of_S, tasks = setup
    eng, M, nc = wl.eng, wl.M, wl.nc
    dev = M.device
    mu = (torch.as_tensor(wl.coef, device=dev) @ M)
    # torch's results must be complete before the engine's stream reads them
    torch.cuda.synchronize()
    # dL/dS at mu, then dL/dmu by the hand-written adjoint
    eng.set_dipole_combination(np.ones(1), mu.data_ptr(), nc)
    S = torch.zeros((len(tasks[0][2]), wl.prob.blocks[(1, 1)].n),
                    dtype=torch.float64, device=dev)
    eng.lines_device(0, 0, tasks[0][2], 1, 1, {"S": S.data_ptr()})
    eng.synchronize()
    Sl = S.reshape(-1).clone().requires_grad_(True)
    (dS,) = torch.autograd.grad(loss_of_S(Sl), Sl)
    dS = dS.contiguous()
    dmu = torch.zeros(3 * nc * nc, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    eng.adjoint(0, 0, tasks[0][2], 1, 1, dS.data_ptr(), dmu.data_ptr())
    eng.synchronize()
    # random damped directions, projected out of span{M_p}
    from gain_engine import workload
    D = workload.basis_matrices_device(3, nc, 77, dev).reshape(3, -1)
    Q, _ = torch.linalg.qr(M.T)             # orthonormal basis of the span
    D = D - (D @ Q) @ Q.T
    assert float((D @ M.T).abs().max()) < 1e-10 * float(D.norm())
    gn = float(dmu.norm())
    for k in range(3):
        # scaled to the norm of mu: a relative step of h on the dipole
        d = (D[k] * (mu.norm() / D[k].norm())).contiguous()
        dn = float(d.norm())
        an = float(dmu @ d)
        pair = torch.stack([mu, d]).contiguous()
        torch.cuda.synchronize()

        def L_at(t):
            eng.set_dipole_combination(np.array([1.0, t]), pair.data_ptr(),
                                       nc)
            eng.lines_device(0, 0, tasks[0][2], 1, 1, {"S": S.data_ptr()})
            eng.synchronize()
            return float(loss_of_S(S.reshape(-1)))
        fd = (L_at(H) - L_at(-H)) / (2 * H)
        print(f"direction {k}: adjoint {an:.9e} central difference "
              f"{fd:.9e}")
        assert abs(fd - an) <= 1e-6 * abs(an) + 1e-8 * gn * dn, (an, fd)
    # dL/dc is the contraction of the same dL/dmu with the M_p; and the
    # state set above reproduces the forward of the autograd harness
    eng.set_dipole_combination(np.ones(1), mu.data_ptr(), nc)
    eng.lines_device(0, 0, tasks[0][2], 1, 1, {"S": S.data_ptr()})
    eng.synchronize()
    assert float(loss_of_S(S.reshape(-1))) == pytest.approx(
        float(setup[1](c0)), rel=1e-12)
    dc = eng.dipole_coeff_grad(M.shape[0], M.data_ptr(), dmu.data_ptr())
    np.testing.assert_allclose(dc, (M @ dmu).cpu().numpy(), rtol=1e-10,
                               atol=1e-12 * float(np.abs(dc).max()))
