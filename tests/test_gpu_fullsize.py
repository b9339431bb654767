"""Parity at the full J_I=20 -> J_F=21, Nc=4000 size (N_I=41000,
N_F=n_F=43000, 14.8 GB of final records in HBM).

The oracle cannot run every pair at this size, so the checks are:
  * transition flags of every pair equal the filter evaluated on the host
    with the reference's floating-point operations (bit-exact);
  * S of sampled pairs against the oracle (double-double) at 1e-12;
  * the sum rule sum_F ls_F^2 = ||P_F^T h||^2 over the complete orthonormal
    final block (wide windows), with h from the oracle;
  * batched FP64 equals plain FP64 (1e-12); the reduced-precision modes
    keep the transition list and their S errors against FP64 are reported
    and bounded.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


@pytest.fixture(scope="module")
def wl():
    from gain_engine import workload
    w = workload.build(nc=4000, ninitial=64)
    yield w
    w.close()


def _oracle_h(w, level):
    prob = w.prob
    bI, bF = prob.bases
    cI = w.CI[level].cpu().numpy()
    tab = orc.threej_table(prob.jmax)
    x = orc.expand(bI, 0, 0, cI)
    return orc.half_ls(prob, bI, bF, x, tab)


def test_fullsize_plain_flags_and_sampled_oracle(wl):
    from gain_engine import workload
    prob, eng = wl.prob, wl.eng
    bF = prob.bases[1]
    eF = wl.blkF.energy
    rng = np.random.default_rng(0)
    for level in (0, 37):
        res = eng.plain_lines(0, 0, level, 1, 1, want_ls=True)
        exp = workload.expected_flags(prob, wl.blkI.energy[level], eF)
        assert np.array_equal((res["flags"] & 1).astype(bool), exp)
        h = _oracle_h(wl, level)
        sample = np.sort(rng.choice(np.nonzero(exp)[0], 300, replace=False))
        rows = wl.CF[sample].cpu().numpy()
        ls_ref = np.array([orc.dot(orc.expand(bF, 1, 0, cF), h)
                           for cF in rows])
        ls_got = res["ls"][sample, 0]
        # condition-normalised: 1e-12 relative plus 1e-14 of the largest
        scale = np.abs(ls_ref).max()
        assert (np.abs(ls_got - ls_ref) <=
                1e-12 * np.abs(ls_ref) + 1e-14 * scale).all()
        s_ref = ls_ref * ls_ref
        strong = s_ref > 1e-6 * s_ref.max()
        rel = np.abs(res["S"][sample] - s_ref)[strong] / s_ref[strong]
        assert rel.max() <= 1e-12, rel.max()


def test_fullsize_sum_rule(wl):
    from gain_engine import synth
    prob, eng = wl.prob, wl.eng
    wide = synth.Params(**{f.name: getattr(prob.params, f.name)
                           for f in __import__("dataclasses").fields(
                               prob.params)})
    wide.freq_window = (-1e9, 1e9)
    wide.erange_upper = (-1e9, 1e9)
    eng.set_params(wide)
    try:
        res = eng.plain_lines(0, 0, 5, 1, 1, want_ls=True)
        assert (res["flags"] & 1).all()
        h = _oracle_h(wl, 5)
        # ||P^T h||^2 for the Hadamard-quartet expansion of irrep A2: case t
        # holds primitives 4t..4t+3 with lambda = r mod 4 and
        # repres[t + lambda*Ntot] the same for every t
        bF = prob.bases[1]
        rep, nt = bF.repres[1], bF.ntot[1]
        coef = np.array([rep[lam * nt] for lam in range(4)])
        g = h.reshape(-1, 4) @ coef
        lhs = math.fsum((res["ls"][:, 0] ** 2).tolist())
        rhs = math.fsum((g * g).tolist())
        assert lhs == pytest.approx(rhs, rel=1e-12)
    finally:
        eng.set_params(prob.params)


def test_fullsize_batched_fp64_matches_plain(wl):
    """Batched FP64 (MFMA GEMMs) against plain FP64 (GEMV) with the same
    condition-normalised bound as the oracle check above: the two paths sum
    the 43000-term products in different orders, so lines with strong
    cancellation (|ls| far below the largest) differ by more than 1e-12 of
    their own value but stay within 1e-14 of the largest |ls|."""
    eng = wl.eng
    levels = list(range(0, 64, 4))
    b64 = eng.batched_lines(0, 0, levels, 1, 1, precision=0, want_ls=True)
    for i, lv in enumerate(levels[:4]):
        p = eng.plain_lines(0, 0, lv, 1, 1, want_ls=True)
        assert np.array_equal(p["flags"], b64["flags"][i])
        ls_ref, ls_got = p["ls"][:, 0], b64["ls"][i][:, 0]
        scale = np.abs(ls_ref).max()
        assert (np.abs(ls_got - ls_ref) <=
                1e-12 * np.abs(ls_ref) + 1e-14 * scale).all()
        ref = p["S"]
        d = np.abs(b64["S"][i] - ref)
        assert (d <= 2e-12 * ref + 2e-14 * np.sqrt(ref * ref.max())).all()
        strong = ref > 1e-2 * ref.max()
        assert (d[strong] <= 1e-12 * ref[strong]).all()


# Reduced-precision products at full size, against batched FP64: measured
# worst strong-line errors are FP32 ~2.0e-4, split-BF16 x3 ~3.0e-4 and
# x4 ~1.1e-6 (median 2e-9).  None of these reaches 1e-6 on every strong
# line here; split-INT8 with 6 or 7 slices does (the 1e-6 target), and so
# does the CRT form (rows scaled by their 2-norm: 11 moduli ~3.9e-8, 12
# ~2.4e-9; 14 moduli, scaled by the row maxima, ~7e-11).  The bounds pin
# the measured behaviour with a margin.
@pytest.mark.parametrize("precision,max_bound,median_bound", [
    (1, 1e-3, 1e-6), (2, 2e-3, 5e-6), (3, 3e-6, 1e-8),
    (4, 1e-4, 1e-7), (5, 1e-6, 1e-9), (6, 1e-8, 1e-11),
    (7, 2e-7, 1e-9), (8, 2e-8, 1e-10), (9, 1e-9, 1e-11)])
def test_fullsize_reduced_precision_error(wl, precision, max_bound,
                                          median_bound):
    eng = wl.eng
    levels = list(range(0, 64, 4))
    b64 = eng.batched_lines(0, 0, levels, 1, 1, precision=0, want_ls=False)
    bp = eng.batched_lines(0, 0, levels, 1, 1, precision=precision,
                           want_ls=False)
    assert np.array_equal(bp["flags"], b64["flags"])
    lines = (b64["flags"] & 1).astype(bool)
    ref = b64["S"][lines]
    rel = np.abs(bp["S"][lines] - ref) / np.where(ref > 0, ref, 1.0)
    strong = ref > 1e-6 * ref.max()
    print(f"precision {precision} at full size: {strong.sum()} strong "
          f"lines, max rel err {rel[strong].max():.3e}, median "
          f"{np.median(rel[strong]):.3e}; {(~strong).sum()} weak lines, "
          f"max {rel[~strong].max():.3e}")
    assert rel[strong].max() <= max_bound
    assert np.median(rel[strong]) <= median_bound
