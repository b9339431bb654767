"""GPU parity: the HIP engine against the CPU oracle on the same seeded
inputs (transition lists bit-exact, FP64 values within 1e-12 relative on
strong lines).  Runs through the C ABI; needs an MI355X."""
import numpy as np
import pytest

from gain_engine import synth
from gain_engine.engine import Engine
from oracle import oracle as orc

from helpers import (compare_lines, engine_lines, irregular_basis,
                     oracle_map, problem_with_bases)

pytestmark = pytest.mark.gpu


def _run_both(prob, initials, mode="plain", precision=0):
    tr, rows = orc.run(prob, initials)
    with Engine.from_problem(prob) as eng:
        got = engine_lines(eng, prob, initials, mode, precision)
    return got, oracle_map(tr)


def test_threej_table_matches_oracle():
    with Engine(0, 31, 1, 1) as eng:
        t = eng.threej()
    ref = orc.threej_table(31)
    np.testing.assert_allclose(t, ref, rtol=2e-16, atol=1e-300)


def test_config0_plain_j0_to_j1():
    """J_I=0 -> J_F=1, Nc=500, C1: the 50 lowest J=0 states x all finals."""
    prob = synth.make_problem([0, 1], 500, "c1")
    rows = prob.state_table()
    initials = [i for i, r in enumerate(rows) if r[1] == 0][:50]
    got, ref = _run_both(prob, initials)
    assert len(ref) > 60000
    compare_lines(got, ref)


@pytest.mark.parametrize("scheme,jvals,nc", [("c2v4", [2, 3], 16),
                                             ("c1", [1, 2], 10),
                                             ("pair2", [3, 4], 8),
                                             ("c3v3", [1, 2], 12)])
def test_plain_every_direction(scheme, jvals, nc):
    """All state pairs of a two-J problem (dJ = -1, 0, +1, same-block pairs,
    degenerate irreps for c3v3)."""
    prob = synth.make_problem(jvals, nc, scheme)
    p = prob.params
    p.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(0, len(rows), max(1, len(rows) // 40)))
    got, ref = _run_both(prob, initials)
    assert len(ref) > 0
    compare_lines(got, ref)


@pytest.mark.parametrize("scheme,jvals,nc", [("c2v4", [2, 3], 16),
                                             ("c1", [1, 2], 10),
                                             ("pair2", [3, 4], 8),
                                             ("c3v3", [1, 2], 12)])
@pytest.mark.parametrize("precision,bound", [(5, 5e-8), (8, 1e-7)])
def test_batched_int8_every_direction(scheme, jvals, nc, precision, bound):
    """Batched split-INT8 x6 and CRT-form (12 moduli) int8 over all state
    pairs of a two-J problem (dJ = -1, 0, +1, same-block pairs, degenerate
    irreps for c3v3): the transition list and the emitted lines equal the
    oracle's, S within `bound` on strong lines."""
    prob = synth.make_problem(jvals, nc, scheme)
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(0, len(rows), max(1, len(rows) // 40)))
    tr, _ = orc.run(prob, initials)
    ref = oracle_map(tr)
    with Engine.from_problem(prob) as eng:
        got = engine_lines(eng, prob, initials, "batched", precision)
    assert len(ref) > 0 and set(got) == set(ref)
    assert {k for k in got if got[k]["emitted"]} == \
        {k for k in ref if ref[k]["emitted"]}
    es, _, ns, _ = _strong_weak_err(got, ref)
    print(f"{scheme} precision {precision}: {ns} strong lines, max rel err "
          f"{es:.3e}")
    assert es <= bound


@pytest.mark.parametrize("seed", [1, 2])
def test_batched_split_int8_irregular_basis(seed):
    """Split-INT8 x6 on random K-block sizes, tau per primitive and
    repeated vib indices (uneven X columns and sliced rows)."""
    rng = np.random.default_rng(seed)
    nc = 7
    bases = [irregular_basis(2, nc, rng, 12), irregular_basis(3, nc, rng, 12)]
    prob = problem_with_bases(bases, nc, seed)
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(0, len(rows), 3))
    tr, _ = orc.run(prob, initials)
    ref = oracle_map(tr)
    with Engine.from_problem(prob) as eng:
        got = engine_lines(eng, prob, initials, "batched", 5)
    assert set(got) == set(ref)
    es, _, ns, _ = _strong_weak_err(got, ref)
    assert ns > 0 and es <= 5e-8, es


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_plain_irregular_basis(seed):
    """Random K-block sizes, tau per primitive and repeated vib indices."""
    rng = np.random.default_rng(seed)
    nc = 7
    bases = [irregular_basis(2, nc, rng, 12), irregular_basis(3, nc, rng, 12)]
    prob = problem_with_bases(bases, nc, seed)
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(0, len(rows), 3))
    got, ref = _run_both(prob, initials)
    assert len(ref) > 0
    compare_lines(got, ref)


@pytest.mark.parametrize("scheme,jI,jF", [("c1", 2, 3), ("c2v4", 3, 3),
                                          ("c3v3", 2, 1)])
def test_half_ls_matches_oracle(scheme, jI, jF):
    nc = 12
    prob = synth.make_problem(sorted({jI, jF}), nc, scheme)
    iI = prob.jvals.index(jI)
    iF = prob.jvals.index(jF)
    bI, bF = prob.bases[iI], prob.bases[iF]
    rng = np.random.default_rng(4)
    tab = orc.threej_table(prob.jmax)
    with Engine.from_problem(prob) as eng:
        for g in range(len(bI.ntot)):
            if bI.ntot[g] == 0:
                continue
            for d in range(bI.sdeg[g]):
                c = rng.standard_normal(bI.ntot[g])
                h = eng.half_ls(iI, g, d, c, iF, bF.nprim)
                ref = orc.half_ls(prob, bI, bF, orc.expand(bI, g, d, c), tab)
                np.testing.assert_allclose(h, ref, rtol=0,
                                           atol=1e-13 * np.abs(ref).max())


def test_batched_fp64_matches_oracle():
    prob = synth.make_problem([3, 4], 16, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(0, len(rows), 2))
    got, ref = _run_both(prob, initials, mode="batched", precision=0)
    compare_lines(got, ref)


def test_dropin_kernels_match_oracle():
    """gain_transform_vector_primitive / gain_compute_half_linestrength /
    gain_ddot with the reference argument meaning and device pointers."""
    import ctypes as C

    import torch

    from gain_engine import _lib
    L = _lib.load()
    dev = torch.device("cuda:0")
    nc = 12
    prob = synth.make_problem([2, 3], nc, "c2v4")
    bI, bF = prob.bases
    jmax = prob.jmax
    tab = orc.threej_table(jmax)
    T = lambda a, dt=None: torch.as_tensor(np.ascontiguousarray(a),
                                           device=dev)
    rng = np.random.default_rng(9)
    g = 1
    c = rng.standard_normal(bI.ntot[g])
    # every device buffer is held by a name until the kernels have run: a
    # temporary tensor would be freed as soon as data_ptr() returns and its
    # block handed to the next allocation
    d_ijt, d_icontr = T(bI.ijterms), T(bI.icontr)
    d_N, d_rep, d_c = T(bI.N[g]), T(bI.repres[g]), T(c)
    d_x = torch.zeros(bI.nprim, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    _lib.check(L.gain_transform_vector_primitive(
        bI.nprim, g, bI.maxsymcoeffs, 0, bI.sdeg[g], bI.ntot[g],
        d_ijt.data_ptr(), d_icontr.data_ptr(), d_N.data_ptr(),
        d_rep.data_ptr(), d_c.data_ptr(), d_x.data_ptr(), None))
    torch.cuda.synchronize()
    x_ref = orc.expand(bI, g, 0, c)
    np.testing.assert_allclose(d_x.cpu().numpy(), x_ref, rtol=0, atol=1e-15)
    # piece [2, 8) of the dipole in the reference piece layout
    s0, s1 = 2, 8
    piece = np.ascontiguousarray(prob.mu[:, :, s0:s1]).reshape(-1)
    _lib.check(L.gain_copy_jmax_constant(jmax))
    _lib.check(L.gain_copy_dipole_constant(nc))
    d_h = torch.zeros(bF.nprim, dtype=torch.float64, device=dev)
    d_tab, d_piece = T(tab), T(piece)
    d_tauI, d_tauF = T(bI.tau), T(bF.tau)
    d_vibI, d_vibF = T(bI.vib), T(bF.vib)
    d_kbI = T(bI.kblock)
    torch.cuda.synchronize()
    for kF in range(bF.J + 1):
        _lib.check(L.gain_compute_half_linestrength(
            int(bF.kblock[kF]), bI.nprim, bI.J, bF.J, kF,
            d_tauI.data_ptr(), d_tauF.data_ptr(), d_vibI.data_ptr(),
            d_vibF.data_ptr(), int(bF.kstart[kF]),
            int(bI.kstart[max(kF - 1, 0)]), s0, s1, s1 - s0,
            int(bF.kblock[kF]), d_kbI.data_ptr(), d_piece.data_ptr(),
            d_x.data_ptr(), d_tab.data_ptr(), d_h.data_ptr(), None))
    torch.cuda.synchronize()
    h_ref = orc.half_ls(prob, bI, bF, x_ref, tab, startF=s0, endF=s1)
    np.testing.assert_allclose(d_h.cpu().numpy(), h_ref, rtol=0,
                               atol=1e-14 * np.abs(h_ref).max())
    d_out = torch.zeros(1, dtype=torch.float64, device=dev)
    y = rng.standard_normal(bF.nprim)
    d_y = T(y)
    torch.cuda.synchronize()
    _lib.check(L.gain_ddot(bF.nprim, d_h.data_ptr(), d_y.data_ptr(),
                           d_out.data_ptr(), None))
    torch.cuda.synchronize()
    ref = orc.dot(h_ref, y)
    assert abs(d_out.item() - ref) <= 1e-14 * np.abs(h_ref * y).sum()


@pytest.mark.parametrize("nI,nF,nlI,nlF,md,piece", [
    (700, 500, 5, 7, 2, (3, 17)), (256, 1, 1, 1, 1, (0, 20)),
    (1500, 2300, 9, 4, 3, (0, 20))])
def test_dropin_rotsym_matches_oracle(nI, nF, nlI, nlF, md, piece):
    """gain_compute_half_linestrength_rotsym against the oracle on random
    symmetric-top data; entries outside the dipole piece keep their value."""
    import torch

    from gain_engine import _lib
    L = _lib.load()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(nI + nF)
    nc = 20
    s0, s1 = piece
    mu = rng.standard_normal((3, nc, nc))
    vibI = rng.integers(0, nc, nI).astype(np.int32)
    vibF = rng.integers(0, nc, nF).astype(np.int32)
    degI = rng.integers(0, md, nI).astype(np.int32)
    degF = rng.integers(0, md, nF).astype(np.int32)
    lvI = rng.integers(0, nlI, nI).astype(np.int32)
    lvF = rng.integers(0, nlF, nF).astype(np.int32)
    W = rng.standard_normal((md, md, nlF, nlI, 3))
    x = rng.standard_normal(nI)
    piece_mu = np.ascontiguousarray(mu[:, :, s0:s1])
    h0 = np.full(nF, -3.5)
    ref = orc.half_ls_rotsym(nlI, nlF, md, degI, degF, lvI, lvF, vibI, vibF,
                             s0, s1, piece_mu, nc, W, x, out=h0)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    bufs = [T(a) for a in (degI, degF, lvI, lvF, vibI, vibF, piece_mu, W, x,
                           h0)]
    torch.cuda.synchronize()
    _lib.check(L.gain_copy_dipole_constant(nc))
    d = [b.data_ptr() for b in bufs]
    _lib.check(L.gain_compute_half_linestrength_rotsym(
        nF, nI, nlI, nlF, md, d[0], d[1], d[2], d[3], d[4], d[5], s0, s1,
        s1 - s0, d[6], d[7], d[8], d[9], None))
    torch.cuda.synchronize()
    got = bufs[9].cpu().numpy()
    np.testing.assert_allclose(got, ref, rtol=0,
                               atol=1e-14 * np.abs(ref).max())
    inside = (vibF >= s0) & (vibF < s1)
    assert np.all(got[~inside] == -3.5)
    rc = L.gain_compute_half_linestrength_rotsym(
        nF, nI, nlI, nlF, md, d[0], d[1], d[2], d[3], d[4], d[5], s0, s1,
        s1 - s0 + 1, d[6], d[7], d[8], d[9], None)
    assert rc == _lib.ERR_ARG


def test_error_paths_return_codes():
    """Missing uploads and bad arguments come back as status codes."""
    import ctypes as C

    from gain_engine import _lib
    L = _lib.load()
    with Engine(0, 4, 1, 1) as eng:
        out = _lib.Lines()
        rc = L.gain_plain_lines(eng._h, 0, 0, 0, None, 0, 0, C.byref(out))
        assert rc == _lib.ERR_STATE  # no params yet
        p = synth.Params(nrepres=1, sym_degen=[1], gns=[1.0],
                         isym_pairs=[1], maxJ=4)
        eng.set_params(p)
        rc = L.gain_plain_lines(eng._h, 0, 0, 0, None, 0, 0, C.byref(out))
        assert rc == _lib.ERR_STATE  # no dipole
        eng.upload_dipole(np.zeros((3, 4, 4)))
        rc = L.gain_plain_lines(eng._h, 0, 0, 0, None, 0, 0, C.byref(out))
        assert rc == _lib.ERR_STATE and b"basis" in L.gain_last_error()
        bad = synth.Params(nrepres=1, sym_degen=[1], gns=[1.0],
                           isym_pairs=[1], maxJ=4, temperature=0.0)
        with pytest.raises(_lib.GainError) as e:
            eng.set_params(bad)
        assert e.value.code == _lib.ERR_ARG
        lv = np.zeros(1, np.int32)
        rc = L.gain_batched_lines(eng._h, 0, 0, lv.ctypes.data, 1, 0, 0, 99,
                                  C.byref(out))
        assert rc == _lib.ERR_ARG and b"precision" in L.gain_last_error()


@pytest.mark.parametrize("mfma", [0, 1])
@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_vector_alu_and_mfma_gemms_agree_with_oracle(mfma, mode):
    """Half line strength and projection GEMMs on the vector ALUs and on
    the FP64 matrix cores; uneven sizes exercise every tile edge."""
    prob = synth.make_problem([4, 5], 20, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(1, len(rows), 3))
    tr, _ = orc.run(prob, initials)
    with Engine.from_problem(prob) as eng:
        eng.set_tuning("mfma", mfma)
        got = engine_lines(eng, prob, initials, mode)
    compare_lines(got, oracle_map(tr))


def _strong_weak_err(got, ref):
    keys = list(ref)
    r = np.array([ref[k]["S"] for k in keys])
    g = np.array([got[k]["S"] for k in keys])
    strong = r > 1e-6 * r.max()
    rel = np.abs(g - r) / np.where(r > 0, r, 1.0)
    return rel[strong].max(), (rel[~strong].max() if (~strong).any()
                               else 0.0), strong.sum(), (~strong).sum()


@pytest.mark.parametrize("precision,bound", [(3, 1e-6), (2, 1e-3)])
def test_batched_split_bf16_error_vs_fp64_oracle(precision, bound):
    """Split-BF16 contractions (4 and 3 slices): transition list identical,
    per-line S error against the FP64 oracle within `bound` on strong lines
    (S > 1e-6 S_max); the FP64 stages (3j, S accumulation, A, intensity)
    are unchanged."""
    prob = synth.make_problem([5, 6], 32, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(0, len(rows), 2))
    tr, _ = orc.run(prob, initials)
    ref = oracle_map(tr)
    with Engine.from_problem(prob) as eng:
        got = engine_lines(eng, prob, initials, "batched", precision)
    assert set(got) == set(ref)
    assert {k for k in got if got[k]["emitted"]} == \
        {k for k in ref if ref[k]["emitted"]}
    es, ew, ns, nw = _strong_weak_err(got, ref)
    print(f"precision {precision}: strong lines {ns} max rel err {es:.3e}; "
          f"weak lines {nw} max rel err {ew:.3e}")
    assert es <= bound


@pytest.mark.parametrize("precision,bound", [(6, 5e-10), (5, 5e-8),
                                             (4, 1e-5)])
def test_batched_split_int8_error_vs_fp64_oracle(precision, bound):
    """Split-INT8 contractions (7, 6 and 5 slices, exact integer products
    on the matrix cores): transition list identical, per-line S error
    against the FP64 oracle within `bound` on strong lines (S > 1e-6
    S_max)."""
    prob = synth.make_problem([5, 6], 32, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(0, len(rows), 2))
    tr, _ = orc.run(prob, initials)
    ref = oracle_map(tr)
    with Engine.from_problem(prob) as eng:
        got = engine_lines(eng, prob, initials, "batched", precision)
    assert set(got) == set(ref)
    assert {k for k in got if got[k]["emitted"]} == \
        {k for k in ref if ref[k]["emitted"]}
    es, ew, ns, nw = _strong_weak_err(got, ref)
    print(f"precision {precision}: strong lines {ns} max rel err {es:.3e}; "
          f"weak lines {nw} max rel err {ew:.3e}")
    assert es <= bound


@pytest.mark.parametrize("precision,bound", [(9, 1e-9), (8, 1e-7),
                                             (7, 1e-6)])
def test_batched_crt_error_vs_fp64_oracle(precision, bound):
    """CRT-form int8 contractions (14, 12 and 11 moduli: exact integer
    products of the operands rounded to ~47, ~40 and ~36 bits): transition
    list identical, per-line S error against the FP64 oracle within
    `bound` on strong lines."""
    prob = synth.make_problem([5, 6], 32, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(0, len(rows), 2))
    tr, _ = orc.run(prob, initials)
    ref = oracle_map(tr)
    with Engine.from_problem(prob) as eng:
        got = engine_lines(eng, prob, initials, "batched", precision)
    assert set(got) == set(ref)
    assert {k for k in got if got[k]["emitted"]} == \
        {k for k in ref if ref[k]["emitted"]}
    es, ew, ns, nw = _strong_weak_err(got, ref)
    print(f"precision {precision}: strong lines {ns} max rel err {es:.3e}; "
          f"weak lines {nw} max rel err {ew:.3e}")
    assert es <= bound


@pytest.mark.parametrize("cfg", list(range(14)))
def test_every_crt_tile_configuration(cfg):
    """Each CRT-form tile shape for both GEMM kinds on sizes with partial
    tiles; a frequency window starts the projected final rows off a 32-row
    block boundary.  Twelve moduli keep S within 1e-7 of the FP64 oracle
    on strong lines."""
    prob = synth.make_problem([4, 5], 20, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    prob.params.freq_window = (1500.0, 1e9)
    rows = prob.state_table()
    initials = list(range(1, len(rows), 2))
    tr, _ = orc.run(prob, initials)
    ref = oracle_map(tr)
    with Engine.from_problem(prob) as eng:
        eng.set_tuning("crt_cfg", cfg)
        got = engine_lines(eng, prob, initials, "batched", 8)
        assert eng.get_tuning("proj_cfg") == cfg
    assert set(got) == set(ref)
    es, _, ns, _ = _strong_weak_err(got, ref)
    print(f"crt12 configuration {cfg}: {ns} strong lines, max rel err "
          f"{es:.3e}")
    assert ns > 0 and es <= 1e-7, es


@pytest.mark.parametrize("cfg", list(range(12)))
@pytest.mark.parametrize("split", [1, 2])
def test_every_int8_tile_configuration(cfg, split):
    """Each split-INT8 tile shape (per-wave reads and LDS-staged tiles)
    for both GEMM kinds, and split-K 2 on the projection,
    on sizes with partial tiles; a frequency window starts the projected
    final rows off a 32-row block boundary.  Six slices keep S within 5e-8
    of the FP64 oracle on strong lines."""
    prob = synth.make_problem([4, 5], 20, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    prob.params.freq_window = (1500.0, 1e9)
    rows = prob.state_table()
    initials = list(range(1, len(rows), 2))
    tr, _ = orc.run(prob, initials)
    ref = oracle_map(tr)
    with Engine.from_problem(prob) as eng:
        eng.set_tuning("i8_cfg", cfg)
        eng.set_tuning("i8_split", split)
        got = engine_lines(eng, prob, initials, "batched", 5)
        assert eng.get_tuning("proj_cfg") == cfg
        assert eng.get_tuning("proj_split") == split
    assert set(got) == set(ref)
    es, _, ns, _ = _strong_weak_err(got, ref)
    print(f"int8 x6 configuration {cfg} split {split}: {ns} strong lines, "
          f"max rel err {es:.3e}")
    assert ns > 0 and es <= 5e-8, es


_LONG_K = {}


def _long_k_case():
    """J 4 and 5 with Nc = 400 (records of 900 and 1100 elements, a 400 x 400
    dipole): the GEMMs have 14 and 36 k blocks of 32 bytes, so every
    pipelined int8 tile form runs its steady loop (the Nc = 20 cases above
    have 2); six initial states against the oracle, computed once."""
    if not _LONG_K:
        prob = synth.make_problem([4, 5], 400, "c2v4")
        prob.params.erange_lower = (0.0, 1e9)
        prob.params.freq_window = (1500.0, 1e9)
        n = len(prob.state_table())
        initials = [1 + k * (n // 6) for k in range(6)]
        tr, _ = orc.run(prob, initials)
        _LONG_K.update(prob=prob, initials=initials, ref=oracle_map(tr))
    return _LONG_K["prob"], _LONG_K["initials"], _LONG_K["ref"]


@pytest.mark.parametrize("cfg", list(range(14)))
def test_every_crt_tile_configuration_long_k(cfg):
    """Each CRT-form tile shape with K long enough for its steady loop
    (stage requests interleaved with the products): S within 1e-7 of the
    FP64 oracle on strong lines, transition list identical."""
    prob, initials, ref = _long_k_case()
    with Engine.from_problem(prob) as eng:
        eng.set_tuning("crt_cfg", cfg)
        got = engine_lines(eng, prob, initials, "batched", 8)
        assert eng.get_tuning("proj_cfg") == cfg
    assert set(got) == set(ref)
    es, _, ns, _ = _strong_weak_err(got, ref)
    print(f"crt12 configuration {cfg}, long K: {ns} strong lines, max rel "
          f"err {es:.3e}")
    assert ns > 0 and es <= 1e-7, es


_BANDED = {}


def _banded_case(zero_tiles=False):
    """The long-K problem with a banded dipole: entries with |v_I - v_F| > 40
    set to zero, and v_F in [320, 352) (one 32-row block of the dipole's
    residue planes) zero for every component, so the dipole's residue planes
    have zero k blocks at both ends of most row blocks and one zero row
    block.  With zero_tiles, v_F in [256, 400) is zero as well: whole
    workgroup tiles (rows 256-511 of the 256-row tiles, 256-383 and 384-511
    of the 128-row ones) have no k block to multiply."""
    if zero_tiles not in _BANDED:
        prob = synth.make_problem([4, 5], 400, "c2v4")
        prob.params.erange_lower = (0.0, 1e9)
        prob.params.freq_window = (1500.0, 1e9)
        a = np.arange(prob.nc)
        band = np.abs(a[:, None] - a[None, :]) <= 40
        prob.mu = np.ascontiguousarray(prob.mu * band[None])
        prob.mu[:, :, 320:352] = 0.0
        if zero_tiles:
            prob.mu[:, :, 256:400] = 0.0
        n = len(prob.state_table())
        initials = [1 + k * (n // 6) for k in range(6)]
        tr, _ = orc.run(prob, initials)
        _BANDED[zero_tiles] = (prob, initials, oracle_map(tr))
    return _BANDED[zero_tiles]


@pytest.mark.parametrize("zero_tiles", [False, True])
@pytest.mark.parametrize("cfg", list(range(14)))
def test_crt_zero_kblock_skip_is_exact(cfg, zero_tiles):
    """The CRT-form dipole products multiply only the k blocks of each
    tile's row blocks that hold a nonzero residue; on a banded dipole with
    a zero row block (and, in the second case, whole zero tiles, whose
    products are skipped entirely) every tile configuration gives S, A, I
    and flags identical bit for bit to the dense products (skipping off),
    fewer k blocks are multiplied, and S is within 1e-7 of the FP64 oracle
    on strong lines."""
    prob, initials, ref = _banded_case(zero_tiles)
    res = {}
    for skip in (1, 0):
        with Engine.from_problem(prob) as eng:
            eng.set_tuning("crt_cfg", cfg)
            eng.set_tuning("crt_kskip", skip)
            assert eng.get_tuning("crt_kskip") == skip
            res[skip] = {"lines": engine_lines(eng, prob, initials,
                                               "batched", 8),
                         "ppm": eng.get_tuning("half_ls_kblocks_ppm")}
    on, off = res[1]["lines"], res[0]["lines"]
    assert set(on) == set(off) == set(ref)
    for k in on:
        for f in ("S", "A", "I", "emitted"):
            assert on[k][f] == off[k][f], (k, f, on[k][f], off[k][f])
    assert res[0]["ppm"] == 1_000_000
    assert res[1]["ppm"] < 900_000, res[1]["ppm"]
    es, _, ns, _ = _strong_weak_err(on, ref)
    print(f"crt12 configuration {cfg}, banded dipole: k blocks "
          f"{res[1]['ppm'] / 1e4:.1f}% of dense, {ns} strong lines, max rel "
          f"err {es:.3e}")
    assert ns > 0 and es <= 1e-7, es


def _fused_vs_two_kernels(prob, initials, precision, key="crt_fuse"):
    res = {}
    for fuse in (1, 0):
        with Engine.from_problem(prob) as eng:
            eng.set_tuning(key, fuse)
            assert eng.get_tuning(key) == fuse
            res[fuse] = engine_lines(eng, prob, initials, "batched",
                                     precision)
    on, off = res[1], res[0]
    assert len(on) > 0 and set(on) == set(off)
    for k in on:
        for f in ("S", "A", "I", "emitted"):
            assert on[k][f] == off[k][f], (k, f, on[k][f], off[k][f])
        assert np.array_equal(on[k]["ls"], off[k]["ls"]), k
    return on


@pytest.mark.parametrize("scheme,jvals,nc", [("c2v4", [2, 3], 16),
                                             ("c1", [1, 2], 10),
                                             ("pair2", [3, 4], 40),
                                             ("c3v3", [1, 2], 12),
                                             ("c2v4", [5, 6], 72)])
@pytest.mark.parametrize("precision", [7, 8, 9])
def test_crt_fused_gather_is_exact(scheme, jvals, nc, precision):
    """The CRT-form half line strength of K-block bases reconstructed and
    gathered by one kernel equals the two-kernel path (reconstruction to
    FP64, then the gather) bit for bit: S, A, I, every degenerate
    component of ls and the flags, over dJ = -1, 0, +1, degenerate irreps
    and rows of the last 32-row block beyond Nc, for 11, 12 and 14
    moduli."""
    prob = synth.make_problem(jvals, nc, scheme)
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(0, len(rows), max(1, len(rows) // 40)))
    _fused_vs_two_kernels(prob, initials, precision)


def test_crt_fused_gather_autotuned():
    """The autotuner times every CRT tile configuration of the half line
    strength on a call whose gather is fused (the product has no FP64
    output, so only the residue products are timed): lines within 1e-7 of
    the oracle on strong lines, the half-ls shape in the tuning report."""
    prob = synth.make_problem([4, 5], 20, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(1, len(rows), 2))
    tr, _ = orc.run(prob, initials)
    ref = oracle_map(tr)
    with Engine.from_problem(prob) as eng:
        eng.set_tuning("autotune", 1)
        got = engine_lines(eng, prob, initials, "batched", 8)
        report = eng.tuning_report()
    assert "crt12 half_ls" in report and "crt12 projection" in report
    assert set(got) == set(ref)
    es, _, ns, _ = _strong_weak_err(got, ref)
    assert ns > 0 and es <= 1e-7, es


@pytest.mark.parametrize("seed", [1, 2])
def test_crt_fused_gather_irregular_basis(seed):
    """Fused CRT gather on random K-block sizes, tau per primitive and
    repeated vib indices (final slots holding several primitives, empty
    slots): identical to the two-kernel path and within 1e-7 of the
    oracle on strong lines."""
    rng = np.random.default_rng(seed)
    nc = 7
    bases = [irregular_basis(2, nc, rng, 12), irregular_basis(3, nc, rng, 12)]
    prob = problem_with_bases(bases, nc, seed)
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(0, len(rows), 3))
    tr, _ = orc.run(prob, initials)
    ref = oracle_map(tr)
    got = _fused_vs_two_kernels(prob, initials, 8)
    assert set(got) == set(ref)
    es, _, ns, _ = _strong_weak_err(got, ref)
    assert ns > 0 and es <= 1e-7, es


def _xslice_problem(case):
    if case.startswith("irregular"):
        seed = int(case[-1])
        rng = np.random.default_rng(seed)
        nc = 7
        bases = [irregular_basis(2, nc, rng, 12),
                 irregular_basis(3, nc, rng, 12)]
        prob = problem_with_bases(bases, nc, seed)
    elif case == "rotsym":
        prob = synth.make_problem([3, 4], 6, "c3v3", rotsym="C3V(M)")
    else:
        scheme, nc = {"c2v4": ("c2v4", 16), "td5": ("td5", 20)}[case]
        prob = synth.make_problem([2, 3], nc, scheme)
    prob.params.erange_lower = (0.0, 1e9)
    return prob


@pytest.mark.parametrize("case", ["c2v4", "td5", "rotsym", "irregular1",
                                  "irregular2"])
@pytest.mark.parametrize("precision", [7, 8, 9])
def test_crt_x_scattered_into_residues_is_exact(case, precision):
    """The CRT-form half line strength's X built straight into its residue
    planes by one kernel (tuning key crt_xfuse) equals the FP64 scatter
    followed by the exponent and slicing kernels: identical S, A, I, flags
    and pair line strengths, on K-block bases, the degenerate Td irreps
    (records' components 1 and 2), a symmetric-top basis and irregular
    bases (slots holding several primitives, empty slots); S within 1e-7
    of the oracle on strong lines."""
    prob = _xslice_problem(case)
    rows = prob.state_table()
    initials = list(range(0, len(rows), 3))
    tr, _ = orc.run(prob, initials)
    ref = oracle_map(tr)
    got = _fused_vs_two_kernels(prob, initials, precision, "crt_xfuse")
    assert set(got) == set(ref)
    es, _, ns, _ = _strong_weak_err(got, ref)
    assert ns > 0 and es <= 1e-7, es


def test_crt_x_scattered_into_residues_long_rows():
    """The same at N_c = 2500 (rows of 2500 dipole columns: the 16-round
    form of the kernel, 79 k blocks): identical to the three-kernel path
    for 11 and 12 moduli."""
    prob = synth.make_problem([1, 2], 2500, "c2v4",
                              blocks={(0, 0), (1, 1)})
    prob.params.erange_lower = (0.0, 1e9)
    n0 = prob.blocks[(0, 0)].n
    levels = list(range(0, n0, 7))
    for precision in (7, 8):
        res = {}
        for fuse in (1, 0):
            with Engine.from_problem(prob) as eng:
                eng.set_tuning("crt_xfuse", fuse)
                res[fuse] = eng.batched_lines(0, 0, levels, 1, 1, precision)
        assert (res[1]["flags"] & 1).sum() > 0
        for f in ("S", "A", "I", "flags", "ls"):
            assert np.array_equal(res[1][f], res[0][f]), (precision, f)


@pytest.mark.parametrize("cfg", list(range(12)))
@pytest.mark.parametrize("split", [1, 2])
def test_every_int8_tile_configuration_long_k(cfg, split):
    """Each split-INT8 tile shape and split-K 1/2 with K long enough for
    the steady loop of the LDS-staged forms: S within 5e-8 of the FP64
    oracle on strong lines, transition list identical."""
    prob, initials, ref = _long_k_case()
    with Engine.from_problem(prob) as eng:
        eng.set_tuning("i8_cfg", cfg)
        eng.set_tuning("i8_split", split)
        got = engine_lines(eng, prob, initials, "batched", 5)
        assert eng.get_tuning("proj_cfg") == cfg
        assert eng.get_tuning("proj_split") == split
    assert set(got) == set(ref)
    es, _, ns, _ = _strong_weak_err(got, ref)
    print(f"int8 x6 configuration {cfg} split {split}, long K: {ns} strong "
          f"lines, max rel err {es:.3e}")
    assert ns > 0 and es <= 5e-8, es


@pytest.mark.parametrize("cfg", list(range(15)))
def test_every_gemm_tile_configuration(cfg):
    """Each tile shape the autotuner can pick, for both GEMM kinds, on
    sizes that leave partial tiles in M, N and K."""
    prob = synth.make_problem([4, 5], 20, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(1, len(rows), 2))
    tr, _ = orc.run(prob, initials)
    with Engine.from_problem(prob) as eng:
        eng.set_tuning("proj_cfg", cfg)
        eng.set_tuning("half_ls_cfg", cfg)
        got = engine_lines(eng, prob, initials, "batched")
        eng.set_tuning("mu_ksplit", 1)
        got1 = engine_lines(eng, prob, initials, "plain")
    compare_lines(got, oracle_map(tr))
    compare_lines(got1, oracle_map(tr))


def test_autotuner_picks_and_reports():
    prob = synth.make_problem([4, 5], 20, "c2v4")
    rows = prob.state_table()
    initials = list(range(0, len(rows), 2))
    tr, _ = orc.run(prob, initials)
    with Engine.from_problem(prob) as eng:
        eng.set_tuning("autotune", 1)
        got = engine_lines(eng, prob, initials, "batched")
        rep = eng.tuning_report()
    assert "projection" in rep and "half_ls" in rep and "->" in rep
    compare_lines(got, oracle_map(tr))


@pytest.mark.parametrize("split", [2, 3, 8])
def test_projection_split_k(split):
    prob = synth.make_problem([4, 5], 20, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(0, len(rows), 2))
    tr, _ = orc.run(prob, initials)
    with Engine.from_problem(prob) as eng:
        eng.set_tuning("proj_cfg", 5)
        eng.set_tuning("proj_split", split)
        got = engine_lines(eng, prob, initials, "batched")
    compare_lines(got, oracle_map(tr))


def test_batched_fp32_mfma_error_vs_fp64_oracle():
    """FP32 matrix-core contractions with FP64 accumulation: transition list
    identical; S error reported (operands rounded to FP32)."""
    prob = synth.make_problem([5, 6], 32, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(0, len(rows), 2))
    tr, _ = orc.run(prob, initials)
    ref = oracle_map(tr)
    with Engine.from_problem(prob) as eng:
        got = engine_lines(eng, prob, initials, "batched", 1)
    assert set(got) == set(ref)
    es, ew, ns, nw = _strong_weak_err(got, ref)
    print(f"precision 1 (fp32): strong lines {ns} max rel err {es:.3e}; "
          f"weak lines {nw} max rel err {ew:.3e}")
    assert es <= 1e-3


@pytest.mark.parametrize("cfg", list(range(4)))
def test_every_f32_tile_configuration(cfg):
    """Each FP32 matrix-core tile shape (operands rounded to FP32, sums of
    32 consecutive k in FP32 added in FP64), both GEMM kinds, on sizes that
    leave partial tiles in M, N and K: transition list and emitted bits
    identical to the FP64 oracle's; the FP32 rounding of the operands
    perturbs each half line strength by a few 2^-24 of the largest one, so
    |S - S_oracle| <= 2e-6 sqrt(S_oracle S_max) + 1e-12 S_max on every
    line."""
    prob = synth.make_problem([4, 5], 20, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(1, len(rows), 2))
    tr, _ = orc.run(prob, initials)
    ref = oracle_map(tr)
    with Engine.from_problem(prob) as eng:
        eng.set_tuning("f32_cfg", cfg)
        assert eng.get_tuning("f32_cfg") == cfg
        got = engine_lines(eng, prob, initials, "batched", 1)
    assert set(got) == set(ref)
    for k in got:
        assert got[k]["emitted"] == ref[k]["emitted"], k
    es, ew, ns, nw = _strong_weak_err(got, ref)
    keys = list(ref)
    r = np.array([ref[k]["S"] for k in keys])
    d = np.abs(np.array([got[k]["S"] for k in keys]) - r)
    bound = 2e-6 * np.sqrt(r * r.max()) + 1e-12 * r.max()
    print(f"fp32 configuration {cfg}: {ns} strong lines, max rel err "
          f"{es:.3e}; {nw} weak lines, max rel err {ew:.3e}; max "
          f"|dS| / bound {(d / bound).max():.3e}")
    assert ns > 0 and (d <= bound).all(), (d / bound).max()


def test_dipole_in_pieces():
    """The dipole uploaded as pieces of final vibrational columns in the
    reference piece layout (TroveDipole.cpp:47-112), out of order: the lines
    equal those of the whole upload bit for bit, FP64 and CRT12; while a
    piece is missing every product returns GAIN_ERR_STATE; a piece outside
    [0, nc) is refused."""
    from gain_engine._lib import GainError
    prob = synth.make_problem([2, 3], 16, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rows = prob.state_table()
    initials = list(range(0, len(rows), 3))
    with Engine.from_problem(prob) as eng:
        ref = {p: engine_lines(eng, prob, initials, "batched", p)
               for p in (0, 8)}
    with Engine.from_problem(prob, dipole_pieces=[(5, 11), (0, 5)]) as eng:
        I = initials[0]
        with pytest.raises(GainError, match="dipole pieces cover 11 of"):
            eng.batched_lines(rows[I][1], rows[I][2], [rows[I][3]],
                              rows[I][1], rows[I][2], 0)
        with pytest.raises(GainError, match="outside"):
            eng.upload_dipole(np.zeros((3, 16, 7)), 10, 17)
        eng.upload_dipole(prob.mu[:, :, 11:16], 11, 16)
        got = {p: engine_lines(eng, prob, initials, "batched", p)
               for p in (0, 8)}
    for p in (0, 8):
        assert set(got[p]) == set(ref[p]) and len(ref[p]) > 0
        for k in ref[p]:
            for f in ("S", "A", "I", "emitted"):
                assert got[p][k][f] == ref[p][k][f], (p, k, f)


@pytest.mark.parametrize("precision,bound", [(7, 1e-6), (8, 1e-7),
                                             (9, 1e-9)])
def test_crt_rows_of_unequal_scale(precision, bound):
    """The CRT form scales every operand row by its own exponent (a bound
    of its 2-norm, or of its maximum for 14 moduli): a dipole whose final
    vibrational rows are scaled by powers of two from 2^-40 to 2^40 gives
    the transition list and emitted bits of the FP64 oracle and S within
    `bound` on strong lines, as an unscaled one does."""
    prob = synth.make_problem([3, 4], 24, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rng = np.random.default_rng(5)
    scale = np.ldexp(1.0, rng.integers(-40, 41, size=prob.nc))
    prob.mu = np.ascontiguousarray(prob.mu * scale[None, None, :])
    rows = prob.state_table()
    initials = list(range(0, len(rows), 2))
    tr, _ = orc.run(prob, initials)
    ref = oracle_map(tr)
    with Engine.from_problem(prob) as eng:
        got = engine_lines(eng, prob, initials, "batched", precision)
    assert set(got) == set(ref)
    for k in got:
        assert got[k]["emitted"] == ref[k]["emitted"], k
    es, _, ns, _ = _strong_weak_err(got, ref)
    print(f"precision {precision}, rows scaled by 2^-40..2^40: {ns} strong "
          f"lines, max rel err {es:.3e}")
    assert ns > 0 and es <= bound, es


@pytest.mark.parametrize("pm,bound", [(11, 1e-6), (14, 1e-7)])
def test_crt_projection_moduli_override(pm, bound):
    """The CRT projection on its own moduli count (tuning key
    proj_crt_moduli; the half line strength keeps 12): transition list
    identical to the oracle's, S within `bound` on strong lines."""
    prob, initials, ref = _long_k_case()
    with Engine.from_problem(prob) as eng:
        eng.set_tuning("proj_crt_moduli", pm)
        assert eng.get_tuning("proj_crt_moduli") == pm
        got = engine_lines(eng, prob, initials, "batched", 8)
    assert set(got) == set(ref)
    es, _, ns, _ = _strong_weak_err(got, ref)
    assert ns > 0 and es <= bound, es
