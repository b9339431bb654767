#!/usr/bin/env python3
"""Line-strength throughput on one MI355X: J_I=20 -> J_F=21, Nc=4000.

Workload (synthetic, seeded): C2v-like 4-irrep basis (Hadamard quartets),
initial block (J=20, A1) with N_I = 41000 functions, final block
(J=21, A2) with N_F = n_F = 43000 states whose 43000 x 43000 FP64
eigenvector records (14.8 GB, orthonormal: Kronecker product of two QR
factors) are generated in HBM; mu is 3 x 4000 x 4000 FP64.  The 2000
lowest initial states are processed against every final state with a
0-15000 cm^-1 frequency window.

A step is one batch of B initial states through the batched engine
(half line strengths, projection GEMM, S/A/I epilogue); the plain FP64
GEMV drop-in (one initial state per call) is timed beside it.  Prints one
JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "gain-mpi_amd")]

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
PRECISION_NAMES = {0: "f64", 1: "f32-mfma/f64-acc",
                   2: "bf16x3-ozaki-mfma/f32-f64-acc",
                   3: "bf16x4-ozaki-mfma/f32-f64-acc"}
PEAKS = {0: ("TFLOP/s", 78.6), 1: ("TFLOP/s", 157.3), 2: ("TFLOP/s", 2500.0),
         3: ("TFLOP/s", 2500.0)}
PRODUCTS = {0: 1, 1: 1, 2: 6, 3: 10}  # hardware products per algorithmic FMA


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_problem(nc, jI, jF):
    from gain_engine import synth
    prob = synth.make_problem([jI, jF], nc, "c2v4", blocks={(0, 0), (1, 1)},
                              kron=True)
    return prob


def rows_in_window(prob, eF, eIs):
    """Final rows the engine streams for these initial energies: the union
    over initials of the states whose energy passes the frequency and upper
    windows (energies are sorted)."""
    import numpy as np
    p = prob.params
    lo, hi = len(eF), 0
    for eI in eIs:
        ok = ((eF - eI) >= p.freq_window[0]) & ((eF - eI) <= p.freq_window[1]) \
            & ((eF - p.zpe) >= p.erange_upper[0]) & \
            ((eF - p.zpe) <= p.erange_upper[1])
        idx = np.nonzero(ok)[0]
        if idx.size:
            lo, hi = min(lo, idx[0]), max(hi, idx[-1] + 1)
    return max(hi - lo, 0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=250)
    ap.add_argument("--plain-states", type=int, default=100)
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--nc", type=int, default=4000)
    ap.add_argument("--ninitial", type=int, default=2000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--variants", type=str, default="",
                    help="comma-separated extra precisions to time")
    ap.add_argument("--copy-calib-gb", type=float, default=0.0,
                    help="also run a device-to-device copy of this many GB "
                         "(reference point for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from gain_engine.engine import Engine

    jI, jF, nc = 20, 21, args.nc
    t0 = time.time()
    prob = build_problem(nc, jI, jF)
    blkI, blkF = prob.blocks[(0, 0)], prob.blocks[(1, 1)]
    nI_all = min(args.ninitial, blkI.n)
    log(f"problem built in {time.time() - t0:.1f}s: N_I={blkI.rec_len} "
        f"N_F={blkF.rec_len} n_F={blkF.n}")

    dev = torch.device("cuda:0")
    qa, qb = (torch.as_tensor(q, device=dev) for q in blkF.kron)
    CF = torch.kron(qa.T.contiguous(), qb.T.contiguous())  # records = rows
    del qa, qb
    qa, qb = blkI.kron
    b = qb.shape[0]
    ka = (nI_all + b - 1) // b
    CI = torch.kron(torch.as_tensor(qa.T[:ka].copy(), device=dev),
                    torch.as_tensor(qb.T.copy(), device=dev))[:nI_all]
    CI = CI.contiguous()
    torch.cuda.synchronize()
    log(f"eigenvectors in HBM: C_F {tuple(CF.shape)} "
        f"({CF.numel() * 8 / 1e9:.1f} GB), C_I {tuple(CI.shape)}")

    eng = Engine(0, prob.jmax, 4, 1)
    eng.set_params(prob.params)
    eng.upload_dipole(prob.mu)
    for jind, bs in enumerate(prob.bases):
        eng.upload_basis(jind, bs)
        eng.upload_inflation(jind, bs)
    eng.bind_block_device(0, 0, blkI.energy[:nI_all], CI.data_ptr(),
                          blkI.rec_len, blkI.rec_len)
    eng.bind_block_device(1, 1, blkF.energy, CF.data_ptr(), blkF.rec_len,
                          blkF.rec_len)
    nF = blkF.n
    stream = torch.cuda.ExternalStream(eng.stream())

    if args.copy_calib_gb > 0:
        n = int(args.copy_calib_gb * 1e9 / 8)
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {n * 8 / 1e9:.2f} GB read + written")
        del src, dst
        torch.cuda.empty_cache()

    # ---------------- plain FP64 GEMV drop-in
    pl = {k: torch.zeros(nF, dtype=torch.float64, device=dev)
          for k in ("S", "A", "I")}
    pflags = torch.zeros((args.plain_states, nF), dtype=torch.uint8,
                         device=dev)
    for w in range(max(args.warmup, 1)):
        eng.lines_device(0, 0, [w], 1, 1,
                         {k: v.data_ptr() for k, v in pl.items()} |
                         {"flags": pflags[0].data_ptr()}, plain=True)
    eng.synchronize()
    eng.set_timing(True)
    for name in ("gemv", "half_ls", "expand_T", "epilogue", "gemm_proj"):
        eng.timing(name, reset=True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), \
        torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for i in range(args.plain_states):
        eng.lines_device(0, 0, [i], 1, 1,
                         {k: v.data_ptr() for k, v in pl.items()} |
                         {"flags": pflags[i].data_ptr()}, plain=True)
    ev1.record(stream)
    eng.synchronize()
    plain_ms = ev0.elapsed_time(ev1)
    plain_trans = int((pflags & 1).sum().item())
    gemv_ms, gemv_n = eng.timing("gemv", reset=True)
    hl_ms, hl_n = eng.timing("half_ls", reset=True)
    plain_tps = plain_trans / (plain_ms / 1e3)
    gemv_avg = gemv_ms / max(gemv_n, 1)
    # algorithmic bytes of one GEMV launch: the final records inside the
    # energy window (rows), the projected vector and the row results
    rows = np.mean([rows_in_window(prob, blkF.energy, [blkI.energy[i]])
                    for i in range(args.plain_states)])
    gemv_bytes = 8.0 * (rows * blkF.rec_len + blkF.rec_len + rows)
    gemv_gbs = gemv_bytes / (gemv_avg / 1e3) / 1e9
    log(f"plain: {args.plain_states} initial states, {plain_trans} "
        f"transitions in {plain_ms:.1f} ms -> {plain_tps:.3e}/s; "
        f"gemv avg {gemv_avg:.3f} ms = {gemv_gbs:.0f} GB/s; half_ls avg "
        f"{hl_ms / max(hl_n, 1):.3f} ms")

    # ---------------- batched engine (headline)
    B = args.batch
    nsteps = args.steps
    outs = {k: torch.zeros((B, nF), dtype=torch.float64, device=dev)
            for k in ("S", "A", "I")}
    bflags = torch.zeros((nsteps, B, nF), dtype=torch.uint8, device=dev)

    def levels_of(step):
        return [(step * B + i) % nI_all for i in range(B)]

    def run_batched(prec):
        for w in range(max(args.warmup, 1)):
            eng.lines_device(0, 0, levels_of(w), 1, 1,
                             {k: v.data_ptr() for k, v in outs.items()} |
                             {"flags": bflags[0].data_ptr()}, precision=prec)
        eng.synchronize()
        for name in ("gemv", "half_ls", "expand_T", "epilogue",
                     "gemm_proj"):
            eng.timing(name, reset=True)
        torch.cuda.synchronize()
        ev0.record(stream)
        for st in range(nsteps):
            eng.lines_device(0, 0, levels_of(st), 1, 1,
                             {k: v.data_ptr() for k, v in outs.items()} |
                             {"flags": bflags[st].data_ptr()}, precision=prec)
        ev1.record(stream)
        eng.synchronize()
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1)
        trans = int((bflags & 1).sum().item())
        gp_ms, gp_n = eng.timing("gemm_proj")
        hb_ms, hb_n = eng.timing("half_ls")
        gp_avg = gp_ms / max(gp_n, 1)
        mrows = np.mean([rows_in_window(prob, blkF.energy,
                                        [blkI.energy[l] for l in
                                         levels_of(st)])
                         for st in range(nsteps)])
        flops = 2.0 * mrows * blkF.rec_len * B
        tfs = flops / (gp_avg / 1e3) / 1e12
        r = {"precision": prec, "dtype": PRECISION_NAMES[prec],
             "transitions": trans, "ms": ms, "tps": trans / (ms / 1e3),
             "gemm_proj_ms": gp_avg, "gemm_proj_tflops": tfs,
             "half_ls_ms": hb_ms / max(hb_n, 1)}
        log(f"batched[{PRECISION_NAMES[prec]}]: {nsteps} x {B} initial "
            f"states, {trans} transitions in {ms:.1f} ms -> "
            f"{r['tps']:.3e}/s; projection avg {gp_avg:.2f} ms = {tfs:.1f} "
            f"algorithmic TF/s; half_ls avg {r['half_ls_ms']:.2f} ms")
        return r

    def s_of_batch0(prec):
        eng.lines_device(0, 0, levels_of(0), 1, 1,
                         {"S": outs["S"].data_ptr(),
                          "flags": bflags[0].data_ptr()}, precision=prec)
        eng.synchronize()
        return outs["S"].cpu().numpy().copy(), \
            bflags[0].cpu().numpy().copy()

    head = run_batched(args.precision)
    variants = [run_batched(int(v)) for v in args.variants.split(",")
                if v.strip() and int(v) != args.precision]
    # per-line error of the measured precision against FP64 on one batch
    s64, f64 = s_of_batch0(0)
    sp, fp = s_of_batch0(args.precision)
    lines = (f64 & 1).astype(bool)
    ref = s64[lines]
    rel = np.abs(sp[lines] - ref) / np.where(ref > 0, ref, 1.0)
    strong = ref > 1e-6 * ref.max()
    prec_check = {"vs": "batched FP64 (checked against the oracle at 1e-12 "
                        "in the GPU tests)",
                  "transition_list_identical": bool(np.array_equal(f64, fp)),
                  "strong_lines": int(strong.sum()),
                  "strong_max_rel_err": float(rel[strong].max()),
                  "weak_lines": int((~strong).sum()),
                  "weak_max_rel_err": float(rel[~strong].max())
                  if (~strong).any() else 0.0}
    log(f"precision check: {prec_check}")
    tps = head["tps"]
    batched_ms = head["ms"]
    gemm_tfs = head["gemm_proj_tflops"]

    # ---------------- CPU baseline (oracle, bounded sample)
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        rows = prob.state_table()
        gidx = {(r[1], r[2], r[3]): i for i, r in enumerate(rows)}
        I0 = gidx[(0, 0, 0)]
        finals = [gidx[(1, 1, f)] for f in range(nF)]
        thr = min(args.cpu_threads, os.cpu_count() or 1)
        del CF, CI
        torch.cuda.empty_cache()
        tc = time.perf_counter()
        tr, _ = orc.run(prob, [I0], nthreads=thr, finals=finals)
        cpu_s = time.perf_counter() - tc
        cpu = {"value": len(tr) / cpu_s, "unit": "transitions/s",
               "cores": thr, "kind": "port",
               "sample": f"1 initial state (lowest J=20 A1) x all {nF} "
                         f"J=21 A2 finals, {len(tr)} transitions, "
                         f"{cpu_s:.1f} s incl. record staging"}
        log(f"cpu baseline: {cpu['value']:.3e} transitions/s on {thr} "
            f"threads ({cpu_s:.1f}s)")

    res = {
        "metric": "transitions/s at J=20->21 (N~4e4): batched MFMA vs plain "
                  "GEMV, rel err<=1e-6",
        "value": tps,
        "unit": "transitions/s",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": args.warmup,
        "ms_per_step": batched_ms / nsteps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": PRECISION_NAMES.get(args.precision, "f64"),
        "data": "synthetic (seeded generator, orthonormal eigenvectors)",
        "config": {"workload": "J_I=20 -> J_F=21 line strengths, Nc=4000, "
                               "N_I=41000, N_F=n_F=43000, 2000 lowest "
                               "initial states, 0-15000 cm-1 window",
                   "batch_initial_states": B, "precision": args.precision},
        "plain": {"value": plain_tps, "unit": "transitions/s",
                  "initial_states": args.plain_states,
                  "ms_per_initial": plain_ms / args.plain_states},
        "speedup_vs_plain": tps / plain_tps,
        "roofline": {"bound": "mfma", "kernel": "gemm_proj",
                     "achieved": gemm_tfs,
                     "peak": PEAKS[args.precision][1] /
                     PRODUCTS[args.precision],
                     "unit": "TFLOP/s",
                     "frac": gemm_tfs * PRODUCTS[args.precision] /
                     PEAKS[args.precision][1],
                     "note": "achieved = algorithmic 2*M*N*K flops / avg "
                             "kernel time; peak = dense MFMA peak of the "
                             "datatype / hardware products per FMA",
                     "traffic": None},
        "precision_check": prec_check,
        "variants": variants,
        "roofline_plain": {"bound": "hbm", "kernel": "gemv",
                           "achieved": gemv_gbs, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": gemv_gbs / HBM_PEAK_GBS,
                           "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
