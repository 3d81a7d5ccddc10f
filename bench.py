"""bench.py -- headline benchmark of the symmetry-adapted Lanczos engine.

Workload: periodic nearest-neighbour Heisenberg chain, L = 1000, Jxy = Jz = 1,
N_down = 4, k = 0 (D = 41,417,250 translation representatives).

A "step" is one full-sector H.v in the FP32-mixed path (complex-float
vectors, operator H - e_zz, FP64 reductions), with inputs resident in HBM.
value = H.v applications per second.  The same JSON line also reports the
FP64 H.v rate; the time to the ground-state energy within 1e-9 of the
Bethe-ansatz value 242.000099091345514 for plain FP64 Lanczos, plain
FP32-mixed Lanczos, the FP32-mixed thick restart (allocation included),
with the thick restart's acceptance certificate, and the FP64 thick restart
with the same basis memory (--NOM64, 400 vectors); the roofline of the H.v
launch (compulsory bytes over event time); and the CPU baseline: the
reference's own ham_to_vec_wave_vector loop (compiled from its sources into
oracle/_ref) timed on a bounded sample of rows with every host core.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-solve] [--no-cpu]
       python bench.py --dsf [--dsf-steps 200]   (batched 64-q DSF workload, its own JSON line)
       python bench.py --sectors                 (280-sector L=40 workload, its own JSON line)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "H·v/s & time-to-E0 @1e-9 rel, FP32-mixed vs FP64, L=1000 N↓=4 k=0"
E0_EXACT = 242.000099091345514  # Bethe ansatz, L=1000, M=4 (real roots I = +-1/2, +-3/2)
HBM_PEAK_GBS = 8000.0
WARMUP_MIN_S = 0.1  # minimum warm-up of the timed H.v loops (see the loop)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sec, budget_s=12.0):
    """H.v/s of the reference on the host cores, extrapolated from a bounded
    sample of rows.  With oracle/_ref built (the reference's own modules
    compiled from its sources), the reference's ham_to_vec_wave_vector loop
    (ham2vec_v3.f90:52-70: all-bond scan, representative_SQ's translation
    loop, binary-search findstate; OpenMP over rows) runs on windows of rows
    spread over the sector, against this sector's lists (kind "reference");
    otherwise the oracle's C restatement of the same loop (kind "port")."""
    from oracle import qs3 as orc
    L, N = 1000, 4
    bonds = orc.chain_bonds(L, [(1, 1.0, 1.0)])
    ls, lr = sec.lists()
    rng = np.random.default_rng(0)
    v = (rng.uniform(-1, 1, sec.dim) + 1j * rng.uniform(-1, 1, sec.dim)).astype(np.complex128)
    chunk = 2048
    starts = np.linspace(0, sec.dim - chunk, 512).astype(np.int64)
    if orc.RefLib.available():
        ref = orc.RefLib()
        ref.setup(L, L, 1, 1, N, (0, 0, 0), bonds)
        ref.set_lists(ls, lr)
        out = np.zeros_like(v)
        run = lambda a, b: ref.hv_rows(v, a, b, out=out)  # noqa: E731
        kind, what = "reference", "the reference's ham_to_vec_wave_vector loop (oracle/_ref, compiled from its sources)"
    else:
        om = orc.Model(L, L, 1, 1, bonds)
        osec = orc.Sector(N, (0.0, 0.0, 0.0), ls, lr, orc.explist(L, 1, 1, 0.0, 0.0, 0.0))
        run = lambda a, b: om.hv(osec, v, rows=(a, b))  # noqa: E731
        kind, what = "port", "the oracle's C restatement of that loop"
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    rows, n = 0, 0
    t0 = time.perf_counter()
    for st in starts:
        run(int(st), int(st) + chunk)
        rows += chunk
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    rate = rows / dt / sec.dim
    return {"value": rate, "unit": "H·v/s", "cores": threads, "kind": kind, "cpu": cpu_model(),
            "sample": f"{rows} rows ({n} windows of {chunk} spread over the sector) of the L=1000 N=4 k=0 sector "
                      f"in {dt:.1f} s with {what}, OMP_NUM_THREADS={threads}, extrapolated to D={sec.dim}"}


def first_within(trace, e_ref, rel, prec=None):
    """(seconds, H.v count) of the first trace row whose Ritz value is within
    rel of e_ref (optionally only rows of the given precision), or None."""
    for t, hv, e, p in trace:
        if (prec is None or int(p) == prec) and abs(e / e_ref - 1.0) < rel:
            return t, int(hv)
    return None


def calibrate_copy(nbytes=1 << 30, reps=5):
    """A plain device copy of nbytes, for scaling the FETCH_SIZE / WRITE_SIZE
    counters (run under rocprofv3 --pmc)."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}), flush=True)


def dsf_workload(args):
    """Dynamical structure factor S^{-+}(q, w) of the L = 1000 NN Heisenberg
    chain: the lowest k = 0 state of N_down = 3 (D = 166,167), S^-_q into the
    N_down = 4, k = q sectors for 64 q = 2 pi n_m / 1000, n_m = round(1000 m / 64),
    and 64 continued fractions of `--dsf-steps` Lanczos steps, all 64 in one
    batched block (FP32 vectors, FP64 coefficients).  Reports the batched time,
    the per-sector time for a sample of q (extrapolated to 64), and the error
    of the FP32 spectra against the FP64 batch (Lorentzian eta = 0.02, 2001
    points of w in [0, 5])."""
    import torch
    from ed_amd import qs3 as q3
    L, Ngs = 1000, 3
    gm = q3.Model(L, L, 1, 1, q3.chain_bonds(L, [(1, 1.0, 1.0)]))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gs = q3.Sector.build(gm, Ngs, (0.0, 0.0, 0.0))
    E0, psi, it = q3.lanczos_routines(gs, maxitr=3000, lnc_ene_conv=1e-13, prec=q3.FP64, seed=1)
    t_gs = time.perf_counter() - t0
    log(f"ground state N_down={Ngs} k=0: D={gs.dim}, E0={E0:.12f}, {it} steps, {t_gs:.2f} s")
    pos = np.array([[i, 0.0, 0.0] for i in range(L)], dtype=float)
    Ks = [int(round(1000 * m / 64)) for m in range(64)]
    steps = args.dsf_steps
    out = {}
    block_tune = {}
    if args.tune:
        q3.set_option("block_autotune", 1)
    q3.set_option("block_fused", args.block_fused)
    for prec, name in ((q3.FP32, "fp32"), (q3.FP64, "fp64")):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        a, b, n, st = q3.dsf_sminus_batch(gs, Ks, pos, psi, steps, prec=prec)
        torch.cuda.synchronize()
        out[name] = (a, b, n, st, time.perf_counter() - t0)
        log(f"batch {name}: {out[name][4]:.2f} s, {st}")
        if args.tune:
            tab = q3.block_tune_last()
            block_tune[name] = {"chosen": [int(x) for x in tab[-1][:3]] + [round(float(tab[-1][3]), 3)],
                                "candidates": [[int(x) for x in r[:3]] + [round(float(r[3]), 3)] for r in tab[:-1]]}
            log(f"block stencil autotune {name}: (lines, threads, slabs, ms) {block_tune[name]['chosen']}")
    q3.set_option("block_autotune", 0)
    a32, b32, n32, st32, t32 = out["fp32"]
    a64, b64, n64, st64, t64 = out["fp64"]
    om = np.linspace(0.0, 5.0, 2001)
    errs, errs_wide, coef = [], [], []
    for j in range(64):
        if n64[j] <= 0:
            continue
        s64 = q3.dsf_analysis(a64[j], b64[j], n64[j], E0, om, 0.02)
        s32 = q3.dsf_analysis(a32[j], b32[j], n32[j], E0, om, 0.02)
        scale = max(np.max(np.abs(s64)), 1e-300)
        errs.append(np.max(np.abs(s32 - s64)) / scale)
        # the same spectra at eta = 0.1 (wider than the spacing of the
        # 200 Lanczos poles, so pole-position shifts average out)
        w64 = q3.dsf_analysis(a64[j], b64[j], n64[j], E0, om, 0.1)
        w32 = q3.dsf_analysis(a32[j], b32[j], n32[j], E0, om, 0.1)
        errs_wide.append(np.max(np.abs(w32 - w64)) / max(np.max(np.abs(w64)), 1e-300))
        # the first 10 coefficients (they fix the first 20 spectral moments)
        k = 10
        coef.append(max(np.max(np.abs(a32[j, :k] - a64[j, :k])) / max(1.0, np.max(np.abs(a64[j, :k]))),
                        np.max(np.abs(b32[j, :k] - b64[j, :k])) / max(1e-300, np.max(np.abs(b64[j, :k])))))
    errs, errs_wide, coef = np.array(errs), np.array(errs_wide), np.array(coef)
    # per-sector path for a sample of q, FP32 (sector build + S^-_q + recurrence)
    sample = [1, 17, 40]
    tps = []
    for j in sample:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rk = q3.wavevector(Ks[j], L)
        tgt = q3.Sector.build(gm, Ngs + 1, (rk, 0.0, 0.0))
        f = q3.calcu_Sq_a(2, gs, tgt, np.array([rk, 0.0, 0.0]), pos, psi)
        q3.dsf_recurrence(tgt, f, steps, prec=q3.FP32)
        torch.cuda.synchronize()
        tps.append(time.perf_counter() - t0)
        del tgt, f
    per_sector = float(np.mean(tps)) * 64
    res = {"metric": "DSF S^{-+}(q,w): 64 q x %d-step continued fractions, L=1000 N_down=3->4" % steps,
           "value": t32, "unit": "s", "higher_is_better": False, "dtype": "complex-fp32 (fp64 coefficients)",
           "data": "synthetic", "config": {"workload": "periodic NN Heisenberg chain L=1000, ground state N_down=3 "
                                                       "k=0, S^-_q into N_down=4 k=q (D~41.4M each), 64 q",
                                           "steps": steps, "eta": 0.02, "omega_points": 2001,
                                           "block_step": "fused" if args.block_fused else "separate"},
           "seconds_fp32_batch": t32, "seconds_fp64_batch": t64, "ms_per_block_step_fp32": st32["ms_per_block_step"],
           "ms_per_block_step_fp64": st64["ms_per_block_step"], "seconds_per_sector_fp32_extrapolated": per_sector,
           "per_sector_sample_seconds": tps, "speedup_batch_vs_per_sector": per_sector / t32,
           "fp32_vs_fp64_S_rel_err_median": float(np.median(errs)), "fp32_vs_fp64_S_rel_err_max": float(np.max(errs)),
           "fp32_vs_fp64_S_rel_err_q_counted": int(len(errs)),
           "fp32_vs_fp64_S_eta0.1_rel_err_median": float(np.median(errs_wide)),
           "fp32_vs_fp64_S_eta0.1_rel_err_max": float(np.max(errs_wide)),
           "fp32_vs_fp64_first10_coef_rel_err_median": float(np.median(coef)),
           "fp32_vs_fp64_first10_coef_rel_err_max": float(np.max(coef)),
           "ground_state_seconds": t_gs, "E0": E0, "block_autotune": block_tune}
    print(json.dumps(res), flush=True)


def sectors_workload(args):
    """Lowest eigenvalue of every (N_down, k) sector of the periodic XXZ chain
    L = 40, Jxy = 1, Jz = 0.8, N_down = 1..7, all 40 k (280 sectors, about
    23.2 M rows): plain Lanczos (FP32-mixed) for all sectors at once, with and
    without hipGraph replay of the step, against sector-by-sector solves; the
    error of every E0 against the FP64 group solve."""
    import torch
    from ed_amd import qs3 as q3
    L = 40
    gm = q3.Model(L, L, 1, 1, q3.chain_bonds(L, [(1, 1.0, 0.8)]))
    t0 = time.perf_counter()
    secs, keys = [], []
    for N in range(1, 8):
        for K in range(L):
            s = q3.Sector.build(gm, N, (q3.wavevector(K, L), 0.0, 0.0))
            if s.dim:
                secs.append(s)
                keys.append((N, K))
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    rows = sum(s.dim for s in secs)
    log(f"{len(secs)} sectors, {rows} rows, built in {t_build:.2f} s")
    kw = dict(maxitr=400, minitr=20, itrint=5, lnc_ene_conv=args.sectors_conv, seed=1)
    res = {}
    for name, prec, graph in (("fp32_graph", q3.FP32, True), ("fp32_nograph", q3.FP32, False),
                              ("fp64_graph", q3.FP64, True)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E, its, st = q3.lanczos_group(secs, prec=prec, use_graph=graph, **kw)
        torch.cuda.synchronize()
        res[name] = (E, its, st, time.perf_counter() - t0)
        log(f"group {name}: {res[name][3]:.3f} s, {st}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E1 = []
    for s in secs:
        e, _, _ = q3.lanczos_routines(s, prec=q3.FP32, want_vector=False, **kw)
        E1.append(e)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    E32, it32, st32, t32 = res["fp32_graph"]
    E64 = res["fp64_graph"][0]
    # |dE| / max(1, |E|): several sector ground states lie near E = 0
    rel = np.abs(E32 - E64) / np.maximum(np.abs(E64), 1.0)
    out = {"metric": "lowest eigenvalue of 280 (N_down, k) sectors, XXZ chain L=40 Jz=0.8, FP32-mixed",
           "value": t32, "unit": "s", "higher_is_better": False, "dtype": "complex-fp32 (fp64 alpha/beta)",
           "data": "synthetic", "config": {"workload": "periodic XXZ chain L=40, Jxy=1, Jz=0.8, N_down=1..7, "
                                                       "all 40 k", "sectors": len(secs), "rows": int(rows),
                                           "conv": args.sectors_conv, "maxitr": 400},
           "seconds_group_fp32_graph": t32, "seconds_group_fp32_no_graph": res["fp32_nograph"][3],
           "seconds_group_fp64_graph": res["fp64_graph"][3], "seconds_sector_by_sector_fp32": t_seq,
           "speedup_group_vs_sector_by_sector": t_seq / t32, "group_steps": st32["steps"],
           "max_iters_fp32": int(np.max(it32)), "fp32_vs_fp64_E0_err_max": float(np.max(rel)),
           "fp32_vs_fp64_E0_err_median": float(np.median(rel)),
           "fp32_vs_fp64_E0_err_definition": "|E32 - E64| / max(1, |E64|)",
           "group_vs_sector_by_sector_fp32_max_rel_diff": float(np.max(np.abs(E32 - np.array(E1)) /
                                                                       np.maximum(np.abs(E32), 1e-300))),
           "sector_build_seconds": t_build}
    print(json.dumps(out), flush=True)


def load_traffic(kernel, launch=None):
    """HBM bytes per launch of the given kernel from the committed counter
    profile (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by a plain copy), or
    None when the profile covers another kernel or launch configuration
    (tile height and task slabs of the stencil)."""
    p = os.path.join(ROOT, "profiles", "hv_stencil_L1000_N4_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        t = json.load(f).get(kernel)
    if t and launch is not None:
        for k, v in t.get("launch", {}).items():
            if launch.get(k) != v:
                return None
    return t


def stencil_kernel_name(prec, form, fused=False):
    """The stencil kernel of a launch: k_hv_stencil4 (slab-window form),
    k_hv_stencil_col (column form) or k_hv_stencil_pair (pair form, FP32)."""
    r = "double" if prec == 0 else "float"
    f = ", fused" if fused else ""
    if form == 2:
        return f"k_hv_stencil_pair<real-phase{f}>"
    return f"{'k_hv_stencil_col' if form == 1 else 'k_hv_stencil4'}<{r}, real-phase{f}>"


def hv_bytes(sec, prec, sv):
    """Algorithmic bytes of one H.v launch, a lower bound: every array the
    launch must read or write counted once -- v_in and v_out (s_v bytes per
    row each) and the index data of the kernel that ran.  Stencil (family 5):
    16 bytes per (n3, n4) line of the line table, and 12 bytes per row left to
    the key search (its list entry and configuration word); nearest-neighbour
    chain kernel (family 0): the 8-byte configuration word of every row and
    the bitmap index (8 bytes per 32 anchored ranks).  Returns (bytes, kernel
    name, launch description)."""
    fam, lds = sec.hv_last_launch()
    D = sec.dim
    tag = "fp64" if prec == 0 else "fp32"
    if fam == 5:
        info = sec.hv_stencil_info(prec)
        b = D * 2 * sv + info["lines"] * 16 + info["slow_rows"] * 12
        name = f"{stencil_kernel_name(prec, info['form'])} + k_hv_slow_rows ({tag})"
        cfg = {"family": "stencil", "form": ["slab window", "column", "pair"][info["form"]],
               "tile_lines": info["tile_lines"], "slabs": info["slabs"],
               "workgroups": info["grid"], "slow_rows": info["slow_rows"], "lines": info["lines"], "lds": lds}
    else:
        import math
        bitmap = 8 * ((math.comb(999, 3) + 31) // 32)
        b = D * (8 + 2 * sv) + bitmap
        name = f"H.v kernel family {fam} ({tag})"
        cfg = {"family": fam, "lds": lds}
    return b, name, cfg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", action="store_true",
                    help="run the H.v launch autotuner first (the timed H.v and the solves then use its choice)")
    ap.add_argument("--lanczos-steps", type=int, default=200,
                    help="plain Lanczos steps timed per precision (ms per step of the solver path)")
    ap.add_argument("--NOK", type=int, default=4)
    # one first cycle holds the whole convergence (the spectrum's low end is a
    # quasi-continuum, level spacing ~1e-4, so small bases restart many
    # times: NOM 128 takes 1.5 s to 1e-9 against 0.37 s; NOM 760 restarts
    # once, 10 steps short of its convergence test); 801 complex-float
    # vectors are 265 GB
    ap.add_argument("--NOM", type=int, default=800,
                    help="basis of the FP32-mixed thick restart")
    # the FP64 thick restart with the same basis memory as the FP32-mixed one
    # (401 complex-double vectors, 265 GB): the same algorithm in FP64
    ap.add_argument("--NOM64", type=int, default=400,
                    help="basis of the FP64 thick-restart leg (0 skips it)")
    ap.add_argument("--reorth", type=int, default=2,
                    help="thick-restart reorthogonalisation: 0 full, 1 locked+local, 2 locked at restart only")
    ap.add_argument("--cgs", type=int, default=1, help="Gram-Schmidt passes per thick-restart step")
    ap.add_argument("--fp64-lanczos", type=int, default=600,
                    help="FP32-mixed thick restart: at most this many FP64 Lanczos steps from the Rayleigh-Ritz "
                         "vector until the residual bound certifies E0 (0: the block-Krylov refinement instead; "
                         "at L=1000 the bound certifies after about 305 steps)")
    ap.add_argument("--fp32-floor", type=float, default=1e-7)
    ap.add_argument("--check-every", type=int, default=5,
                    help="thick restart: convergence check every n steps inside cycles (0: at restarts)")
    ap.add_argument("--conv-plain", type=float, default=1e-11)
    # the thick restarts' tolerance: the FP32 cycles' relative-change test and
    # the Kato-Temple bound r^2 / gap <= conv |E| that certifies the energy
    # (1e-9: the target accuracy)
    ap.add_argument("--conv-mixed", type=float, default=1e-9)
    # the FP64 thick restart stops on the reference's relative-change test
    # only, so it needs a margin to land within 1e-9
    ap.add_argument("--conv-fp64-tr", type=float, default=3e-10)
    ap.add_argument("--calibrate-copy", action="store_true", help="only run a 1 GiB device copy (counter scaling)")
    ap.add_argument("--dsf", action="store_true", help="run the batched 64-q DSF workload instead of the headline")
    ap.add_argument("--dsf-steps", type=int, default=200)
    ap.add_argument("--block-fused", type=int, default=1, choices=(0, 1),
                    help="batched DSF: 1 the fused block step (default), 0 block stencil H.v then block update")
    ap.add_argument("--sectors", action="store_true", help="run the 280-sector L=40 workload instead of the headline")
    ap.add_argument("--sectors-conv", type=float, default=1e-10)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X; runs on several GPUs are not supported (the H.v would "
              "shard by n4 slabs, with one halo slab exchanged with each neighbour).", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.dsf:
        dsf_workload(args)
        return
    if args.sectors:
        sectors_workload(args)
        return

    import torch
    from ed_amd import qs3 as q3

    L, N = 1000, 4
    gm = q3.Model(L, L, 1, 1, q3.chain_bonds(L, [(1, 1.0, 1.0)]))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sec = q3.Sector.build(gm, N, (0.0, 0.0, 0.0))
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    D = sec.dim
    nnz = sec.nnz()
    log(f"sector D={D} built in {t_build:.3f} s, nnz={nnz} ({nnz / D:.3f} hops/row)")

    # launch autotuner: kernel, tile schedule and grid per precision (the
    # timed H.v and the solvers below use the chosen configuration)
    tune = {}
    if args.tune:
        log("autotuning the H.v launch (kernel family, then its tile / schedule parameters) ...")
        for name, prec in (("fp32", q3.FP32), ("fp64", q3.FP64)):
            best, bms, table = sec.hv_autotune(prec, n_rep=5)
            tune[name] = dict(best, ms=bms, candidates=[[int(x) for x in r[:11]] + [round(float(r[11]), 4)]
                                                        for r in table])
            log(f"autotune {name}: {best}: {bms:.4f} ms (default config {float(table[0][11]):.4f} ms, "
                f"{len(table)} candidates)")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    res = {}
    for name, prec, dt, sv in (("fp32", q3.FP32, torch.complex64, 8), ("fp64", q3.FP64, torch.complex128, 16)):
        g = torch.Generator(device=dev)
        g.manual_seed(1)
        x = torch.randn(D, dtype=dt, device=dev, generator=g)
        y = torch.empty_like(x)
        # W warm-up launches, continued until WARMUP_MIN_S of them have run:
        # the launch time settles over the first ~20 ms of back-to-back H.v
        # after an idle device (power management), and the timed steps should
        # see the steady rate the solvers run at.
        nwarm = 0
        tw = time.perf_counter()
        while nwarm < args.warmup or time.perf_counter() - tw < WARMUP_MIN_S:
            for _ in range(8 if nwarm >= args.warmup else 1):
                sec.hv_dev(y.data_ptr(), x.data_ptr(), prec, stream.cuda_stream)
                nwarm += 1
            if nwarm >= args.warmup:
                torch.cuda.synchronize()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            sec.hv_dev(y.data_ptr(), x.data_ptr(), prec, stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms_ev = e0.elapsed_time(e1) / args.steps
        B, kname, kcfg = hv_bytes(sec, prec, sv)
        res[name] = {"ms": ms_ev, "wall_ms": wall * 1e3 / args.steps, "hv_per_s": args.steps / wall,
                     "warmup_launches": nwarm, "alg_bytes": B, "alg_GBs": B / (ms_ev * 1e-3) / 1e9,
                     "kernel": kname, "launch": kcfg,
                     # the gather model: every hop's target element charged to HBM
                     "all_gathers_bytes": D * (16 + 2 * sv) + nnz * sv}
        log(f"{name}: {ms_ev:.4f} ms/H.v (events), {wall * 1e3 / args.steps:.4f} ms (wall), "
            f"{res[name]['alg_GBs']:.1f} GB/s of compulsory bytes ({kname})")
        del x, y
        torch.cuda.empty_cache()

    # A fixed number of plain Lanczos steps per precision, without the
    # convergence checks' effect on the count (ms per step of the path the
    # solvers take: the fused step in FP64, the H.v and update kernels in FP32)
    lz = {}
    for name, prec in (("fp32", q3.FP32), ("fp64", q3.FP64)):
        kw = dict(minitr=10 ** 6, itrint=10 ** 6, lnc_ene_conv=1e-30, prec=prec, seed=1, want_vector=False)
        q3.lanczos_routines(sec, maxitr=20, **kw)  # buffers, plans
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, _, n = q3.lanczos_routines(sec, maxitr=args.lanczos_steps, **kw)
        torch.cuda.synchronize()
        fam = sec.hv_last_launch()[0]
        lz[name] = {"ms_per_step": (time.perf_counter() - t0) * 1e3 / n, "steps": n,
                    "path": (f"fused step ({stencil_kernel_name(prec, sec.hv_stencil_info(prec)['form'], True)}"
                             " + k_hv_slow_rows fused + k_finalize_fused)") if fam == 6 else
                            "H.v + k_update3 (+ finalize)"}
        log(f"plain Lanczos {name}: {lz[name]['ms_per_step']:.4f} ms per step over {n} steps, {lz[name]['path']}")

    solve = {}
    if not args.no_solve:
        q3.set_option("verbose", 1)  # one stderr line per restart / 100 Lanczos steps
        q3.set_option("tr_reorth", args.reorth)
        q3.set_option("tr_cgs_passes", args.cgs)
        q3.set_option("fp32_floor", args.fp32_floor)
        q3.set_option("tr_check_every", args.check_every)
        q3.set_option("fp64_lanczos", args.fp64_lanczos)

        def thick(prec, nom, conv):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ene, _, st = q3.my_trlanczos_routines_complex(sec, NOE=1, NOK=args.NOK, NOM=nom, maxitr=400,
                                                           lnc_ene_conv=conv, prec=prec, seed=1,
                                                           want_vectors=False)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            tr = q3.solver_trace()
            alloc = st.get("seconds_alloc") or 0.0  # allocating the Krylov bases
            freed = q3.solver_certificate()["seconds_free"]  # freeing them, after the energies are known
            hit = first_within(tr, E0_EXACT, 1e-9)
            hit64v = first_within(tr, E0_EXACT, 1e-9, prec=64)
            E = float(ene[0])
            return {"E0": E, "rel_err": abs(E / E0_EXACT - 1), "NOM": nom, "wall_s": wall,
                    "seconds_solve": st["seconds"], "seconds_alloc": alloc, "seconds_free": freed,
                    "seconds_including_alloc": st["seconds"] + alloc,
                    # first Ritz value within 1e-9 (solver trace, allocation excluded; seconds, H.v)
                    "time_to_1e-9": hit, "time_to_1e-9_fp64_value": hit64v,
                    "time_to_1e-9_including_alloc": (hit[0] + alloc) if hit else None,
                    "time_to_1e-9_fp64_value_including_alloc": (hit64v[0] + alloc) if hit64v else None,
                    "stats": st}

        # the thick restart first: its basis is the process's first large
        # allocation, as in a driver that solves once
        try:
            tr32 = thick(q3.FP32, args.NOM, args.conv_mixed)
        except q3.QS3XError as e:  # its basis not available on this device: a smaller one
            nom = args.NOM * 3 // 4
            log(f"FP32-mixed thick restart (NOM {args.NOM}) failed ({e}); NOM {nom} instead")
            args.NOM = nom
            tr32 = thick(q3.FP32, args.NOM, args.conv_mixed)
        cert = q3.solver_certificate()
        log(f"FP32-mixed thick restart (NOM {args.NOM}): E0={tr32['E0']:.15f}, {tr32}, certificate {cert}")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E64, _, it64 = q3.lanczos_routines(sec, maxitr=5000, minitr=20, itrint=5, lnc_ene_conv=args.conv_plain,
                                           prec=q3.FP64, seed=1, want_vector=False)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        tr64 = q3.solver_trace()
        hit64 = first_within(tr64, E0_EXACT, 1e-9)
        log(f"plain FP64 Lanczos: E0={E64:.15f} in {t64:.2f} s, {it64} steps, rel err {abs(E64 / E0_EXACT - 1):.2e}, "
            f"first within 1e-9 {hit64}")
        # the same three-term recurrence on complex-float vectors (FP64 alpha/beta)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E32p, _, it32p = q3.lanczos_routines(sec, maxitr=5000, minitr=20, itrint=5, lnc_ene_conv=args.conv_plain,
                                             prec=q3.FP32, seed=1, want_vector=False)
        torch.cuda.synchronize()
        t32p = time.perf_counter() - t0
        hit32p = first_within(q3.solver_trace(), E0_EXACT, 1e-9)
        log(f"plain FP32-mixed Lanczos: E0={E32p:.15f} in {t32p:.2f} s, {it32p} steps, "
            f"rel err {abs(E32p / E0_EXACT - 1):.2e}, first within 1e-9 {hit32p}")
        solve = {"time_to_E0_fp64_plain_s": t64, "E0_fp64_plain": E64, "iters_fp64_plain": it64,
                 "rel_err_fp64_plain": abs(E64 / E0_EXACT - 1), "time_to_1e-9_fp64_plain": hit64,
                 "plain_fp32_mixed": {"E0": E32p, "iters": it32p, "seconds": t32p,
                                      "rel_err": abs(E32p / E0_EXACT - 1), "time_to_1e-9": hit32p},
                 "fp32_mixed_thick_restart": tr32,
                 # Kato-Temple bound r^2 / gap of the FP64-refined value (gap: the FP32
                 # projection's next level); accepted 1 = certified at conv, 2 = by
                 # agreement of the FP32 and FP64 values only
                 "fp32_mixed_certificate": cert,
                 "fp32_mixed_settings": {"NOK": args.NOK, "NOM": args.NOM, "tr_reorth": args.reorth,
                                         "cgs_passes": args.cgs, "fp32_floor": args.fp32_floor,
                                         "fp64_lanczos": args.fp64_lanczos,
                                         "conv": args.conv_mixed, "check_every": args.check_every},
                 "fp64_plain_settings": {"conv": args.conv_plain, "itrint": 5, "minitr": 20},
                 "fp64_thick_restart_settings": {"NOM": args.NOM64, "conv": args.conv_fp64_tr}}
        if args.NOM64 > 0:
            try:
                tr64 = thick(q3.FP64, args.NOM64, args.conv_fp64_tr)
                log(f"FP64 thick restart (NOM {args.NOM64}): E0={tr64['E0']:.15f}, {tr64}")
                solve["fp64_thick_restart"] = tr64
            except q3.QS3XError as e:  # e.g. its 252 GB basis not available on this device
                log(f"FP64 thick restart (NOM {args.NOM64}) skipped: {e}")
        # time to the first Ritz value within 1e-9 of E0, allocation included:
        # plain FP64 Lanczos (the reference's ALG=1 path; its three vectors are
        # allocated inside the timed call) over each FP32-mixed solver
        ratio = {}
        if hit64:
            d = hit64[0]
            if hit32p:
                ratio["fp32_plain"] = d / hit32p[0]
            if tr32["time_to_1e-9_including_alloc"]:
                ratio["fp32_thick_restart"] = d / tr32["time_to_1e-9_including_alloc"]
            if tr32["time_to_1e-9_fp64_value_including_alloc"]:
                ratio["fp32_thick_restart_fp64_value"] = d / tr32["time_to_1e-9_fp64_value_including_alloc"]
        solve["speedup_to_1e-9_vs_fp64_plain"] = ratio
        # the same solver in FP64 with the same basis memory, solve times (its
        # basis allocation follows the free of the FP32 one and is reported
        # apart, in fp64_thick_restart)
        if "fp64_thick_restart" in solve and solve["fp64_thick_restart"]["time_to_1e-9"]:
            d = solve["fp64_thick_restart"]["time_to_1e-9"][0]
            r2 = {}
            if tr32["time_to_1e-9"]:
                r2["fp32_thick_restart"] = d / tr32["time_to_1e-9"][0]
            if tr32["time_to_1e-9_fp64_value"]:
                r2["fp32_thick_restart_fp64_value"] = d / tr32["time_to_1e-9_fp64_value"][0]
            solve["speedup_to_1e-9_vs_fp64_thick_restart"] = r2
            log(f"speed-up to 1e-9 over the FP64 thick restart (NOM {args.NOM64}, solve times): {r2}")
        solve["time_to_1e-9_including_build"] = {
            "fp64_plain": (hit64[0] + t_build) if hit64 else None,
            "fp32_plain": (hit32p[0] + t_build) if hit32p else None,
            "fp32_thick_restart": (tr32["time_to_1e-9_including_alloc"] + t_build)
            if tr32["time_to_1e-9_including_alloc"] else None}
        log(f"speed-up to 1e-9 over plain FP64 (allocation included): {ratio}")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(sec)
            log(f"cpu baseline: {cpu}")
        except Exception as e:  # the oracle library may be absent on a stripped tree
            log(f"cpu baseline unavailable: {e}")

    f32 = res["fp32"]
    traffic = load_traffic(f32["kernel"], f32["launch"])
    tb = (traffic or {}).get("bytes_per_launch")
    roof = {"bound": "hbm", "achieved": f32["alg_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": f32["alg_GBs"] / HBM_PEAK_GBS, "traffic": tb,
            "alg_bytes_per_launch": f32["alg_bytes"],
            "alg_bytes_rule": "lower bound: v_in and v_out once (8 B each per row in FP32) plus the kernel's "
                              "index data once (16 B per (n3, n4) line, 12 B per key-search row)",
            "traffic_over_alg": (tb / f32["alg_bytes"]) if tb else None,
            "counter_GBs": (tb / (f32["ms"] * 1e-3) / 1e9) if tb else None,
            "counter_frac": (tb / (f32["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS) if tb else None,
            "all_gathers_model_bytes": f32["all_gathers_bytes"],
            "kernel": f32["kernel"], "launch_config": f32["launch"], "avg_launch_ms": f32["ms"]}
    f64 = res["fp64"]
    tb64 = (load_traffic(f64["kernel"], f64["launch"]) or {}).get("bytes_per_launch")
    out = {
        "metric": METRIC,
        "value": f32["hv_per_s"],
        "unit": "H·v/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_launches": {"fp32": f32["warmup_launches"], "fp64": f64["warmup_launches"],
                            "min_seconds": WARMUP_MIN_S},
        "ms_per_step": f32["wall_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex-fp32 (fp64 reductions)",
        "data": "synthetic",
        "config": {"workload": "periodic NN Heisenberg chain L=1000, Jxy=Jz=1, N_down=4, k=0 "
                               f"(D={D}), one full-sector H.v per step", "dim": D, "nnz": nnz,
                   "parallelism": "1 GPU"},
        "hv_per_s_fp64": f64["hv_per_s"],
        "ms_per_hv_fp64": f64["ms"],
        "roofline_fp64": {"achieved": f64["alg_GBs"], "frac": f64["alg_GBs"] / HBM_PEAK_GBS,
                          "alg_bytes_per_launch": f64["alg_bytes"], "kernel": f64["kernel"],
                          "launch_config": f64["launch"], "traffic": tb64,
                          "counter_GBs": (tb64 / (f64["ms"] * 1e-3) / 1e9) if tb64 else None},
        "sector_build_s": t_build,
        "hv_autotune": tune,
        "plain_lanczos_step": lz,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(solve)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
