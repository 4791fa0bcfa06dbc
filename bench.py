#!/usr/bin/env python3
"""Benchmark: mixed-precision vs plain FP64 velocity solve on one MI355X.

Workload: the synthetic ice-stream system at 1024x1024 vertices x 11 levels
(23,068,672 unknowns), assembled on the GPU from a plug-plus-shear velocity
guess.  A "step" is one complete linear solve of that system from x0 = 0 to
a relative residual of 1e-10 with the Chronopoulos-Gear method and the SIA
column-tridiagonal preconditioner:

* mixed: FP32 matrix and FP32 inner Krylov iterations inside FP64 iterative
  refinement (FP64 residual on the FP64 matrix, compensated FP64 sums);
* fp64:  the plain FP64 drop-in solver (reference semantics).

value = mixed solves per second; the FP64 rate, the speedup, the error of
each solution against the oracle's FP64 solve of the same exported system
(the CPU restatement on all host cores, timed as the CPU baseline), the
SpMV bandwidth against the HBM roofline, the BF16-storage variant, the
512^2 x 11 FP64 solve, the batched ensemble, the marine-sheet Picard solve
and the depth-integrated 2-D solves are reported beside it.  Prints one
JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0   # MI355X HBM3E, 8 TB/s


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1024, help="owned vertices per horizontal direction")
    ap.add_argument("--nz", type=int, default=11)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--precision", choices=["fp32", "bf16"], default="fp32")
    ap.add_argument("--inner-tol", type=float, default=1e-6)
    ap.add_argument("--cpu-iters", type=int, default=10, help="C-G iterations of each one-core CPU sample")
    ap.add_argument("--cpu-repeats", type=int, default=3, help="one-core samples (median taken)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-all", action="store_true", help="skip the all-cores oracle solve")
    ap.add_argument("--ensemble", type=int, default=256, help="members of the batched-ensemble run (0: skip)")
    ap.add_argument("--picard-n", type=int, default=1024, help="marine-sheet Picard run size (0: skip)")
    ap.add_argument("--picard-maxiters", type=int, default=1000, help="linear_maxiters of the Picard legs")
    ap.add_argument("--no-picard-fp64", action="store_true", help="skip the FP64-inner Picard beside the mixed one")
    ap.add_argument("--no-twod", action="store_true", help="skip the depth-integrated 2-D solver leg")
    ap.add_argument("--no-bf16", action="store_true", help="skip the BF16-storage solve leg")
    ap.add_argument("--config2-n", type=int, default=512, help="FP64 C-G solve leg size (0: skip)")
    ap.add_argument("--only-ensemble", action="store_true", help="run the ensemble leg only (profiling)")
    ap.add_argument("--out", default=None, help="also write the JSON line to this file")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the single-GPU share of the path); --gpus {args.gpus} is not "
              f"supported: the multi-GPU halo/all-reduce design is described but not built.", file=sys.stderr)
        sys.exit(2)

    if args.only_ensemble:
        # the STREAM probe runs too: its copy kernel of known size calibrates
        # the hardware counters of a profiled run of this leg
        from cism_amd import _lib
        hbm = {k: round(v, 1) for k, v in _lib.bench_stream(2 << 30, 3).items()}
        print(json.dumps({"ensemble": run_ensemble(args), "hbm_measured_gbps": hbm}), flush=True)
        return

    import numpy as np

    from cism_amd import pcg, synthetic
    from cism_amd.velo import BPSolver

    t0 = time.time()
    log(f"generating synthetic ice stream {args.n}x{args.n}x{args.nz} seed={args.seed}")
    prob = synthetic.ice_stream(args.n, nz=args.nz, seed=args.seed)
    g = prob.grid
    nodes = g.n_owned_nodes
    log(f"grid nx={g.nx} ny={g.ny} nz={g.nz}: {nodes} owned nodes, {2 * nodes} unknowns")
    bp = BPSolver(prob)
    ainfo = bp.assemble(prob.uvel, prob.vvel)
    s = bp.sys
    log(f"GPU assembly: element kernel {ainfo['assemble_ms']:.1f} ms, total {ainfo['total_ms']:.1f} ms, "
        f"nnz {ainfo['nnz']}")
    lowp = pcg.MIXED_FP32 if args.precision == "fp32" else pcg.MIXED_BF16

    def fp64_step(time_spmv=False):
        s.zero_x()
        return s.solve(precond=2, tolerance=args.tol, maxiters=50000, time_spmv=time_spmv)

    def mixed_step(precision=lowp, time_spmv=False):
        s.zero_x()
        return s.solve(precond=2, tolerance=args.tol, maxiters=50000, precision=precision, inner_tol=args.inner_tol,
                       max_outer=40, time_spmv=time_spmv)

    # The first solve after an assembly checks exact symmetry and builds the
    # half-storage and low-precision copies; later solves of the same matrix
    # reuse them.  A Picard iteration pays this once per assembly.
    setup = {}
    for w in range(max(args.warmup, 1)):
        r = fp64_step()
        if w == 0:
            setup["fp64_ms"] = round(r["setup_ms"], 3)
        log(f"warmup fp64 {w}: {r['niters']} it {r['solve_ms']:.1f} ms (set-up {r['setup_ms']:.2f} ms)")
        r = mixed_step()
        if w == 0:
            setup["mixed_ms"] = round(r["setup_ms"], 3)
        log(f"warmup mixed {w}: {r['niters']} inner it, {r['outer_iters']} outer, {r['solve_ms']:.1f} ms "
            f"(set-up {r['setup_ms']:.2f} ms)")
    setup["note"] = ("host wall time of the per-assembly set-up done by the first solve after an assembly "
                     "(exact-symmetry checks, FP64 half-storage copy; mixed: FP32 copy and preconditioner cast); "
                     "cached and outside the timed solves below, paid once per Picard assembly")

    # ---- timed FP64 solves
    t64, it64, sp64 = [], [], []
    for k in range(args.steps):
        tw = time.perf_counter()
        r = fp64_step(time_spmv=True)
        t64.append(time.perf_counter() - tw)
        it64.append(r["niters"])
        sp64.append((r["spmv_ms"], r["spmv_count"]))
        r64_kind = r["spmv_kind"]
        log(f"fp64 step {k}: {r['niters']} it, device {r['solve_ms']:.1f} ms, wall {t64[-1] * 1e3:.1f} ms")
    x64 = np.stack(s.get_x())

    # ---- timed mixed solves (the metric)
    tm, itm, outm, spm = [], [], [], []
    for k in range(args.steps):
        tw = time.perf_counter()
        r = mixed_step(time_spmv=True)
        tm.append(time.perf_counter() - tw)
        itm.append(r["niters"])
        outm.append(r["outer_iters"])
        spm.append((r["spmv_ms"], r["spmv_count"]))
        log(f"mixed step {k}: {r['niters']} inner it, {r['outer_iters']} outer, device {r['solve_ms']:.1f} ms, "
            f"wall {tm[-1] * 1e3:.1f} ms, err {r['err']:.2e}")
    spmv_kind = r["spmv_kind"]
    xm = np.stack(s.get_x())

    ms_mixed = 1e3 * sum(tm) / len(tm)
    ms_fp64 = 1e3 * sum(t64) / len(t64)
    # dominant kernel of the mixed solve: the low-precision SpMV, timed live
    sp_ms = sum(a for a, _ in spm)
    sp_n = sum(b for _, b in spm)
    # storage of the inner SpMV: 1 FP32 quad, 2 BF16 quad, 3 FP32 half (symmetric) storage
    coef_bytes = {1: 108 * 4, 2: 108 * 2, 3: 56 * 4}[spmv_kind]
    bytes_launch = nodes * (coef_bytes + 4 * 4)
    spmv_avg_ms = sp_ms / max(sp_n, 1)
    achieved = bytes_launch / (spmv_avg_ms * 1e-3) / 1e9
    # HBM bytes per launch of the same kernel from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    # passes of this workload, scaled by the counters' reading of a copy kernel of known size
    traffic = None
    kname = {1: "cg::k_spmv_q<float, 1>", 2: "cg::k_spmv_q<__hip_bfloat16, 1>", 3: "cg::k_spmv_symx<1>"}[spmv_kind]
    pmc_path = os.path.join(ROOT, "profiles", "spmv_pmc_icestream1024.json")
    if args.n == 1024 and args.nz == 11 and os.path.exists(pmc_path):
        with open(pmc_path) as f:
            traffic = json.load(f)["kernels"].get(kname, {}).get("traffic_bytes")
    sp64_ms = sum(a for a, _ in sp64) / max(sum(b for _, b in sp64), 1)
    coef64 = 56 * 8 if r64_kind == 4 else 108 * 8      # FP64 half (symmetric) or full rows
    ach64 = nodes * (coef64 + 4 * 8) / (sp64_ms * 1e-3) / 1e9

    # ---- BF16-storage variant of the mixed solve (same system, same tolerance)
    bf16 = None
    xb = None
    if not args.no_bf16 and lowp != pcg.MIXED_BF16:
        r = mixed_step(pcg.MIXED_BF16)
        bsetup = r["setup_ms"]
        tb, itb, spb = [], [], []
        for k in range(max(3, args.steps // 4)):
            tw = time.perf_counter()
            r = mixed_step(pcg.MIXED_BF16, time_spmv=True)
            tb.append(time.perf_counter() - tw)
            itb.append(r["niters"])
            spb.append((r["spmv_ms"], r["spmv_count"]))
        xb = np.stack(s.get_x())
        sbm = sum(a for a, _ in spb) / max(sum(b for _, b in spb), 1)
        byb = nodes * (108 * 2 + 4 * 4)
        bf16 = dict(ms_per_solve=round(1e3 * float(np.mean(tb)), 2), solves_per_s=round(1.0 / float(np.mean(tb)), 4),
                    inner_iterations=itb, outer_iterations=r["outer_iters"], err=r["err"],
                    speedup_vs_fp64=round(ms_fp64 / (1e3 * float(np.mean(tb))), 3), setup_ms=round(bsetup, 3),
                    spmv_kernel="cg::k_spmv_q<__hip_bfloat16, 1> (quad rows)", spmv_avg_ms=round(sbm, 4),
                    spmv_algorithmic_bytes=byb, spmv_gbps=round(byb / (sbm * 1e-3) / 1e9, 1),
                    spmv_frac=round(byb / (sbm * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4))
        log(f"bf16 storage: {bf16}")

    # isolated SpMV rates (back-to-back launches, HIP events)
    spmv_iso = {}
    for kind, name in ((0, "fp64"), (1, "fp32"), (2, "bf16")):
        ms, by = s.bench_spmv(kind, 20)
        spmv_iso[name] = dict(avg_ms=round(ms, 4), algorithmic_bytes=int(by), gbps=round(by / (ms * 1e-3) / 1e9, 1),
                              frac_of_peak=round(by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 3))
    log(f"isolated SpMV: {spmv_iso}")
    # internal stress of the mixed solution (compute_internal_stress): one kernel over
    # the elements and the six (nz-1, nx, ny) fields copied to the host
    stress = None
    try:
        s.set_x(np.asfortranarray(xm[0]), np.asfortranarray(xm[1]))
        bp.internal_stress()
        tw = time.perf_counter()
        tau = bp.internal_stress()
        stress = dict(ms_with_host_copy=round(1e3 * (time.perf_counter() - tw), 2),
                      max_tau_eff_kpa=round(float(tau["tau_eff"].max()) / 1e3, 2),
                      note="kernel alone: profiles/kernel_stats_icestream1024.md (k_internal_stress)")
        del tau
        log(f"internal stress: {stress}")
    except Exception as e:
        log(f"internal stress leg failed: {e!r}")
    # basal friction heat flux and traction of the same solution
    # (compute_basal_friction_heatflx, both options; one kernel and the cell halo)
    bfric = None
    try:
        bfric = {}
        for which, name in ((1, "local"), (0, "standard")):
            bp.basal_friction(which)
            tw = time.perf_counter()
            bf, _, _ = bp.basal_friction(which)
            bfric[name] = dict(ms_with_host_copy=round(1e3 * (time.perf_counter() - tw), 2),
                               max_w_per_m2=round(float(bf.max()), 4))
        bfric["note"] = "kernel alone: profiles/kernel_stats_icestream1024.md (k_basal_friction)"
        log(f"basal friction: {bfric}")
    except Exception as e:
        log(f"basal friction leg failed: {e!r}")
    # velocity-dependent basal traction on the same grid (glissade_calcbeta, Schoof law)
    cbeta, cb_sample = None, None
    try:
        cbeta, cb_sample = run_calcbeta(args, bp, s, xm)
        log(f"calcbeta: {cbeta}")
    except Exception as e:
        log(f"calcbeta leg failed: {e!r}")
    # measured HBM ceilings on this box (STREAM-style copy / triad / read, 2 GiB arrays)
    from cism_amd import _lib
    hbm = {k: round(v, 1) for k, v in _lib.bench_stream(2 << 30, 10).items()}
    log(f"measured HBM GB/s: {hbm}")

    # ---- the oracle on the exported system: errors against it and the CPU baseline
    cpu = None
    dropin = None
    twod_sys = None
    oracle_cmp = None
    if not args.no_twod:
        try:
            (Auu, Auv, Avu, Avv), (bu, bv) = s.get_matrix()
            twod_sys = synthetic.depth_integrated_system(Auu, Auv, Avu, Avv, bu, bv)
            del Auu, Auv, Avu, Avv, bu, bv
        except Exception as e:
            log(f"2-D system export failed: {e!r}")
    if not args.no_cpu:
        try:
            cpu, oracle_cmp, dropin = run_cpu_legs(args, g, bp.active_vertex, s, x64, xm, xb, it64)
        except Exception as e:   # the CPU leg must not hide the GPU result
            log(f"CPU baseline failed: {e!r}")

    if cpu is not None and cb_sample is not None:
        try:
            cpu["calcbeta"] = calcbeta_oracle_check(cb_sample)
        except Exception as e:
            log(f"calcbeta oracle check failed: {e!r}")
    del cb_sample
    twod = run_twod(args, g, bp.active_vertex, twod_sys) if twod_sys is not None else None
    del s, bp, twod_sys
    cfg2 = run_config2(args) if args.config2_n > 0 else None
    ens_sample = [] if not args.no_cpu else None
    ens = run_ensemble(args, ens_sample) if args.ensemble > 0 else None
    if ens_sample and cpu is not None:
        try:
            cpu["ensemble_members"] = ensemble_oracle_check(ens_sample)
        except Exception as e:
            log(f"ensemble oracle check failed: {e!r}")
    del ens_sample
    pic = run_marine_picard(args) if args.picard_n > 0 else None

    errs = oracle_cmp or {}
    out = {
        "metric": "velocity solves/sec mixed vs FP64 at 1e-6 rel err; SpMV HBM GB/s vs peak",
        "value": round(1e3 / ms_mixed, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mixed, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if lowp == pcg.MIXED_FP32 else "bf16+f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"synthetic ice stream {args.n}x{args.n}x{args.nz} (mixed-precision C-G solve, FP64 refinement)",
            "n": args.n, "nz": args.nz, "owned_nodes": nodes, "unknowns": 2 * nodes, "seed": args.seed,
            "precond": "SIA column tridiagonal (which_ho_precond=2)", "tolerance": args.tol,
            "inner_tol": args.inner_tol, "matrix_storage": args.precision, "parallelism": "single GPU",
        },
        "fp64": {"solves_per_s": round(1e3 / ms_fp64, 4), "ms_per_solve": round(ms_fp64, 2),
                 "iterations": it64, "rel_err_vs_oracle": errs.get("fp64"),
                 "spmv_avg_ms": round(sp64_ms, 4), "spmv_gbps": round(ach64, 1),
                 "spmv_frac": round(ach64 / HBM_PEAK_GBPS, 4),
                 "spmv_storage": "half (symmetric)" if r64_kind == 4 else "full rows"},
        "mixed": {"inner_iterations": itm, "outer_iterations": outm, "rel_err_vs_oracle": errs.get("mixed"),
                  "meets_1e-6": None if errs.get("mixed") is None else bool(errs["mixed"] <= 1e-6)},
        "speedup_mixed_vs_fp64": round(ms_fp64 / ms_mixed, 3),
        "per_assembly_setup": setup,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "traffic_source": "profiles/spmv_pmc_icestream1024.json" if traffic else None,
                     "kernel": f"{kname} (inner SpMV + (Az,z) partials; "
                               f"{'half (symmetric) storage' if spmv_kind == 3 else 'quad layout'})",
                     "algorithmic_bytes_per_launch": bytes_launch, "avg_launch_ms": round(spmv_avg_ms, 4),
                     "launches_timed": sp_n},
        "bf16_storage": bf16,
        "oracle_comparison": oracle_cmp,
        "spmv_isolated": spmv_iso,
        "dropin_host_call": dropin,
        "hbm_measured_gbps": hbm,
        "assembly": {"element_kernel_ms": round(ainfo["assemble_ms"], 2), "total_ms": round(ainfo["total_ms"], 2)},
        "calcbeta": cbeta,
        "internal_stress": stress,
        "basal_friction": bfric,
        "cpu_baseline": cpu,
        "fp64_config2": cfg2,
        "ensemble": ens,
        "picard_marine": pic,
        "twod": twod,
        "wall_s": round(time.time() - t0, 1),
    }
    line = json.dumps(out)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")


# algorithmic bytes per vertex of k_calcbeta for the Schoof law with a
# c_space_factor field: u, v (16) + f_ground, N, coulomb_c, powerlaw_c,
# c_space_factor (40) read, beta (8) written
CALCBETA_BYTES_PER_VERTEX = 64


def run_calcbeta(args, bp, s, xm, ncall=20):
    """glissade_calcbeta (which_ho_babc 11, Schoof, with c_space_factor) from
    the resident basal velocity of the headline grid: wall time per call
    (kernel, invalid-value check and stream synchronisation) and the sample
    the CPU leg checks against the oracle."""
    import numpy as np

    from cism_amd import basal, geometry as ge
    p = bp.p
    g = p.grid
    rng = np.random.default_rng(11)
    shp = (g.nvx, g.nvy)
    stag = 0.25 * (p.thck[:-1, 1:] + p.thck[1:, 1:] + p.thck[:-1, :-1] + p.thck[1:, :-1])
    fr = basal.Friction(11, powerlaw_m=3.0,
                        effecpress_stag=np.asfortranarray(910.0 * 9.81 * stag * rng.uniform(0.05, 0.3, shp)),
                        coulomb_c=np.asfortranarray(rng.uniform(0.3, 0.6, shp)),
                        powerlaw_c=np.asfortranarray(rng.uniform(1.0e4, 3.0e4, shp)),
                        c_space_factor_stag=np.asfortranarray(rng.uniform(0.8, 1.0, shp)))
    fg = ge.grounded_fraction(bp.ice_mask, bp.floating_mask, bp.land_mask)
    basal.device_setup(s, fr, f_ground=fg, topg=p.topg, thck=p.thck, eus=p.eus,
                       flwa_basal=np.asarray(p.flwa)[g.nz - 2])
    s.set_x(np.asfortranarray(xm[0]), np.asfortranarray(xm[1]))
    for _ in range(3):
        basal.device_calcbeta(s)
    tw = time.perf_counter()
    for _ in range(ncall):
        basal.device_calcbeta(s)
    ms = 1e3 * (time.perf_counter() - tw) / ncall
    beta = basal.device_calcbeta(s, want=True)
    nv = g.nvx * g.nvy
    out = dict(law="which_ho_babc 11 (Schoof) with c_space_factor", vertices=nv, ms_per_call=round(ms, 4),
               algorithmic_bytes=nv * CALCBETA_BYTES_PER_VERTEX,
               note="wall time per cismgpu_system_calcbeta call: k_calcbeta, the invalid-value readback and the "
                    "stream synchronisation; the kernel alone is in profiles/kernel_stats_calcbeta1024.md")
    sample = dict(problem=p, friction=fr, f_ground=fg, ubas=np.asfortranarray(xm[0][g.nz - 1]),
                  vbas=np.asfortranarray(xm[1][g.nz - 1]), beta=beta)
    return out, sample


def calcbeta_oracle_check(sample):
    """The oracle's calcbeta of the same basal velocity (1 thread): error of
    the device field against it and the CPU time."""
    import types

    import numpy as np

    from oracle import pyoracle as po
    p = sample["problem"]
    q = types.SimpleNamespace(grid=p.grid, friction=sample["friction"], eus=p.eus, topg=p.topg, thck=p.thck,
                              flwa=p.flwa)
    tw = time.perf_counter()
    ref = po.calcbeta(q, {"f_ground": sample["f_ground"]}, sample["ubas"], sample["vbas"])
    cpu_s = time.perf_counter() - tw
    b = sample["beta"]
    scale = np.where(ref != 0, np.abs(ref), 1.0)
    return dict(cpu_s=round(cpu_s, 3), cores=1, max_rel_err_vs_oracle=float(np.max(np.abs(b - ref) / scale)),
                kind="port", sample=f"the whole {p.grid.nvx}x{p.grid.nvy} vertex field, one call of orc_calcbeta "
                                    f"on 1 thread")


def run_cpu_legs(args, g, act, s, x64, xm, xb, it64):
    """The oracle (C restatement of pcg_solver_chrongear_3d) on the exported
    system.  All host cores: one full solve to the bench tolerance, timed,
    which is the CPU baseline and the FP64 reference every GPU result is
    compared with.  One core (the reference runs one rank single-threaded):
    the median of a few fixed-length samples, scaled to the all-cores
    solve's iteration count.  Also the host-array drop-in call on the same
    system (upload + FP64 solve + download)."""
    import numpy as np

    from cism_amd import pcg
    from oracle import pyoracle as po
    log("exporting the assembled system for the oracle")
    (Auu, Auv, Avu, Avv), (bu, bv) = s.get_matrix()
    A = (Auu, Auv, Avu, Avv)
    ncores = min(os.cpu_count() or 1, 16)
    full = None
    cmp = None
    if not args.no_cpu_all:
        ou, ov = g.zeros_vec(), g.zeros_vec()
        tc = time.perf_counter()
        st, oerr, onit = po.pcg_solve(g, act, *A, bu, bv, ou, ov, precond=2, tolerance=args.tol, maxiters=50000,
                                      nthreads=ncores)
        tsolve = time.perf_counter() - tc
        xo = np.stack([ou, ov])
        del ou, ov
        nrm = float(np.linalg.norm(xo))
        rel = lambda x: None if x is None else float(np.linalg.norm(x - xo) / nrm)
        cmp = dict(oracle_iterations=int(onit), oracle_err=oracle_float(oerr), oracle_status=int(st),
                   gpu_fp64_iterations=int(it64[-1]), fp64=rel(x64), mixed=rel(xm), bf16=rel(xb),
                   note=f"relative 2-norm velocity error of each GPU solution against the oracle's FP64 "
                        f"Chronopoulos-Gear solve of the same exported system to {args.tol:g}")
        full = dict(value=round(1.0 / tsolve, 6), unit="solves/s", cores=ncores, kind="port",
                    sample=(f"one full FP64 Chronopoulos-Gear solve (SIA preconditioner) of the same "
                            f"{args.n}x{args.n}x{args.nz} assembled system to {args.tol:g} by the C restatement in "
                            f"oracle/ on {ncores} OpenMP threads: {onit} iterations in {tsolve:.1f} s"),
                    seconds_per_solve=round(tsolve, 2), iterations=int(onit))
        log(f"oracle all-cores solve: {full['sample']}; errors {cmp}")
    # one core: median of fixed-length samples
    per = []
    for rep in range(args.cpu_repeats):
        xu, xv = g.zeros_vec(), g.zeros_vec()
        tc = time.perf_counter()
        po.pcg_solve(g, act, *A, bu, bv, xu, xv, precond=2, tolerance=0.0, maxiters=args.cpu_iters)
        per.append((time.perf_counter() - tc) / args.cpu_iters)
    per_iter = float(np.median(per))
    n_it = full["iterations"] if full else int(it64[-1])
    one = dict(value=round(1.0 / (per_iter * n_it), 6), unit="solves/s", cores=1, kind="port",
               sample=(f"median of {args.cpu_repeats} samples of {args.cpu_iters} Chronopoulos-Gear iterations of "
                       f"the same system by the C restatement on 1 thread ({', '.join(f'{p * 1e3:.0f}' for p in per)} "
                       f"ms/iteration) x {n_it} iterations"),
               seconds_per_solve=round(per_iter * n_it, 2))
    log(f"CPU one core: {one['sample']}")
    if full:
        cpu = full
        cpu["one_core"] = one
    else:
        cpu = one
    cpu["host"] = host_cpu()
    cpu["dome_picard"] = dome_picard_legs()
    # the host-array drop-in, as the Fortran shim calls it: the first call
    # makes the library's device system for this grid, a second call (the
    # next Picard iteration) reuses it
    calls = []
    for _ in range(2):
        xu, xv = g.zeros_vec(), g.zeros_vec()
        td = time.perf_counter()
        derr, dnit = pcg.pcg_solver_chrongear_3d(g, act, *A, bu, bv, xu, xv, 2, 5, args.tol, 50000)
        calls.append(round(time.perf_counter() - td, 3))
    dropin = dict(seconds=calls[0], seconds_next_call=calls[1], iterations=int(dnit), err=derr,
                  matrix_bytes_uploaded=int(4 * Auu.nbytes))
    from cism_amd import _lib
    _lib.load().cismgpu_dropin_release()       # give the drop-in's device memory back
    log(f"host-array drop-in call: {dropin}")
    return cpu, cmp, dropin


def ensemble_oracle_check(sample):
    """Part of the CPU-baseline leg: the oracle's FP64 Chronopoulos-Gear
    solve (all host cores) of a few ensemble members' exported systems, each
    member's batched mixed-precision solution compared with it, and the
    oracle's time per member beside the batched launch."""
    import numpy as np

    from oracle import pyoracle as po
    ncores = min(os.cpu_count() or 1, 16)
    rows, worst = [], 0.0
    for m, g, act, A, b, x in sample:
        ou, ov = g.zeros_vec(), g.zeros_vec()
        tc = time.perf_counter()
        st, oerr, onit = po.pcg_solve(g, act, *A, *b, ou, ov, precond=2, tolerance=1e-12, maxiters=50000,
                                      nthreads=ncores)
        t = time.perf_counter() - tc
        own = g.owned_slice()
        xo = np.stack([ou[own], ov[own]])
        err = float(np.linalg.norm(np.stack([x[0][own], x[1][own]]) - xo) / np.linalg.norm(xo))
        worst = max(worst, err)
        rows.append(dict(member=m, nx=g.nx, ny=g.ny, nz=g.nz, owned_nodes=int(g.n_owned_nodes),
                         oracle_iterations=int(onit), oracle_status=int(st), oracle_s=round(t, 3),
                         rel_err_vs_oracle=err))
    out = dict(members=rows, max_rel_err_vs_oracle=worst, meets_1e6=bool(worst <= 1e-6), cores=ncores,
               note="ensemble members' batched FP64-refined FP32 solutions (tolerance 1e-10) against the oracle's "
                    "FP64 solve of the same member system to 1e-12")
    log(f"ensemble members against the oracle: {out}")
    return out


def oracle_float(v):
    return None if v != v else float(v)


def run_config2(args):
    """Plain FP64 Chronopoulos-Gear solve of one pre-assembled ice-stream
    system at 512 x 512 x 11 (2.9 M nodes) with the column-tridiagonal
    preconditioner to 1e-8, as the drop-in runs it; the FP64 SpMV timed live
    inside the solves against the HBM roofline."""
    import numpy as np

    from cism_amd import synthetic
    from cism_amd.velo import BPSolver
    try:
        n = args.config2_n
        prob = synthetic.ice_stream(n, nz=args.nz, seed=5)
        bp = BPSolver(prob)
        bp.assemble(prob.uvel, prob.vvel)
        s = bp.sys
        nodes = prob.grid.n_owned_nodes
        s.zero_x()
        r = s.solve(precond=2, tolerance=1e-8, maxiters=50000)
        setup_ms = r["setup_ms"]
        ts, its, sp = [], [], []
        for _ in range(max(3, args.steps // 2)):
            s.zero_x()
            tw = time.perf_counter()
            r = s.solve(precond=2, tolerance=1e-8, maxiters=50000, time_spmv=True)
            ts.append(time.perf_counter() - tw)
            its.append(r["niters"])
            sp.append((r["spmv_ms"], r["spmv_count"]))
        spavg = sum(a for a, _ in sp) / max(sum(b for _, b in sp), 1)
        coef = 56 * 8 if r["spmv_kind"] == 4 else 108 * 8
        by = nodes * (coef + 4 * 8)
        out = dict(workload=f"synthetic ice stream {n}x{n}x{args.nz} seed 5, FP64 C-G, SIA preconditioner, tol 1e-8",
                   owned_nodes=nodes, ms_per_solve=round(1e3 * float(np.mean(ts)), 2),
                   solves_per_s=round(1.0 / float(np.mean(ts)), 4), iterations=its, err=r["err"],
                   setup_ms=round(setup_ms, 3),
                   spmv_kernel=("cg::k_spmv_sym64<1> (half storage)" if r["spmv_kind"] == 4 else
                                "cg::k_spmv<double, double, 1> (full rows)"),
                   spmv_avg_ms=round(spavg, 4), spmv_algorithmic_bytes=by,
                   spmv_gbps=round(by / (spavg * 1e-3) / 1e9, 1),
                   spmv_frac=round(by / (spavg * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4))
        s.close()
        log(f"config 2: {out}")
        return out
    except Exception as e:
        log(f"config 2 leg failed: {e!r}")
        return {"error": repr(e)}


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(model=model, nproc=os.cpu_count())


def dome_picard_legs():
    """The dome (31x31 cells, 10 levels, no-slip base) Picard solve with the
    standard PCG and diagonal preconditioner to 1e-8, Picard to 1e-4: the C
    restatement on one host core against the device path, same iterations."""
    from cism_amd import synthetic
    from cism_amd.velo import BPSolver
    from oracle import pyoracle as po
    p = synthetic.dome()
    kw = dict(linear_tol=1e-8, linear_maxiters=200, resid_target=1e-4)
    t = time.perf_counter()
    ref = po.picard(p, algo="standard", precond=1, **kw)
    t_cpu = time.perf_counter() - t
    bp = BPSolver(p)
    bp.picard(algo=2, precond=1, **kw)          # warm-up
    t = time.perf_counter()
    r = bp.picard(algo=2, precond=1, **kw)
    t_gpu = time.perf_counter() - t
    import numpy as np
    rel = float(np.linalg.norm(np.stack([r.uvel - ref["uvel"], r.vvel - ref["vvel"]])) /
                np.linalg.norm(np.stack([ref["uvel"], ref["vvel"]])))
    return dict(cpu_s=round(t_cpu, 3), gpu_s=round(t_gpu, 3), cpu_iterations=ref["iterations"],
                gpu_iterations=r.iterations, rel_diff=rel, cores=1)


def run_twod(args, g, act, system):
    """Depth-integrated 2-D solves (pcg_solver_chrongear_2d path) on the
    P^T A P collapse of the assembled ice-stream system: FP64 drop-in solves
    with the diagonal (which_ho_precond 1, and 2 as replaced for 2-D solves)
    and local tridiagonal preconditioners, the mixed FP32 solve, isolated
    9-point SpMV rates, and the oracle's 2-D C-G on the host as the CPU
    sample."""
    import numpy as np

    from cism_amd import pcg
    try:
        (A2, bu, bv) = system
        nv = (g.ihi - g.ilo + 1) * (g.jhi - g.jlo + 1)
        s2 = pcg.System(g, two_d=True)
        s2.set_matrix(act, *A2, bu, bv)
        s2.zero_x()
        s2.solve(precond=1, tolerance=1e-12, maxiters=200000)
        xref = np.stack(s2.get_x())

        def err():
            x = np.stack(s2.get_x())
            return float(np.linalg.norm(x - xref) / np.linalg.norm(xref))

        legs = {}
        for name, kw in (("fp64_diag", dict(precond=1)), ("fp64_tridiag_local", dict(precond=3)),
                         ("mixed_fp32_diag", dict(precond=1, precision=pcg.MIXED_FP32, inner_tol=1e-6,
                                                  max_outer=40))):
            s2.zero_x()
            s2.solve(tolerance=args.tol, maxiters=200000, **kw)       # warm-up
            ms, its = [], []
            for _ in range(max(args.steps, 3)):
                s2.zero_x()
                r = s2.solve(tolerance=args.tol, maxiters=200000, time_spmv=True, **kw)
                ms.append(r["solve_ms"])
                its.append(r["niters"])
            legs[name] = dict(ms_per_solve=round(float(np.mean(ms)), 3), solves_per_s=round(float(1e3 / np.mean(ms)), 2),
                              iterations=its, outer=r["outer_iters"], rel_err_vs_tight_fp64=err(),
                              spmv_avg_ms=round(r["spmv_ms"] / max(r["spmv_count"], 1), 4))
            log(f"2-D {name}: {legs[name]}")
        spmv = {}
        for kind, name in ((0, "fp64"), (1, "fp32"), (2, "bf16")):
            ms, by = s2.bench_spmv(kind, 50)
            spmv[name] = dict(avg_ms=round(ms, 4), algorithmic_bytes=int(by), gbps=round(by / (ms * 1e-3) / 1e9, 1),
                              frac_of_peak=round(by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 3))
        log(f"2-D isolated SpMV: {spmv}")
        s2.close()
        cpu = None
        if not args.no_cpu:
            from oracle import pyoracle as po
            k = 20
            xu, xv = np.zeros((g.nvx, g.nvy), order="F"), np.zeros((g.nvx, g.nvy), order="F")
            tc = time.perf_counter()
            po.pcg_solve_2d(g, act, *A2, bu, bv, xu, xv, algo="chrongear", precond=1, tolerance=0.0, maxiters=k)
            per = (time.perf_counter() - tc) / k
            n_it = int(np.mean(legs["fp64_diag"]["iterations"]))
            cpu = dict(value=round(1.0 / (per * n_it), 4), unit="solves/s", cores=1, kind="port",
                       sample=f"{k} Chronopoulos-Gear iterations (diagonal preconditioner) of the same 2-D system by "
                              f"the C restatement in oracle/ on 1 thread; {per * 1e3:.1f} ms/iteration x {n_it} "
                              f"iterations", seconds_per_solve=round(per * n_it, 3))
            log(f"2-D CPU baseline: {cpu}")
        return dict(workload=f"depth-integrated (P^T A P) ice stream {args.n}x{args.n}, 9-point 2x2-block stencil",
                    owned_vertices=nv, unknowns=2 * nv, tolerance=args.tol, legs=legs, spmv_isolated=spmv,
                    cpu_baseline=cpu)
    except Exception as e:      # a failing side leg must not hide the headline result
        log(f"2-D leg failed: {e!r}")
        return {"error": repr(e)}


def ensemble_bytes(probs, E, storage):
    """Algorithmic HBM bytes of one ensemble solve, summed over members from
    their own iteration counts: per inner iteration the FP32 product (half
    storage 224 B of matrix, full quad rows 432, BF16 quad rows 216, plus z
    read and Az written, 16 B) and the column update (z, d, Az, q, e, r and
    three SIA coefficients read, d, q, e, r, z written: 112 B); per
    refinement step the FP64 residual on full FP64 rows (864 B of matrix, x
    and b read, r written: 912 B), the start of the inner solve (r read, the
    SIA coefficients read, e, d, q, r, z written: 80 B) and the fold x += e
    (40 B); the final residual of each member adds one more 912 B pass."""
    coef = {1: 432, 2: 216, 3: 224}[storage]
    inner = coef + 16 + 112
    outer = 912 + 80 + 40
    total = 0
    for m, p in enumerate(probs):
        mi = E.member_info(m)
        n = p.grid.n_owned_nodes
        total += n * (inner * mi["total_inner"] + outer * mi["outer"] + 912)
    return total, inner, outer


def run_ensemble(args, sample=None):
    """Batched solve of a seeded ensemble of uneven ice streams (side 32..192
    owned vertices, nz in {5,7,11,15}) in one persistent launch, against the
    same members solved one after another by the single-system mixed solver;
    every member's error is measured against its own FP64 solve.  Bandwidth:
    the members' algorithmic bytes (ensemble_bytes) over the launch time,
    against the HBM roofline."""
    import numpy as np

    from cism_amd import pcg, synthetic
    from cism_amd.ensemble import Ensemble
    try:
        log(f"ensemble: generating and assembling {args.ensemble} members")
        probs = synthetic.ensemble(n_members=args.ensemble, seed=2024)
        E = Ensemble(probs, keep_members=True)
        nodes = sum(p.grid.n_owned_nodes for p in probs)
        split, tune_ms = E.autotune()
        log(f"ensemble autotune: split per bucket {split}")
        # inner tolerance 1e-4: the fewest lock-step steps of the sweep 1e-3 .. 1e-6
        opts = dict(tolerance=1e-10, inner_tol=1e-4, inner_maxiters=600, max_outer=30)
        E.solve(**opts)                                   # warm-up (graph capture)
        runs = [E.solve(**opts) for _ in range(3)]
        info = runs[-1]
        solve_ms = float(np.median([r["solve_ms"] for r in runs]))
        log(f"ensemble solve: {solve_ms:.1f} ms, {info['steps']} steps, converged {info['converged']}")
        by, inner_b, outer_b = ensemble_bytes(probs, E, info["matrix_storage"])
        gbps = by / (solve_ms * 1e-3) / 1e9
        # HBM bytes per launch from the committed counter runs of this leg
        traffic = None
        try:
            with open(os.path.join(ROOT, "profiles", "ensemble_pmc_256.json")) as f:
                traffic = json.load(f)["kernels"]["cg::k_ensemble<16>"]["traffic_bytes"]
        except (OSError, KeyError, ValueError):
            pass
        out = dict(members=args.ensemble, owned_nodes=int(nodes), solve_ms=round(solve_ms, 2),
                   solve_ms_runs=[round(r["solve_ms"], 2) for r in runs],
                   member_solves_per_s=round(args.ensemble / (solve_ms * 1e-3), 1),
                   steps=info["steps"], converged=info["converged"], max_outer=info["max_outer"],
                   max_inner=info["max_inner"], total_inner=int(info["total_inner"]), grid_blocks=info["grid_blocks"],
                   matrix_storage={1: "fp32 quad rows", 2: "bf16 quad rows", 3: "fp32 half (symmetric)"}.get(
                       info["matrix_storage"], info["matrix_storage"]),
                   split_per_bucket=split, autotune_ms=np.round(tune_ms, 3).tolist(),
                   roofline={"bound": "hbm", "kernel": "cg::k_ensemble<16>",
                             "algorithmic_bytes_per_solve": int(by), "achieved": round(gbps, 1),
                             "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBPS, 4),
                             "bytes_per_node_inner_iteration": inner_b, "bytes_per_node_refinement_step": outer_b,
                             "traffic": traffic,
                             "traffic_source": "profiles/ensemble_pmc_256.json" if traffic else None})
        if not args.only_ensemble:
            seq_ms, max_err = 0.0, 0.0
            # members checked against the oracle by the CPU leg: the largest
            # of each quarter of the ensemble
            sizes = [p.grid.n_owned_nodes for p in probs]
            q = max(1, len(probs) // 4)
            picks = {max(range(a, min(a + q, len(probs))), key=lambda k: sizes[k]) for a in range(0, len(probs), q)}
            for m, sv in enumerate(E.solvers):
                sv.sys.zero_x()
                r = sv.sys.solve(precond=2, tolerance=1e-10, maxiters=20000, precision=pcg.MIXED_FP32,
                                 inner_tol=1e-4, max_outer=30)
                seq_ms += r["solve_ms"]
                sv.sys.zero_x()
                sv.sys.solve(precond=2, tolerance=1e-12, maxiters=20000)
                ru, rv = sv.sys.get_x()
                xu, xv = E.get_x(m)
                g = probs[m].grid
                own = g.owned_slice()
                den = np.sqrt((ru[own] ** 2).sum() + (rv[own] ** 2).sum())
                num = np.sqrt(((xu - ru)[own] ** 2).sum() + ((xv - rv)[own] ** 2).sum())
                max_err = max(max_err, float(num / den) if den > 0 else 0.0)
                if sample is not None and m in picks:
                    (Auu, Auv, Avu, Avv), (bu, bv) = sv.sys.get_matrix()
                    sample.append((m, g, sv.active_vertex, (Auu, Auv, Avu, Avv), (bu, bv), (xu, xv)))
            out.update(sequential_mixed_ms=round(seq_ms, 2), speedup_vs_sequential=round(seq_ms / solve_ms, 3),
                       max_rel_err_vs_member_fp64=max_err, meets_1e6=bool(max_err <= 1e-6))
        E.close()
        log(f"ensemble: {out}")
        return out
    except Exception as e:      # a failing side leg must not hide the headline result
        log(f"ensemble leg failed: {e!r}")
        return {"error": repr(e)}


def run_marine_picard(args):
    """Full Picard solve on the GPU of a seeded marine ice sheet (grounded
    interior on a retrograde bed, floating shelf, calving front with the
    lateral ocean-pressure BC, f_ground-weighted friction): accelerated
    Picard to a 1e-4 maximum relative speed change or 20 iterations, with
    the local tridiagonal preconditioner (which_ho_precond = 3), linear
    tolerance 1e-8 and linear_maxiters bounding each linear solve (mixed
    solves: the inner iterations of all refinement steps together).  The
    mixed-precision run and the same Picard run with FP64 linear solves,
    each with its per-iteration linear status, error and iteration count."""
    from cism_amd import pcg, synthetic
    from cism_amd.velo import BPSolver

    def one(prob, precision):
        bp = BPSolver(prob)
        t = time.perf_counter()
        r = bp.picard(maxiter=20, resid_target=1e-4, whichresid=0, precond=3, linear_tol=1e-8,
                      linear_maxiters=args.picard_maxiters, precision=precision, inner_tol=1e-3, max_outer=20,
                      accelerate=True)
        wall = time.perf_counter() - t
        bp.sys.close()
        return dict(iterations=r.iterations, converged=bool(r.converged),
                    resid_velo=[float(f"{x:.3e}") for x in r.resid_velo], alpha=[round(a, 2) for a in r.alpha],
                    linear_iters=r.linear_iters, linear_err=[float(f"{x:.3e}") for x in r.linear_err],
                    linear_status=r.linear_status, linear_outer=r.linear_outer,
                    assemble_ms=round(sum(r.assemble_ms), 1), solve_ms=round(sum(r.solve_ms), 1),
                    wall_s=round(wall, 2)), r

    try:
        import numpy as np
        n = args.picard_n
        log(f"marine Picard: generating {n}x{n}x{args.nz}")
        prob = synthetic.marine_sheet(n, nz=args.nz, dx=1000.0, seed=11)
        mixed, rm = one(prob, pcg.MIXED_FP32)
        log(f"marine Picard (mixed inner solves): {mixed}")
        out = dict(n=n, nz=args.nz, precond="which_ho_precond=3 (horizontal line solves + SIA)", linear_tol=1e-8,
                   linear_maxiters=args.picard_maxiters, mixed=mixed)
        if not args.no_picard_fp64:
            fp64, r64 = one(prob, pcg.FP64)
            log(f"marine Picard (FP64 inner solves): {fp64}")
            out["fp64"] = fp64
            x64 = np.stack([r64.uvel, r64.vvel])
            out["rel_diff_mixed_vs_fp64"] = float(np.linalg.norm(np.stack([rm.uvel, rm.vvel]) - x64) /
                                                  np.linalg.norm(x64))
            out["speedup_mixed_vs_fp64_solve_time"] = round(fp64["solve_ms"] / max(mixed["solve_ms"], 1e-9), 3)
        return out
    except Exception as e:
        log(f"marine Picard leg failed: {e!r}")
        return {"error": repr(e)}


if __name__ == "__main__":
    main()
