"""TreePM force-evaluation benchmark on one MI355X.

One step = one full periodic TreePM force evaluation (refresh on, all N active):
unwrap, path keys + sort, octree build, multipoles, CIC deposit, r2c FFT,
Green's function, three gradient c2r FFTs, fused PM interpolation and the
short-range residual walk. Inputs are resident in HBM before the timed region.

Workload: 256^3 particles in a 200 Mpc/h box, 20% in 1000 NFW halos, 512^3
mesh, r_s = 1.25 cells, r_cut = 4.5 r_s, theta = 0.5, leaf 16, eps = spacing/30.
value = mixed-precision particles/s; the FP64 GPU form is timed on the same
data and reported beside it, with the force error of both against each other
and (on sample targets) against an FP64 Ewald sum.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for p in (str(ROOT / "chui-dev_amd"), str(ROOT)):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK = 8.0e12          # B/s, MI355X
FP32_VALU_PEAK = 157.3e12  # FLOP/s, MI355X vector FP32
FLOP_PER_PAIR = 40         # accounting constant per residual pair
FP64_VALU_PEAK = 78.6e12   # FLOP/s, MI355X vector FP64
GROUP_FLOP_PER_PAIR = 60   # FP64 Hermite-table pair of the group walk (accounting constant)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def error_metrics(value, ref, floor_fraction=1e-3):
    err = np.sqrt(np.sum((value - ref) ** 2, axis=0))
    rn = np.sqrt(np.sum(ref**2, axis=0))
    rms_ref = np.sqrt(np.mean(rn**2))
    norm = np.sort(err / np.maximum(rn, floor_fraction * rms_ref))
    p99 = float(norm[min(int(np.ceil(0.99 * norm.size)), norm.size) - 1])
    return {"relative_l2": float(np.sqrt(np.sum((value - ref) ** 2) / np.sum(ref**2))), "p99": p99}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-side", type=int, default=256)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--ewald-samples", type=int, default=8192)
    ap.add_argument("--cpu-n-side", type=int, default=64,
                    help="particles per side of the CPU baseline's sample of the 256^3 workload law")
    ap.add_argument("--cpu-repeats", type=int, default=5,
                    help="timed CPU solves (median reported) after one warm-up")
    ap.add_argument("--accuracy-steps", type=int, default=2,
                    help="timed mixed solves per accuracy profile (TSC + deconvolution settings measured against "
                         "the Ewald sum); 0 skips")
    ap.add_argument("--other-configs", type=int, default=1,
                    help="1: also time the 128^3 workload (mixed and FP64) and state the Ewald force error of "
                         "the 32^3, 128^3, 64^3 (adjoint) and ensemble workloads")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--grad-n-side", type=int, default=64,
                    help="particles per side of the differentiable KDK workload (0 skips the grad/fwd figure)")
    ap.add_argument("--grad-directions", type=int, default=16,
                    help="random unit directions of the central-difference check of the gradient")
    ap.add_argument("--adaptive-steps", type=int, default=2,
                    help="timed FP64 solves of the adaptive policy with the group walk (the reference's fast deck "
                         "settings), history from the FP64 strict forces; 0 skips")
    ap.add_argument("--ensemble-steps", type=int, default=5,
                    help="KDK steps of the 48-box ensemble (16 x 32^3, 16 x 64^3, 12 x 96^3, 4 x 128^3); 0 skips")
    ap.add_argument("--ensemble-gradients", type=int, default=1,
                    help="1: also time the FP64 gradients of every ensemble box")
    ap.add_argument("--hierarchical-rungs", type=int, default=3,
                    help="one hierarchical KDK block with this many rungs on the 128^3 workload (mixed); 0 skips")
    ap.add_argument("--profile", default="",
                    help="run the timed solves with one of the accuracy profiles (e.g. tsc_deconv_rs1.25_rcut5rs, "
                         "TSC + deconvolution, r_cut = 6.25 cells) instead of the headline's CIC settings; for "
                         "kernel traces of the setting that meets the metric's 1e-3 force error")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="one plain device-to-device copy of this size before the timed region; its HBM counters "
                         "scale the walk's FETCH_SIZE/WRITE_SIZE readings under rocprofv3 --pmc")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import cosmosim._lib as L
    from cosmosim import ForceAccumulator, TreePmCoordinator, ics

    t0 = time.time()
    w = ics.config(2, n_side=args.n_side)
    n = w.x.size
    M = w.mesh**3
    log(f"workload {w.name}: N={n} mesh={w.mesh}^3 box={w.box} eps={w.softening:.4g} ({time.time()-t0:.1f}s)")
    dev = [torch.from_numpy(a).to("cuda:0") for a in (w.x, w.y, w.z, w.mass)]
    opts = L.default_options(box_size=w.box, gravitational_constant=ics.G_CODE, opening_theta=0.5,
                             rcut_cells=4.5 * 1.25, asmth_cells=1.25, softening_epsilon=w.softening,
                             max_leaf_size=16)
    asmth_used, rcut_rs_used = 1.25, 4.5
    if args.profile:
        prof = {p["name"]: p for p in ACCURACY_PROFILES}.get(args.profile)
        if prof is None:
            print(f"unknown --profile {args.profile}; one of {[p['name'] for p in ACCURACY_PROFILES]}",
                  file=sys.stderr)
            sys.exit(2)
        opts = L.default_options(box_size=w.box, gravitational_constant=ics.G_CODE, opening_theta=0.5,
                                 asmth_cells=prof["asmth"], rcut_cells=prof["rcut_over_rs"] * prof["asmth"],
                                 softening_epsilon=w.softening, max_leaf_size=16,
                                 assignment_scheme=prof["assignment"], window_deconvolution=prof["deconv"])
        asmth_used, rcut_rs_used = prof["asmth"], prof["rcut_over_rs"]
        log(f"timed solves use accuracy profile {args.profile}")
    coord = TreePmCoordinator(w.mesh)

    def run(precision, steps, warmup):
        opts.precision = precision
        dt = torch.float32 if precision == L.PRECISION_MIXED else torch.float64
        acc = tuple(torch.zeros(n, dtype=dt, device="cuda:0") for _ in range(3))
        fa = ForceAccumulator(acc)
        for i in range(warmup):
            coord.solve_active_set(*dev, fa, opts)
            torch.cuda.synchronize()
            log(f"warmup {i + 1}/{warmup} precision={precision}")
        diag = coord.solve_active_set(*dev, fa, opts, want_diagnostics=True).diagnostics
        coord.profiling(True)
        torch.cuda.synchronize()
        start = time.perf_counter()
        for _ in range(steps):
            coord.solve_active_set(*dev, fa, opts)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - start
        prof = coord.profile_read()
        coord.profiling(False)
        log(f"precision={precision}: {steps} steps in {elapsed:.3f}s, stages {prof}")
        return elapsed, prof, diag, np.stack([a.double().cpu().numpy() for a in acc])

    if args.copy_calibration_mib > 0:
        nbytes = args.copy_calibration_mib << 20
        src_buf = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0").fill_(1)
        dst_buf = torch.empty_like(src_buf)
        torch.cuda.synchronize()
        dst_buf.copy_(src_buf)
        torch.cuda.synchronize()
        log(f"calibration copy: {nbytes} bytes read and written once")
        del src_buf, dst_buf

    t_mixed, prof_mixed, diag, acc_mixed = run(L.PRECISION_MIXED, args.steps, args.warmup)
    value = n * args.steps / t_mixed
    fp64 = None
    err_mix_vs_fp64 = None
    acc_fp64 = None
    if args.fp64_steps > 0:
        t_fp64, prof_fp64, diag64, acc_fp64 = run(L.PRECISION_FP64, args.fp64_steps, 1)
        value_fp64 = n * args.fp64_steps / t_fp64
        err_mix_vs_fp64 = error_metrics(acc_mixed, acc_fp64)
        fp64 = {"value": value_fp64, "ms_per_step": 1e3 * t_fp64 / args.fp64_steps,
                "mixed_over_fp64": value / value_fp64,
                "stages_ms": {k: v / max(prof_fp64["solves"], 1) for k, v in prof_fp64.items() if k.endswith("_ms")}}
        log(f"mixed {value:.4g} p/s, fp64 {value_fp64:.4g} p/s, ratio {value / value_fp64:.3f}, "
            f"mixed-vs-fp64 {err_mix_vs_fp64}")

    # Adaptive policy + group walk (FP64) with the fast deck's TreePM settings
    # (configs/treepm_evolved_snapshot_fast.param.txt:61-73): relative_monopole,
    # alpha 0.0025, maximum angle 1.0, com_distance cutoff, Gaussian pair lookup;
    # |a_prev| = |a| of the FP64 strict solve (or the mixed one).
    adaptive = None
    acc_adaptive = None
    if args.adaptive_steps > 0:
        adaptive, acc_adaptive = measure_adaptive(L, coord, dev, w, acc_fp64 if acc_fp64 is not None else acc_mixed,
                                                  args.adaptive_steps,
                                                  "fp64 strict" if acc_fp64 is not None else "mixed strict")
        log(f"adaptive {adaptive}")

    # FP64 Ewald on sample targets (periodic_ewald_reference.cpp semantics).
    ewald = None
    if args.ewald_samples > 0:
        rng = np.random.default_rng(ics.BASE_SEED + 99)
        tgt = np.sort(rng.choice(n, min(args.ewald_samples, n), replace=False)).astype(np.uint32)
        dt_ = torch.from_numpy(tgt.view(np.int32)).to("cuda:0")
        out = [torch.zeros(tgt.size, dtype=torch.float64, device="cuda:0") for _ in range(3)]
        te = time.time()
        alpha = 11.8 / w.box
        coord.ewald(*dev, dt_, (w.box,) * 3, ics.G_CODE, alpha, 0, 24, w.softening, *out)
        ref = np.stack([o.cpu().numpy() for o in out])
        ewald = {"samples": int(tgt.size), "alpha_box": 11.8, "modes": 24, "seconds": round(time.time() - te, 2),
                 "fp64_vs_ewald": error_metrics(acc_fp64[:, tgt], ref) if acc_fp64 is not None else None,
                 "mixed_vs_ewald": error_metrics(acc_mixed[:, tgt], ref),
                 "adaptive_group_vs_ewald": error_metrics(acc_adaptive[:, tgt], ref) if acc_adaptive is not None
                 else None}
        log(f"ewald: {ewald}")

    # Settings that meet the metric's accuracy bar, timed beside the headline.
    accuracy = None
    if args.accuracy_steps > 0 and ewald is not None:
        accuracy = measure_accuracy_profiles(L, coord, dev, w, tgt, ref, args.accuracy_steps, value)
        log(f"accuracy profiles {accuracy}")
    opts.assignment_scheme = 0
    opts.window_deconvolution = 0
    opts.asmth_cells = 1.25
    opts.rcut_cells = 4.5 * 1.25

    other = None
    if args.other_configs:
        other = measure_other_configs(L, ics)
        log(f"other configs {other}")

    # Roofline of the dominant kernel (the short-range walk with fused PM
    # interpolation): algorithmic bytes per launch / measured launch time.
    s_p, s_a, s_m = 4, 4, 4
    nodes = diag["node_count"]
    walk_bytes = n * (3 * s_p + 3 * s_a) + n * (3 * s_p + 0) + 3 * s_m * M + nodes * 80
    walk_ms = prof_mixed["walk_ms"] / max(prof_mixed["solves"], 1)
    achieved = walk_bytes / (walk_ms * 1e-3)
    pairs = diag["direct_pair_evaluations"]
    path_bytes = 185 * n + 64 * M
    step_ms = 1e3 * t_mixed / args.steps

    grad = None
    if args.grad_n_side > 0:
        grad = measure_gradient(args.grad_n_side, args.grad_directions)
        log(f"grad/fwd {grad}")

    hierarchical = None
    if args.hierarchical_rungs > 0:
        hierarchical = measure_hierarchical(args.hierarchical_rungs)
        log(f"hierarchical {hierarchical}")

    ensemble = None
    if args.ensemble_steps > 0:
        ensemble = measure_ensemble(args.ensemble_steps, args.ensemble_gradients != 0)
        log(f"ensemble {ensemble}")

    cpu = None
    cpu_config1 = None
    if not args.no_cpu:
        cpu, cpu_config1 = cpu_baselines(L, ics, args.cpu_n_side, args.cpu_repeats)
        log(f"cpu baseline {cpu}; config 1 {cpu_config1}")

    line = {
        "metric": "TreePM accel particles/s at 256^3 (mixed vs FP64) at RMS force err <=1e-3; grad/fwd ratio",
        "value": value,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fixed32 (mixed); fp64 drift/kick",
        "data": "synthetic (seeded Zel'dovich lattice + NFW halos)",
        "config": {"workload": w.name, "particles": n, "mesh": w.mesh, "box_mpc_h": w.box,
                   "halo_fraction": w.halo_fraction, "halos": w.halo_count, "theta": 0.5, "leaf": 16,
                   "asmth_cells": asmth_used, "rcut_over_rs": rcut_rs_used, "softening": w.softening,
                   "parallelism": "single GPU", **({"profile": args.profile} if args.profile else {})},
        "fp64_gpu": fp64,
        "force_error": {"mixed_vs_fp64": err_mix_vs_fp64, "ewald": ewald},
        "accuracy_profiles": accuracy,
        "other_configs": other,
        "cpu_config1_force_kdk": cpu_config1,
        "grad_fwd_ratio": grad,
        "adaptive_group_walk_fp64": adaptive,
        "ensemble": ensemble,
        "hierarchical_kdk": hierarchical,
        "stages_ms_mixed": {k: v / max(prof_mixed["solves"], 1) for k, v in prof_mixed.items() if k.endswith("_ms")},
        "roofline": {"bound": "hbm", "kernel": "walkmix_kernel", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": measured_traffic(w.name),
                     "algorithmic_bytes_per_launch": walk_bytes, "launch_ms": walk_ms,
                     "valu": {"pairs": pairs, "flop_per_pair": FLOP_PER_PAIR,
                              "achieved_tflops": pairs * FLOP_PER_PAIR / (walk_ms * 1e-3) / 1e12,
                              "peak_tflops": FP32_VALU_PEAK / 1e12},
                     "whole_path": {"algorithmic_bytes": path_bytes,
                                    "frac": path_bytes / (step_ms * 1e-3) / HBM_PEAK}},
        "cpu_baseline": cpu,
        "counters": {k: diag[k] for k in ("node_count", "max_depth", "direct_pair_evaluations",
                                          "accepted_internal_multipoles", "accepted_leaves", "visited_nodes",
                                          "wave_node_visits", "wave_leaf_source_slots",
                                          "wave_sum_flushes")},
    }
    coord.close()
    print(json.dumps(line), flush=True)


def solve_timed(L, coord, dev, opts, precision, steps, warmup=1):
    """Median seconds of `steps` solves (each synchronised) after `warmup`,
    and the accelerations (3, N) FP64 of the last one."""
    import torch

    from cosmosim import ForceAccumulator

    opts.precision = precision
    n = int(dev[0].shape[0])
    dt = torch.float32 if precision == L.PRECISION_MIXED else torch.float64
    acc = tuple(torch.zeros(n, dtype=dt, device="cuda:0") for _ in range(3))
    fa = ForceAccumulator(acc)
    times = []
    for i in range(warmup + steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        coord.solve_active_set(*dev, fa, opts)
        torch.cuda.synchronize()
        if i >= warmup:
            times.append(time.perf_counter() - t0)
    return float(np.median(times)), np.stack([a.double().cpu().numpy() for a in acc])


def ewald_reference(coord, dev, w, samples, seed, G):
    """FP64 Ewald accelerations (3, samples) on seeded sample particles (alpha
    11.8 / L, minimum real image, +-24 modes: converged to 1e-10 against the
    oracle's splitting) and the sample rows."""
    import torch

    n = int(dev[0].shape[0])
    rng = np.random.default_rng(seed)
    tgt = np.sort(rng.choice(n, min(samples, n), replace=False)).astype(np.uint32)
    dt_ = torch.from_numpy(tgt.view(np.int32)).to("cuda:0")
    out = [torch.zeros(tgt.size, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    coord.ewald(*dev, dt_, (w.box,) * 3, G, 11.8 / w.box, 0, 24, w.softening, *out)
    coord.synchronize()
    return tgt, np.stack([o.cpu().numpy() for o in out])


# Accuracy profiles: the reference's certified PM settings (TSC + window
# deconvolution, test_tree_pm_ewald_accuracy.cpp:10-14) with wider cutoffs
# and split scales, in mixed precision, each timed and measured against the
# same Ewald samples as the headline.
ACCURACY_PROFILES = (
    {"name": "tsc_deconv_rs1.25_rcut5rs", "assignment": 1, "deconv": 1, "asmth": 1.25, "rcut_over_rs": 5.0},
    {"name": "tsc_deconv_rs1.25_rcut6rs", "assignment": 1, "deconv": 1, "asmth": 1.25, "rcut_over_rs": 6.0},
    {"name": "tsc_deconv_rs1.5_rcut6rs", "assignment": 1, "deconv": 1, "asmth": 1.5, "rcut_over_rs": 6.0},
    {"name": "tsc_deconv_rs2_rcut6rs", "assignment": 1, "deconv": 1, "asmth": 2.0, "rcut_over_rs": 6.0},
)


def measure_accuracy_profiles(L, coord, dev, w, tgt, ref, steps, headline_value):
    n = int(dev[0].shape[0])
    rows = []
    for prof in ACCURACY_PROFILES:
        opts = L.default_options(box_size=w.box, gravitational_constant=43.0091, opening_theta=0.5,
                                 asmth_cells=prof["asmth"], rcut_cells=prof["rcut_over_rs"] * prof["asmth"],
                                 softening_epsilon=w.softening, max_leaf_size=16,
                                 assignment_scheme=prof["assignment"], window_deconvolution=prof["deconv"])
        t, acc = solve_timed(L, coord, dev, opts, L.PRECISION_MIXED, steps)
        err = error_metrics(acc[:, tgt], ref)
        rows.append({**prof, "ms_per_step": 1e3 * t, "value": n / t, "cost_vs_headline": headline_value / (n / t),
                     "mixed_vs_ewald": err, "meets_1e-3": err["relative_l2"] <= 1e-3})
    passing = [r for r in rows if r["meets_1e-3"]]
    best = min(passing, key=lambda r: r["ms_per_step"]) if passing else None
    return {"profiles": rows, "fastest_meeting_1e-3": best["name"] if best else None,
            "value_at_1e-3": best["value"] if best else None}


def measure_other_configs(L, ics):
    """The 128^3 workload timed in both precisions, and the Ewald force error
    of the 32^3 (FP64), 128^3, 64^3 adjoint (FP64, quadrupoles, its initial
    state) and ensemble (mixed, one box per size) workloads."""
    import torch

    from cosmosim import TreePmCoordinator
    from cosmosim import ensemble as E

    out = {}
    w1 = ics.config(0)
    with TreePmCoordinator(w1.mesh) as c1:
        d1 = [torch.from_numpy(a).to("cuda:0") for a in (w1.x, w1.y, w1.z, w1.mass)]
        o1 = opts_fp64_for(L, w1, ics)
        t64, a64 = solve_timed(L, c1, d1, o1, L.PRECISION_FP64, 3)
        tgt, ref = ewald_reference(c1, d1, w1, 4096, ics.BASE_SEED + 90, ics.G_CODE)
        out[w1.name] = {"particles": w1.x.size, "fp64_ms": 1e3 * t64, "fp64_value": w1.x.size / t64,
                        "fp64_vs_ewald": error_metrics(a64[:, tgt], ref), "ewald_samples": int(tgt.size)}
    w2 = ics.config(1)
    with TreePmCoordinator(w2.mesh) as c2:
        d2 = [torch.from_numpy(a).to("cuda:0") for a in (w2.x, w2.y, w2.z, w2.mass)]
        o2 = opts_fp64_for(L, w2, ics)
        tmx, amx = solve_timed(L, c2, d2, o2, L.PRECISION_MIXED, 5)
        t64, a64 = solve_timed(L, c2, d2, o2, L.PRECISION_FP64, 3)
        tgt, ref = ewald_reference(c2, d2, w2, 4096, ics.BASE_SEED + 91, ics.G_CODE)
        out[w2.name] = {"particles": w2.x.size, "mixed_ms": 1e3 * tmx, "mixed_value": w2.x.size / tmx,
                        "fp64_ms": 1e3 * t64, "fp64_value": w2.x.size / t64, "mixed_over_fp64": t64 / tmx,
                        "mixed_vs_fp64": error_metrics(amx, a64),
                        "fp64_vs_ewald": error_metrics(a64[:, tgt], ref),
                        "mixed_vs_ewald": error_metrics(amx[:, tgt], ref), "ewald_samples": int(tgt.size)}
    from cosmosim import adjoint

    prob, psi, _ = adjoint.adjoint_problem(n_side=64)
    try:
        x0 = prob.wrap(prob.lattice + psi.to(torch.float64))
        dev4 = [x0[0].contiguous(), x0[1].contiguous(), x0[2].contiguous(), prob.mass]
        a4 = prob.forces(x0).cpu().numpy()

        class _W:
            box = prob.box
            softening = prob.softening

        tgt, ref = ewald_reference(prob.coord, dev4, _W, 4096, ics.BASE_SEED + 93, ics.G_CODE)
        out["adjoint_64^3_box50_mesh128"] = {"precision": "fp64, quadrupoles (the differentiated forward)",
                                            "fp64_vs_ewald": error_metrics(a4[:, tgt], ref),
                                            "ewald_samples": int(tgt.size)}
    finally:
        prob.close()
    ens = {}
    specs = E.ensemble_specs()
    for n_side in (32, 64, 96, 128):
        s = next(sp for sp in specs if sp.n_side == n_side)
        lattice, psi = E.box_initial_conditions(s)
        pos = np.mod(lattice + psi, s.box)

        class _B:
            box = s.box
            softening = s.box / s.n_side / 30.0

        with TreePmCoordinator(s.mesh) as cb:
            db = [torch.from_numpy(np.ascontiguousarray(pos[a])).to("cuda:0") for a in range(3)]
            db.append(torch.full((s.n,), 1.0, dtype=torch.float64, device="cuda:0"))
            ob = E.box_options(s, L.PRECISION_MIXED)
            ob.gravitational_constant = 1.0
            _, ab = solve_timed(L, cb, db, ob, L.PRECISION_MIXED, 1)
            tgt, ref = ewald_reference(cb, db, _B, 2048, s.seed, 1.0)
            ens[f"{n_side}^3_mesh{s.mesh}"] = {"mixed_vs_ewald": error_metrics(ab[:, tgt], ref),
                                               "ewald_samples": int(tgt.size)}
    out["ensemble_boxes"] = ens
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baselines(L, ics, n_side, repeats):
    """The oracle's FP64 force evaluation on the host cores: the 256^3 workload
    law at n_side^3 (median of `repeats` after one warm-up), and the 32^3
    workload's force evaluation plus one global KDK step."""
    from oracle import oracle as O

    threads = O.set_threads(min(os.cpu_count() or 1, 16))
    cw = ics.config(2, n_side=n_side)
    oc = opts_fp64_for(L, cw, ics)
    times = []
    for i in range(repeats + 1):
        tc = time.perf_counter()
        O.solve(cw.x, cw.y, cw.z, cw.mass, oc, cw.mesh)
        if i > 0:
            times.append(time.perf_counter() - tc)
    tcpu = float(np.median(times))
    model = cpu_model()
    cpu = {"value": cw.x.size / tcpu, "unit": "particles/s", "cores": threads, "kind": "port",
           "cpu_model": model,
           "sample": f"oracle FP64 CPU restatement (own FFT, no FFTW), one full force evaluation of the "
                     f"256^3 workload law sampled at {n_side}^3 particles / {cw.mesh}^3 mesh (box {cw.box:g} "
                     f"Mpc/h, same density and mesh-to-particle ratio); median of {repeats} after one "
                     f"warm-up, {tcpu:.2f}s each on {threads} threads of {model}; particles/s taken as the "
                     f"256^3 rate (per-particle work is set by the density, which the sample keeps)"}
    w1 = ics.config(0)
    o1 = opts_fp64_for(L, w1, ics)
    from cosmosim import kdk_factors

    f = kdk_factors(w1.scale_factor, w1.scale_factor * 1.02, ics.H0_CODE, ics.OMEGA_M, 1.0 - ics.OMEGA_M)
    times = []
    for i in range(repeats + 1):
        x, y, z = w1.x.copy(), w1.y.copy(), w1.z.copy()
        vx, vy, vz = w1.vx.copy(), w1.vy.copy(), w1.vz.copy()
        tc = time.perf_counter()
        a = O.solve(x, y, z, w1.mass, o1, w1.mesh)
        O.kick(vx, vy, vz, *a, f["first_kick_factor"], f["first_hubble_drag"])
        O.drift(x, y, z, vx, vy, vz, f["drift_factor"], (w1.box,) * 3)
        a = O.solve(x, y, z, w1.mass, o1, w1.mesh)
        O.kick(vx, vy, vz, *a, f["second_kick_factor"], f["second_hubble_drag"])
        if i > 0:
            times.append(time.perf_counter() - tc)
    t1 = float(np.median(times))
    cfg1 = {"workload": w1.name, "seconds": t1, "value": w1.x.size / t1, "unit": "particles/s",
            "what": "force evaluation + one global KDK step (kick, drift, force evaluation, kick), FP64 oracle",
            "cores": threads, "median_of": repeats}
    return cpu, cfg1


def measure_hierarchical(max_bin, n_side=128):
    """Hierarchical KDK blocks (mixed precision) on the 128^3 workload law
    at a = 0.5 (velocities from its linear field), block span 2^M times the
    smallest gravity interval: median wall time of three consecutive blocks
    after two warm-up blocks, fine-tick solves and the active fraction of the
    last, beside the cost of 2^M global steps at the quantum (full solves)."""
    import torch

    import cosmosim._lib as L
    from cosmosim import TreePmCoordinator, ForceAccumulator, hierarchical as H, ics

    w = ics.config(1, n_side=n_side)
    cosmo = H.Cosmology(ics.H0_CODE, ics.OMEGA_M, 1.0 - ics.OMEGA_M)
    a0 = w.scale_factor
    o = L.default_options(box_size=w.box, gravitational_constant=ics.G_CODE, opening_theta=0.5,
                          rcut_cells=4.5 * 1.25, asmth_cells=1.25, softening_epsilon=w.softening,
                          max_leaf_size=16, precision=L.PRECISION_MIXED)
    dev = lambda a: [torch.tensor(c, dtype=torch.float64, device="cuda:0") for c in a]  # noqa: E731
    x = dev([w.x, w.y, w.z])
    u = dev([w.vx / a0, w.vy / a0, w.vz / a0])
    m = torch.tensor(w.mass, device="cuda:0")
    kdk = H.HierarchicalKdk(w.mesh, x, u, m, o, cosmo, max_bin)
    timed = 3
    try:
        plan = kdk.block(a0)  # warm-up blocks: size the buffers and advance the state
        plan = kdk.block(plan.scale[-1])
        solves0, sr0 = kdk.solves, kdk.short_range_solves
        times = []
        for _ in range(timed):
            torch.cuda.synchronize()
            t = time.perf_counter()
            plan = kdk.block(plan.scale[-1])
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
        el = float(np.median(times))  # a block that first grows a buffer pays hipMalloc once
    finally:
        kdk.close()
    counts = np.bincount(plan.bins.cpu().numpy(), minlength=max_bin + 1)
    periods = 1 << max_bin
    active_targets = sum(int(sum(counts[b] for b in plan.closing[k])) for k in range(periods - 1))
    # cost of one full mixed solve for comparison (2^M global steps at the quantum)
    acc = tuple(torch.empty(w.x.size, dtype=torch.float32, device="cuda:0") for _ in range(3))
    o.short_range_only = 0
    with TreePmCoordinator(w.mesh) as coord:
        coord.solve_active_set(*x, m, ForceAccumulator(acc), o)
        coord.synchronize()
        t = time.perf_counter()
        for _ in range(timed):
            coord.solve_active_set(*x, m, ForceAccumulator(acc), o)
        coord.synchronize()
        full = (time.perf_counter() - t) / timed
    return {"workload": w.name, "particles": int(w.x.size), "max_bin": max_bin, "block_seconds": el,
            "blocks_timed": timed, "block_seconds_each": times,
            "bins": counts.tolist(), "quantum_code_time": plan.quantum,
            "synchronized_solves": (kdk.solves - solves0) // timed,
            "short_range_solves": (kdk.short_range_solves - sr0) // timed,
            "fine_tick_targets": active_targets, "full_solve_seconds": full,
            "global_steps_at_quantum_seconds": periods * full,
            "speedup_vs_global_at_quantum": periods * full / el}


def measure_ensemble(steps, gradients, workers=8):
    """The 48-box ensemble: leaf size per box size from the autotuner, one
    batched mixed-precision KDK forward of `steps` steps (initial forces plus
    one batched solve per step), the same forward from one host worker, and the
    FP64 gradients of every box."""
    import torch

    import cosmosim._lib as L
    from cosmosim import ensemble as E

    specs = E.ensemble_specs()
    tuned = {}
    for n_side in sorted({s.n_side for s in specs}):
        spec = next(s for s in specs if s.n_side == n_side)
        tuned[n_side] = E.autotune(spec)
        log(f"autotune {n_side}^3: leaf {tuned[n_side]['leaf']}, fft lines {tuned[n_side]['fft_lines']}, "
            f"walk waves {tuned[n_side]['walk_waves']}")
    leaves = {k: v["leaf"] for k, v in tuned.items()}
    shapes = {k: (v["fft_lines"], v["walk_waves"]) for k, v in tuned.items()}
    particles = sum(s.n for s in specs)
    out = {"boxes": len(specs), "particles": particles, "kdk_steps": steps, "precision": "mixed",
           "autotuned": tuned, "meshes": "2 n_side per axis (64, 128, 192, 256)"}
    def timed(fn, start, fwd, repeats=3):
        els = []
        for _ in range(repeats):
            fwd.set_state(start)
            t = time.perf_counter()
            fn()
            fwd.synchronize()
            els.append(time.perf_counter() - t)
        return float(np.median(els)), els

    for workers_ in (workers, 1):
        with E.Ensemble(specs, workers=workers_) as ens:
            fwd = E.EnsembleForward(ens, steps, L.PRECISION_MIXED, leaves, shape_by_n_side=shapes)
            start = fwd.state()
            ens.solve(fwd.x, fwd.m, fwd.g, fwd.options)  # warm-up: sizes every box's buffers
            fwd.synchronize()
            el, els = timed(fwd.run, start, fwd)
            key = "forward" if workers_ == workers else "forward_one_worker"
            out[key] = {"workers": workers_, "seconds": el, "seconds_each": els,
                        "force_evaluations_per_box": steps + 1,
                        "value": particles * (steps + 1) / el, "unit": "particle force evaluations/s"}
            log(f"ensemble forward with {workers_} workers: {el:.3f} s (median of {len(els)})")
            if workers_ == workers:
                # the whole forward of all 48 boxes recorded as one HIP graph
                # (capture-safe solves, each box a branch), replayed
                fwd.set_state(start)
                t = time.perf_counter()
                graph = fwd.capture()
                t_capture = time.perf_counter() - t
                try:
                    el_g, els_g = timed(graph.launch, start, fwd)
                finally:
                    graph.close()
                out["forward_graph"] = {"seconds": el_g, "seconds_each": els_g, "capture_seconds": t_capture,
                                        "force_evaluations_per_box": steps + 1,
                                        "value": particles * (steps + 1) / el_g,
                                        "unit": "particle force evaluations/s",
                                        "what": "one hipGraphLaunch of the recorded forward (initial forces, "
                                                f"{steps} KDK steps of every box), median of {len(els_g)}"}
                log(f"ensemble forward as one graph: {el_g:.3f} s (capture {t_capture:.2f} s)")
    out["worker_speedup"] = out["forward"]["value"] / out["forward_one_worker"]["value"]
    out["graph_speedup"] = out["forward_graph"]["value"] / out["forward"]["value"]
    if gradients:
        E.box_gradient(specs[0], steps, explicit=True)  # warm-up (tables, allocator)
        t = time.perf_counter()
        nf = nv = 0
        for s in specs:
            _, _, f, v = E.box_gradient(s, steps)
            nf += f
            nv += v
        torch.cuda.synchronize()
        serial = time.perf_counter() - t
        t = time.perf_counter()
        E.ensemble_gradients(specs, steps, workers=workers)
        threaded = time.perf_counter() - t
        out["gradients"] = {"seconds": threaded, "workers": workers, "serial_seconds": serial,
                            "precision": "fp64", "force_evaluations": nf, "force_vjps": nv, "checkpoint_every": 2,
                            "what": f"d loss / d psi of all {len(specs)} boxes; boxes on {workers} host threads, "
                                    "each with its own torch stream and context (explicit adjoint sweeps); "
                                    "serial_seconds: one box after another through torch.autograd"}
        log(f"ensemble gradients: {threaded:.2f} s on {workers} threads, {serial:.2f} s serial")
        # forward + adjoint sweeps of all 48 boxes recorded as one graph
        # (torch.cuda.graph; every box a branch on its own stream)
        t = time.perf_counter()
        gg = E.GradientGraph(specs, steps)
        t_build = time.perf_counter() - t
        try:
            els = []
            for _ in range(3):
                t = time.perf_counter()
                gg.replay()
                els.append(time.perf_counter() - t)
        finally:
            gg.close()
        out["gradients_graph"] = {"seconds": float(np.median(els)), "seconds_each": els,
                                  "prepare_and_record_seconds": t_build, "precision": "fp64",
                                  "what": f"one replay of the recorded forward and adjoint sweeps of all {len(specs)} "
                                          "boxes (FP64 capture-safe solves and VJPs, CIC density and transpose, "
                                          "element-wise KDK work), median of 3; the losses are summed outside it"}
        log(f"ensemble gradients as one graph: {np.median(els):.2f} s (prepared and recorded in {t_build:.1f} s)")
    return out


def measure_adaptive(L, coord, dev, w, acc_prev, steps, prev_name):
    """FP64 adaptive-policy solves with the group walk (and one per-target
    adaptive solve for comparison); returns (summary, group-walk accelerations)."""
    import torch

    from cosmosim import ForceAccumulator, ics

    n = w.x.size
    o = L.default_options(box_size=w.box, gravitational_constant=ics.G_CODE, opening_theta=0.7,
                          rcut_cells=4.5 * 1.25, asmth_cells=1.25, softening_epsilon=w.softening, max_leaf_size=16)
    o.acceptance_policy = 1
    o.relative_force_tolerance = 0.0025
    prev = torch.from_numpy(np.sqrt((acc_prev ** 2).sum(axis=0))).to("cuda:0")
    acc = tuple(torch.zeros(n, dtype=torch.float64, device="cuda:0") for _ in range(3))
    fa = ForceAccumulator(acc, previous_acceleration=prev)
    out = {"settings": {"acceptance": "adaptive relative_monopole", "alpha": 0.0025, "maximum_opening_angle": 1.0,
                        "cutoff_node_policy": "com_distance", "gaussian_pair_lookup": True, "group_walk": True,
                        "history": f"|a| of the {prev_name} solve"}}
    result = None
    for group in (1, 0):
        ad = L.default_adaptive_options(error_estimator=L.ESTIMATOR_RELATIVE_MONOPOLE, maximum_opening_angle=1.0,
                                        cutoff_node_policy=L.CUTOFF_COM_DISTANCE, gaussian_pair_lookup=1,
                                        group_walk=group)
        reps = steps if group else 1
        counters = coord.solve_active_set(*dev, fa, o, want_diagnostics=True, adaptive_options=ad).diagnostics
        coord.profiling(True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            coord.solve_active_set(*dev, fa, o, adaptive_options=ad)
        torch.cuda.synchronize()
        elapsed = (time.perf_counter() - t) / reps
        prof = coord.profile_read()
        coord.profiling(False)
        walk_ms = prof["walk_ms"] / max(prof["solves"], 1)
        key = "group_walk" if group else "per_target_walk"
        out[key] = {"value": n / elapsed, "unit": "particles/s", "ms_per_step": 1e3 * elapsed, "walk_ms": walk_ms,
                    "pairs_per_target": counters["direct_pair_evaluations"] / n,
                    "visited_nodes_per_target": counters["visited_nodes"] / n,
                    "accepted_multipoles": counters["accepted_internal_multipoles"]}
        if group:
            acc_group = np.stack([a.cpu().numpy() for a in acc])
            pairs = counters["direct_pair_evaluations"]
            out[key].update({"groups": counters["group_walk_groups"],
                             "fallback_targets": counters["group_walk_fallback_targets"],
                             "list_entries_max": counters["group_walk_list_entries_max"],
                             "valu": {"bound": "fp64 valu", "pairs": pairs, "flop_per_pair": GROUP_FLOP_PER_PAIR,
                                      "achieved_tflops": pairs * GROUP_FLOP_PER_PAIR / (walk_ms * 1e-3) / 1e12,
                                      "peak_tflops": FP64_VALU_PEAK / 1e12}})
            out["group_vs_strict"] = {"reference": prev_name, **error_metrics(acc_group, acc_prev)}
            result = acc_group
    out["group_over_per_target"] = out["group_walk"]["value"] / out["per_target_walk"]["value"]
    # the same group walk in mixed precision: FP64 groups, lists and decisions,
    # FP32 listed-source pairs, FP32 PM and output lanes
    o.precision = L.PRECISION_MIXED
    acc32 = tuple(torch.zeros(n, dtype=torch.float32, device="cuda:0") for _ in range(3))
    fa32 = ForceAccumulator(acc32, previous_acceleration=prev)
    ad = L.default_adaptive_options(error_estimator=L.ESTIMATOR_RELATIVE_MONOPOLE, maximum_opening_angle=1.0,
                                    cutoff_node_policy=L.CUTOFF_COM_DISTANCE, gaussian_pair_lookup=1, group_walk=1)
    counters = coord.solve_active_set(*dev, fa32, o, want_diagnostics=True, adaptive_options=ad).diagnostics
    coord.profiling(True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        coord.solve_active_set(*dev, fa32, o, adaptive_options=ad)
    torch.cuda.synchronize()
    elapsed = (time.perf_counter() - t) / steps
    prof = coord.profile_read()
    coord.profiling(False)
    o.precision = L.PRECISION_FP64
    acc_mix = np.stack([a.cpu().numpy().astype(np.float64) for a in acc32])
    out["group_walk_mixed"] = {
        "value": n / elapsed, "unit": "particles/s", "ms_per_step": 1e3 * elapsed,
        "walk_ms": prof["walk_ms"] / max(prof["solves"], 1),
        "pairs_per_target": counters["direct_pair_evaluations"] / n,
        "accepted_multipoles": counters["accepted_internal_multipoles"],
        "fallback_targets": counters["group_walk_fallback_targets"],
        "mixed_vs_fp64_group": error_metrics(acc_mix, result),
        "over_fp64_group": (n / elapsed) / out["group_walk"]["value"]}
    return out, result


def measure_gradient(n_side, directions):
    """grad/fwd ratio of the differentiable workload: n_side^3 particles, box
    50 (n_side/64) Mpc/h, mesh 2 n_side, ten KDK steps a = 0.05 -> 0.5 with
    checkpoints every two steps, loss = sum (delta_final - delta_target)^2 on an
    n_side^3 CIC mesh; FP64 throughout (the force VJP is FP64). Also reports
    central differences of the loss along random unit directions against the
    gradient, at the 1e-3 spacing step and at 1e-6 spacings."""
    import torch

    from cosmosim import adjoint

    prob, psi, target = adjoint.adjoint_problem(n_side=n_side, steps=10, a_begin=0.05, a_end=0.5,
                                                checkpoint_every=2)
    try:
        psi64 = psi.to(torch.float64)

        def run(with_grad):
            p = psi64.clone().requires_grad_(with_grad)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            loss = adjoint.density_loss(p, target, prob)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            grad = None
            if with_grad:
                (grad,) = torch.autograd.grad(loss, p)
                torch.cuda.synchronize()
            t2 = time.perf_counter()
            return float(loss), grad, t1 - t0, t2 - t1

        run(True)  # warm-up: buffers and tables
        prob.force_evaluations = prob.vjp_evaluations = 0
        loss, grad, t_fwd_g, t_bwd = run(True)
        evals = {"forward_force_evaluations": prob.force_evaluations, "force_vjps": prob.vjp_evaluations}
        _, _, t_fwd, _ = run(False)
        prob.record_pair_sets = True
        prob.pair_sets = []
        with torch.no_grad():
            adjoint.density_loss(psi64, target, prob)
        accepted = int(sum(rec[2] for rec in prob.pair_sets))
        prob.record_pair_sets = False
        prob.pair_sets = []
        spacing = prob.box / n_side
        fd = {}
        for label, h in (("step_1e-3_spacing", 1e-3 * spacing), ("step_1e-6_spacing", 1e-6 * spacing)):
            rows = adjoint.central_differences(prob, psi64, target, grad, directions=directions, step=h)
            errs = np.array([r["relative_error"] for r in rows])
            kept = np.array([not r["crosses_cutoff"] for r in rows])
            fd[label] = {"directions": directions, "median_relative_error": float(np.median(errs)),
                         "max_relative_error": float(np.max(errs)),
                         "directions_keeping_pair_sets": int(kept.sum()),
                         "kept_median_relative_error": float(np.median(errs[kept])) if kept.any() else None,
                         "kept_max_relative_error": float(np.max(errs[kept])) if kept.any() else None,
                         "rule": "a direction whose +h and -h trajectories sum different pairs inside r_cut or "
                                 "accept a different number of nodes in any force evaluation (pair counts, pair-set "
                                 "checksums and accepted-node counts of the FP64 walk) moves a pair across the "
                                 "cutoff or flips an acceptance, where the loss jumps; it is excluded"}
        return {"value": (t_fwd_g + t_bwd) / t_fwd, "backward_over_forward": t_bwd / t_fwd,
                "forward_ms": 1e3 * t_fwd, "gradient_ms": 1e3 * (t_fwd_g + t_bwd), "loss": loss,
                "workload": f"adjoint_{n_side}^3_box{prob.box:g}_mesh{prob.mesh}, 10 KDK steps a 0.05->0.5, "
                            f"checkpoints every 2, FP64",
                "forward": "strict walk with quadrupole moments (multipole_order 2)",
                "accepted_nodes_over_trajectory": accepted,
                **evals, "central_differences": fd}
    finally:
        prob.close()


def measured_traffic(workload):
    """HBM bytes per walk launch from the committed rocprofv3 --pmc summary of
    this workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    path = ROOT / "profiles" / "walkmix_kernel_treepm256_traffic.txt"
    if not path.exists():
        return None
    fields = {}
    for line in path.read_text().splitlines():
        if ":" in line and not line.startswith(" "):
            k, v = line.split(":", 1)
            fields[k.strip()] = v.strip()
    if fields.get("workload") != workload:
        return None
    try:
        return float(fields["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def opts_fp64_for(L, w, ics):
    return L.default_options(box_size=w.box, gravitational_constant=ics.G_CODE, opening_theta=0.5,
                             rcut_cells=4.5 * 1.25, asmth_cells=1.25, softening_epsilon=w.softening,
                             max_leaf_size=16)


if __name__ == "__main__":
    main()
