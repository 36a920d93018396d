#!/usr/bin/env python3
"""Benchmark of the MI355X lattice-Boltzmann BGK step (one JSON line).

Workload (the headline configuration): D3Q19, 256^3 lattice, walls on the two
faces normal to the wall axis (the reference's up/down faces), periodic in the
other two axes, seeded Bernoulli(0.10) interior solid voxels with halfway
bounce-back, tau = 0.55, body force 5e-7 along x, FP32 shifted-population
storage (f - w_k) with FP32 collision arithmetic.  A step is the fused
collide + stream + bounce-back update of the whole lattice; with 4-byte
storage one launch advances two steps (scheme 'pull2': one read and one
write of the populations per two steps, the values of two one-step launches
bit for bit).

Reported:
  value        MLUPS over every lattice cell (walls and solids included, as the
               reference counts its loop), host wall clock around K steps;
  roofline     algorithmic HBM bytes per launch / average launch duration
               (HIP events on the state's stream), against 8 TB/s.  Bytes per
               lattice update = 2*Q*sizeof(storage) + 1 flag byte (153 B for
               D3Q19 FP32), times every interior cell and every step of a
               launch (two for a two-step launch); traffic =
               measured HBM bytes per launch (rocprofv3 FETCH_SIZE +
               WRITE_SIZE, calibrated on a copy kernel) where a profile of
               this workload is in profiles/;
  variants     the same workload with FP64, FP32 and FP16-shifted storage;
  literal      the reference-staged step (bit-exact FP64, and FP32 with the
               reference's rounding) on the same lattice without the solids
               the reference lacks, MLUPS;
  errors       velocity / density L2 error of each variant against the FP64
               run after --error-steps steps;
  cpu_baseline the reference's FP32 staged CPU step (restated with OpenMP and
               AVX2) on this host's cores, bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--error-steps S]
                       [--no-errors] [--no-cpu] [--no-variants] [--profile]
                       [--copy-only]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS + % HBM roofline, D3Q19 256³: FP32/FP16 vs FP64 GPU at stated error"
HBM_PEAK_GBS = 8000.0
N = 256
Q = 19
TAU = 0.55
FORCE = (5e-7, 0.0, 0.0)
SOLID_FRACTION = 0.10
SOLID_SEED = 20250303
STORAGE_BYTES = {"f64": 8, "f32": 4, "f32s": 4, "f16s": 2, "f64a32": 8}
DTYPE = {"f64": "f64", "f32": "f32", "f32s": "f32", "f16s": "f32", "f64a32": "f32"}


def solid_mask(fraction=SOLID_FRACTION):
    if fraction <= 0:
        return None
    rng = np.random.default_rng(SOLID_SEED)
    return (rng.random((N, N, N)) < fraction).astype(np.uint8)


def make_state(P, variant, mask):
    bc = P.channel()  # periodic east/west + north/south, no-slip up/down
    st = P.State(Q, N, N, N, variant, P.FUSED)
    st.set_boundary(bc, mask)
    st.initialize(1.0, (0.0, 0.0, 0.0), None, bc)
    st.configure(TAU, FORCE, bc)
    return st


def fluid_cells(mask):
    m = np.zeros((N, N, N), np.uint8) if mask is None else mask.copy()
    m[0] = 1
    m[-1] = 1  # wall layers
    return int((m == 0).sum())


def bytes_per_update(variant):
    """Bytes per lattice update: every population read once and written once,
    plus the 1-byte cell flag."""
    return 2 * Q * STORAGE_BYTES[variant] + 1


def algorithmic_bytes(variant):
    """Bytes per launch: one update of every interior cell, wall and solid
    cells included (the reference collides and streams them all)."""
    return N ** 3 * bytes_per_update(variant)


def fluid_bytes(variant, nfluid):
    """The same count over fluid cells only (the flag byte over all cells)."""
    return nfluid * 2 * Q * STORAGE_BYTES[variant] + N ** 3


def time_variant(P, variant, mask, steps, warmup, graph):
    st = make_state(P, variant, mask)
    st.step(warmup)
    st.synchronize()
    # 20 steps per graph: ten two-step launches where the step runs two per launch
    spg = 20 if graph else 0
    if graph:
        st.time_steps(spg, spg)  # build the graph outside the timed region
    t0 = time.perf_counter()
    ev_ms = st.time_steps(steps, spg)
    st.synchronize()
    wall = time.perf_counter() - t0
    scheme = st.scheme()
    st.close()
    return wall, ev_ms, scheme


def time_literal(P, steps=20, warmup=2):
    """MLUPS of the literal (reference-staged) step, FP64 and FP32, on the
    headline lattice without interior solids (the reference has none)."""
    out = {"config": f"D3Q19 {N}^3 channel, walls up/down, periodic x/y, tau {TAU}, g {FORCE[0]}; "
                     f"plb_step in literal mode, {steps} timed steps"}
    for prec in ("f64", "f32"):
        bc = P.channel()
        with P.State(Q, N, N, N, prec, P.LITERAL) as st:
            st.initialize(1.0, (0.0, 0.0, 0.0), None, bc)
            st.configure(TAU, FORCE, bc)
            st.step(warmup)
            st.synchronize()
            ms = st.time_steps(steps) / steps
        out[prec] = {"mlups": round(N ** 3 / (ms * 1e-3) / 1e6, 1), "ms_per_step": round(ms, 4)}
    return out


def errors_vs_f64(P, variants, mask, steps):
    ref = make_state(P, "f64", mask)
    ref.step(steps)
    rho_r, u_r = ref.moments()
    gm_r = ref.global_moments()
    ref.close()
    fluid = (mask == 0) if mask is not None else np.ones((N, N, N), bool)
    fluid[0] = False
    fluid[-1] = False
    out = {}
    for v in variants:
        st = make_state(P, v, mask)
        gm0 = st.global_moments()
        st.step(steps)
        rho, u = st.moments()
        gm = st.global_moments()
        st.close()
        du = u[fluid] - u_r[fluid]
        dr = rho[fluid] - rho_r[fluid]
        out[v] = {
            "steps": steps,
            "u_l2_rel": float(np.sqrt((du ** 2).sum() / (u_r[fluid] ** 2).sum())),
            "u_max_abs": float(np.abs(du).max()),
            "rho_l2_rel": float(np.sqrt((dr ** 2).sum() / (rho_r[fluid] ** 2).sum())),
            "rho_dev_l2_rel": float(np.sqrt((dr ** 2).sum() / ((rho_r[fluid] - 1.0) ** 2).sum())),
            "mass_drift_fp64": float((gm[0] - gm0[0]) / gm0[0]),
        }
    out["f64_reference"] = {"u_max": float(np.abs(u_r[fluid]).max()),
                            "mass": float(gm_r[0]), "momentum_x": float(gm_r[1])}
    return out


def pattern_ceiling(capi, variant, solids, achieved):
    """The step's memory pattern without its arithmetic (plb_bench_pattern),
    each rate counted in the step's algorithmic bytes: a Q-stream copy, the
    row kernel's pulls, and its pulls around interior solids (own reverse
    rows and neighbour masks as well).  The kernel's own pattern is the last
    with solids, the second without."""
    sb = STORAGE_BYTES[variant]
    # the widths the step's kernel uses: 4 cells per lane around solids for
    # 4-byte storage, 2 for 8-byte, 1 without solids and for 2-byte storage
    vec = {4: 4, 8: 2}.get(sb, 1) if solids else 1
    out = {"unit": "GB/s", "cells_per_lane": vec, "storage_bytes": sb}
    for name, pat in (("q_stream_copy", capi.PATTERN_COPY), ("pull", capi.PATTERN_PULL),
                      ("pull_bounce_back", capi.PATTERN_PULL_SOLIDS)):
        g, _ = capi.bench_pattern(Q, N, N, N, sb, vec, pat, iters=20)
        out[name] = round(g, 1)
        out[name + "_frac"] = round(g / HBM_PEAK_GBS, 4)
    out["kernel_pattern"] = "pull_bounce_back" if solids else "pull"
    # each pattern makes one pass over the populations per step; the two-step
    # kernel makes one pass per two steps, so it can exceed these ceilings
    out["kernel_frac_of_pattern"] = round(achieved / out[out["kernel_pattern"]], 4)
    return out


def cpu_baseline(sample_steps=150, single_core_steps=8):
    """About 10 s of work on all host cores and 7 s on one (at about 250 and
    20 MLUPS): a bounded sample of the same lattice, reported, not targeted."""
    from oracle.pyoracle import CpuBaseline  # the oracle's restatement of the reference CPU path
    cb = CpuBaseline()
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    nu = (TAU - 0.5) / 3.0
    t_all, ph_all = cb.run(Q, N, N, N, nu, FORCE, 2, sample_steps, cores)
    t_one, _ = cb.run(Q, N, N, N, nu, FORCE, 1, single_core_steps, 1)
    mlups_all = N ** 3 * sample_steps / t_all / 1e6
    mlups_one = N ** 3 * single_core_steps / t_one / 1e6
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": round(mlups_all, 2), "unit": "MLUPS", "cores": cores, "kind": "port",
        "sample": (f"D3Q19 {N}^3 channel (walls up/down, periodic x/y), the reference's staged FP32 "
                   f"step (collide AVX2+FMA, push, no-slip, periodic, moments, reset; OpenMP over z), "
                   f"{sample_steps} timed steps after 2 warm-up; no interior solids (the reference has none)"),
        "single_core_mlups": round(mlups_one, 2), "single_core_steps": single_core_steps,
        "phase_seconds": {k: round(v, 4) for k, v in zip(
            ["collide", "stream", "noslip", "periodic", "moments", "reset"], ph_all.tolist())},
        "cpu_model": cpu_model,
        "nproc": len(os.sched_getaffinity(0)),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--variant", default="f32s")
    ap.add_argument("--solid-fraction", type=float, default=SOLID_FRACTION)
    ap.add_argument("--error-steps", type=int, default=10000)
    ap.add_argument("--no-errors", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--profile", action="store_true", help="timed steps only (for rocprofv3)")
    ap.add_argument("--copy-only", action="store_true", help="run only the copy calibration kernel")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only; the multi-GPU slab decomposition is a design "
              "(no figure for N > 1).", file=sys.stderr)
        sys.exit(2)

    import plbpd_amd as P
    from plbpd_amd import capi

    if args.copy_only:
        gbs, ms = capi.bench_copy(1 << 32, max(args.steps, 1))
        print(json.dumps({"copy_gbps": gbs, "copy_ms": ms, "bytes": 1 << 32}))
        return

    mask = solid_mask(args.solid_fraction)
    nfluid = fluid_cells(mask)
    print(f"[bench] D3Q19 {N}^3, {nfluid} fluid cells, variant {args.variant}", file=sys.stderr, flush=True)
    wall, ev_ms, scheme = time_variant(P, args.variant, mask, args.steps, args.warmup, not args.no_graph)
    ms_per_step = wall / args.steps * 1e3
    mlups = N ** 3 * args.steps / wall / 1e6
    # one launch of the 'pull2' kernel advances two steps: per-launch time and
    # bytes both count two steps
    spl = 2 if scheme == "pull2" else 1
    kernel_ms = ev_ms / (args.steps / spl)
    bytes_launch = spl * algorithmic_bytes(args.variant)
    achieved = bytes_launch / (kernel_ms * 1e-3) / 1e9

    traffic = None
    tag = f"d3q19_256_{args.variant}" + ("_solids" if mask is not None else "_walls")
    tpath = os.path.join(ROOT, "profiles", f"traffic_{tag}.json")
    if os.path.exists(tpath):
        try:
            tj = json.load(open(tpath))
            # only a profile of the kernel this run launches
            if tj.get("steps_per_launch", 1) == spl:
                traffic = tj.get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    line = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.variant],
        "data": "synthetic",
        "config": {
            "workload": (f"D3Q19 {N}^3 channel: walls on the up/down faces, periodic x and y, "
                         f"{args.solid_fraction:.0%} Bernoulli solid voxels (seed {SOLID_SEED}), tau {TAU}, "
                         f"gx {FORCE[0]}, {args.variant} storage, fused {scheme} step"),
            "lattice": "D3Q19", "size": [N, N, N], "storage": args.variant, "scheme": scheme,
            "tau": TAU, "force": list(FORCE), "solid_fraction": args.solid_fraction, "solid_seed": SOLID_SEED,
            "fluid_cells": nfluid, "hip_graph": not args.no_graph,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "kernel_ms": round(kernel_ms, 5),
        "steps_per_launch": spl,
        "fluid_mlups": round(nfluid * args.steps / wall / 1e6, 1),
        "bytes_per_update": bytes_per_update(args.variant),
        "bytes_per_launch_algorithmic": bytes_launch,
        "fluid_only_gbps": round(spl * fluid_bytes(args.variant, nfluid) / (kernel_ms * 1e-3) / 1e9, 1),
    }
    if args.profile:
        print(json.dumps(line))
        return

    gbs, _ = capi.bench_copy(1 << 32, 20)
    line["copy_gbps"] = round(gbs, 1)
    line["roofline"]["frac_of_copy"] = round(achieved / gbs, 4)
    line["roofline"]["pattern_ceiling"] = pattern_ceiling(capi, args.variant, mask is not None, achieved)
    try:
        line["device"] = P.device_info(0)
    except Exception:
        pass

    if not args.no_variants:
        var = {}
        for v in ("f64", "f32", "f16s"):
            if v == args.variant:
                continue
            w, e, sch = time_variant(P, v, mask, max(args.steps // 2, 50), args.warmup, not args.no_graph)
            steps_v = max(args.steps // 2, 50)
            spl_v = 2 if sch == "pull2" else 1
            kms = e / (steps_v / spl_v)
            ach = spl_v * algorithmic_bytes(v) / (kms * 1e-3) / 1e9
            var[v] = {"mlups": round(N ** 3 * steps_v / w / 1e6, 1), "kernel_ms": round(kms, 5),
                      "steps_per_launch": spl_v, "achieved_gbps": round(ach, 1),
                      "frac": round(ach / HBM_PEAK_GBS, 4), "scheme": sch}
        var[args.variant] = {"mlups": line["value"], "kernel_ms": line["kernel_ms"], "steps_per_launch": spl,
                             "achieved_gbps": line["roofline"]["achieved"], "frac": line["roofline"]["frac"],
                             "scheme": scheme}
        line["variants"] = var
        if "f64" in var:
            line["speedup_vs_f64"] = round(var[args.variant]["mlups"] / var["f64"]["mlups"], 3)
        line["literal"] = time_literal(P)

    if not args.no_errors:
        print(f"[bench] errors vs FP64 after {args.error_steps} steps", file=sys.stderr, flush=True)
        line["errors_vs_f64"] = errors_vs_f64(P, ["f32s", "f32", "f64a32", "f16s"],
                                              mask, args.error_steps)

    if not args.no_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        try:
            line["cpu_baseline"] = cpu_baseline()
        except Exception as e:  # the oracle library may be missing on a stripped tree
            line["cpu_baseline"] = {"value": None, "unit": "MLUPS", "error": str(e)}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
