"""Per-configuration reports of the MI355X lattice-Boltzmann step.

    python -m plbpd_amd.reports c1 c2 c3 c4 c5 [--out DIR] [--quick]

Each report is one JSON file (DIR/report_<config>.json, default profiles/):

  c1  D2Q9 Poiseuille channel 256 x 64 (walls on the up/down faces, periodic
      x), tau 0.8, g 1e-6, 10,000 steps, FP64: the x-averaged physical
      velocity of the literal and fused runs against the transient Fourier
      series of the start-up flow, and their difference (stale seam reads).
  c2  D3Q19 channel 128^3 (walls up/down, periodic x and y), tau 0.6,
      g 1e-7, seeded 1e-4 perturbation, 1,000 steps, FP64: literal and fused
      throughput and their difference.  (Bit-level parity of both against the
      CPU oracle at these sizes is tests/test_gpu_fullsize.py.)
  c3  D3Q19 256^3 with 10% solids, tau 0.55, g 5e-7, 10,000 steps: velocity
      and density error of each storage variant against the FP64 run, MLUPS.
  c4  D3Q27 384^3 Taylor-Green start (u0 0.02, k = 2 pi / 384), walls
      up/down, tau 0.52, 5,000 steps: kinetic-energy decay of FP16-shifted,
      FP32-shifted and FP32 storage against FP64, per-field error, FP64 mass
      drift.
  c5  512 batched channel domains (edges log-uniform in [16, 96], tau in
      [0.55, 1.5], g in [1e-7, 1e-6]), FP32-shifted, 2,000 steps: autotuned
      tiling, aggregate MLUPS of one launch per step (HIP graph) against one
      launch per domain, per-domain velocity error against an FP64 batch.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import plbpd_amd as P  # noqa: E402


def _rel(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))


def _l2(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def poiseuille_series(yt, H, g, nu, t, nmax=399):
    us = g / (2 * nu) * yt * (H - yt)
    s = np.zeros_like(yt)
    for n in range(1, nmax + 1, 2):
        s += 4 * g * H * H / (nu * math.pi ** 3 * n ** 3) * np.sin(n * math.pi * yt / H) * \
            math.exp(-n * n * math.pi ** 2 * nu * t / (H * H))
    return us - s


def report_c1(quick=False):
    X, Y, Z, tau, g = 256, 1, 64, 0.8, 1e-6
    steps = 2000 if quick else 10000
    pb = P.channel(periodic_y=False)
    out = {"config": "D2Q9 256x64 Poiseuille, tau 0.8, g 1e-6, FP64", "steps": steps}
    with P.State(9, X, Y, Z, "f64", P.LITERAL) as st:
        st.initialize(1.0, (0, 0, 0), None, pb)
        st.configure(tau, (g, 0, 0), pb)
        st.step(steps)
        rho_l, u_l = st.download(P.FIELD_RHO), st.download(P.FIELD_U)
    with P.State(9, X, Y, Z, "f64", P.FUSED) as st:
        st.set_boundary(pb)
        st.initialize(1.0, (0, 0, 0), None, pb)
        st.configure(tau, (g, 0, 0), pb)
        st.step(steps)
        rho_f, u_f = st.moments()
        mass = st.global_moments()[0]
    # transient Fourier series; walls halfway between the wall layer and the first fluid row
    H, nu = Z - 2, (tau - 0.5) / 3
    yt = np.arange(H) + 0.5
    ua = poiseuille_series(yt, H, g, nu, steps)
    ux_f = (u_f[1:-1, 0, :, 0] + g / (2 * rho_f[1:-1, 0, :])).mean(axis=1)
    cols = slice(1 + 16, X + 1 - 16)
    ux_l = (u_l[2:Z, 1, cols, 0] + g / (2 * rho_l[2:Z, 1, cols])).mean(axis=1)
    out["analytic"] = {"u_max_series": float(ua.max()), "u_max_steady": float(g / (8 * nu) * H * H),
                       "fused_max_err_rel": float(np.abs(ux_f - ua).max() / ua.max()),
                       "literal_max_err_rel_interior_columns": float(np.abs(ux_l - ua).max() / ua.max()),
                       "fused_profile": ux_f.tolist(), "series_profile": ua.tolist()}
    # literal minus fused: the stale seam reads of the staged step
    du = u_l[1:-1, 1, 1:-1, 0] - u_f[:, 0, :, 0]
    out["literal_minus_fused"] = {"u_max_abs": float(np.abs(du[1:-1]).max()),
                                  "u_max_abs_seam_columns": float(np.abs(du[1:-1][:, [0, -1]]).max()),
                                  "u_max_abs_interior_columns": float(np.abs(du[1:-1][:, 16:-16]).max())}
    out["fused_mass_fp64"] = float(mass)
    return out


def report_c2(quick=False):
    N = 64 if quick else 128
    steps = 200 if quick else 1000
    tau, g = 0.6, 1e-7
    init = (P.INIT_PERTURBED, 2024, 1e-4)
    pb = P.channel()
    out = {"config": f"D3Q19 {N}^3 channel, tau 0.6, g 1e-7, 1e-4 seeded perturbation, FP64", "steps": steps}
    res = {}
    for mode in (P.LITERAL, P.FUSED):
        with P.State(19, N, N, N, "f64", mode) as st:
            if mode == P.FUSED:
                st.set_boundary(pb)
            st.initialize(1.0, (0, 0, 0), init, pb)
            st.configure(tau, (g, 0, 0), pb)
            t0 = time.perf_counter()
            st.step(steps)
            st.synchronize()
            res[mode] = (time.perf_counter() - t0, st.download(P.FIELD_RHO, P.LAYOUT_COMPACT),
                         st.download(P.FIELD_U, P.LAYOUT_COMPACT))
    out["literal_seconds"], out["fused_seconds"] = res[P.LITERAL][0], res[P.FUSED][0]
    out["literal_mlups"] = N ** 3 * steps / res[P.LITERAL][0] / 1e6
    out["fused_mlups"] = N ** 3 * steps / res[P.FUSED][0] / 1e6
    fl = (slice(1, -1),)
    out["literal_minus_fused"] = {"u_l2_rel": _l2(res[P.LITERAL][2][fl], res[P.FUSED][2][fl]),
                                  "rho_max_abs": float(np.abs(res[P.LITERAL][1][fl] - res[P.FUSED][1][fl]).max())}
    return out


def report_c3(quick=False):
    import bench
    mask = bench.solid_mask()
    steps = 1000 if quick else 10000
    out = {"config": "D3Q19 256^3, 10% solids, tau 0.55, g 5e-7", "steps": steps,
           "errors_vs_f64": bench.errors_vs_f64(P, ["f32s", "f32", "f64a32", "f16s"], mask, steps)}
    nfluid = bench.fluid_cells(mask)
    perf = {}
    for v in ("f64", "f64a32", "f32", "f32s", "f16s"):
        wall, ev, scheme = bench.time_variant(P, v, mask, 200, 10, True)
        kms = ev / 200
        perf[v] = {"mlups": 256 ** 3 * 200 / wall / 1e6, "kernel_ms": kms, "scheme": scheme,
                   "achieved_gbps": bench.algorithmic_bytes(v) / (kms * 1e-3) / 1e9,
                   "fluid_only_gbps": bench.fluid_bytes(v, nfluid) / (kms * 1e-3) / 1e9}
    out["performance"] = perf
    out["speedup_f32s_vs_f64"] = perf["f32s"]["mlups"] / perf["f64"]["mlups"]
    return out


def _tg_state(variant, L, bc):
    st = P.State(27, L, L, L, variant, P.FUSED)
    st.set_boundary(bc)
    st.initialize(1.0, (0.02, 0.0, 0.0), (P.INIT_TAYLOR_GREEN, 0, 0.0), bc)
    st.configure(0.52, (0.0, 0.0, 0.0), bc)
    return st


def report_c4(quick=False):
    L = 128 if quick else 384
    steps = 1000 if quick else 5000
    every = steps // 10
    bc = P.channel()
    out = {"config": f"D3Q27 {L}^3 Taylor-Green (u0 0.02, k 2pi/L), walls up/down, tau 0.52",
           "steps": steps, "sample_every": every, "variants": {}}
    fluid = np.ones((L, L, L), bool)
    fluid[0] = fluid[-1] = False
    ref = None
    for v in ("f64", "f32s", "f32", "f16s"):
        st = _tg_state(v, L, bc)
        g0 = st.global_moments()
        ke, t_gpu = [], 0.0
        for s in range(0, steps + 1, every):
            if s:
                t0 = time.perf_counter()
                st.step(every)
                st.synchronize()
                t_gpu += time.perf_counter() - t0
            rho, u = st.moments()
            ke.append(float(0.5 * (rho[fluid] * (u[fluid] ** 2).sum(-1)).sum()))
        g1 = st.global_moments()
        rho, u = st.moments()
        st.close()
        rec = {"kinetic_energy": ke, "ke_ratio_end": ke[-1] / ke[0], "mass_drift_fp64": float((g1[0] - g0[0]) / g0[0]),
               "mlups": L ** 3 * steps / t_gpu / 1e6}
        if v == "f64":
            ref = (rho, u, ke)
        else:
            rr, ur, ker = ref
            rec["u_l2_rel_vs_f64"] = _l2(u[fluid], ur[fluid])
            rec["rho_dev_l2_rel_vs_f64"] = float(np.sqrt(((rho[fluid] - rr[fluid]) ** 2).sum() /
                                                         ((rr[fluid] - 1.0) ** 2).sum()))
            rec["ke_max_rel_dev_vs_f64"] = float(max(abs(a - b) / b for a, b in zip(ke, ker)))
        out["variants"][v] = rec
    nu = (0.52 - 0.5) / 3
    out["periodic_tg_energy_factor"] = math.exp(-2 * nu * 3 * (2 * math.pi / L) ** 2 * steps)
    return out


def c5_domains(n=512, seed=5):
    rng = np.random.default_rng(seed)
    doms = []
    for _ in range(n):
        x, y, z = (int(round(math.exp(v))) for v in rng.uniform(math.log(16), math.log(96), 3))
        doms.append({"x": x, "y": y, "z": z, "tau": float(rng.uniform(0.55, 1.5)),
                     "force": (float(rng.uniform(1e-7, 1e-6)), 0.0, 0.0)})
    return doms


def _rank_correlations(rows):
    """Spearman rank correlation of the per-domain velocity error with each
    domain attribute (the steady channel velocity g H^2 / (8 nu), H the
    fluid height between the walls, among them)."""
    def ranks(v):
        order = np.argsort(v, kind="stable")
        r = np.empty(len(v))
        r[order] = np.arange(len(v))
        return r
    err = ranks(np.array([r["u_l2"] for r in rows]))
    attrs = {
        "tau": [r["tau"] for r in rows], "g": [r["g"] for r in rows],
        "cells": [r["x"] * r["y"] * r["z"] for r in rows], "wall_gap_z": [r["z"] - 2 for r in rows],
        "u_max_f64": [r["u_max_f64"] for r in rows],
        "steady_u": [r["g"] * (r["z"] - 2) ** 2 / (8 * (r["tau"] - 0.5) / 3) for r in rows],
        "diffusion_time_over_steps": [(r["z"] - 2) ** 2 / ((r["tau"] - 0.5) / 3) for r in rows],
    }
    return {k: float(np.corrcoef(err, ranks(np.array(v)))[0, 1]) for k, v in attrs.items()}


def report_c5(quick=False):
    doms = c5_domains(64 if quick else 512)
    steps = 200 if quick else 2000
    bc = P.channel()
    init = (P.INIT_PERTURBED, 31, 1e-4)
    out = {"config": f"{len(doms)} channel domains, edges log-uniform [16, 96], FP32-shifted", "steps": steps}
    b = P.Batch(19, "f32s", doms, bc)
    total, fluid = b.cells()
    out["total_cells"], out["fluid_cells"] = total, fluid
    b.initialize(1.0, (0, 0, 0), init)
    res = b.autotune(steps=6)
    out["autotune"] = [{"tile_cells": int(r[0]), "grid_blocks": int(r[1]), "xcd_runs": int(r[2]),
                        "ms_per_step": float(r[3])} for r in res]
    out["tiling"] = dict(zip(("tile_cells", "grid_blocks"), b.tiling()))
    out["tiling"]["xcd_runs"] = b.mapping()
    b.initialize(1.0, (0, 0, 0), init)  # autotuning advanced the state
    b.synchronize()
    t0 = time.perf_counter()
    b.step(steps, steps_per_graph=20)
    b.synchronize()
    t_batch = time.perf_counter() - t0
    out["batched_mlups"] = total * steps / t_batch / 1e6
    out["batched_fluid_mlups"] = fluid * steps / t_batch / 1e6
    bref = P.Batch(19, "f64", doms, bc)
    bref.initialize(1.0, (0, 0, 0), init)
    bref.step(steps, steps_per_graph=20)
    errs, per = [], []
    for d in range(len(doms)):
        _, u = b.moments(d)
        _, ur = bref.moments(d)
        fl = (slice(1, -1),)
        e = _l2(u[fl], ur[fl])
        errs.append(e)
        dm = doms[d]
        per.append({"domain": d, "x": dm["x"], "y": dm["y"], "z": dm["z"], "tau": dm["tau"],
                    "g": dm["force"][0], "u_max_f64": float(np.abs(ur[fl][..., 0]).max()), "u_l2": e})
    out["per_domain_u_l2_vs_f64"] = {"max": max(errs), "median": float(np.median(errs)), "min": min(errs)}
    # While a domain's bulk still accelerates uniformly (u = g t there), every
    # bulk cell holds the same values, so the FP32 rounding of each step has
    # the same sign in all of them and the error grows with the step count
    # instead of averaging out: at most about 2^-24 per step.
    out["error_bound_2^-24_x_steps"] = 2.0 ** -24 * steps
    out["largest_errors"] = sorted(per, key=lambda r: -r["u_l2"])[:12]
    out["error_rank_correlation"] = _rank_correlations(per)
    bref.close()
    # the same steps with one launch per domain
    b.initialize(1.0, (0, 0, 0), init)
    nb = min(steps, 200)
    b.synchronize()
    t0 = time.perf_counter()
    b.step_per_domain(nb)
    b.synchronize()
    t_dom = time.perf_counter() - t0
    out["per_domain_launch_mlups"] = total * nb / t_dom / 1e6
    out["per_domain_launch_steps"] = nb
    out["batched_over_per_domain"] = out["batched_mlups"] / out["per_domain_launch_mlups"]
    b.close()
    return out


REPORTS = {"c1": report_c1, "c2": report_c2, "c3": report_c3, "c4": report_c4, "c5": report_c5}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("configs", nargs="*", default=list(REPORTS))
    ap.add_argument("--out", default=os.path.join(ROOT, "profiles"))
    ap.add_argument("--quick", action="store_true", help="reduced sizes and step counts")
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    for c in args.configs:
        print(f"[report] {c}", flush=True)
        t0 = time.perf_counter()
        rep = REPORTS[c](args.quick)
        rep["seconds"] = time.perf_counter() - t0
        rep["quick"] = args.quick
        try:
            rep["device"] = P.device_info(0)
        except Exception:
            pass
        name = f"report_{c}{'_quick' if args.quick else ''}.json"
        with open(os.path.join(args.out, name), "w") as f:
            json.dump(rep, f, indent=1)
        print(json.dumps(rep)[:2000], flush=True)


if __name__ == "__main__":
    main()
