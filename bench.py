#!/usr/bin/env python3
"""Benchmark of the MI355X transport step on the single-forecast workload.

Workload (default): spherical lon-lat grid 0.25 deg, 360x240x50 cells,
dz = 0.5 km, 15 grain bins (fall speeds log-spaced 0.01-5 m/s), zonal jet
with seeded correlated perturbations, Kh = Kz = 500 m2/s, Crank-Nicolson,
superbee, FP64.  One step = the reference's hot-path sequence
(Set_BC(1); AdvectHorz; AdvectVert; Set_BC(2); DiffuseHorz; DiffuseVert)
plus the driver glue, on device-resident state.

Prints ONE JSON line (stdout): FP64 bin-cell-updates/s, the HBM roofline of
the dominant kernel (timed live with HIP events on the library's stream),
the whole-step fraction of HBM peak on the compulsory bytes, and the CPU
baseline (the C oracle, a port of the reference Fortran, on the host cores).
"""
import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E
METRIC = "FP64 bin-cell-updates/s & % HBM roofline; adjoint/forward time ratio"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(kind, st):
    """Compulsory HBM bytes of one launch of a kernel kind: every array the
    routine reads or writes, touched once (concen read+write per bin-cell,
    vf per bin-cell for z, per-cell geometry/wind/K once, the deposit and
    DepositGranularity per column)."""
    ncell = st.nx * st.ny * st.nz
    ns = st.ns
    if kind in ("advect_x", "advect_y"):
        return (16 * ns + 24) * ncell          # q r/w; vel, sigma, kappa
    if kind == "advect_z":
        return (24 * ns + 32) * ncell + 24 * ns * st.nx * st.ny  # + vf; vz, vh, sigma, kappa; deposit
    if kind.startswith("diffuse"):
        return (16 * ns + 24) * ncell          # q r/w; kappa, sigma, K
    return None


def step_bytes_per_bincell(st, use_diffusion):
    """B_step / ns: compulsory bytes per bin-cell-update of the whole step."""
    per_cell = st.ns * 24 + 8 * 4 + 8 * 4 + (24 if use_diffusion else 0) + st.ns * 24.0 / st.nz
    return per_cell / st.ns


def cpu_baseline(st, use_diffusion, steps):
    """Time the oracle (C port of the reference path, OpenMP rows in the
    advection sweeps as in the reference) on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    o = oracle.Oracle(st)
    o.step_glue()
    o.transport_step(1, use_diffusion)  # warm-up
    t0 = time.perf_counter()
    for it in range(2, 2 + steps):
        o.step_glue()
        o.transport_step(it, use_diffusion)
    dt = time.perf_counter() - t0
    rate = st.ns * st.nx * st.ny * st.nz * steps / dt
    return {"value": rate, "unit": "bin-cell-updates/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": f"{steps} full transport steps (after 1 warm-up) of the same "
                      f"{st.nx}x{st.ny}x{st.nz}x{st.ns} workload, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", choices=("forecast", "cartesian"), default="forecast")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch autotuner")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the transport step shards "
              "by grain bin or ensemble member with no per-step exchange, but no multi-GPU path is built.",
              file=sys.stderr)
        sys.exit(2)

    pkg = importlib.import_module("volcano-ash3d_amd")
    tr = importlib.import_module("volcano-ash3d_amd.transport")
    tr.load()
    t0 = time.perf_counter()
    if args.config == "forecast":
        st = pkg.workloads.config2_state()
        use_diffusion = True
        workload = ("single forecast: lon-lat 0.25 deg 360x240x50, dz 0.5 km, 15 bins (vf 0.01-5 m/s), "
                    "zonal jet + seeded perturbations, Kh=Kz=500 m2/s, Crank-Nicolson, superbee")
    else:
        st = pkg.workloads.config1_state()
        use_diffusion = False
        workload = "projected 100x100x30, 5 km, uniform wind, 1 bin, no diffusion, superbee"
    st.concen_pd[..., 1] = st.concen_pd[..., 0]  # the glue copies ts1 into ts0 each step
    cpu_state = None if args.no_cpu_baseline else st.copy()
    log(f"[bench] state built in {time.perf_counter() - t0:.1f} s, dt={st.dt} h, "
        f"{st.nbytes() / 1e9:.2f} GB host arrays")

    T = tr.Transport(st)
    itime = 0
    for _ in range(args.warmup):
        itime += 1
        T.step_glue()
        T.transport_step(itime, use_diffusion)
    T.synchronize()
    log(f"[bench] warm-up {args.warmup} steps done")
    tuning = None
    if not args.no_tune:
        at = importlib.import_module("volcano-ash3d_amd.autotune")
        tuning = at.autotune(T, use_diffusion)
        log("[bench] autotune: " + json.dumps({k: [v[0], {str(a): round(b, 4) for a, b in (v[1] or {}).items()}]
                                                for k, v in tuning.items()}))
        itime += 1000

    T.profile(True)
    T.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        itime += 1
        T.step_glue()
        T.transport_step(itime, use_diffusion)
    T.synchronize()
    wall = time.perf_counter() - t0
    prof = T.profile_read()
    T.profile(False)
    log(f"[bench] timed {args.steps} steps in {wall:.3f} s")

    bincells = st.ns * st.nx * st.ny * st.nz
    rate = bincells * args.steps / wall
    ms_per_step = 1e3 * wall / args.steps
    bpb = step_bytes_per_bincell(st, use_diffusion)
    step_frac = bpb * rate / (HBM_PEAK_GBPS * 1e9)

    kernels = {}
    for kind, (ms, n) in prof.items():
        if n == 0:
            continue
        avg_ms = ms / n
        ab = algorithmic_bytes(kind, st)
        kernels[kind] = {"avg_ms": round(avg_ms, 5), "launches": n, "share": None,
                         "gbps": round(ab / (avg_ms * 1e-3) / 1e9, 1) if ab else None}
    tot = sum(v["avg_ms"] * v["launches"] for v in kernels.values())
    for v in kernels.values():
        v["share"] = round(v["avg_ms"] * v["launches"] / tot, 3)
    dominant = max((k for k in kernels if algorithmic_bytes(k, st)),
                   key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches"])
    dk = kernels[dominant]
    achieved = dk["gbps"]
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                "avg_ms": dk["avg_ms"], "algorithmic_bytes": algorithmic_bytes(dominant, st)}
    log("[bench] kernels: " + json.dumps(kernels))

    T.close()
    cpu = None
    if cpu_state is not None:
        log("[bench] timing the CPU oracle ...")
        cpu = cpu_baseline(cpu_state, use_diffusion, args.cpu_steps)
        log(f"[bench] cpu baseline {cpu['value']:.3e} bin-cell-updates/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC, "value": rate, "unit": "bin-cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload, "nx": st.nx, "ny": st.ny, "nz": st.nz, "bins": st.ns,
                   "dt_h": st.dt, "diffusion": "crank-nicolson" if use_diffusion else "off",
                   "parallelism": "single gpu",
                   "tuning": {k: v[0] for k, v in tuning.items()} if tuning else "defaults"},
        "step_hbm": {"bytes_per_bincell": round(bpb, 2), "frac_of_peak": round(step_frac, 4)},
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (rate / cpu["value"]) if cpu else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
