#!/usr/bin/env python3
"""Benchmark of the MI355X transport step on the large global workload.

Workload (default): periodic spherical lon-lat grid 0.25 deg, 1440x720x80
cells from pole to pole, dz = 0.3 km, 20 grain bins (fall speeds
log-spaced 0.005-8 m/s), two zonal jets with seeded correlated
perturbations vanishing toward the poles, Kh = 2000 m2/s, Kz = 10 m2/s,
Crank-Nicolson, superbee, FP64 (--config forecast: the 360x240x50x15
single forecast).  One step = the reference's hot-path sequence
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
    if kind == "advect_hz":
        # the fused second horizontal + vertical pass: q r/w and vf per
        # bin-cell; vel, sigma (horizontal), vz, vh, sigma_z, dt/kappa per
        # cell; deposit and outflows per column
        return (24 * ns + 48) * ncell + 24 * ns * st.nx * st.ny
    if kind.startswith("diffuse"):
        return (16 * ns + 24) * ncell          # q r/w; kappa, sigma, K
    return None


def step_bytes_per_bincell(st, use_diffusion):
    """B_step / ns: compulsory bytes per bin-cell-update of the whole step."""
    per_cell = st.ns * 24 + 8 * 4 + 8 * 4 + (24 if use_diffusion else 0) + st.ns * 24.0 / st.nz
    return per_cell / st.ns


def traffic_file(st):
    """The committed FETCH_SIZE / WRITE_SIZE summary of this grid."""
    name = {(1440, 720, 80, 20): "global", (360, 240, 50, 15): "forecast"}.get((st.nx, st.ny, st.nz, st.ns), "grid")
    return os.path.join(ROOT, "profiles", f"hbm_traffic_{name}_{st.nx}x{st.ny}x{st.nz}x{st.ns}.md")


def effective_knobs(kind, st, tuning):
    """The launch settings that decide which kernel a sweep kind runs and
    how (from the context's settings after tuning)."""
    if kind == "advect_hz":
        return {"adv_fused": tuning["adv_fused"]}
    if kind == "advect_x":
        return {"adv_x_rows": tuning["adv_x_rows"], "adv_fused": tuning["adv_fused"]}
    if kind in ("advect_y", "advect_z"):
        return {"nb": tuning["nb_" + kind], "adv_groups": tuning["adv_groups_" + kind[-1]],
                "adv_fused": tuning["adv_fused"]}
    d = kind[-1]
    n = {"x": st.nx, "y": st.ny, "z": st.nz}[d]
    nmax = tuning["part_max_line"] if d == "x" else tuning["part_max_line_strided"]
    if tuning["line_solver"] == 1 and 3 <= n <= nmax:
        part = {"solver": "partitioned", "part_shape": tuning["part_shape_" + d], "part_bins": tuning["part_bins"]}
        if d == "x" and tuning.get("jacobi_last_x", 0) > 0:
            part |= {"jacobi_x": tuning["jacobi_x"], "sweeps": tuning["jacobi_last_x"]}
        return part
    if d != "x" and tuning.get("jacobi_last_" + d, 0) > 0:
        return {"solver": "jacobi", "sweeps": tuning["jacobi_last_" + d], "nb": min(tuning["nb_diffuse_" + d], 4)}
    return {"solver": "dgtsv-order", "nb": tuning["nb_diffuse_" + d]} | ({"x_tile": tuning["x_tile"]} if d == "x" else {})


def measured_traffic(kind, st, tuning):
    """HBM bytes per launch of a kernel kind from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this grid (copy-calibrated), when it
    was measured with the same effective launch settings; else None."""
    path = traffic_file(st)
    if not os.path.exists(path):
        return None
    text = open(path).read()
    block = text.split("```json", 1)[1].split("```", 1)[0]
    rec = json.loads(block).get(kind)
    if rec is None or rec.get("knobs") != effective_knobs(kind, st, tuning):
        return None
    return rec["traffic_bytes"]


def cpu_model():
    """(CPU model name, logical CPUs this process may run on)."""
    name = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                name = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count()
    return name, ncpu


def cpu_baseline(make_state, use_diffusion, steps, steps_1t, sample):
    """Time the oracle (a C port of the reference path: OpenMP rows in the
    advection sweeps and serial diffusion, as in the reference) on a
    bounded sample of the workload: 1 warm-up step, then `steps` timed
    steps on every core this process may use and `steps_1t` timed steps on
    one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    st = make_state()
    model, ncpu = cpu_model()
    oracle.set_num_threads(ncpu)
    o = oracle.Oracle(st)
    o.transport_step(1, use_diffusion)  # warm-up (first touch of every array)
    t0 = time.perf_counter()
    for it in range(2, 2 + steps):
        o.step_glue()
        o.transport_step(it, use_diffusion)
    dt = time.perf_counter() - t0
    cores = oracle.num_threads()
    rate = st.ns * st.nx * st.ny * st.nz * steps / dt
    oracle.set_num_threads(1)
    t0 = time.perf_counter()
    for it in range(2 + steps, 2 + steps + steps_1t):
        o.step_glue()
        o.transport_step(it, use_diffusion)
    dt1 = time.perf_counter() - t0
    oracle.set_num_threads(ncpu)
    rate1 = st.ns * st.nx * st.ny * st.nz * steps_1t / dt1
    return {"value": rate, "unit": "bin-cell-updates/s", "cores": cores, "kind": "port",
            "cpu_model": model, "host_logical_cpus": ncpu,
            "value_1thread": rate1,
            "restates": "oracle/ash3d_oracle.c: Set_BC, AdvectHorz (advect_x/y), AdvectVert (advect_z), Set_BC(2), "
                        "DiffuseHorz/DiffuseVert (diffCN_x/y/z with dgtsv) of src/Ash3d.F90:491-510, "
                        "advection rows threaded with OpenMP and diffusion serial as in the reference",
            "sample": f"{sample}: 1 warm-up step, then {steps} timed steps ({dt:.1f} s) on {cores} threads "
                      f"and {steps_1t} timed step(s) ({dt1:.1f} s) on 1 thread"}


def timed_gradient(inv, s, w, reps=1):
    """(forward seconds, gradient seconds, gradient) after one warm-up pair."""
    inv.forward(s)
    inv.gradient(w)
    inv.T.synchronize()
    t_f = t_g = 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        inv.forward(s)
        inv.T.synchronize()
        t_f += time.perf_counter() - t0
        t0 = time.perf_counter()
        g = inv.gradient(w)
        inv.T.synchronize()
        t_g += time.perf_counter() - t0
    return t_f / reps, t_g / reps, g


def inversion_leg(pkg, line_solver, tuning, revolve_slots=40):
    """The adjoint source inversion at its full size (240x160x50, 8 bins,
    40 x 24 source table, 12 h = 552 steps, column loads every 30 min,
    squared misfit): forward run and checkpointed gradient with the
    segment plan of the default 64 GB device budget, and with binomial
    checkpointing in revolve_slots slots."""
    adj = importlib.import_module("volcano-ash3d_amd.adjoint")
    import numpy as np
    st, (iv, jv), spp, frac, dz, s_true = pkg.workloads.inversion_case()
    P, H = s_true.shape
    out = {"workload": f"source inversion 240x160x50, 8 bins, {H} heights x {P} periods of 30 min "
                       f"({P * spp} steps of {st.dt * 3600:.1f} s, 12 h), column loads every 30 min, "
                       "J = sum W (L - obs)^2 with 10 % noise", "line_solver": line_solver}
    rng = np.random.default_rng(11)
    for label, kw in (("segments", {}), ("revolve", {"schedule": "revolve", "revolve_slots": revolve_slots})):
        inv = adj.SourceInversion(st, iv, jv, n_heights=H, n_periods=P, steps_per_period=spp, frac=frac, dz=dz,
                                  obs_every=spp, use_diffusion=True, **kw)
        inv.T.tune(line_solver=line_solver, **{k: v[0] for k, v in (tuning or {}).items()})
        if label == "segments":
            truth = inv.forward(s_true)
            obs = truth * (1.0 + 0.1 * rng.standard_normal(truth.shape))
            weight = 1.0 / (0.1 * truth.max()) ** 2
            s0 = 0.6 * s_true + 0.3 * s_true.max() * rng.random(s_true.shape)
            L = inv.forward(s0)
            w = 2.0 * weight * (L - obs)
        t_f, t_g, g = timed_gradient(inv, s0, w)
        rec = {"forward_s": round(t_f, 4), "gradient_s": round(t_g, 4), "ratio": round(t_g / t_f, 3),
               "checkpoint_states": inv.n_slots,
               "state_gb": round(inv.state_bytes / 1e9, 3)}
        if label == "segments":
            rec.update(segments=inv.n_seg, segment_steps=inv.K, budget_gb=64)
        else:
            rec.update(free_slots=inv.S, recomputed_steps=inv.forward_steps)
        out[label] = rec
        inv.close()
    return out


def device_memory():
    """(free, total) bytes of the current HIP device from the HIP runtime
    that the library uses, or None."""
    import ctypes
    for name in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
        try:
            hip = ctypes.CDLL(name)
            break
        except OSError:
            hip = None
    if hip is None:
        return None
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        return None
    return free.value, total.value


def global_adjoint_leg(pkg, bins, n_steps, slots, line_solver, tuning):
    """Binomial checkpointing on the global grid (a bin subset of its full
    1440x720x80 grid): a source at 20-24 km over n_steps steps, loads every
    n_steps/4 steps, reverse sweep in `slots` free slots (0: as many as the
    free HBM holds next to the context, measured on a probe inversion after
    one forward run, less 8 GB)."""
    adj = importlib.import_module("volcano-ash3d_amd.adjoint")
    import numpy as np
    st = pkg.workloads.global_state(bins=bins)
    st.concen_pd[...] = 0.0
    dz = np.ascontiguousarray(st.dz_vec_pd[2:st.nz + 2])
    P = 4
    spp = n_steps // P
    H = min(60, st.nz)
    rng = np.random.default_rng(0)
    s = 1e6 * rng.random((P, H))
    knobs = {k: v[0] for k, v in (tuning or {}).items()}

    def make(S):
        inv = adj.SourceInversion(st, st.nx // 2, st.ny // 2, n_heights=H, n_periods=P, steps_per_period=spp,
                                  frac=np.full(st.ns, 1.0 / st.ns), dz=dz, obs_every=spp, use_diffusion=True,
                                  schedule="revolve", revolve_slots=S)
        inv.T.tune(line_solver=line_solver, **knobs)
        return inv

    free_gb = None
    if slots <= 0:
        probe = make(0)
        probe.forward(s)
        probe.T.synchronize()
        mem = device_memory()
        state_bytes = probe.state_bytes
        probe.close()
        del probe
        if mem is None:
            slots = 150
        else:
            free_gb = mem[0] / 1e9
            slots = int(max(1, min(n_steps - 1, (mem[0] - 8e9) // state_bytes)))
        log(f"[bench] global adjoint: {slots} free checkpoint slots"
            + (f" ({free_gb:.1f} GB free next to the context)" if free_gb is not None else ""))
    inv = make(slots)
    w = rng.random((len(inv.obs_steps), st.nx, st.ny))
    t_f, t_g, g = timed_gradient(inv, s, w)
    out = {"workload": f"global 1440x720x80 grid, bins {bins} of 20, {P * spp} steps, {H} source heights x {P} periods, "
                       f"binomial checkpointing in {slots} free slots of {inv.state_bytes / 1e9:.2f} GB",
           "forward_s": round(t_f, 4), "gradient_s": round(t_g, 4), "ratio": round(t_g / t_f, 3),
           "recomputed_steps": inv.forward_steps, "steps": inv.n_steps}
    if free_gb is not None:
        out["free_hbm_gb_next_to_context"] = round(free_gb, 1)
    inv.close()
    return out


def ensemble_leg(pkg, tr, members, steps, size_scale, line_solver, streams=0):
    """Ensemble throughput: one context per member on the one GPU (grids of
    120-400 x 80-300 x 40-60 cells, 10 bins, each member its own dt), all
    members stepped together, four ways: each member's own steps (eager
    launches, then replays of its captured glue + step graphs), and the
    batched ensemble step (ash3d_tr_ensemble_step: one launch per sweep
    kind and kernel variant over all members, eager, then as one graph)."""
    t0 = time.perf_counter()
    states = [pkg.workloads.ensemble_member(m, size_scale=size_scale) for m in range(members)]
    log(f"[bench] ensemble: {members} members built in {time.perf_counter() - t0:.1f} s")
    Ts = [tr.Transport(st) for st in states]
    for T in Ts:
        T.tune(line_solver=line_solver)
    work = sum(st.ns * st.nx * st.ny * st.nz for st in states)
    step_bytes = sum(step_bytes_per_bincell(st, True) * st.ns * st.nx * st.ny * st.nz for st in states)
    for T in Ts:
        T.st.concen_pd[..., 1] = T.st.concen_pd[..., 0]
        T.sync_from_host(tr.SYNC_CONCEN)
        T.step_glue()
        T.transport_step(1, True)
    for T in Ts:
        T.synchronize()
    E = None

    def run(mode, it0):
        t = time.perf_counter()
        for s_ in range(steps):
            it = it0 + s_
            if mode.startswith("batched"):
                E.step(it, use_graph=mode == "batched_graph")
                continue
            for T in Ts:
                if mode == "graph":
                    T.graph_launch(it)
                else:
                    T.step_glue()
                    T.transport_step(it, True)
        if E is not None:
            E.synchronize()
        for T in Ts:
            T.synchronize()
        return time.perf_counter() - t

    it = 2
    t = {}
    for mode in ("eager", "graph", "batched_eager", "batched_graph"):
        if mode == "graph":
            for T in Ts:
                T.graph_capture(2, True)
                T.graph_capture(3, True)
        if mode == "batched_eager":
            E = tr.Ensemble(Ts, streams=streams)
        run(mode, it)  # warm-up
        it += steps
        t[mode] = run(mode, it)
        it += steps
        log(f"[bench] ensemble {mode}: {1e3 * t[mode] / steps:.2f} ms per ensemble step")
    launches = E.info(2)
    E.close()
    sizes = [(st.nx, st.ny, st.nz) for st in states]
    for T in Ts:
        T.close()
    out = {"members": members, "steps": steps, "bin_cells_per_step": work,
           "grid_sizes": f"{min(sizes)}..{max(sizes)}, 10 bins",
           "algorithmic_bytes_per_step": step_bytes,
           "launches_per_step": {"batched": launches[0], "members_own": launches[1]}}
    for mode, v in t.items():
        out[f"{mode}_rate"] = work * steps / v
        out[f"{mode}_ms_per_ensemble_step"] = 1e3 * v / steps
        out[f"{mode}_hbm_frac"] = step_bytes * steps / v / 8.0e12
    return out


WORKLOADS = {
    "global": ("large global run: periodic lon-lat 0.25 deg 1440x720x80 (90S-90N), dz 0.3 km, 20 bins "
               "(vf 0.005-8 m/s), two jets + seeded correlated noise vanishing at the poles, Kh=2000 m2/s, "
               "Kz=10 m2/s, Crank-Nicolson, superbee, broad cloud at 20-24 km"),
    "forecast": ("single forecast: lon-lat 0.25 deg 360x240x50, dz 0.5 km, 15 bins (vf 0.01-5 m/s), "
                 "zonal jet + seeded perturbations, Kh=Kz=500 m2/s, Crank-Nicolson, superbee"),
    "cartesian": "projected 100x100x30, 5 km, uniform wind, 1 bin, no diffusion, superbee",
}


def build_state(pkg, config, bins=None):
    if config == "global":
        return pkg.workloads.global_state(bins=bins), True
    if config == "forecast":
        st = pkg.workloads.config2_state()
        if bins is not None:
            st = pkg.workloads.config2_state(ns=len(bins))
        return st, True
    return pkg.workloads.config1_state(), False


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", choices=tuple(WORKLOADS), default="global")
    ap.add_argument("--cpu-steps", type=int, default=2, help="timed CPU steps on all cores")
    ap.add_argument("--cpu-steps-1t", type=int, default=1, help="timed CPU steps on one thread")
    ap.add_argument("--cpu-bins", type=int, default=1,
                    help="bins of the CPU sample on the full grid of the global workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch autotuner")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward timing")
    ap.add_argument("--no-profile", action="store_true",
                    help="no HIP-event kernel timing (for hardware-counter runs); no per-kernel figures")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="hardware-counter runs: start with one step glue, which copies one concentration level "
                         "(the host's ts1, zeros) device to device, a copy of known size to scale FETCH_SIZE / "
                         "WRITE_SIZE by; the timed steps then start from a zero state (same memory traffic)")
    ap.add_argument("--ensemble-members", type=int, default=64,
                    help="members of the ensemble leg (0: skip); each on its own context and stream")
    ap.add_argument("--ensemble-steps", type=int, default=5)
    ap.add_argument("--ensemble-scale", type=float, default=1.0, help="horizontal size factor of the members")
    ap.add_argument("--only-ensemble", action="store_true",
                    help="run the ensemble leg alone and print its record (no bench line)")
    ap.add_argument("--ensemble-streams", type=int, default=0,
                    help="stream groups of the batched ensemble (0: the library's default)")
    ap.add_argument("--line-solver", choices=("exact", "partitioned"), default="partitioned",
                    help="theta-scheme line solves: dgtsv order (bitwise the reference) or partitioned "
                         "(lanes of a wave per line; agrees to rounding)")
    ap.add_argument("--adj-bins", type=int, default=2, help="bins of the global adjoint leg's state")
    ap.add_argument("--adj-steps", type=int, default=240, help="steps of the global adjoint leg")
    ap.add_argument("--adj-slots", type=int, default=0,
                    help="free checkpoint slots of the global adjoint leg (0: as many as the free HBM holds)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the transport step shards "
              "by grain bin or ensemble member with no per-step exchange, but no multi-GPU path is built.",
              file=sys.stderr)
        sys.exit(2)

    pkg = importlib.import_module("volcano-ash3d_amd")
    tr = importlib.import_module("volcano-ash3d_amd.transport")
    tr.load()
    if args.only_ensemble:
        # the ensemble leg alone (for traces of it); no bench line
        ens = ensemble_leg(pkg, tr, args.ensemble_members or 64, args.ensemble_steps, args.ensemble_scale,
                           args.line_solver, args.ensemble_streams)
        log("[bench] ensemble: " + json.dumps(ens))
        return
    t0 = time.perf_counter()
    st, use_diffusion = build_state(pkg, args.config)
    workload = WORKLOADS[args.config]
    log(f"[bench] state built in {time.perf_counter() - t0:.1f} s, dt={st.dt} h, "
        f"{st.nbytes() / 1e9:.2f} GB host arrays")

    t0 = time.perf_counter()
    T = tr.Transport(st)
    T.tune(line_solver=args.line_solver)
    log(f"[bench] device context created (upload) in {time.perf_counter() - t0:.1f} s")
    # ts0 holds the cloud, as after the driver glue: the first step runs
    # without it, every later one is glue + step (src/Ash3d.F90:376-378, 491-510)
    itime = 1
    if args.copy_calibration:
        T.step_glue()
        log(f"[bench] calibration copy: {8 * st.ns * (st.nx + 4) * (st.ny + 4) * (st.nz + 4)} bytes device to device")
    T.transport_step(itime, use_diffusion)
    for _ in range(args.warmup):
        itime += 1
        T.step_glue()
        T.transport_step(itime, use_diffusion)
    T.synchronize()
    log(f"[bench] warm-up {args.warmup + 1} steps done")
    tuning = None
    if not args.no_tune:
        at = importlib.import_module("volcano-ash3d_amd.autotune")
        tuning = at.autotune(T, use_diffusion, line_solver=args.line_solver)
        log("[bench] autotune: " + json.dumps({k: [v[0], {str(a): round(b, 4) for a, b in (v[1] or {}).items()}]
                                                for k, v in tuning.items()}))
        itime += 1000

    # the timed steps run without kernel timers; a separate pass below
    # brackets each kernel with HIP events for the per-kernel figures
    T.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        itime += 1
        T.step_glue()
        T.transport_step(itime, use_diffusion)
    T.synchronize()
    wall = time.perf_counter() - t0
    log(f"[bench] timed {args.steps} steps in {wall:.3f} s")
    if args.no_profile:
        T.close()
        print(json.dumps({"steps": args.steps, "ms_per_step": 1e3 * wall / args.steps, "profile": None}))
        return
    prof_steps = max(4, min(args.steps, 10))
    T.profile(True)
    for _ in range(prof_steps):
        itime += 1
        T.step_glue()
        T.transport_step(itime, use_diffusion)
    T.synchronize()
    prof = T.profile_read()
    T.profile(False)
    # the step's mass budget after the timed steps (device totals)
    aloft, deposit, outflow = T.mass_totals()
    negatives = None

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
    # the launch settings in force (library defaults, or the autotuner's)
    settings = T.tuning()
    traffic = measured_traffic(dominant, st, settings)
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "avg_ms": dk["avg_ms"], "algorithmic_bytes": algorithmic_bytes(dominant, st),
                "traffic_source": os.path.relpath(traffic_file(st), ROOT) if traffic else None}
    for k, v in kernels.items():
        t = measured_traffic(k, st, settings) if algorithmic_bytes(k, st) else None
        if t:
            v["traffic_gbps"] = round(t / (v["avg_ms"] * 1e-3) / 1e9, 1)
            v["traffic_over_algorithmic"] = round(t / algorithmic_bytes(k, st), 3)
    log("[bench] kernels: " + json.dumps(kernels))
    T.close()
    del T
    config = {"workload": workload, "nx": st.nx, "ny": st.ny, "nz": st.nz, "bins": st.ns,
              "dt_h": st.dt, "diffusion": "crank-nicolson" if use_diffusion else "off",
              "line_solver": args.line_solver if use_diffusion else None,
              "winds": "static within the timed steps (the hourly refresh and per-step interpolation of "
                       "MesoInterpolater are driver work outside the transport step)",
              "parallelism": "single gpu",
              "tuning": {k: v[0] for k, v in tuning.items()} if tuning else "defaults",
              "settings": settings}
    budget = {"aloft_kg": float(aloft.sum()), "deposit_kg": float(deposit.sum()), "outflow_kg": float(outflow.sum())}
    del st

    ensemble = None
    if args.ensemble_members > 0:
        log("[bench] ensemble leg ...")
        ensemble = ensemble_leg(pkg, tr, args.ensemble_members, args.ensemble_steps, args.ensemble_scale,
                                args.line_solver, args.ensemble_streams)
        log("[bench] ensemble: " + json.dumps(ensemble))
    adjoint = None
    if not args.no_adjoint:
        log("[bench] adjoint legs ...")
        adjoint = {"inversion": inversion_leg(pkg, args.line_solver, tuning)}
        log("[bench] inversion: " + json.dumps(adjoint["inversion"]))
        if args.config == "global":
            bins = list(range(0, 20, max(1, 20 // args.adj_bins)))[:args.adj_bins]
            adjoint["global"] = global_adjoint_leg(pkg, bins, args.adj_steps, args.adj_slots, args.line_solver,
                                                   tuning)
            log("[bench] global adjoint: " + json.dumps(adjoint["global"]))
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] timing the CPU oracle ...")
        if args.config == "global":
            cbins = list(range(0, 20, max(1, 20 // args.cpu_bins)))[:args.cpu_bins]
            sample = (f"the global workload's full 1440x720x80 grid with {len(cbins)} of its 20 bins "
                      f"(bins {cbins})")
            make = lambda: build_state(pkg, args.config, bins=cbins)[0]  # noqa: E731
        else:
            sample = f"the {args.config} workload"
            make = lambda: build_state(pkg, args.config)[0]  # noqa: E731
        cpu = cpu_baseline(make, use_diffusion, args.cpu_steps, args.cpu_steps_1t, sample)
        log(f"[bench] cpu baseline {cpu['value']:.3e} bin-cell-updates/s on {cpu['cores']} threads, "
            f"{cpu['value_1thread']:.3e} on 1")

    out = {
        "metric": METRIC, "value": rate, "unit": "bin-cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": config,
        "step_hbm": {"bytes_per_bincell": round(bpb, 2), "frac_of_peak": round(step_frac, 4)},
        "roofline": roofline,
        "kernels": kernels,
        "mass_budget": budget,
        "adjoint": adjoint,
        "ensemble": ensemble,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (rate / cpu["value"]) if cpu else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
