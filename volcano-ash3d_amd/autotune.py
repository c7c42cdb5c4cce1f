"""Launch autotuner for the transport step.

For each sweep kind it times the launch settings that do not change the
results (bins one GPU thread carries along a line, bin groups per
workgroup, row kernels, tile shapes of the partitioned line solves, the
longest line those take) with the library's HIP-event kernel timers, on the
live context.  The library's default is timed first and kept unless another
setting is faster by more than `margin` (2 % by default), so that timing
noise does not flip the choice.  The tuning steps advance the context's
state like ordinary steps; results of a step do not depend on the setting.
Choices can be cached in a JSON file keyed by grid shape and bin count.
"""
from __future__ import annotations

import json
import os

PARAMS = (("adv_fused", ("advect_x", "advect_y", "advect_z", "advect_hz")), ("adv_x_rows", "advect_x"), ("nb_advect_y", "advect_y"), ("nb_advect_z", "advect_z"), ("adv_groups_y", "advect_y"),
          ("adv_groups_z", "advect_z"), ("nb_diffuse_x", "diffuse_x"), ("x_tile", "diffuse_x"),
          ("nb_diffuse_y", "diffuse_y"), ("nb_diffuse_z", "diffuse_z"))
X_TILES = (8, 16, 32)
# streamed advection: line segments per workgroup (0) or up to 4 bin groups
# of the tile's lines per workgroup
ADV_GROUPS = (0, 4)
# advection of the step with diffusion: the fused second horizontal +
# vertical pass (1) or three split sweeps (0), timed as a whole step's
# advection
# x advection rows: one wavefront per row (1) or one workgroup per row (0)
ADV_X_ROWS = (1, 0)
# the partitioned line solves have one setting for the three directions:
# bins per workgroup (0: all bins of the tile's lines in one workgroup)
PART_BINS = (0, 8, 4)
# longest line of the partitioned solves: single-wave lines only (512) or
# lines over several waves too (2048); only tried where a line is longer
PART_MAX_LINE = (512, 2048)
# most Jacobi sweeps for strided lines beyond the partitioned solves (0: never)
JACOBI_SWEEPS = 12


def _key(st):
    return f"{st.nx}x{st.ny}x{st.nz}x{st.ns}"


def autotune(T, use_diffusion: bool, steps: int = 2, candidates=(1, 2, 3, 4, 8), itime0: int = 1000,
             cache_path: str | None = None, line_solver: str = "exact", margin: float = 0.02) -> dict:
    """line_solver: the solver the context runs ("exact" tunes the
    dgtsv-order kernels' bins per thread and x tile, "partitioned" also the
    tile shapes, bins per workgroup and longest lines of the partitioned
    solves; the dgtsv-order settings stay tuned there for the lines the
    partitioned solves leave to them)."""
    st = T.st
    if cache_path and os.path.exists(cache_path):
        cached = json.load(open(cache_path)).get(_key(st))
        if cached:
            T.tune(**{p: v for p, v in cached.items()})
            return {p: (v, None) for p, v in cached.items()}
    chosen = {}
    itime = itime0
    params = list(PARAMS)
    if line_solver == "partitioned":
        params += [("part_shape_x", "diffuse_x"), ("part_shape_y", "diffuse_y"), ("part_shape_z", "diffuse_z"),
                   ("part_bins", ("diffuse_x", "diffuse_y", "diffuse_z"))]
        if st.nx > 512:
            params.append(("part_max_line", "diffuse_x"))
        if st.ny > 512 or st.nz > 512:
            params.append(("part_max_line_strided", ("diffuse_y", "diffuse_z")))
        if st.nx >= 256:
            # the x lines that Jacobi sweeps solve off the partitioned solves
            params.append(("jacobi_x", "diffuse_x"))
        if max(st.nx, st.ny, st.nz) > 512:
            # Jacobi sweeps: strided lines left to dgtsv order, x lines as above
            params.append(("jacobi_sweeps", ("diffuse_x", "diffuse_y", "diffuse_z")))
    for param, kind in params:
        kinds = kind if isinstance(kind, tuple) else (kind,)
        if kinds[0].startswith("diffuse") and not use_diffusion:
            continue
        if param == "adv_fused" and not use_diffusion:
            continue  # the fused pass belongs to the step with diffusion
        if param in ("nb_advect_z", "adv_groups_z") and use_diffusion and T.tune_get("adv_fused"):
            continue  # the vertical sweep runs inside the fused pass
        values = (X_TILES if param == "x_tile" else PART_BINS if param == "part_bins" else
                  (1, 0) if param == "adv_fused" else
                  (JACOBI_SWEEPS, 0) if param == "jacobi_sweeps" else
                  (1, 0) if param == "jacobi_x" else
                  ADV_GROUPS if param.startswith("adv_groups") else
                  ADV_X_ROWS if param == "adv_x_rows" else
                  PART_MAX_LINE if param.startswith("part_max_line") else
                  (0, 1) if param.startswith("part_shape") else candidates)
        default = T.tune_get(param)
        values = (default,) + tuple(v for v in values if v != default)
        times = {}
        for nb in values:
            if param.startswith("nb_") and nb > 1 and nb // 2 >= st.ns:
                continue
            if param.startswith("nb_diffuse") and nb == 3:
                continue  # 3 bins per thread: the advection sweeps only
            T.tune(**{param: nb})
            T.profile(True)
            for _ in range(steps):
                itime += 1
                T.step_glue()
                T.transport_step(itime, use_diffusion)
            prof = T.profile_read()
            T.profile(False)
            if param == "adv_fused":
                # the whole advection of a step (the kinds it launches differ)
                times[nb] = sum(prof[k][0] for k in kinds) / steps
            else:
                times[nb] = sum(prof[k][0] / max(prof[k][1], 1) for k in kinds)
        best = min(times, key=times.get)
        if times[best] > (1.0 - margin) * times[default]:
            best = default
        T.tune(**{param: best})
        chosen[param] = (best, times)
    if cache_path:
        data = json.load(open(cache_path)) if os.path.exists(cache_path) else {}
        data[_key(st)] = {p: v[0] for p, v in chosen.items()}
        json.dump(data, open(cache_path, "w"), indent=1)
    return chosen
