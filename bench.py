"""Throughput of the MI355X GRMHD fluid step (zone-updates/s) on one GPU.

Workload of `value`: the mixed-precision step on the torus on a single-level
384x192x192 grid (r 1.2-1000 M log-spaced, Kerr a = 0.9375, Gamma = 13/9, FM
torus r_in = 6, r_max = 12, poloidal loop normalised to min beta = 100, seeded
+-4% perturbation of u, floors 1e-5 r^-1.5 and 1e-7 r^-2.5) as 12x6x6 blocks
of 32^3 interior zones plus 3 ghost zones, all resident in HBM: FP32
reconstruction, fluxes and EMFs; FP64 metric, face-field update, compensated
conserved update and inversions.  A step is one predictor + corrector over
every block (PPM + HLL + CT + Noble 2D + ghost fill, Courant dt kept on the
device), replayed from a HIP graph.  A zone-update is one interior zone
advanced one full step.

Beside it, on the same grid and state: the FP64 step (the reference's
arithmetic) and the speed-up; the error of the mixed state against the FP64
state after 100 steps whatever --steps is (both advancing with the FP64 run's
dt sequence; per variable max and L1, and the same figures for an FP64 run
from the FP32-rounded initial state, the flow's own response to perturbations
of FP32 size).  On the 256x128x128 grid (8x4x4 blocks of 32^3): the FP64
batched step and the reference-structured FP64 step (the C host's per-block
call sequence with a host dt read-back per block), and the FP64 step with
HLLC fluxes.  Also the multi-level torus (3 levels, 722 blocks of uneven
sizes, both arithmetics, and its local adaptive time stepping: time levels
from the block dts, cycles against single-level steps), the adjoint of one stage on the flat 128^3 box, and
the restart records of the whole grid (pack and unpack, device to device).

Launch shapes come from the autotuner (hamr_autotune) for each arithmetic.
Prints ONE JSON line with the contract's fields plus `roofline` (the mixed
step's dominant kernel: algorithmic bytes over its mean launch time, HIP
events on the library's stream) and `cpu_baseline` (the FP64 CPU oracle on the
host cores, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-updates/s on 1 MI355X: mixed/batched vs FP64 step at stated err; %HBM peak"
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # name: (block size, blocks per direction)
    "torus256": ((32, 32, 32), (8, 4, 4)),
    "torus384": ((32, 32, 32), (12, 6, 6)),
    "torus64": ((16, 8, 8), (4, 4, 4)),
}

STAGES = ["flux_x1", "flux_x2", "flux_x3", "emf", "pole", "ct_update", "update", "bound"]


KERNELS = ["flux1", "flux2", "flux3", "emf_centre", "emf_corner", "ct_update", "cons", "invert", "fallback",
           "bound1", "bound2", "bound3", "exchange"]


def kernel_bytes(kernel, bs, mixed=False):
    """Algorithmic HBM bytes per interior zone of one launch of a kernel (each
    array element it must read or write once; halo re-reads excluded).  FP64
    values are 8 B; the fluxes and EMFs of the mixed step 4 B."""
    b1, b2, b3 = bs
    w = 4 if mixed else 8
    nz = b1 * b2 * b3
    ghosts = {"bound1": 2 * 3 * (b2 + 6) * (b3 + 6), "bound2": 2 * 3 * (b1 + 6) * (b3 + 6),
              "bound3": 2 * 3 * (b1 + 6) * (b2 + 6)}
    table = {
        "flux1": 10 * 8 + 9 * w, "flux2": 10 * 8 + 9 * w, "flux3": 10 * 8 + 9 * w,
        "emf_centre": 6 * 8 + 3 * w,        # U1-3, B1-3 read; 3 centre EMFs written
        "emf_corner": 15 * w,               # 3 centre EMFs, 6 flux B-components, 3 mass fluxes read; 3 EMFs written
        "ct_update": 3 * w + 6 * 8,         # 3 EMFs + 3 face B read, 3 face B written
        # k_update (predictor form): 9 prims + 3 face B, 27 fluxes read; U^n, 9 prims and a flag written
        "cons": 12 * 8 + 27 * w + 18 * 8 + 4,
        "invert": 18 * 8 + 9 * 8 + 4,       # k_invert (two-kernel form): U' and the guess read; 9 prims, flag written
        "fallback": 4,                      # the flags (the few failed cells are negligible)
        "exchange": 0,
    }
    if kernel in ghosts:
        return ghosts[kernel] * 12 * 8 * 2 / nz
    return table[kernel]


def read_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` measured with rocprofv3
    --pmc (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel; committed
    summary profiles/<kernel>_<workload>_traffic.txt), or None."""
    path = os.path.join(ROOT, "profiles", "%s_%s_traffic.txt" % (kernel, workload))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(desc, nsteps):
    """FP64 CPU oracle on the same initial state, on the host's cores."""
    import numpy as np

    import hamr_gamma_wind_amd as H
    from tests import gridtools as GT
    from tests import oracle_ctypes as O

    # the threads OpenMP runs with: OMP_NUM_THREADS where the host sets it (a
    # GPU host's CPU share for one GPU), else every core
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(cores)
    grid = GT.oracle_grid(O.load(), desc)
    p, ps = H.desc_torus(desc, H.torus_desc())
    P = GT.blocks_to_global(desc, p)
    PS = GT.blocks_to_global(desc, ps)
    del p, ps
    grid.bound(P, PS)
    F1, d1 = grid.flux(1, P, PS)
    F2, d2 = grid.flux(2, P, PS)
    F3, d3 = grid.flux(3, P, PS)
    dt = grid.block_dt(d1, d2, d3)
    del F1, F2, F3, d1, d2, d3
    nz = int(np.prod([desc.bs[k] * desc.nb[k] for k in range(3)]))
    t0 = time.perf_counter()
    for _ in range(nsteps):
        dt, _ = grid.step(P, PS, dt)
    el = time.perf_counter() - t0
    return nz * nsteps / el, cores, el


def run_mode(H, ctx, p0, ps0, mode, args):
    """Time `steps` graph-replayed steps after `warmup` steps from (p0, ps0),
    with the launch shapes the autotuner picks for this arithmetic; return
    (seconds, final state, per-stage ms, dt, failure counters, tune table)."""
    ctx.set_precision(mode)
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)
    ctx.clear_flags()
    tune = ctx.autotune(3) if not args.no_tune else None
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)
    ctx.clear_flags()
    graph = not args.no_graph
    dt0 = ctx.estimate_dt()
    dt = ctx.step(args.warmup, dt0, graph=graph) if args.warmup > 0 else dt0
    ctx.sync()
    t0 = time.perf_counter()
    dt = ctx.step(args.steps, None, graph=graph)  # synchronises at the end (reads dt back)
    ctx.sync()
    el = time.perf_counter() - t0
    p = ctx.get(H.HAMR_P)
    if isinstance(ctx, H.AmrContext):
        fi = np.array([f[1] for f in ctx.all_flags()])
    else:
        _, fi = ctx.all_flags()
    stage_ms = {s: ctx.time_stage(i, 5) for i, s in enumerate(STAGES)}
    kern_ms = {k: ctx.time_kernel(i, 5) for i, k in enumerate(KERNELS)}
    return el, p, stage_ms, dt, fi, tune, kern_ms


def host_sequence(H, ctx, p0, ps0, nsteps):
    """The reference-structured FP64 step: the C host (hamr_host_advance in
    libhamr_host.so) calling the per-block entry points, each block's dt read
    back to the host; seconds for nsteps steps."""
    ctx.set_precision(H.HAMR_PREC_FP64)
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)
    dt = ctx.estimate_dt()
    dt = H.host_advance(ctx, dt)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        dt = H.host_advance(ctx, dt)
    return time.perf_counter() - t0


def _amr_torus():
    from hamr_gamma_wind_amd import workloads as W

    return W.AMR_TORUS


def amr_bench(H, args):
    """Both arithmetics on the multi-level block set: tuned launches, graph replay."""
    import numpy as np

    A = _amr_torus()
    desc = H.make_desc(A["bs"], A["nb"])
    amr = H.make_amr_desc(A["boxes"], seed=A["seed"])
    ctx = H.AmrContext(desc, amr)
    try:
        info = ctx.info()
        ctx.set_grid()
        ctx.init_torus(H.torus_desc())
        ctx.boundprim(1)
        p0, ps0 = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
        steps = args.amr_steps
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup = steps, args.warmup
        out = {"workload": "torus on 3 levels (ratio 2; level 1: r <= 80 M, 30-150 deg; level 2: r <= 34.6 M, "
                           "60-120 deg), %d blocks of 64x32x32, 32x16x16 and 16^3, multi-level ghost exchange"
                           % info["blocks"],
               "blocks": info["blocks"], "zones": info["zones"],
               "exchange_entries": sum(info[k] for k in ("ex1", "ex2", "ex3", "pro1", "pro2", "pro3")),
               "exchange_entries_after_update": sum(info[k] for k in ("ex1", "ex2", "ex3", "pro1", "pro2", "pro3"))
               - info["push3"],
               "steps": steps}
        res = {}
        for name, mode in (("fp64", H.HAMR_PREC_FP64), ("mixed", H.HAMR_PREC_MIXED)):
            el, p, stage_ms, dt, fi, tune, kern_ms = run_mode(H, ctx, p0, ps0, mode, sub)
            res[name] = p
            out[name] = {"value": info["zones"] * steps / el, "ms_per_step": el / steps * 1e3,
                         "stages_ms_per_step": {k: round(2 * v, 4) for k, v in stage_ms.items()},
                         "kernels_ms_per_launch": {k: round(v, 4) for k, v in kern_ms.items()},
                         "inversion_failures_2d": int(fi[:, 0].sum()),
                         "tuned": {k: [v["threads"], v["grid"], v["xcd_order"], v["second_form"]] for k, v in tune.items()}
                         if tune else None}
        out["speedup_mixed_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
        err = {"steps": args.warmup + steps}
        for name, k in (("rho", 0), ("B1", 5), ("B2", 6), ("B3", 7), ("u", 1)):
            d = max(float(np.abs(a[k] - b[k])[3:-3, 3:-3, 3:-3].max()) for a, b in zip(res["mixed"], res["fp64"]))
            m = max(float(np.abs(b[k])[3:-3, 3:-3, 3:-3].max()) for b in res["fp64"])
            err[name + "_max"] = d / m if m > 0 else d
        out["error_vs_fp64"] = err
        out["local_time_stepping"] = las_bench(H, ctx, p0, ps0, out)
        return out
    finally:
        ctx.close()


def las_bench(H, ctx, p0, ps0, amr_out, maxtl=4, cycles=10):
    """Local adaptive time stepping on the multi-level set: time levels from
    the block dts of one step (set_timelevel's rule, maxtl 4), then cycles of
    2 maxtl sub-iterations replayed from a captured graph; each cycle advances
    every block by maxtl dt, which the single-level step needs maxtl steps
    for."""
    import numpy as np

    zones = np.array([int(np.prod(r[4:7])) for r in ctx.layout], dtype=np.int64)
    out = {"maxtl": maxtl, "cycles": cycles}
    for name, mode in (("fp64", H.HAMR_PREC_FP64), ("mixed", H.HAMR_PREC_MIXED)):
        ctx.set_precision(mode)
        ctx.set(H.HAMR_P, p0)
        ctx.set(H.HAMR_PS, ps0)
        dt = ctx.step(1)
        tl = ctx.timelevels_from_dt(maxtl)
        dt = ctx.step_las(maxtl, 1, dt=dt, graph=True)  # capture + warm-up
        t0 = time.perf_counter()
        ctx.step_las(maxtl, cycles, graph=True)
        el = time.perf_counter() - t0
        ms_cycle = el / cycles * 1e3
        updates = int((zones * (maxtl // tl)).sum())
        out[name] = {"ms_per_cycle": ms_cycle,
                     "blocks_per_time_level": {str(k): int((tl == k).sum()) for k in (1, 2, 4, 8, 16) if (tl == k).any()},
                     "zone_updates_per_cycle": updates,
                     "zone_updates_single_level": int(zones.sum()) * maxtl,
                     "zone_updates_per_s": updates * cycles / el,
                     "speedup_vs_single_level_steps": maxtl * amr_out[name]["ms_per_step"] / ms_cycle}
    ctx.set_precision(H.HAMR_PREC_FP64)
    return out


def adjoint_bench(H, n=128, reps=5):
    """Forward stage and its adjoint on the flat periodic n^3 box of one block
    (the adjoint workload), device-resident; ms per call (wall time between
    device synchronisations)."""
    import torch

    from hamr_gamma_wind_amd import workloads as W

    gam = 13.0 / 9.0
    desc = H.make_desc((n, n, n), (1, 1, 1), gam=gam, minkowski=1, rho_floor=0.0, u_floor=0.0)
    ctx = H.Context(desc)
    try:
        ctx.set_grid()
        prims, faces = W.minkowski_box((n, n, n), gam=gam, seed=1)
        p, ps = W.to_block(prims, faces)
        ctx.set(H.HAMR_P, p)
        ctx.set(H.HAMR_PS, ps)
        _, dx = H.desc_coords(desc)
        Dt = 0.3 * min(dx)
        dev = torch.device("cuda:0")
        D = tuple(k + 6 for k in (n, n, n))
        w = torch.tensor(np.random.default_rng(2).normal(size=(1, 9) + D), device=dev)
        gp = torch.empty((1, 9) + D, dtype=torch.float64, device=dev)
        gps = torch.empty((1, 3) + D, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        out = {}
        for name, fn in (("forward_ms", lambda: ctx.stage_forward(Dt)),
                         ("adjoint_ms", lambda: ctx.stage_adjoint_dev(Dt, w.data_ptr(), gp.data_ptr(), gps.data_ptr()))):
            fn()
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            ctx.sync()
            out[name] = (time.perf_counter() - t0) / reps * 1e3
        out["zones"] = n ** 3
        out["adjoint_zones_per_s"] = n ** 3 / (out["adjoint_ms"] * 1e-3)
        out["workload"] = ("flat periodic %d^3 box (one block), random state: rho lognormal, u = rho lognormal, "
                           "Lorentz factor <= 5, max b^2/rho = 10; L = sum w . U', FP64" % n)
        return out
    finally:
        ctx.close()


def hllc_bench(H, ctx, p0, ps0, nz, args, steps=20, mixed=False):
    """The step with the HLLC Riemann solver (hamr_set_riemann) on the same
    state, FP64 or mixed (FP32 fluxes): tuned launches, graph replay, 3 warm-up
    and `steps` timed steps, and the flux kernels' ms per launch."""
    ctx.set_precision(H.HAMR_PREC_MIXED if mixed else H.HAMR_PREC_FP64)
    ctx.set_riemann(H.HAMR_RIEMANN_HLLC)
    try:
        ctx.set(H.HAMR_P, p0)
        ctx.set(H.HAMR_PS, ps0)
        if not args.no_tune:
            ctx.autotune(3)
        ctx.set(H.HAMR_P, p0)
        ctx.set(H.HAMR_PS, ps0)
        ctx.clear_flags()
        graph = not args.no_graph
        dt = ctx.step(3, ctx.estimate_dt(), graph=graph)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(steps, None, graph=graph)
        ctx.sync()
        el = time.perf_counter() - t0
        kern = {k: round(ctx.time_kernel(i, 5), 4) for i, k in enumerate(KERNELS[:3])}
    finally:
        ctx.set_riemann(H.HAMR_RIEMANN_HLL)
        ctx.set_precision(H.HAMR_PREC_FP64)
    return {"value": nz * steps / el, "unit": "zone-updates/s", "ms_per_step": el / steps * 1e3, "steps": steps,
            "flux_ms_per_launch": kern,
            "what": ("mixed step (FP32 fluxes)" if mixed else "FP64 step") +
                    " with HLLC fluxes in the face's orthonormal frame (the reference's HLLC = 1)"}


def rdump_bench(H, ctx, bs, reps=5):
    """Restart records of every block (rdump_block_write / rdump_block_read,
    restart.c:52, :178) packed from / unpacked into the resident state, and the
    dump records (dump_block, dump.c:276), device to device in one launch each:
    ms per call (wall time between device synchronisations, launch included)
    and the algorithmic rate.  Write: each ghost-inclusive cell reads 12
    doubles and writes 13; read: each cell of [NG, NG+BS]^3 reads 12 and
    writes 12; dump: each interior cell reads 9 doubles and writes 9 floats
    (the column metric aside)."""
    import torch

    off = ctx.rdump_offsets()
    buf = torch.empty(int(off[-1]), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    nb = ctx.nblk
    D = [b + 6 for b in bs]
    cells_w = nb * D[0] * D[1] * D[2]
    cells_r = nb * (bs[0] + 1) * (bs[1] + 1) * (bs[2] + 1)
    doff = ctx.dump_offsets()
    dbuf = torch.empty(int(doff[-1]), dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    cells_d = nb * bs[0] * bs[1] * bs[2]
    out = {"record_bytes": int(off[-1]) * 8, "dump_bytes": int(doff[-1]) * 4}
    for name, fn, nbytes in (("write", lambda: ctx.rdump_write_dev(buf.data_ptr()), cells_w * (12 + 13) * 8),
                             ("read", lambda: ctx.rdump_read_dev(buf.data_ptr()), cells_r * 24 * 8),
                             ("dump", lambda: ctx.dump_write_dev(dbuf.data_ptr()), cells_d * (9 * 8 + 9 * 4))):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ms = (time.perf_counter() - t0) / reps * 1e3
        out[name + "_ms"] = ms
        out[name + "_algorithmic_GBps"] = nbytes / (ms * 1e-3) / 1e9
    return out


def error_run(H, desc, p0, ps0, nsteps=100):
    """The mixed step's error against the FP64 step after nsteps steps from the
    same state (both with the FP64 run's dt sequence), and the same figures for
    an FP64 run from the FP32-rounded state: per variable the max error
    relative to the variable's largest magnitude and the L1 error (sum |d| /
    sum |ref|) over the interior; for velocity and field the max error relative
    to the vector's largest component."""
    ctxs = {}
    try:
        for name, mode, q, qs in (("fp64", H.HAMR_PREC_FP64, p0, ps0), ("mixed", H.HAMR_PREC_MIXED, p0, ps0),
                                  ("fp32_state", H.HAMR_PREC_FP64, p0.astype(np.float32).astype(np.float64),
                                   ps0.astype(np.float32).astype(np.float64))):
            c = H.Context(desc)
            ctxs[name] = c
            c.set_grid()
            c.set(H.HAMR_P, q)
            c.set(H.HAMR_PS, qs)
            c.boundprim(1)
            c.set_precision(mode)
        dt = ctxs["fp64"].estimate_dt()
        for _ in range(nsteps):
            ndt = ctxs["fp64"].step(1, dt)
            ctxs["mixed"].step(1, dt)
            ctxs["fp32_state"].step(1, dt)
            dt = ndt
        I = slice(3, -3)
        ref = ctxs["fp64"].get(H.HAMR_P)[:, :, I, I, I]
        out = {"steps": nsteps}
        for run in ("mixed", "fp32_state"):
            a = ctxs[run].get(H.HAMR_P)[:, :, I, I, I]
            e = {}
            for name, k in (("rho", 0), ("u", 1), ("B1", 5), ("B2", 6), ("B3", 7)):
                d = np.abs(a[:, k] - ref[:, k])
                e[name + "_max"] = float(d.max() / np.abs(ref[:, k]).max())
                e[name + "_l1"] = float(d.sum() / np.abs(ref[:, k]).sum())
            for name, k0, k1 in (("velocity", 2, 5), ("field", 5, 8)):
                d = np.abs(a[:, k0:k1] - ref[:, k0:k1])
                e[name + "_max_of_vector"] = float(d.max() / np.abs(ref[:, k0:k1]).max())
                e[name + "_l1"] = float(d.sum() / np.abs(ref[:, k0:k1]).sum())
            del a
            if run == "mixed":
                out.update(e)
            else:
                out["fp64_from_fp32_rounded_state"] = e
        return out
    finally:
        for c in ctxs.values():
            c.close()


def torus_context(H, config):
    bs, nb = CONFIGS[config]
    desc = H.make_desc(bs, nb)
    ctx = H.Context(desc)
    ctx.set_grid()
    ctx.init_torus(H.torus_desc())
    ctx.boundprim(1)
    return desc, ctx, bs, nb, bs[0] * bs[1] * bs[2] * nb[0] * nb[1] * nb[2]


def grid_name(bs, nb):
    return "%dx%dx%d" % (bs[0] * nb[0], bs[1] * nb[1], bs[2] * nb[2])


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=97)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="torus384", choices=sorted(CONFIGS),
                    help="grid of the headline (mixed) and FP64 batched steps")
    ap.add_argument("--fp64-config", default="torus256", choices=sorted(CONFIGS),
                    help="grid of the FP64 batched, per-block and HLLC steps beside them")
    ap.add_argument("--error-steps", type=int, default=100, help="steps of the mixed-vs-FP64 error run (0: skip)")
    ap.add_argument("--no-graph", action="store_true", help="enqueue each step instead of replaying a HIP graph")
    ap.add_argument("--cpu-steps", type=int, default=3, help="steps of the CPU oracle baseline (0: skip)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one streaming copy of this many MiB (counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--only", choices=["fp64", "mixed"], default=None, help="run one arithmetic only")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint measurement (128^3 flat box)")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes instead of the autotuner's")
    ap.add_argument("--no-amr", action="store_true", help="skip the multi-level block set")
    ap.add_argument("--no-hllc", action="store_true", help="skip the steps with HLLC fluxes (FP64 and mixed)")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the headline grid (no second grid, AMR, adjoint, HLLC, restart records)")
    ap.add_argument("--amr-steps", type=int, default=40)
    ap.add_argument("--amr-only", action="store_true",
                    help="measure only the multi-level block set (prints its JSON object; for profiling)")
    ap.add_argument("--host-steps", type=int, default=3,
                    help="steps of the reference-structured per-block FP64 sequence (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the step is measured on one MI355X (one process, one GPU); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.no_extra:
        args.no_adjoint = args.no_amr = args.no_hllc = True

    import hamr_gamma_wind_amd as H

    if args.amr_only:
        print(json.dumps({"amr": amr_bench(H, args)}), flush=True)
        return
    desc, ctx, bs, nb, nz = torus_context(H, args.config)
    p0 = ctx.get(H.HAMR_P)
    ps0 = ctx.get(H.HAMR_PS)

    modes = {"fp64": H.HAMR_PREC_FP64, "mixed": H.HAMR_PREC_MIXED}
    if args.only:
        modes = {args.only: modes[args.only]}
    res = {name: run_mode(H, ctx, p0, ps0, m, args) for name, m in modes.items()}
    head = "mixed" if "mixed" in res else "fp64"
    el, p_head, stage_ms, dt, fi, tune, kern_ms = res[head]
    value = nz * args.steps / el
    mixed = head == "mixed"

    # one step runs every stage and kernel twice (predictor, corrector); the
    # roofline is that of the kernel with the largest share of the step
    per_step = {s: 2 * ms for s, ms in stage_ms.items()}
    dom = max(kern_ms, key=kern_ms.get)
    bpz = kernel_bytes(dom, bs, mixed)
    achieved = nz * bpz / (kern_ms[dom] * 1e-3) / 1e9 if bpz > 0 else 0.0
    traffic = read_traffic(dom, "%s_torus%s" % ("mixed" if mixed else "fp64", grid_name(bs, nb)))
    copy = None
    if args.copy_probe_mib > 0:
        nbytes = args.copy_probe_mib << 20
        ms = ctx.copy_probe(nbytes, 1)
        copy = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms, "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}
    rdump = rdump_bench(H, ctx, bs) if not args.no_extra else None
    ctx.close()
    del p_head, res[head]

    fp64 = None
    if "fp64" in res:
        el64 = res["fp64"][0]
        fp64 = {"value": nz * args.steps / el64, "ms_per_step": el64 / args.steps * 1e3,
                "stages_ms_per_step": {k: round(2 * v, 4) for k, v in res["fp64"][2].items()},
                "kernels_ms_per_launch": {k: round(v, 4) for k, v in res["fp64"][6].items()},
                "workload": "FP64 step on the same grid and state"}
    res.clear()
    err = error_run(H, desc, p0, ps0, args.error_steps) if (args.error_steps > 0 and not args.only) else None
    del p0, ps0

    second = None
    if not args.no_extra:
        # the FP64 GPU step on the second grid: batched (graph, tuned), the
        # reference-structured per-block sequence and HLLC fluxes
        desc2, ctx2, bs2, nb2, nz2 = torus_context(H, args.fp64_config)
        try:
            q0, qs0 = ctx2.get(H.HAMR_P), ctx2.get(H.HAMR_PS)
            el2, _, st2, _, _, tune2, k2 = run_mode(H, ctx2, q0, qs0, H.HAMR_PREC_FP64, args)
            second = {"workload": "FP64 GRMHD step, torus %s as %dx%dx%d blocks of %dx%dx%d + 3 ghosts"
                                  % (grid_name(bs2, nb2), nb2[0], nb2[1], nb2[2], bs2[0], bs2[1], bs2[2]),
                      "zones": nz2, "value": nz2 * args.steps / el2, "ms_per_step": el2 / args.steps * 1e3,
                      "stages_ms_per_step": {k: round(2 * v, 4) for k, v in st2.items()},
                      "kernels_ms_per_launch": {k: round(v, 4) for k, v in k2.items()}}
            if args.host_steps > 0:
                el_host = host_sequence(H, ctx2, q0, qs0, args.host_steps)
                second["per_block_host"] = {
                    "value": nz2 * args.host_steps / el_host, "ms_per_step": el_host / args.host_steps * 1e3,
                    "steps": args.host_steps,
                    "what": "FP64, the C host's per-block call sequence of advance_GPU (libhamr_host.so), "
                            "each block's Courant dt read back to the host"}
                second["speedup_batched_vs_per_block_host"] = second["value"] / second["per_block_host"]["value"]
            if not args.no_hllc:
                second["hllc"] = hllc_bench(H, ctx2, q0, qs0, nz2, args)
                second["hllc_mixed"] = hllc_bench(H, ctx2, q0, qs0, nz2, args, mixed=True)
            del q0, qs0
        finally:
            ctx2.close()

    adj = None
    if not args.no_adjoint:
        adj = adjoint_bench(H)
    amr_res = None
    if not args.no_amr:
        amr_res = amr_bench(H, args)

    cpu = None
    if args.cpu_steps > 0:
        v, cores, el_cpu = cpu_baseline(desc, args.cpu_steps)
        cpu = {"value": v, "unit": "zone-updates/s", "cores": cores, "kind": "port",
               "sample": "%d FP64 steps of the same torus %s state by the oracle (OpenMP, %d threads, %.1f s; "
                         "the host reports %d cores)" % (args.cpu_steps, grid_name(bs, nb), cores, el_cpu,
                                                        os.cpu_count() or 0)}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (mixed)" if mixed else "f64",
        "data": "synthetic (seeded Fishbone-Moncrief torus initial state)",
        "config": {"workload": "%s GRMHD step, torus %s as %dx%dx%d blocks of %dx%dx%d + 3 ghosts, "
                               "PPM+HLL+CT+Noble2D, %s"
                               % ("mixed-precision" if mixed else "FP64", grid_name(bs, nb), nb[0], nb[1], nb[2],
                                  bs[0], bs[1], bs[2], "HIP graph replay" if not args.no_graph else "stream"),
                   "zones": nz, "dt_final": dt, "inversion_failures_2d": int(fi[:, 0].sum()),
                   "floor_hits": int(fi[:, 3].sum())},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": nz * bpz, "avg_launch_ms": kern_ms[dom]},
        "kernels_ms_per_launch": {k: round(v, 4) for k, v in kern_ms.items()},
        "kernels_hbm_frac": {k: round(nz * kernel_bytes(k, bs, mixed) / (v * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                             for k, v in kern_ms.items() if v > 0 and kernel_bytes(k, bs, mixed) > 4},
        "stages_ms_per_step": {k: round(v, 4) for k, v in per_step.items()},
        "fp64": fp64,
        "speedup_vs_fp64": (value / fp64["value"]) if fp64 else None,
        "error_vs_fp64": err,
        "tuned": {k: [v["threads"], v["grid"], v["xcd_order"], v["second_form"]] for k, v in tune.items()} if tune else None,
        "fp64_second_grid": second,
        "amr": amr_res,
        "adjoint": adj,
        "copy_probe": copy,
        "rdump": rdump,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
