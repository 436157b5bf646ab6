"""Throughput of the MI355X GRMHD fluid step (zone-updates/s) on one GPU.

Workload (default): the FP64 torus step on a single-level 256x128x128 grid
(r 1.2-1000 M log-spaced, Kerr a = 0.9375, Gamma = 13/9, FM torus r_in = 6,
r_max = 12, poloidal loop normalised to min beta = 100, seeded +-4% perturbation
of u, floors 1e-5 r^-1.5 and 1e-7 r^-2.5) as 8x4x4 blocks of 32^3 interior
zones plus 3 ghost zones, all resident in HBM.  A step is one predictor +
corrector over every block (PPM + HLL + CT + Noble 2D + ghost fill, Courant dt
kept on the device).  A zone-update is one interior zone advanced one full step.

Prints ONE JSON line with the contract's fields plus `roofline` (the dominant
kernel, timed with HIP events on the step's stream) and `cpu_baseline` (the
FP64 CPU oracle on the host cores, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-updates/s on 1 MI355X: mixed/batched vs FP64 step at stated err; %HBM peak"
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # name: (block size, blocks per direction)
    "c2": ((32, 32, 32), (8, 4, 4)),
    "c3": ((32, 32, 32), (12, 6, 6)),
    "c1": ((16, 8, 8), (4, 4, 4)),
}

# Algorithmic HBM bytes per interior zone per launch (FP64, each array element
# once per stage; halo re-reads and intermediates excluded):
#   flux (one direction): 9 prims + 1 face B read, 9 fluxes written = 19 values
#   EMF: U1-3, B1-3, 6 flux B-components, 3 mass fluxes read, 3 EMFs written = 18
#   face-B update: 3 EMFs + 3 face B read, 3 written = 9
#   update + inversion: P 9 + U^n 9 + fluxes 27 + face B 3 read, P 9 written = 57 (+ 4 B flag)
#   ghost fill: (NPR+3) values per ghost cell, both ways, 24*g with g the ghost/interior ratio
STAGES = ["flux_x1", "flux_x2", "flux_x3", "emf", "pole", "ct_update", "update", "bound"]


def stage_bytes(stage, bs):
    b1, b2, b3 = bs
    g = ((b1 + 6) * (b2 + 6) * (b3 + 6) - b1 * b2 * b3) / (b1 * b2 * b3)
    per_zone = {"flux_x1": 19 * 8, "flux_x2": 19 * 8, "flux_x3": 19 * 8, "emf": 18 * 8, "pole": 0.0,
                "ct_update": 9 * 8, "update": 57 * 8 + 4, "bound": 24 * g * 8}
    return per_zone[stage]


def read_traffic(kernel, workload):
    """HBM bytes per launch measured with rocprofv3 --pmc (committed summary)."""
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

    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
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


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--no-graph", action="store_true", help="enqueue each step instead of replaying a HIP graph")
    ap.add_argument("--cpu-steps", type=int, default=2, help="steps of the CPU oracle baseline (0: skip)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one streaming copy of this many MiB (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (each GPU box holds one MI355X); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import hamr_gamma_wind_amd as H

    bs, nb = CONFIGS[args.config]
    desc = H.make_desc(bs, nb)
    nz = bs[0] * bs[1] * bs[2] * nb[0] * nb[1] * nb[2]
    ctx = H.Context(desc)
    ctx.set_grid()
    ctx.init_torus(H.torus_desc())
    ctx.boundprim(1)
    dt0 = ctx.estimate_dt()
    graph = not args.no_graph
    dt = ctx.step(args.warmup, dt0, graph=graph) if args.warmup > 0 else dt0
    ctx.sync()
    t0 = time.perf_counter()
    dt = ctx.step(args.steps, None, graph=graph)  # synchronises at the end (reads dt back)
    ctx.sync()
    el = time.perf_counter() - t0
    value = nz * args.steps / el
    _, fi = ctx.all_flags()

    # per-stage device time (HIP events on the step's stream)
    stage_ms = {s: ctx.time_stage(i, 5) for i, s in enumerate(STAGES)}
    # one step runs every stage twice (predictor, corrector)
    per_step = {s: 2 * ms for s, ms in stage_ms.items()}
    dom = max(per_step, key=per_step.get)
    bpz = stage_bytes(dom, bs)
    achieved = nz * bpz / (stage_ms[dom] * 1e-3) / 1e9 if bpz > 0 else 0.0
    traffic = read_traffic(dom, args.config)
    copy = None
    if args.copy_probe_mib > 0:
        nbytes = args.copy_probe_mib << 20
        ms = ctx.copy_probe(nbytes, 1)
        copy = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms, "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}
    ctx.close()

    cpu = None
    if args.cpu_steps > 0:
        v, cores, el_cpu = cpu_baseline(desc, args.cpu_steps)
        cpu = {"value": v, "unit": "zone-updates/s", "cores": cores, "kind": "port",
               "sample": "%d FP64 steps of the same %dx%dx%d torus state by the oracle (OpenMP, %.1f s)"
                         % (args.cpu_steps, bs[0] * nb[0], bs[1] * nb[1], bs[2] * nb[2], el_cpu)}

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
        "dtype": "f64",
        "data": "synthetic (seeded Fishbone-Moncrief torus initial state)",
        "config": {"workload": "FP64 GRMHD step, torus %dx%dx%d as %dx%dx%d blocks of %dx%dx%d + 3 ghosts, "
                               "PPM+HLL+CT+Noble2D, %s"
                               % (bs[0] * nb[0], bs[1] * nb[1], bs[2] * nb[2], nb[0], nb[1], nb[2], bs[0], bs[1],
                                  bs[2], "HIP graph replay" if graph else "stream"),
                   "zones": nz, "dt_final": dt, "inversion_failures_2d": int(fi[:, 0].sum()),
                   "floor_hits": int(fi[:, 3].sum())},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": nz * bpz, "avg_launch_ms": stage_ms[dom]},
        "stages_ms_per_step": {k: round(v, 4) for k, v in per_step.items()},
        "copy_probe": copy,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
