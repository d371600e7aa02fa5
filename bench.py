#!/usr/bin/env python3
"""Benchmark: RK-stage throughput of the periodic Hall problem on one MI355X.

Workload (default): nonlinear Hall induction equation on the periodic cube
[0, 2pi)^3 at 128^3, 6th-order central SBP + order-matched high-order dissipation,
u = 0, Hall coefficient 1 (rho = 1), B0 = (0,0,1) + divergence-free random Fourier
modes |k| <= 4 (seed 1), CarpenterKennedy2N54 (low-storage RK(5,4)), FP64.
One "step" is one full RK step (5 stages).  The timed region runs exactly --steps
steps through the C ABI (ieq_step) with the fields resident in HBM.

Printed: one JSON line with value = Mcell-updates/s per RK stage
(N^3 * stages / time), the roofline of the fused RK stage (96 algorithmic bytes per
cell in FP64: read B and dB, write B' and dB) against the 8 TB/s HBM peak, and the
CPU baseline (the reference-schedule C oracle on the host cores, bounded sample).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--integrator", default="CarpenterKennedy2N54")
    ap.add_argument("--cpu-steps", type=int, default=15, help="RK steps of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (calibrates the HBM byte counters)")
    return ap.parse_args()


def copy_probe(mib):
    """A plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE give the counter scale."""
    import torch

    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": 5}))


def main():
    a = parse()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (the workload is single-GPU; multi-GPU domain "
              "decomposition is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from inductioneq_amd.problems import hall_dt, random_modes_field
    from inductioneq_amd.program import Program
    from inductioneq_amd.schedules import STAGES_PER_STEP
    from inductioneq_amd.settings import ProgramConfig

    n = a.n
    L = 2 * math.pi
    real = "double" if a.precision == "fp64" else "float"
    cfg = ProgramConfig(real=real, opset="6thOrder", nx=n, ny=n, nz=n, xmin=0, xmax=L, ymin=0, ymax=L, zmin=0, zmax=L,
                        periodic=True, hall=True, dissipation="HIGH_ORDER", testcase="uniform_flow", w_size=256)
    B0 = random_modes_field(n, seed=1)
    cfg.dt = hall_dt(B0, (cfg.dx, cfg.dy, cfg.dz), 6, c_hall=1.0, cfl=0.5)
    prog = Program(cfg)
    prog.use_current_stream()
    b1 = prog.vector()
    b1[:, : cfg.num_nodes] = torch.as_tensor(B0.T, dtype=prog.torch_dtype, device="cuda")
    b2 = prog.vector()
    J, u, rho, tb = prog.vector(), prog.vector(), prog.scalar(1.0), prog.time_buffer()
    args = [b1, b2, J, u, rho, tb]
    stages = STAGES_PER_STEP[a.integrator]

    print("bench: warmup %d steps (n=%d, %s, %s)" % (a.warmup, n, a.precision, a.integrator), file=sys.stderr, flush=True)
    prog.step(a.integrator, a.warmup, args)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    prog.step(a.integrator, a.steps, args)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    sec = ms * 1e-3
    ncell = n ** 3
    value = ncell * stages * a.steps / sec / 1e6
    stage_s = sec / (stages * a.steps)
    rb = 8 if real == "double" else 4
    alg_bytes = 4 * 3 * rb * ncell  # read B, dB; write B', dB (3 components each)
    achieved = alg_bytes / stage_s / 1e9
    finite = bool(torch.isfinite(b1[:, : cfg.num_nodes]).all().item())
    print("bench: %.3f ms for %d steps, %.1f us/stage, finite=%s" % (ms, a.steps, stage_s * 1e6, finite), file=sys.stderr,
          flush=True)

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(cfg, B0, a.integrator, a.cpu_steps, stages)

    out = {
        "metric": "Mcell-updates/s per RK stage",
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms / a.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if real == "double" else "f32",
        "data": "synthetic",
        "config": {"workload": "hall_periodic_%d^3_6th_HOD_%s" % (n, a.integrator), "grid": [n, n, n],
                   "order": 6, "dissipation": "high_order", "u": [0, 0, 0], "hall_coefficient": 1.0, "dt": cfg.dt,
                   "stages_per_step": stages, "seed": 1},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "one fused RK stage (k_emf + k_stage), %d B/cell algorithmic" % (4 * 3 * rb),
                     "us_per_stage": round(stage_s * 1e6, 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(cfg, B0, integrator, nsteps, stages):
    """Reference-schedule C oracle (AoS REAL4, per-node loops, curl + stage kernels) on host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle

    from inductioneq_amd.problems import to_aos4
    from inductioneq_amd.schedules import rk_schedule

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    orc = Oracle(cfg, nthreads=threads)
    names, nf = rk_schedule(integrator, True, False)
    b = [to_aos4(B0, cfg.num_nodes_pad, orc.dtype)] + [orc.vec() for _ in range(nf - 1)]
    args = b + [orc.vec(), orc.vec(), orc.scal(1.0), orc.time()]
    print("bench: cpu baseline %d step(s) on %d threads" % (nsteps, threads), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    orc.run(names * nsteps, args)
    dt = time.perf_counter() - t0
    v = cfg.num_nodes * stages * nsteps / dt / 1e6
    return {"value": round(v, 3), "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": "%d RK step(s) (%d stages) of the same %dx%dx%d workload, %.2f s" % (nsteps, stages * nsteps, cfg.nx,
                                                                                        cfg.ny, cfg.nz, dt)}


if __name__ == "__main__":
    main()
