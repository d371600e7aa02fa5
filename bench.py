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
    ap.add_argument("--no-secondary", action="store_true",
                    help="hall128: skip the other measurements (256^3 FP64, 128^3 FP32, ensemble, adjoint, CG)")
    ap.add_argument("--workload", choices=["hall128", "ensemble64", "adjoint96", "rotation40", "cg192"],
                    default="hall128",
                    help="hall128: one periodic Hall solve (the headline metric); ensemble64: 64 batched FP32 solves "
                         "against the same solves run one by one in FP64")
    ap.add_argument("--members", type=int, default=64, help="ensemble64: number of members")
    ap.add_argument("--traffic-from", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="only summarise rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE counter_collection.csv files of this "
                         "benchmark and of --copy-probe into HBM bytes per launch (JSON on stdout)")
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


def _counter_means(path, counter):
    import csv
    import re

    acc = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        name = r["Kernel_Name"].replace("void ", "", 1).replace("ieq::(anonymous namespace)::", "")
        m = re.match(r"([\w:]+(?:<[^()]*>)?)", name)
        k = m.group(1) if m else name
        acc.setdefault(k, []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def traffic_summary(fetch_csv, write_csv, copyf_csv, copyw_csv, copy_mib=512):
    """HBM-side bytes per dispatch: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what the same
    counters read for a plain device copy of copy_mib MiB (bench.py --copy-probe)."""
    f, w = _counter_means(fetch_csv, "FETCH_SIZE"), _counter_means(write_csv, "WRITE_SIZE")
    cf = max(_counter_means(copyf_csv, "FETCH_SIZE").values())
    cw = max(_counter_means(copyw_csv, "WRITE_SIZE").values())
    kib = copy_mib * 1024.0
    ff, wf = kib / cf, kib / cw
    kern = {k: {"fetch_kib_raw": f[k], "write_kib_raw": w.get(k, 0.0), "bytes": (f[k] * ff + w.get(k, 0.0) * wf) * 1024}
            for k in f if k.startswith("k_")}
    return {"fetch_factor": ff, "write_factor": wf, "kernels": kern}


def stored_traffic(name, real):
    """Traffic per RK stage of REAL type `real` (the k_hall_stage dispatch, or k_emf + k_stage of the two-kernel
    form) from a committed summary, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    d = json.load(open(path))
    kern = {k: v for k, v in d["kernels"].items() if ("<%s," % real) in k}
    one = [v["bytes"] for k, v in kern.items() if k.startswith("k_hall_stage")]
    total = sum(one) if one else sum(v["bytes"] for k, v in kern.items() if k.startswith(("k_emf", "k_stage")))
    return {"bytes_per_stage": total, "source": "profiles/" + name} if total else None


def main():
    a = parse()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.traffic_from:
        print(json.dumps(traffic_summary(*a.traffic_from), indent=1))
        return
    if a.gpus != 1:
        print("bench.py: measures one GPU only (the workload is single-GPU; multi-GPU domain "
              "decomposition is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if a.workload == "ensemble64":
        bench_ensemble(a)
        return
    if a.workload == "adjoint96":
        bench_adjoint(a)
        return
    if a.workload == "rotation40":
        bench_rotation(a)
        return
    if a.workload == "cg192":
        bench_cg(a)
        return
    import torch

    from inductioneq_amd.schedules import STAGES_PER_STEP

    n = a.n
    real = "double" if a.precision == "fp64" else "float"
    stages = STAGES_PER_STEP[a.integrator]
    h = run_hall(n, real, a.integrator, a.steps, a.warmup)
    cfg, B0 = h["cfg"], h["B0"]
    rb = 8 if real == "double" else 4
    err64 = None
    if real == "float":
        # the FP32 state against the FP64 path (equal to the CPU oracle to 1e-12 per step) after the same steps
        err64 = fp32_vs_fp64(cfg, B0, a.integrator, a.warmup + a.steps, h["b1"])
        print("bench: fp32 vs fp64 after %d steps: max rel err %.3e" % (a.warmup + a.steps, err64), file=sys.stderr,
              flush=True)
    del h["b1"]
    torch.cuda.empty_cache()

    secondary = {}
    if not a.no_secondary:
        secondary = run_secondary(a)
    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(cfg, B0, a.integrator, a.cpu_steps, stages)
    traffic = None
    if n == 128 and a.integrator == "CarpenterKennedy2N54":
        traffic = stored_traffic("hall128_%s_traffic.json" % a.precision, real)

    out = {
        "metric": "Mcell-updates/s per RK stage",
        "value": h["value"],
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if real == "double" else "f32",
        "data": "synthetic",
        "config": {"workload": "hall_periodic_%d^3_6th_HOD_%s" % (n, a.integrator), "grid": [n, n, n],
                   "order": 6, "dissipation": "high_order", "u": [0, 0, 0], "hall_coefficient": 1.0, "dt": cfg.dt,
                   "stages_per_step": stages, "seed": 1},
        "roofline": {"bound": "hbm", "achieved": h["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": h["frac"],
                     "traffic": round(traffic["bytes_per_stage"]) if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "kernel": "one RK stage = one k_hall_stage launch, %d B/cell algorithmic (read B, dB; write B', dB)"
                               % (4 * 3 * rb),
                     "us_per_stage": h["us_per_stage"]},
        "cpu_baseline": cpu,
    }
    if err64 is not None:
        out["max_rel_err_vs_fp64"] = err64
    out.update(secondary)
    print(json.dumps(out), flush=True)


def run_hall(n, real, integrator, steps, warmup):
    """The periodic Hall workload at n^3 (6th order + HO dissipation, u = 0, rho = 1, B0 = (0,0,1) + random modes,
    seed 1): `steps` timed RK steps through ieq_step_uniform (inputs resident in HBM, no host synchronisation
    inside the timed region)."""
    import torch

    from inductioneq_amd.problems import hall_dt, random_modes_field
    from inductioneq_amd.program import Program
    from inductioneq_amd.schedules import STAGES_PER_STEP
    from inductioneq_amd.settings import ProgramConfig

    L = 2 * math.pi
    cfg = ProgramConfig(real=real, opset="6thOrder", nx=n, ny=n, nz=n, xmin=0, xmax=L, ymin=0, ymax=L, zmin=0, zmax=L,
                        periodic=True, hall=True, dissipation="HIGH_ORDER", testcase="uniform_flow", w_size=256)
    B0 = random_modes_field(n, seed=1)
    cfg.dt = hall_dt(B0, (cfg.dx, cfg.dy, cfg.dz), 6, c_hall=1.0, cfl=0.5)
    prog = Program(cfg)
    prog.use_current_stream()
    b1 = prog.vector()
    b1[:, : cfg.num_nodes] = torch.as_tensor(B0.T, dtype=prog.torch_dtype, device="cuda")
    args = [b1, prog.vector(), prog.vector(), prog.vector(), prog.scalar(1.0), prog.time_buffer()]
    uni = (0.0, 0.0, 0.0, 1.0)
    stages = STAGES_PER_STEP[integrator]
    print("bench: warmup %d steps (n=%d, %s, %s)" % (warmup, n, real, integrator), file=sys.stderr, flush=True)
    prog.step(integrator, warmup, args, uniform=uni)
    torch.cuda.synchronize()
    assert prog.last_step_fused, "the fused kernels did not run"
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    prog.step(integrator, steps, args, uniform=uni)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    sec = ms * 1e-3
    ncell = n ** 3
    stage_s = sec / (stages * steps)
    rb = 8 if real == "double" else 4
    alg_bytes = 4 * 3 * rb * ncell  # read B, dB; write B', dB (3 components each)
    achieved = alg_bytes / stage_s / 1e9
    finite = bool(torch.isfinite(b1[:, : cfg.num_nodes]).all().item())
    print("bench: n=%d %s: %.3f ms for %d steps, %.2f us/stage, finite=%s" % (n, real, ms, steps, stage_s * 1e6, finite),
          file=sys.stderr, flush=True)
    if not finite:
        raise RuntimeError("non-finite state after the timed steps")
    return {"cfg": cfg, "B0": B0, "b1": b1, "value": round(ncell * stages * steps / sec / 1e6, 3),
            "ms_per_step": round(ms / steps, 4), "us_per_stage": round(stage_s * 1e6, 2),
            "achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4)}


def fp32_vs_fp64(cfg, B0, integrator, nsteps, b1_32):
    """Max relative difference of an FP32 state to the FP64 path after the same steps."""
    from dataclasses import replace

    import torch

    from inductioneq_amd.program import Program

    p64 = Program(replace(cfg, real="double"))
    p64.use_current_stream()
    c1 = p64.vector()
    c1[:, : cfg.num_nodes] = torch.as_tensor(B0.T, dtype=torch.float64, device="cuda")
    p64.step(integrator, nsteps, [c1, p64.vector(), p64.vector(), p64.vector(), p64.scalar(1.0), p64.time_buffer()],
             uniform=(0.0, 0.0, 0.0, 1.0))
    torch.cuda.synchronize()
    ref = c1[:, : cfg.num_nodes]
    return float(((b1_32[:, : cfg.num_nodes].double() - ref).abs().max() / ref.abs().max()).item())


def run_secondary(a):
    """The secondary measurements, each on its own configuration: the Hall stage at 256^3 FP64 (beyond the
    256 MB cache) and at 128^3 FP32, the batched ensemble against per-instance FP64 solves, the adjoint gradient
    cost at 96^3 x 200 steps with 15 binomial snapshots, and CG divergence cleaning at 192^3."""
    import torch

    out = {}
    h = run_hall(256, "double", "CarpenterKennedy2N54", 10, 2)
    out["hall256_fp64"] = {k: h[k] for k in ("value", "us_per_stage", "achieved", "frac")}
    out["hall256_fp64"]["unit"] = "Mcell-updates/s"
    t256 = stored_traffic("hall256_fp64_traffic.json", "double")
    if t256:  # HBM-side bytes per stage (the 768 MiB working set does not fit the 256 MB cache)
        out["hall256_fp64"]["traffic"] = round(t256["bytes_per_stage"])
        out["hall256_fp64"]["traffic_source"] = t256["source"]
    del h
    torch.cuda.empty_cache()
    h = run_hall(128, "float", "CarpenterKennedy2N54", 40, 5)
    out["hall128_fp32"] = {k: h[k] for k in ("value", "us_per_stage", "achieved", "frac")}
    out["hall128_fp32"]["max_rel_err_vs_fp64"] = fp32_vs_fp64(h["cfg"], h["B0"], "CarpenterKennedy2N54", 45, h["b1"])
    del h
    torch.cuda.empty_cache()
    e = run_ensemble(a, 200)
    out["ensemble64"] = {"value": e["value"], "unit": e["unit"], "speedup_vs_per_instance_fp64": e["speedup_vs_per_instance_fp64"],
                         "max_rel_err_vs_fp64": e["max_rel_err_vs_fp64"], "roofline_frac": e["roofline"]["frac"],
                         "traffic": e["roofline"]["traffic"],
                         "per_instance_fp64": e["per_instance_fp64"], "launch": e["config"]["launch"]}
    torch.cuda.empty_cache()
    d = run_adjoint(a, 200, 96)
    out["adjoint96"] = {"cost_x_forward_15_snapshots": d["binomial_checkpointing"]["cost_x_forward"],
                        "forward_steps_15_snapshots": d["binomial_checkpointing"]["forward_steps"],
                        "lower_bound_x_forward_15_snapshots": d["lower_bound_x_forward_15_snapshots"],
                        "cost_x_forward_all_stages_stored": d["all_stages_stored"]["cost_x_forward"],
                        "bitwise_same_gradient": d["all_stages_stored"]["bitwise_same_gradient_as_binomial"],
                        "fd_rel_err_fp64": {k: v["rel_err"] for k, v in d["fd_check_fp64"].items()},
                        "fp32_vs_fp64_grad_rel_err": d["fp32_vs_fp64_grad_rel_err"]}
    torch.cuda.empty_cache()
    c = run_cg(a, 192)
    out["cg192"] = {"iterations_per_s": c["value"], "iterations": c["fp32"]["iterations"],
                    "seconds_fp32_stencil_fp64_dots": c["fp32"]["seconds"], "seconds_fp64": c["fp64"]["seconds"],
                    "divergence_norm_before": c["fp64"]["divergence_norm_initial"],
                    "divergence_norm_after": c["fp64"]["divergence_norm_final"],
                    "roofline_frac": c["roofline"]["frac"]}
    return out


def bench_ensemble(a):
    print(json.dumps(run_ensemble(a, a.steps if a.steps != 40 else 200)), flush=True)


def run_ensemble(a, steps):
    """64 periodic Hall solves (n in {32,48,64,96}, c_H in U(0.1, 2), B0 random modes with the member's seed,
    SSPRK33, dt from the Hall CFL 0.4), 6th order + HO dissipation: FP32 ensemble (one persistent launch per
    fused kernel per stage, hipGraph replay) vs the same solves one by one in FP64 (fused single-solve path)."""
    import numpy as np
    import torch

    from inductioneq_amd.ensemble import Ensemble, MemberSpec
    from inductioneq_amd.problems import hall_dt, random_modes_field
    from inductioneq_amd.program import Program
    from inductioneq_amd.settings import ProgramConfig

    rng = np.random.default_rng(2024)
    sizes = rng.choice([32, 48, 64, 96], size=a.members)
    c_h = rng.uniform(0.1, 2.0, size=a.members)
    L = 2 * math.pi
    integ = "SSPRK33"

    def cfg_for(n, real, dt=1e-3):
        return ProgramConfig(real=real, opset="6thOrder", nx=n, ny=n, nz=n, xmin=0, xmax=L, ymin=0, ymax=L, zmin=0,
                             zmax=L, periodic=True, hall=True, dissipation="HIGH_ORDER", testcase="uniform_flow",
                             w_size=256, dt=dt)

    B0s, specs = [], []
    for i, (n, c) in enumerate(zip(sizes, c_h)):
        n = int(n)
        B0 = random_modes_field(n, seed=i)
        h = L / n
        dt = hall_dt(B0, (h, h, h), 6, c_hall=float(c), cfl=0.4)
        B0s.append(B0)
        specs.append(MemberSpec(n, n, n, h, h, h, dt, rho=1.0 / float(c)))
    ncell = int(sum(int(n) ** 3 for n in sizes))
    print("bench: ensemble of %d members, %d cells, %d steps" % (a.members, ncell, steps), file=sys.stderr, flush=True)

    prog32 = Program(cfg_for(32, "float"))
    ens = Ensemble(prog32, specs, integ)

    def load():  # B0 is node-major (x fastest): the member layout of b1 directly
        for i, B0 in enumerate(B0s):
            ens.fields[i][0][:, : B0.shape[0]] = torch.as_tensor(B0.T, dtype=torch.float32, device="cuda")

    load()
    ens.step(1)  # tunes the launch shape (fields restored), then one step
    load()
    torch.cuda.synchronize()
    sec32 = ens.step(steps)
    cfg_e = ens.config()
    print("bench: fp32 ensemble %.3f s %s" % (sec32, cfg_e), file=sys.stderr, flush=True)

    sec64, err = 0.0, 0.0
    for i, B0 in enumerate(B0s):
        n = specs[i].nx
        prog = Program(cfg_for(n, "double", specs[i].dt))
        prog.use_current_stream()
        b1 = prog.vector()
        b1[:, : n ** 3] = torch.as_tensor(B0.T, dtype=torch.float64, device="cuda")
        args = [b1, prog.vector(), prog.vector(), prog.vector(), prog.vector(), prog.scalar(specs[i].rho),
                prog.time_buffer()]
        torch.cuda.synchronize()
        sec64 += prog.step(integ, steps, args, timed=True, uniform=(0.0, 0.0, 0.0, specs[i].rho))
        ref = b1[:, : n ** 3].double()
        got = ens.fields[i][0][:, : n ** 3].double()
        err = max(err, float(((got - ref).abs().max() / ref.abs().max()).item()))
        del prog, b1, args
    stages = 3 * steps
    value = ncell * stages / sec32 / 1e6
    rb = 4
    alg = 64 // 2 * ncell  # SSPRK33 average 8/3 vector fields per stage: 32 B/cell FP32
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ensemble64_fp32_traffic.json")
    if a.members == 64 and os.path.exists(tpath):
        t = json.load(open(tpath))
        if t["launch"]["tile"] == cfg_e["tile"] and t["launch"]["z_chunk"] == cfg_e["z_chunk"]:
            # HBM-side bytes per stage of the profiled launch form: SSPRK33 runs stage 1 in update mode 3 and
            # stages 2-3 in mode 2
            by = {int(k.rsplit(",", 1)[1].strip(" >")): v["bytes"] for k, v in t["kernels"].items()}
            traffic = round((by[3] + 2 * by[2]) / 3)
    out = {
        "metric": "Mcell-updates/s per RK stage",
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": round(sec32 * 1e3 / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "ensemble_%d_hall_periodic_6th_HOD_SSPRK33" % a.members, "members": a.members,
                   "cells": ncell, "sizes": [int(v) for v in sizes], "launch": cfg_e},
        "roofline": {"bound": "hbm", "achieved": round(alg / (sec32 / stages) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg / (sec32 / stages) / 1e9 / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": "profiles/ensemble64_fp32_traffic.json" if traffic else None,
                     "kernel": "one batched RK stage (%s), 32 B/cell algorithmic"
                               % ("k_hall_stage_batched" if cfg_e["tile"].startswith("hall") else "k_emf_batched + k_stage_batched")},
        "per_instance_fp64": {"seconds": round(sec64, 4), "value": round(ncell * stages / sec64 / 1e6, 3)},
        "speedup_vs_per_instance_fp64": round(sec64 / sec32, 3),
        "max_rel_err_vs_fp64": err,
        "cpu_baseline": None,
    }
    return out


def bench_adjoint(a):
    print(json.dumps(run_adjoint(a, a.steps if a.steps != 40 else 200, 96 if a.n == 128 else a.n)), flush=True)


def run_adjoint(a, steps, n):
    """Gradient of 0.5 h^3 |B(T) - B_target|^2 after 200 SSPRK33 steps of the periodic Hall problem at 96^3
    (6th order + HO dissipation, u = (0.2, -0.1, 0.15), c_H = 1), w.r.t. B0 (FP32 state), u and c_H; B_target
    is the same solve from a different seed.  `value` is the gradient time with 15 binomial snapshots (checkpointed
    forward, then the backward sweep with its recomputation) over the plain forward time, reported beside the
    3x target and the lower bound the checkpoint schedule itself sets (forward steps / steps + 2/3 for the two
    stage inputs each reversed step recomputes + 1 for three transposed stages at the cost of forward ones).
    Also reported: the same gradient with every stage input kept (601 fields, 6.4 GB in FP32 at 96^3, no
    recomputation, bitwise the same result), and central finite differences (eps = 1e-4, FP64) of the FP64
    gradient."""
    import numpy as np
    import torch

    from inductioneq_amd.adjoint import AdjointSolver
    from inductioneq_amd.problems import hall_dt, random_modes_field
    from inductioneq_amd.program import Program
    from inductioneq_amd.settings import ProgramConfig

    snaps = 15
    L = 2 * math.pi
    u, ch = [0.2, -0.1, 0.15], 1.0
    B0 = random_modes_field(n, seed=1)
    Bt = random_modes_field(n, seed=2)
    h = L / n
    dt = hall_dt(B0, (h, h, h), 6, c_hall=ch, cfl=0.4)
    res = {}
    for real in ("float", "double"):
        cfg = ProgramConfig(real=real, opset="6thOrder", nx=n, ny=n, nz=n, xmin=0, xmax=L, ymin=0, ymax=L, zmin=0, zmax=L,
                            periodic=True, hall=True, dissipation="HIGH_ORDER", testcase="uniform_flow", w_size=256,
                            dt=dt)
        prog = Program(cfg)
        prog.use_current_stream()
        N3 = n ** 3
        b0, tgt0 = prog.vector(), prog.vector()
        b0[:, :N3] = torch.as_tensor(B0.T, dtype=prog.torch_dtype)
        tgt0[:, :N3] = torch.as_tensor(Bt.T, dtype=prog.torch_dtype)
        solver = AdjointSolver(prog, steps, snaps)
        tgt = solver.forward(tgt0, u, ch)  # B_target: the different-seed forward run
        # plain forward (fused ieq_step, same kernels) for the cost ratio
        b1 = b0.clone()
        uf = prog.vector()
        for k in range(3):
            uf[k, :N3] = u[k]
        args = [b1, prog.vector(), prog.vector(), prog.vector(), uf, prog.scalar(1.0 / ch), prog.time_buffer()]
        prog.step("SSPRK33", 2, args)  # warm-up
        t_plain = 1e30
        for _ in range(3):  # best of three plain solves
            b1.copy_(b0)
            t_plain = min(t_plain, prog.step("SSPRK33", steps, args, timed=True))
        bT = solver.forward(b0, u, ch)
        t_fwd = solver.last_forward_seconds
        # the cost ratio is taken against the faster of the two forward solves the library offers (ieq_step,
        # and the adjoint solver's own forward sweep, which also stores the snapshots)
        t_plain = min(t_plain, t_fwd)
        lam = (bT - tgt) * (h ** 3)
        g, gc, gu = solver.backward(lam)
        t_bwd = solver.last_backward_seconds
        nf, na = solver.counts()
        d = (bT - tgt)[:, :N3].double()
        loss = 0.5 * h ** 3 * float((d * d).sum())
        res[real] = {"t_plain": t_plain, "t_fwd": t_fwd, "t_bwd": t_bwd, "grad": g[:, :N3].double().cpu(), "gc": gc,
                     "gu": gu, "loss": loss, "counts": (nf, na)}
        print("bench: adjoint %s plain %.3f s, forward %.3f s, backward %.3f s, steps fwd %d adj %d" %
              (real, t_plain, t_fwd, t_bwd, nf, na), file=sys.stderr, flush=True)
        if real == "double":  # central finite differences of the FP64 gradient
            eps = 1e-4

            def L_of(bb, uu, cc):
                d = (solver.forward(bb, uu, cc) - tgt)[:, :N3].double()
                return 0.5 * h ** 3 * float((d * d).sum())

            v = prog.vector()
            v[:, :N3] = torch.as_tensor(np.random.default_rng(7).standard_normal((3, N3)), dtype=torch.float64)
            fd = {"b0": (L_of(b0 + eps * v, u, ch) - L_of(b0 - eps * v, u, ch)) / (2 * eps)}
            ad = {"b0": float((g[:, :N3] * v[:, :N3]).sum())}
            for k in range(3):
                up, um = list(u), list(u)
                up[k] += eps
                um[k] -= eps
                fd["u%d" % k] = (L_of(b0, up, ch) - L_of(b0, um, ch)) / (2 * eps)
                ad["u%d" % k] = gu[k]
            fd["c_hall"] = (L_of(b0, u, ch + eps) - L_of(b0, u, ch - eps)) / (2 * eps)
            ad["c_hall"] = gc
            res[real]["fd"] = {k: {"fd": fd[k], "adjoint": ad[k], "rel_err": abs(fd[k] - ad[k]) / abs(fd[k])}
                               for k in fd}
        if real == "float":  # the same gradient with every stage input kept (no recomputation)
            full = AdjointSolver(prog, steps, steps)
            full.forward(tgt0, u, ch)
            bT2 = full.forward(b0, u, ch)
            g2, gc2, gu2 = full.backward((bT2 - tgt) * (h ** 3))
            res["float_stored"] = {"t_fwd": full.last_forward_seconds, "t_bwd": full.last_backward_seconds,
                                   "counts": full.counts(),
                                   "same_gradient": bool(torch.equal(g2, g)) and gc2 == gc and gu2 == gu}
            del full
        del solver, prog
    r32, r64 = res["float"], res["double"]
    rs = res["float_stored"]
    gerr = float((r32["grad"] - r64["grad"]).abs().max() / r64["grad"].abs().max())
    ratio_ckpt = (r32["t_fwd"] + r32["t_bwd"]) / r32["t_plain"]
    ratio = (rs["t_fwd"] + rs["t_bwd"]) / r32["t_plain"]
    stages = 3 * steps
    floor = r32["counts"][0] / steps + 2.0 / 3.0 + 1.0
    out = {
        "metric": "adjoint gradient cost / forward solve",
        "value": round(ratio_ckpt, 3),
        "unit": "x forward",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 2,
        "ms_per_step": round((r32["t_fwd"] + r32["t_bwd"]) * 1e3 / steps, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "adjoint_hall_periodic_%d^3_6th_HOD_SSPRK33" % n, "grid": [n, n, n], "steps": steps,
                   "snapshots": snaps, "u": u, "c_hall": ch, "dt": dt, "seeds": [1, 2]},
        "target_x_forward": 3.0,
        "lower_bound_x_forward_15_snapshots": round(floor, 3),
        "forward_Mcell_per_s": round(n ** 3 * stages / r32["t_plain"] / 1e6, 3),
        "timings_s": {k: {"plain_forward": round(v["t_plain"], 5), "checkpointed_forward": round(v["t_fwd"], 5),
                          "backward": round(v["t_bwd"], 5), "forward_steps": v["counts"][0],
                          "adjoint_steps": v["counts"][1]} for k, v in res.items() if k in ("float", "double")},
        "all_stages_stored": {"cost_x_forward": round(ratio, 3),
                              "checkpointed_forward": round(rs["t_fwd"], 5), "backward": round(rs["t_bwd"], 5),
                              "forward_steps": rs["counts"][0], "adjoint_steps": rs["counts"][1],
                              "bitwise_same_gradient_as_binomial": rs["same_gradient"]},
        "binomial_checkpointing": {"snapshots": snaps, "cost_x_forward": round(ratio_ckpt, 3),
                                   "forward_steps": r32["counts"][0], "adjoint_steps": r32["counts"][1]},
        "fp32_vs_fp64_grad_rel_err": gerr,
        "fd_check_fp64": r64["fd"],
        "cpu_baseline": None,
    }
    return out


def bench_rotation(a):
    """Linear induction equation on [-1,1]^3 at 40^3, 4th-order SBP with SAT boundaries, u = (-y, x, 0),
    B0 = curl(0, 0, exp(-20((x-1/2)^2 + y^2 + z^2))), SSPRK33 at CFL 0.5 to t = pi/2 (reference-schedule
    kernels, FP64).  Reports throughput, the error against the rotated analytic field, and the full-run
    difference to the CPU oracle (which is also the CPU baseline)."""
    import torch

    from inductioneq_amd.harness import HipExecutor, compute_numerical_solution, initialize
    from inductioneq_amd.settings import ProgramConfig

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    n = 40 if a.n == 128 else a.n
    T = math.pi / 2
    cfg = ProgramConfig(real="double", opset="4thOrder", nx=n, ny=n, nz=n, xmin=-1, xmax=1, ymin=-1, ymax=1, zmin=-1,
                        zmax=1, periodic=False, hall=False, dissipation=None, testcase="gaussian_rotation_3D",
                        form_uibj="PRODUCT", form_source="CENTRAL", form_ujbi="CENTRAL", w_size=64)
    init = initialize(cfg, HipExecutor, 0.5, T)
    cfg.dt = init["dt"]
    ex = HipExecutor(cfg)
    ex.prog.use_current_stream()
    r = compute_numerical_solution(cfg, ex, "SSPRK33", init["b"], init["u"], init["rho"], 2, 2 * cfg.dt, fused=True)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    res = compute_numerical_solution(cfg, ex, "SSPRK33", init["b"], init["u"], init["rho"], init["num_steps"], T,
                                     fused=True)
    e1.record()
    torch.cuda.synchronize()
    stages = 3 * init["num_steps"]
    sec = res["runtime"]
    value = n ** 3 * stages / sec / 1e6
    cpu, parity = None, None
    if not a.no_cpu_baseline:
        from pyoracle import Oracle

        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        orc = Oracle(cfg, nthreads=threads)
        t0 = time.perf_counter()
        ro = compute_numerical_solution(cfg, orc, "SSPRK33", init["b"], init["u"], init["rho"], init["num_steps"], T)
        dt_cpu = time.perf_counter() - t0
        cpu = {"value": round(n ** 3 * stages / ro["runtime"] / 1e6, 3), "unit": "Mcell-updates/s", "cores": threads,
               "kind": "port", "sample": "the whole run (%d steps, %d stages), %.2f s" % (init["num_steps"], stages,
                                                                                         dt_cpu)}
        import numpy as np

        d = np.abs(res["field_b"] - ro["field_b"]).max() / np.abs(ro["field_b"]).max()
        parity = {"max_rel_diff_final_field": float(d), "oracle_rel_err": ro["rel_err"]}
    out = {
        "metric": "Mcell-updates/s per RK stage",
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": init["num_steps"],
        "warmup": 2,
        "ms_per_step": round(sec * 1e3 / init["num_steps"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "rotation_%d^3_4th_SAT_SSPRK33" % n, "dt": cfg.dt, "final_time": T},
        "rel_err_vs_analytic": res["rel_err"],
        "divergence_norm": res["divergence_norm"],
        "oracle_parity": parity,
        # 88 B/cell: the SSPRK33 stages read b1, u / b1, b2, u / b1, b3, u and write one field each (average over
        # the three stages, FP64); at 40^3 the 6 MB working set is cache-resident, so frac is not an HBM figure
        "roofline": {"bound": "hbm", "achieved": round(88 * n ** 3 * stages / sec / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(88 * n ** 3 * stages / sec / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "reference-schedule kernels (one launch per SSPRK33 stage), 88 B/cell "
                                                "algorithmic", "us_per_stage": round(sec * 1e6 / stages, 2),
                     "cache_resident": 88 * n ** 3 < 256e6},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_cg(a):
    print(json.dumps(run_cg(a, 192 if a.n == 128 else a.n)), flush=True)


def run_cg(a, n):
    """Divergence cleaning at 192^3 on [-1,1]^3 (4th-order SBP, WS_D0 Laplacian): B = the rotation40 field plus
    grad psi (Fourier |k| <= 6, seed 3), CG to relative residual 1e-10; FP32 stencils with FP64 dot products
    against all-FP64.  Reports iterations, time per iteration and the divergence norm before and after."""
    import numpy as np
    import torch

    from inductioneq_amd.problems import gaussian_rotation_field, gradient_potential
    from inductioneq_amd.program import Program
    from inductioneq_amd.settings import ProgramConfig

    B = gaussian_rotation_field(n) + gradient_potential(n, seed=3)
    res = {}
    for real in ("double", "float"):
        cfg = ProgramConfig(real=real, opset="4thOrder", nx=n, ny=n, nz=n, xmin=-1, xmax=1, ymin=-1, ymax=1, zmin=-1,
                            zmax=1, periodic=False, testcase="rotation_3D", laplace="WIDE_STENCIL_DIRICHLET", w_size=256)
        prog = Program(cfg)
        prog.use_current_stream()
        b = prog.vector()
        b[:, : n ** 3] = torch.as_tensor(B.T, dtype=prog.torch_dtype)
        warm = prog.vector()
        warm.copy_(b)
        prog.clean_div_cg(warm, rel_tolerance=1e-10, max_iterations=20)
        r = prog.clean_div_cg(b, rel_tolerance=1e-10, max_iterations=20000)
        r["b"] = b[:, : n ** 3].double().cpu()
        res[real] = r
        print("bench: cg %s %d iterations, %.3f s, div %.3e -> %.3e" % (real, r["iterations"], r["seconds"],
              r["divergence_norm_initial"], r["divergence_norm_final"]), file=sys.stderr, flush=True)
        del prog
    r64, r32 = res["double"], res["float"]
    it_s = r32["seconds"] / max(r32["iterations"], 1)
    cells = (n - 2) ** 3
    alg = 40 * cells  # FP32 fields: k_cg_lap reads u, r, writes u', q; k_cg_update reads phi, u', r, q, writes phi, r
    diff = float((r32["b"] - r64["b"]).abs().max() / r64["b"].abs().max())
    out = {
        "metric": "CG iterations/s",
        "value": round(1.0 / it_s, 1),
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": r32["iterations"],
        "warmup": 20,
        "ms_per_step": round(it_s * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 stencil / f64 dots",
        "data": "synthetic",
        "config": {"workload": "divclean_%d^3_4th_WS_D0_rel1e-10" % n, "seed": 3},
        "fp32": {k: r32[k] for k in ("iterations", "seconds", "divergence_norm_initial", "divergence_norm_final",
                                     "residual_norm")},
        "fp64": {k: r64[k] for k in ("iterations", "seconds", "divergence_norm_initial", "divergence_norm_final",
                                     "residual_norm")},
        "fp32_vs_fp64_cleaned_field_max_rel_diff": diff,
        "roofline": {"bound": "hbm", "achieved": round(alg / it_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(alg / it_s / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "one CG iteration (k_cg_lap + k_cg_update, FP64 dot partials summed inside both), 40 B/cell algorithmic FP32"},
        "cpu_baseline": None,
    }
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(cfg, B0, integrator, nsteps, stages):
    """Reference-schedule C oracle (AoS REAL4, per-node loops, curl + stage kernels) on the host cores: `nsteps`
    steps on every thread OpenMP is given, and one step on a single core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle

    from inductioneq_amd.problems import to_aos4
    from inductioneq_amd.schedules import rk_schedule

    names, nf = rk_schedule(integrator, True, False)

    def timed(threads, k):
        orc = Oracle(cfg, nthreads=threads)
        b = [to_aos4(B0, cfg.num_nodes_pad, orc.dtype)] + [orc.vec() for _ in range(nf - 1)]
        args = b + [orc.vec(), orc.vec(), orc.scal(1.0), orc.time()]
        print("bench: cpu baseline %d step(s) on %d thread(s)" % (k, threads), file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        orc.run(names * k, args)
        return time.perf_counter() - t0

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    dt = timed(threads, nsteps)
    dt1 = timed(1, 1)
    v = cfg.num_nodes * stages * nsteps / dt / 1e6
    return {"value": round(v, 3), "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": "%d RK step(s) (%d stages) of the same %dx%dx%d workload, %.2f s" % (nsteps, stages * nsteps, cfg.nx,
                                                                                        cfg.ny, cfg.nz, dt),
            "cpu_model": cpu_model(), "logical_cpus_visible": os.cpu_count(),
            "single_core": {"value": round(cfg.num_nodes * stages / dt1 / 1e6, 3), "cores": 1,
                            "sample": "1 RK step (%d stages), %.2f s" % (stages, dt1)}}


if __name__ == "__main__":
    main()
