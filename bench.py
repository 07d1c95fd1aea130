#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the MG-PICOLA COLA step on one MI355X.

Workload: f(R) Hu-Sawicki (fR0 = 1e-5, n = 1) with scale-dependent growth and
chameleon screening, P(k) every step (512 linear bins), 512^3 particles on a
1024^3 mesh in a 1024 Mpc/h box, 30 COLA steps z = 19 -> 0, synthetic Gaussian
2LPT initial conditions (Eisenstein-Hu no-wiggle P(k), seed 5001).

A "step" is one iteration of the C host's time loop (libmgpicola, the
reference's src/main.c:405-610): PtoMesh + P(k), the f(R) fifth force with
screening, Forces, MtoParticles, the two scale-dependent LPT fields, Kick and
Drift, all on the GPU through libmgpu.  W warm-up steps, then K timed steps
between device synchronisations.  One further step runs with per-kernel HIP
event timing to give the roofline of the dominant kernel.

The CPU baseline is the FP64 oracle (a C restatement of the reference's step)
timed on the host cores for one step of the same path on a bounded sample.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mg-picola-public_amd"))

METRIC = "particle-steps/s at 512^3/Nmesh 1024; learned vs 3x-mesh P(k) err, k<1"
HBM_PEAK_GBS = 8000.0


def gpu_bench(steps, warmup, workload):
    import mgpicola as M
    import workloads as W

    out = tempfile.mkdtemp(prefix="mgpicola_bench_")
    model, vals, nsteps = W.values_for(workload, out)
    if warmup + steps + 1 > nsteps:
        raise SystemExit(f"warmup + steps + 1 must be <= {nsteps} (the run has {nsteps} steps)")
    with open(vals["OutputRedshiftFile"], "w") as f:
        f.write(f"0, {nsteps}\n")
    pf = os.path.join(out, "params.txt")
    M.write_paramfile(pf, vals, model=model)
    hp = M.read_params(pf, model=model)
    t0 = time.time()
    sim = M.Sim(hp)
    t_tables = time.time() - t0
    t0 = time.time()
    sim.ic()
    sim.gpu.sync()
    t_ic = time.time() - t0
    print(f"[bench] tables {t_tables:.1f} s, initial conditions {t_ic:.1f} s", file=sys.stderr, flush=True)
    for _ in range(warmup):
        sim.advance()
    sim.gpu.sync()
    t0 = time.perf_counter()
    for i in range(steps):
        sim.advance()
        print(f"[bench] step {i + 1}/{steps} a = {sim.a:.4f}", file=sys.stderr, flush=True)
    sim.gpu.sync()
    elapsed = time.perf_counter() - t0
    # one more step with per-kernel event timing
    M.enable_kernel_timing(sim.gpu, True)
    sim.advance()
    sim.gpu.sync()
    kt = M.kernel_timing(sim.gpu)
    M.enable_kernel_timing(sim.gpu, False)
    sim.close()
    npart = hp.Nsample ** 3
    return dict(elapsed=elapsed, npart=npart, nmesh=hp.Nmesh, box=hp.Box, kernels=kt, t_tables=t_tables, t_ic=t_ic)


def cpu_baseline(ns=256, nmesh=512, box=512.0):
    """One COLA step of the same f(R) + screening + scale-dependent path in the
    FP64 oracle: deposit, r2c, P(k), fifth force (2 FFTs), Forces (3 c2r),
    gather, 4 LPT fields (3 c2r each), Kick, Drift."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle as O
    from helpers import lattice_particles

    O.lib()
    pos = lattice_particles(ns, box, amp=2.0, seed=7, clustered=True)
    vel = np.zeros_like(pos)
    c = O.cfg(nmesh, ns, box, 0.3, model=O.FOFR, mg_active=1, include_screening=1)
    mg = O.MG(a=0.5, coupling=1.0 / 3.0, mass2=7e5, phi_crit=5e-6, dgp_fac=0, rsmooth=1, geff=1)
    cd = O.new_mesh(nmesh)
    g = np.ones(3 * (nmesh // 2) ** 2 + 1)
    d1 = np.zeros_like(pos)
    t0 = time.perf_counter()
    d = O.deposit(c, pos)
    dx = d.copy()
    O.fft_r2c_3d(nmesh, d)
    O.pofk(c, d, 512, 0, 0.0, float(nmesh), 1)
    phik = O.fifth_force(c, mg, d, dx)
    f = O.forces(c, d, phik)
    disp, sD = O.gather(c, pos, f)
    for _ in range(4):
        O.lpt_field(c, cd, g)
    sV = O.kick(vel, disp, d1, d1, 0, 0, 1, 0.3, 0.1, 0.5, 1, sD)
    O.drift(pos, vel, d1, d1, 0, 0, 1, box, 0.1, 1, sV)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=ns ** 3 / dt, unit="particle-steps/s", cores=cores, kind="port",
                sample=f"one f(R)+screening+scale-dependent COLA step of the FP64 oracle at {ns}^3 particles on a "
                       f"{nmesh}^3 mesh ({dt:.1f} s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="fofr_sd_512_1024")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    r = gpu_bench(args.steps, args.warmup, args.workload)
    value = r["npart"] * args.steps / r["elapsed"]
    kt = r["kernels"]
    dom = max(kt, key=lambda k: kt[k][1])
    n, ms, by = kt[dom]
    achieved = (by / n) / (ms / n * 1e-3) / 1e9 if n and ms > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
            "launches_per_step": n, "ms_per_launch": round(ms / max(n, 1), 4)}
    per_kernel = {k: {"launches": v[0], "ms": round(v[1], 3),
                      "GBps": round(v[2] / (v[1] * 1e-3) / 1e9, 1) if v[1] > 0 else None}
                  for k, v in kt.items() if v[0]}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline()
        cpu["value"] = round(cpu["value"], 1)
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(r["elapsed"] / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "particles": r["npart"], "nmesh": r["nmesh"], "box": r["box"],
                   "model": "f(R) Hu-Sawicki fR0=1e-5 n=1, scale-dependent, screening, P(k) every step"},
        "roofline": roof, "cpu_baseline": cpu, "kernels_one_step": per_kernel,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
