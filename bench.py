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
between device synchronisations (defaults 1 + 29: the whole schedule, whose
late clustered steps are the slow ones).  A further run of the whole schedule
with per-kernel HIP event timing gives every kernel class's device time per
step and the roofline of the dominant kernel.

The CPU baseline is the FP64 oracle (a C restatement of the reference's step)
timed on the host cores for one step of the same path on a bounded sample.

The "learned" object is the metric's second part: the P(k) error at z = 0 for
k < 1 h/Mpc of a coarse run (Nmesh = Nsample = 512) with the learned
Green's-function multiplier against the plain run on the 3x finer mesh (1536),
beside the plain coarse run's error and the time per step of each run.  The
"ensemble" object times a slice of the batched nDGP ensemble run one
realization after the other and concurrently (one host thread and stream each).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mg-picola-public_amd"))

METRIC = "particle-steps/s at 512^3/Nmesh 1024; learned vs 3x-mesh P(k) err, k<1"
HBM_PEAK_GBS = 8000.0

# kernel-timing class -> the kernel (rocprofv3 name) it times, for the f(R)
# scale-dependent workload's mesh size N and lattice size Ns
KERNEL_SYMBOL = {
    "deposit": "k_dep_tiles<false, 16>", "mean": "k_sum3", "dep_sort": "k_dep_bin + k_dep_scatter",
    "fft_r2c": "k_fft_z_r2c<{N}>", "fft_c2r": "k_fft_z_c2r<{N}, *>", "fft_col_fwd": "k_fft_col<{N}, 16, -1, 2>",
    "fft_col_inv": "k_fft_col<{N}, 16, 1, 2>", "fft_green": "k_col_green<{N}, 16, 0>",
    "fft_pot": "k_col_green<{N}, 16, 1>", "fft_dual": "k_col_dual<{N}, 16>", "fft_screen": "k_z_screen<{N}, 0>",
    "fft_epi": "k_col_epi<{N}, 16>", "fft_pk": "k_col_pofk<{N}, 16>", "fold": "k_lpt_fold_gs<{m}, 2>",
    "gather": "k_gather_tile<16>", "kick": "k_kick", "drift": "k_drift", "pofk": "k_pofk",
    "lpt": "k_fft_col<{Ns}, ...> + k_fft_z_c2r<{Ns}, false>",
    "fft_xdens": "k_col_fused<{N}, 16, 0, 1>", "fft_xforce": "k_col_fused<{N}, 8, 1, 0>",
}


def gpu_bench(steps, warmup, workload):
    """Time `steps` COLA steps after `warmup` untimed ones.  The workload's run
    has a fixed schedule (30 steps z = 19 -> 0); when more steps are asked for
    than one run holds, further runs start from fresh initial conditions (the
    initial conditions themselves are never timed).  Then one more full run
    with per-kernel HIP-event timing gives each kernel class's device time per
    step, averaged over the whole schedule."""
    import mgpicola as M
    import workloads as W

    out = tempfile.mkdtemp(prefix="mgpicola_bench_")
    model, vals, nsteps = W.values_for(workload, out)
    with open(vals["OutputRedshiftFile"], "w") as f:
        f.write(f"0, {nsteps}\n")
    pf = os.path.join(out, "params.txt")
    M.write_paramfile(pf, vals, model=model)
    hp = M.read_params(pf, model=model)

    def new_sim():
        t0 = time.time()
        sim = M.Sim(hp)
        t1 = time.time()
        sim.ic()
        sim.gpu.sync()
        print(f"[bench] tables {t1 - t0:.1f} s, initial conditions {time.time() - t1:.1f} s", file=sys.stderr,
              flush=True)
        return sim

    sim = new_sim()
    for _ in range(warmup):
        if sim.finished:
            sim.close()
            sim = new_sim()
        sim.advance()
    timed, elapsed = 0, 0.0
    while timed < steps:
        if sim.finished:
            sim.close()
            sim = new_sim()
        sim.gpu.sync()
        t0 = time.perf_counter()
        n = 0
        while timed + n < steps and not sim.finished:
            sim.advance()
            n += 1
        sim.gpu.sync()
        elapsed += time.perf_counter() - t0
        timed += n
        print(f"[bench] {timed}/{steps} steps timed, a = {sim.a:.4f}", file=sys.stderr, flush=True)
    sim.close()
    # per-kernel event timing over one whole run
    sim = new_sim()
    M.enable_kernel_timing(sim.gpu, True)
    M.kernel_timing(sim.gpu)  # reset (the initial conditions are not counted)
    nrun = 0
    while not sim.finished:
        sim.advance()
        nrun += 1
        if nrun % 10 == 0:
            print(f"[bench] kernel-timed run: {nrun} steps", file=sys.stderr, flush=True)
    sim.gpu.sync()
    kt = M.kernel_timing(sim.gpu)
    M.enable_kernel_timing(sim.gpu, False)
    sim.close()
    kt = {k: (v[0] / nrun, v[1] / nrun, v[2] / nrun) for k, v in kt.items()}
    npart = hp.Nsample ** 3
    return dict(elapsed=elapsed, npart=npart, nmesh=hp.Nmesh, nsample=hp.Nsample, box=hp.Box, kernels=kt,
                steps_per_run=nsteps)


def learned_leg(np_eval=512, box_eval=1024.0, nsteps=30, np_train=128, box_train=256.0, iterations=40):
    """The metric's second part: the learned Green's-function multiplier on the
    benchmark's model (f(R) fR0 = 1e-5, scale-dependent, screening, 30 steps).
    A network is trained on the GPU on coarse (Nmesh = Np) / 3x-fine pairs at
    128^3 particles in 256 Mpc/h (the same 2 Mpc/h particle spacing and coarse
    Nyquist k as the evaluation; seeds 101-103, LCDM and f(R) with f_R0 = 1e-6,
    1e-5 and 1e-4; the iterate kept is the best on held-out seed-104 pairs).
    Then at 512^3 in 1024 Mpc/h the plain fine run (Nmesh 1536), the plain
    coarse run and the learned coarse run (Nmesh 512) give max |P/P_fine - 1|
    over the bins with k < 1 h/Mpc at z = 0 (64 linear bins up to the coarse
    Nyquist k, shot noise subtracted) and their time per step (steps 2-30, as
    in the throughput measurement: the first step carries one-time costs and
    the output with its snapshot write is not a step).  The same network's
    errors on the LCDM run of the same initial conditions (a model in its
    training set) go in "lcdm", and on f(R) with f_R0 = 3e-5 (in no training or
    validation pair) in "fofr_3e-5"."""
    import math

    import learned as LR
    import mgpicola as M
    import workloads as W

    out = tempfile.mkdtemp(prefix="mgpicola_learned_")

    def params(tag, nsample, nmesh, box, seed, fofr0):
        d = os.path.join(out, tag)
        os.makedirs(d, exist_ok=True)
        over = dict(Nsample=nsample, Nmesh=nmesh, Box=box, Seed=seed, pofk_nbins=64, pofk_kmin=0.0,
                    pofk_kmax=math.pi * nsample / box, pofk_compute_every_step=1)
        if not fofr0:
            over.update(modified_gravity_active=0, include_screening=0)
        else:
            over.update(fofr0=fofr0)
        model, vals, _ = W.values_for("fofr_sd_512_1024", d, **over)
        with open(vals["OutputRedshiftFile"], "w") as f:
            f.write(f"0, {nsteps}\n")
        pf = os.path.join(d, "params.txt")
        M.write_paramfile(pf, vals, model=model)
        return M.read_params(pf, model=model)

    t0 = time.time()

    def pair(s, f):
        return (params(f"c{s}_{f}", np_train, np_train, box_train, s, f),
                params(f"f{s}_{f}", np_train, 3 * np_train, box_train, s, f))

    # LCDM and f(R) with f_R0 spread over two decades (log-uniform nodes); sample weights modes^(1/4), and
    # ln M = 0 anchors at k = 0 (the coarse mesh is exact at the largest scales)
    pairs = [pair(s, f) for s in (101, 102, 103) for f in (0.0, 1e-6, 1e-5, 1e-4)]
    # the kept iterate is chosen on held-out pairs (seed 104), never sampled
    val = [pair(104, f) for f in (0.0, 1e-5)]
    mlp = M.MLP(seed=7)
    hist = LR.train(pairs, mlp, iterations=iterations, iters=1500, lr=3e-3, eta=0.6, eta_decay=0.95,
                    kmax_select=1.0, val_pairs=val, wpow=0.25, anchor=4.0,
                    log=lambda m: print(f"[bench] learned: {m}", file=sys.stderr, flush=True))
    t_train = time.time() - t0

    def timed_run(hp, m=None):
        sim = M.Sim(hp)
        try:
            sim.set_learned(m, True)
            sim.ic()
            sim.gpu.sync()
            ts = []
            while not sim.finished:
                t = time.perf_counter()
                sim.advance()
                sim.gpu.sync()
                ts.append(time.perf_counter() - t)
            # the first step carries one-time costs (first launches, P(k) mode counts) and the
            # last loop iteration is the output (snapshot write), outside the metric
            return sim.log(), sum(ts[1:-1]) / len(ts[1:-1]) * 1e3
        finally:
            sim.close()

    fine, ms_fine = timed_run(params("eval_fine", np_eval, 3 * np_eval, box_eval, 5001, 1e-5))
    print(f"[bench] learned: fine run {ms_fine:.1f} ms/step", file=sys.stderr, flush=True)
    hc = params("eval_coarse", np_eval, np_eval, box_eval, 5001, 1e-5)
    plain, ms_plain = timed_run(hc)
    lrn, ms_lrn = timed_run(hc, mlp)
    e_plain, e_lrn = LR.pofk_error(fine, plain)[0], LR.pofk_error(fine, lrn)[0]
    print(f"[bench] learned: P(k) error plain {e_plain:.4f}, learned {e_lrn:.4f}", file=sys.stderr, flush=True)
    # the same network on the LCDM run of the same initial conditions (a model it was trained on) and on f(R)
    # with f_R0 = 3e-5, which no training or validation pair has
    fine_l, _ = timed_run(params("eval_fine_lcdm", np_eval, 3 * np_eval, box_eval, 5001, 0.0))
    hcl = params("eval_coarse_lcdm", np_eval, np_eval, box_eval, 5001, 0.0)
    el_plain = LR.pofk_error(fine_l, timed_run(hcl)[0])[0]
    el_lrn = LR.pofk_error(fine_l, timed_run(hcl, mlp)[0])[0]
    print(f"[bench] learned: LCDM P(k) error plain {el_plain:.4f}, learned {el_lrn:.4f}", file=sys.stderr, flush=True)
    fine_o, _ = timed_run(params("eval_fine_3e-5", np_eval, 3 * np_eval, box_eval, 5001, 3e-5))
    hco = params("eval_coarse_3e-5", np_eval, np_eval, box_eval, 5001, 3e-5)
    eo_plain = LR.pofk_error(fine_o, timed_run(hco)[0])[0]
    eo_lrn = LR.pofk_error(fine_o, timed_run(hco, mlp)[0])[0]
    print(f"[bench] learned: f_R0 = 3e-5 P(k) error plain {eo_plain:.4f}, learned {eo_lrn:.4f}", file=sys.stderr,
          flush=True)
    mlp.close()
    return {"pofk_err_learned": round(e_lrn, 4), "pofk_err_plain_coarse": round(e_plain, 4), "kmax": 1.0,
            "lcdm": {"pofk_err_learned": round(el_lrn, 4), "pofk_err_plain_coarse": round(el_plain, 4),
                     "note": "in distribution: LCDM pairs are in the training set"},
            "fofr_3e-5": {"pofk_err_learned": round(eo_lrn, 4), "pofk_err_plain_coarse": round(eo_plain, 4),
                          "note": "out of distribution: no training or validation pair has f_R0 = 3e-5"},
            "ms_per_step_fine": round(ms_fine, 2), "ms_per_step_coarse_learned": round(ms_lrn, 2),
            "ms_per_step_coarse_plain": round(ms_plain, 2), "speedup_vs_fine": round(ms_fine / ms_lrn, 2),
            "eval": f"{np_eval}^3 particles, Box {box_eval:g}, coarse Nmesh {np_eval} vs fine {3 * np_eval}, "
                    f"f(R) fR0=1e-5 scale-dependent + screening, {nsteps} steps, seed 5001",
            "train": f"{len(pairs)} pairs at {np_train}^3/Box {box_train:g} (coarse {np_train} vs {3 * np_train}), "
                     f"{iterations} fixed-point iterations (gain 0.6 x 0.95^i), the iterate with the smallest k < 1 "
                     f"error on 2 "
                     f"held-out pairs (seed 104) kept, {t_train:.0f} s",
            "train_history": [[h[0], round(h[2], 4)] for h in hist]}


def fft_leg(n=1024, reps=3):
    """Our 3D r2c of an n^3 real mesh (z r2c pass, y pass, x pass:
    mgpu_pm_transform without P(k); device time of its kernels by HIP events)
    against rocFFT's out-of-place single-precision r2c plan of the same mesh
    (librocfft of the same ROCm, called through its C API; HIP events around
    `reps` executions after one untimed one), and the largest difference of the
    two spectra (rocFFT's from its first execution) relative to the largest
    mode."""
    import ctypes as C

    import mgpicola as M

    rng = np.random.default_rng(11)
    x = rng.standard_normal((n, n, n), dtype=np.float32)
    g = M.MGPU(n, 8, float(n), 0.3)
    ours = []
    for r in range(reps + 1):
        g.set_mesh(M.MESH_DENSITY_X, x)
        M.enable_kernel_timing(g, True)
        M.kernel_timing(g)
        g.pm_transform()
        g.sync()
        kt = M.kernel_timing(g)
        M.enable_kernel_timing(g, False)
        if r:
            ours.append(sum(v[1] for k, v in kt.items() if k.startswith("fft")))
    got = g.get_mesh(M.MESH_DENSITY_K)
    g.close()

    hip = C.CDLL("libamdhip64.so")
    fft = C.CDLL("librocfft.so")
    vp = C.c_void_p
    nin, nout = 4 * n ** 3, 8 * n * n * (n // 2 + 1)
    din, dout, work = vp(), vp(), vp()
    plan, info = vp(), vp()

    def ck(rc, what):
        if rc:
            raise RuntimeError(f"{what} returned {rc}")

    ck(hip.hipMalloc(C.byref(din), C.c_size_t(nin)), "hipMalloc")
    ck(hip.hipMalloc(C.byref(dout), C.c_size_t(nout)), "hipMalloc")
    try:
        ck(hip.hipMemcpy(din, x.ctypes.data_as(vp), C.c_size_t(nin), 1), "hipMemcpy")  # host to device
        ck(fft.rocfft_setup(), "rocfft_setup")
        lengths = (C.c_size_t * 3)(n, n, n)
        # not in place, real forward, single precision, 3 dimensions, one transform, default layout
        ck(fft.rocfft_plan_create(C.byref(plan), 1, 2, 0, C.c_size_t(3), lengths, C.c_size_t(1), vp()),
           "rocfft_plan_create")
        wsz = C.c_size_t()
        ck(fft.rocfft_plan_get_work_buffer_size(plan, C.byref(wsz)), "rocfft_plan_get_work_buffer_size")
        ck(fft.rocfft_execution_info_create(C.byref(info)), "rocfft_execution_info_create")
        if wsz.value:
            ck(hip.hipMalloc(C.byref(work), wsz), "hipMalloc")
            ck(fft.rocfft_execution_info_set_work_buffer(info, work, wsz), "rocfft_execution_info_set_work_buffer")
        ins, outs = (vp * 1)(din), (vp * 1)(dout)
        ck(fft.rocfft_execute(plan, ins, outs, info), "rocfft_execute")
        ck(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
        # the spectrum of the first execution (a transform may use its input as scratch)
        ref = np.empty((n, n, n // 2 + 1), np.complex64)
        ck(hip.hipMemcpy(ref.ctypes.data_as(vp), dout, C.c_size_t(nout), 2), "hipMemcpy")  # device to host
        e0, e1 = vp(), vp()
        hip.hipEventCreate(C.byref(e0))
        hip.hipEventCreate(C.byref(e1))
        hip.hipEventRecord(e0, vp())
        for _ in range(reps):
            ck(fft.rocfft_execute(plan, ins, outs, info), "rocfft_execute")
        hip.hipEventRecord(e1, vp())
        ck(hip.hipEventSynchronize(e1), "hipEventSynchronize")
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        roc_ms = ms.value / reps
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        if info:
            fft.rocfft_execution_info_destroy(info)
        if plan:
            fft.rocfft_plan_destroy(plan)
        for b in (din, dout, work):
            if b:
                hip.hipFree(b)
    err = float(np.abs(got - ref).max() / np.abs(ref).max())
    ours_ms = float(np.mean(ours))
    return {"n": n, "ours_ms": round(ours_ms, 3), "rocfft_ms": round(roc_ms, 3),
            "rocfft_over_ours": round(roc_ms / ours_ms, 3), "max_rel_diff": float(f"{err:.2e}"),
            "what": f"3D r2c of a {n}^3 FP32 mesh: libmgpu in place (padded rows) vs a rocFFT out-of-place plan, "
                    "device time per transform"}


def ensemble_leg(per_size=32, threads=12, nsteps=20):
    """The batched nDGP ensemble (rc H0 = 1, density screening, Box = 2 Np,
    Nmesh = 2 Np, 20 steps): per_size realizations at each of 96^3, 128^3 and
    192^3 particles (all 96 by default), in fixed-amplitude pairs (delta and -delta),
    run one after the other and then concurrently on `threads` host threads
    (one libmgpu context and stream each).  Whole-ensemble wall time, set-up
    and initial conditions included, and the particle-steps/s it gives."""
    import ensemble as E
    import mgpicola as M
    import workloads as W

    out = tempfile.mkdtemp(prefix="mgpicola_ens_")
    hps = []
    for name in ("dgp_96_192", "dgp_128_256", "dgp_192_384"):
        for k in range(per_size):
            d = os.path.join(out, f"{name}_{k}")
            os.makedirs(d, exist_ok=True)
            model, vals, _ = W.values_for(name, d, Seed=7000 + 10 * len(hps) // 2, amplitude_fixed_initial_condition=1,
                                          inverted_initial_condition=k % 2)
            with open(vals["OutputRedshiftFile"], "w") as f:
                f.write(f"0, {nsteps}\n")
            pf = os.path.join(d, "params.txt")
            M.write_paramfile(pf, vals, model=model)
            hp = M.read_params(pf, model=model)
            hp.write_particles = 0  # no snapshot files at the z = 0 output
            hp.device_step = 1  # the steps between outputs keep the means on the device (mgpu_step)
            hps.append(hp)
    work = sum(h.Nsample ** 3 for h in hps) * nsteps
    seq, t_seq = E.run_ensemble([M.HostParams.from_buffer_copy(h) for h in hps], threads=1)
    print(f"[bench] ensemble: {len(hps)} runs one after the other {t_seq:.2f} s", file=sys.stderr, flush=True)
    con, t_con = E.run_ensemble([M.HostParams.from_buffer_copy(h) for h in hps], threads=threads)
    print(f"[bench] ensemble: concurrently on {threads} threads {t_con:.2f} s", file=sys.stderr, flush=True)
    bag, t_bag = E.run_batch([M.HostParams.from_buffer_copy(h) for h in hps], graph=True)
    print(f"[bench] ensemble: batched (one launch per kernel per size class), one hipGraph per step {t_bag:.2f} s",
          file=sys.stderr, flush=True)
    bat, t_bat = E.run_batch([M.HostParams.from_buffer_copy(h) for h in hps], threads=3)
    print(f"[bench] ensemble: batched, the size classes on 3 threads {t_bat:.2f} s", file=sys.stderr, flush=True)
    # time loops per size class: the sum over its realizations run alone, its batch's lockstep loop
    loops = {}
    for c, name in enumerate(("96^3/192", "128^3/256", "192^3/384")):
        sl = slice(c * per_size, (c + 1) * per_size)
        loops[name] = {"sequential": round(sum(r["seconds"] for r in seq[sl]), 3),
                       "batched_graph": round(bag[sl][0]["seconds"], 3)}
    return {"realizations": len(hps), "steps": nsteps, "threads": threads,
            "seconds_sequential": round(t_seq, 3), "seconds_concurrent": round(t_con, 3),
            "seconds_batched_graph": round(t_bag, 3), "seconds_batched_threads3": round(t_bat, 3),
            "time_loop_seconds_by_size_class": loops,
            "particle_steps_per_s_sequential": round(work / t_seq, 1),
            "particle_steps_per_s_concurrent": round(work / t_con, 1),
            "particle_steps_per_s_batched": round(work / min(t_bat, t_bag), 1),
            "speedup": round(t_seq / t_con, 2), "speedup_batched": round(t_seq / min(t_bat, t_bag), 2),
            "workload": f"nDGP rcH0=1 + screening, {per_size} each of 96^3/192, 128^3/256, 192^3/384 "
                        f"(Box = 2 Np), fixed-amplitude pairs, {nsteps} steps, device-resident steps; wall time "
                        f"incl. set-up and ICs; batched: the realizations of one size class as one batch "
                        f"(mgpu_batch_step), the classes one after the other with a hipGraph per step, or on 3 "
                        f"host threads without graphs"}


def cpu_sample(ns, nmesh, box):
    """Phase times [s] of one COLA step of the metric's path (f(R) with screening,
    scale-dependent LPT fields, P(k) with 512 bins) in the FP64 oracle at ns^3
    particles on an nmesh^3 mesh: deposit, r2c + P(k), fifth force (2 FFTs and
    the screening), Forces (3 c2r), gather, the 4 LPT fields (3 c2r each and the
    lattice read-out), Kick and Drift.  Returns (phases, threads)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle as O
    from helpers import lattice_particles

    O.lib()
    pos = lattice_particles(ns, box, amp=2.0 * box / ns / 2, seed=7, clustered=True)
    vel = np.zeros_like(pos)
    c = O.cfg(nmesh, ns, box, 0.3, model=O.FOFR, mg_active=1, include_screening=1)
    mg = O.MG(a=0.5, coupling=1.0 / 3.0, mass2=7e5, phi_crit=5e-6, dgp_fac=0, rsmooth=1, geff=1)
    cd = O.new_mesh(nmesh)
    g = np.ones(3 * (nmesh // 2) ** 2 + 1)
    d1 = np.zeros_like(pos)
    t = {}
    t0 = time.perf_counter()
    d = O.deposit(c, pos)
    t["deposit"] = time.perf_counter() - t0
    dx = d.copy()
    t0 = time.perf_counter()
    O.fft_r2c_3d(nmesh, d)
    O.pofk(c, d, 512, 0, 0.0, float(nmesh), 1)
    t["r2c_pofk"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    phik = O.fifth_force(c, mg, d, dx)
    t["fifth_force"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    f = O.forces(c, d, phik)
    t["forces"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    disp, sD = O.gather(c, pos, f)
    t["gather"] = time.perf_counter() - t0
    del f, phik, dx, d
    t0 = time.perf_counter()
    for _ in range(4):
        O.lpt_field(c, cd, g)
    t["lpt_fields"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    sV = O.kick(vel, disp, d1, d1, 0, 0, 1, 0.3, 0.1, 0.5, 1, sD)
    O.drift(pos, vel, d1, d1, 0, 0, 1, box, 0.1, 1, sV)
    t["kick_drift"] = time.perf_counter() - t0
    return t, O.num_threads()


# phases whose work scales with the particle count; the rest scale with the mesh as G log G
CPU_PARTICLE_PHASES = ("deposit", "gather", "kick_drift")


def cpu_extrapolate(t, ns0, n0, ns1, n1):
    """Seconds per step at (ns1, n1) from phase times at (ns0, n0): particle phases
    scale with Ns^3, mesh phases (FFT-dominated) with N^3 log N^3."""
    mesh = (n1 ** 3 * np.log(n1 ** 3)) / (n0 ** 3 * np.log(n0 ** 3))
    part = (ns1 / ns0) ** 3
    return sum(v * (part if k in CPU_PARTICLE_PHASES else mesh) for k, v in t.items())


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ns=256, nmesh=512, box=512.0, ns1=128, nmesh1=256, target=(512, 1024)):
    """The oracle on this host's cores: one step at ns^3/nmesh^3 on all the
    threads OpenMP is given, and (in a child process with OMP_NUM_THREADS=1)
    one step at ns1^3/nmesh1^3 on one core; both extrapolated phase by phase to
    the metric's 512^3 particles on a 1024^3 mesh."""
    import subprocess
    t, threads = cpu_sample(ns, nmesh, box)
    sec = cpu_extrapolate(t, ns, nmesh, *target)
    value = target[0] ** 3 / sec
    code = (f"import sys, json; sys.path.insert(0, {ROOT!r}); import bench; "
            f"t, n = bench.cpu_sample({ns1}, {nmesh1}, {float(nmesh1)}); print(json.dumps([t, n]))")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    one = None
    try:
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
        t1, n1 = json.loads(r.stdout.strip().splitlines()[-1])
        if n1 == 1:
            one = dict(value=round(target[0] ** 3 / cpu_extrapolate(t1, ns1, nmesh1, *target), 1),
                       sample_seconds=round(sum(t1.values()), 2),
                       sample=f"one step at {ns1}^3 particles on a {nmesh1}^3 mesh, 1 thread")
    except Exception as e:  # the all-core figure stands on its own
        print(f"[bench] one-core CPU sample failed: {e}", file=sys.stderr, flush=True)
    return dict(value=round(value, 1), unit="particle-steps/s", cores=threads, kind="port",
                sample=(f"one f(R)+screening+scale-dependent COLA step of the FP64 oracle (-O3, OpenMP) at {ns}^3 "
                        f"particles on a {nmesh}^3 mesh ({sum(t.values()):.1f} s), extrapolated phase by phase to "
                        f"{target[0]}^3 on {target[1]}^3 (particle phases x Ns^3, mesh phases x N^3 log N^3)"),
                seconds_per_step=round(sec, 2), phases_s={k: round(v, 3) for k, v in t.items()},
                one_core=one, nproc=os.cpu_count(), cpu_model=cpu_model(), fft="the oracle's own radix FFT "
                "(no FFTW on the host)")


def traffic_of(workload, symbol):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by the same counter's
    reading for a plain device copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t.get("kernels", {}).get(symbol)
    return None if k is None else k.get("bytes_per_dispatch")


def copy_probe(nbytes):
    """A plain device-to-device copy of nbytes (and a fill of the source), to
    calibrate the FETCH_SIZE / WRITE_SIZE counters of a profiling run."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise SystemExit("copy probe: hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[bench] copy probe: {nbytes} bytes copied device to device", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=29)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fofr_sd_512_1024")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-multiplier P(k) error leg")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the concurrent-ensemble leg")
    ap.add_argument("--no-fft-compare", action="store_true", help="skip the rocFFT comparison")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first copy this many MiB device to device (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib << 20)
    r = gpu_bench(args.steps, args.warmup, args.workload)
    value = r["npart"] * args.steps / r["elapsed"]
    kt = r["kernels"]
    # the dominant kernel: the single-kernel class with the most device time in one step
    single = {k: v for k, v in kt.items() if k not in ("lpt", "dep_sort", "fft_c2r")}
    dom = max(single, key=lambda k: single[k][1])
    n, ms, by = kt[dom]
    achieved = (by / n) / (ms / n * 1e-3) / 1e9 if n and ms > 0 else 0.0
    symbol = KERNEL_SYMBOL.get(dom, dom).format(N=r["nmesh"], Ns=r["nsample"], m=r["nmesh"] // r["nsample"])
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_of(args.workload, symbol),
            "kernel": symbol, "class": dom, "launches_per_step": round(n, 2), "ms_per_launch": round(ms / n, 4),
            "algorithmic_bytes_per_launch": round(by / n)}
    per_kernel = {k: {"launches": round(v[0], 2), "ms": round(v[1], 3),
                      "GBps": round(v[2] / (v[1] * 1e-3) / 1e9, 1) if v[1] > 0 else None}
                  for k, v in kt.items() if v[0]}
    lrn = None
    if not args.no_learned:
        try:
            lrn = learned_leg()
        except Exception as e:  # the throughput line still prints; the failure is reported in it
            lrn = {"error": f"{type(e).__name__}: {e}"}
    ens = None
    if not args.no_ensemble:
        try:
            ens = ensemble_leg()
        except Exception as e:  # the throughput line still prints; the failure is reported in it
            ens = {"error": f"{type(e).__name__}: {e}"}
    fft = None
    if not args.no_fft_compare:
        try:
            fft = fft_leg()
        except Exception as e:  # the throughput line still prints; the failure is reported in it
            fft = {"error": f"{type(e).__name__}: {e}"}
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
        "roofline": roof, "cpu_baseline": cpu, "learned": lrn, "ensemble": ens, "fft_vs_rocfft": fft,
        "kernels_per_step": per_kernel,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
