"""Ensembles of independent realizations (paired, fixed-amplitude initial
conditions of uneven Np and Nmesh) advanced concurrently on one GPU.

Each realization is the C host's simulation (libmgpicola) with its own libmgpu
context, and so its own HIP stream and device buffers.  A small run leaves most
of the chip idle between its launches and host synchronisations (the means of
Disp and of the velocities come back to the host every step), so one host
thread per realization keeps several streams busy at once: the kernels of
different realizations overlap on the GPU.  ctypes releases the interpreter
lock for every foreign call, and neither library holds mutable global state
(errors are per thread or per context), so the threads do not serialize.

Every realization stays deterministic on its own: its results do not depend on
how many run beside it.

run_batch takes the other route: the realizations of one configuration (the
pairs of one size class) are one mgpu batch, advanced in lockstep from one
host thread, and every step between outputs is one launch per kernel over all
of them (mgpu_batch_step), optionally one hipGraph per step.
"""
import concurrent.futures as cf
import time

import mgpicola as M


def run_one(hp, keep_particles=False):
    """Run one realization to the end; returns its step count, time stepping
    seconds and (optionally) the final particles."""
    sim = M.Sim(hp)
    try:
        sim.ic()
        sim.gpu.sync()
        t0 = time.perf_counter()
        n = 0
        while not sim.finished:
            sim.advance()
            n += 1
        sim.gpu.sync()
        out = {"steps": n, "seconds": time.perf_counter() - t0, "particles": hp.Nsample ** 3}
        if keep_particles:
            out["pos"], out["vel"] = sim.gpu.get_particles()
        return out
    finally:
        sim.close()


def run_ensemble(hps, threads=8, keep_particles=False):
    """Run the realizations hps (HostParams) on `threads` host threads; returns
    (per-realization results in the order of hps, wall seconds of the whole
    ensemble)."""
    t0 = time.perf_counter()
    if threads > 1:
        # a hipGraph capture needs the process to itself: concurrent runs keep the device-resident step
        # without graphs
        for hp in hps:
            hp.device_step = min(hp.device_step, 1)
    if threads <= 1:
        res = [run_one(hp, keep_particles) for hp in hps]
    else:
        with cf.ThreadPoolExecutor(max_workers=threads) as ex:
            res = list(ex.map(lambda hp: run_one(hp, keep_particles), hps))
    return res, time.perf_counter() - t0


def _run_group(hps, keep_particles):
    """One batch (the realizations hps of one configuration) in lockstep; per-realization results."""
    sims = M.Sim.batch(hps)
    try:
        for s in sims:
            s.ic()
        sims[0].gpu.sync()
        t1 = time.perf_counter()
        n = 0
        while not all(s.finished for s in sims):
            for s in sims:
                if not s.finished:
                    s.advance()
            n += 1
        sims[0].gpu.sync()
        dt = time.perf_counter() - t1
        res = []
        for hp, s in zip(hps, sims):
            out = {"steps": n, "seconds": dt, "particles": hp.Nsample ** 3}
            if keep_particles:
                out["pos"], out["vel"] = s.gpu.get_particles()
            res.append(out)
        return res
    finally:
        for s in sims:
            s.close()


def run_batch(hps, keep_particles=False, graph=False, threads=1):
    """Run the realizations hps as one batch per configuration (Nmesh, Nsample, Box), each batch's members in
    lockstep; with threads > 1 the batches run on that many host threads (one stream each; no graphs then,
    since a capture needs the process to itself).  Returns (per-realization results in the order of hps, wall
    seconds of the whole ensemble); a result's "seconds" is its batch's time loop."""
    t0 = time.perf_counter()
    groups = {}
    for i, hp in enumerate(hps):
        hp.device_step = 2 if graph and threads <= 1 else 1
        groups.setdefault((hp.Nmesh, hp.Nsample, hp.Box), []).append(i)
    res = [None] * len(hps)
    idxs = list(groups.values())
    if threads <= 1:
        outs = [_run_group([hps[i] for i in idx], keep_particles) for idx in idxs]
    else:
        with cf.ThreadPoolExecutor(max_workers=threads) as ex:
            outs = list(ex.map(lambda idx: _run_group([hps[i] for i in idx], keep_particles), idxs))
    for idx, out in zip(idxs, outs):
        for i, r in zip(idx, out):
            res[i] = r
    return res, time.perf_counter() - t0
