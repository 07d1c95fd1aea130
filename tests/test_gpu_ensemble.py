"""Concurrent realizations (ensemble.py): paired fixed-amplitude nDGP runs of
uneven sizes on host threads give, run by run, the same particles bit for bit
as the same runs one after the other."""
import os

import numpy as np
import pytest

import ensemble as E
import mgpicola as M
import workloads as W

pytestmark = pytest.mark.gpu


def params(tmp_path, tag, nsample, seed, inverted, nsteps=4):
    d = tmp_path / tag
    d.mkdir()
    model, vals, _ = W.values_for("dgp_96_192", str(d), Nsample=nsample, Nmesh=2 * nsample, Box=2.0 * nsample,
                                  Seed=seed, amplitude_fixed_initial_condition=1, inverted_initial_condition=inverted)
    with open(vals["OutputRedshiftFile"], "w") as f:
        f.write(f"0, {nsteps}\n")
    pf = str(d / "params.txt")
    M.write_paramfile(pf, vals, model=model)
    return M.read_params(pf, model=model)


def test_concurrent_runs_equal_sequential_runs(tmp_path):
    hps = [params(tmp_path, f"r{i}", ns, 100 + i // 2, i % 2) for i, ns in enumerate((16, 16, 24, 24, 32, 32))]
    seq, _ = E.run_ensemble(hps, threads=1, keep_particles=True)
    hps2 = [M.HostParams.from_buffer_copy(h) for h in hps]
    con, _ = E.run_ensemble(hps2, threads=6, keep_particles=True)
    for a, b in zip(seq, con):
        assert a["steps"] == b["steps"] == 5
        np.testing.assert_array_equal(a["pos"], b["pos"])
        np.testing.assert_array_equal(a["vel"], b["vel"])
        assert np.all(np.isfinite(a["pos"]))
    # the two members of a pair start from delta and -delta: different runs
    assert not np.array_equal(seq[0]["pos"], seq[1]["pos"])


def test_fixed_amplitude_pair_fields_are_exact_opposites():
    """The two members of a pair draw the same phases with |delta| = sqrt(P) and delta -> -delta: the
    Zel'dovich field of one is exactly minus the other's (negation commutes with every FP operation of
    the transforms) and the 2LPT field, quadratic in delta, is bit-identical (src/2LPT.c:389-412)."""
    n, ns = 96, 48
    kf = 2 * np.pi / (2.0 * ns)
    n2 = np.arange(M.n2_table_len(n))
    k = kf * np.sqrt(n2)
    pk = np.where(n2 > 0, 3e3 * k / (1 + (k / 0.05) ** 2.5), 0.0)
    out = []
    for inv in (0, 1):
        g = M.MGPU(n, ns, 2.0 * ns, 0.3, model=M.DGP, mg_active=1, include_screening=1)
        g.ic_displacement_fields(777, pk, sphere_mode=0, amplitude_fixed=1, inverted=inv)
        out.append((g.get_lpt(0), g.get_lpt(1)))
        g.close()
    assert np.abs(out[0][0]).max() > 0 and np.abs(out[0][1]).max() > 0
    np.testing.assert_array_equal(out[1][0], -out[0][0])
    np.testing.assert_array_equal(out[1][1], out[0][1])


@pytest.mark.parametrize("inverted", [0, 1])
def test_ensemble_member_matches_oracle(tmp_path, inverted):
    """One realization of the ensemble's shape (nDGP rc H0 = 1 with density screening, Box = 2 Np,
    Nmesh = 2 Np, fixed amplitude, either member of the pair) at 32^3 for 6 steps against the oracle's run:
    positions within 1e-4 of the box."""
    from e2e import oracle_run, periodic_diff
    hp = params(tmp_path, f"m{inverted}", 32, 301, inverted, nsteps=6)
    res, _ = E.run_ensemble([hp], threads=1, keep_particles=True)
    gpos = res[0]["pos"]
    cosmo = M.Cosmo(hp)
    rpos, _, _ = oracle_run(hp, cosmo, 6)
    d = np.abs(periodic_diff(gpos, rpos, hp.Box))
    print(f"inverted {inverted}: max |dpos| / Box = {d.max() / hp.Box:.2e}", flush=True)
    assert d.max() <= 1e-4 * hp.Box


@pytest.mark.parametrize("model", ["lcdm", "dgp"])
def test_device_step_equals_separate_calls(model):
    """mgpu_step (the whole step with the means on the device), eager and as a per-step hipGraph replayed with
    new scalars each step, gives the particles of mgpu_get_displacements + mgpu_kick_drift bit for bit over five
    steps, and its device-resident means equal the host ones."""
    from helpers import lattice_particles
    n, ns, box = 48, 24, 48.0
    pos0 = lattice_particles(ns, box, amp=1.0, seed=9)
    rng = np.random.default_rng(3)
    vel0 = (rng.normal(size=pos0.shape) * 2.0).astype(np.float32)
    kw = dict(model=M.DGP, mg_active=1, include_screening=1, filter=1) if model == "dgp" else {}
    mg = M.MG(a=0.6, coupling=0.4, mass2=0.0, phi_crit=0.0, dgp_fac=1.7, rsmooth=1.0, geff=1.0) if model == "dgp" \
        else None
    coeffs = [(0.1 + 0.01 * s, 0.5 + 0.05 * s, 0.3, -0.2, 0.07 + 0.01 * s, 0.02, -0.01) for s in range(5)]
    out = []
    for mode in ("host", "device", "graph"):
        g = M.MGPU(n, ns, box, 0.3, **kw)
        g.set_particles(pos0, vel0)
        sD, _ = g.get_displacements(mg=mg)
        sV = g.kick(0.1, 0.5, 0.3, -0.2, sD)  # the velocity sums land on the device
        g.drift(0.05, 0.01, 0.0, sV)
        for dda, A, s1k, s2k, dyyy, s1d, s2d in coeffs:
            if mode == "host":
                sD, _ = g.get_displacements(mg=mg)
                sV = g.kick_drift(dda, A, s1k, s2k, sD, dyyy, s1d, s2d, sV)
            else:
                g.step(dda, A, s1k, s2k, dyyy, s1d, s2d, mg=mg, graph=mode == "graph")
        p, v = g.get_particles()
        means = (sD, sV) if mode == "host" else g.get_means()
        out.append((p, v, means))
        g.close()
    for p, v, (md, mv) in out[1:]:
        np.testing.assert_array_equal(p, out[0][0])
        np.testing.assert_array_equal(v, out[0][1])
        np.testing.assert_array_equal(md, out[0][2][0])
        np.testing.assert_array_equal(mv, out[0][2][1])
    assert not np.array_equal(out[0][0], pos0)


def _member_start(g, mg):
    sD, _ = g.get_displacements(mg=mg)
    sV = g.kick(0.1, 0.5, 0.3, -0.2, sD)  # the velocity sums land on the device
    g.drift(0.05, 0.01, 0.0, sV)


@pytest.mark.parametrize("model,n,ns,count", [("lcdm", 48, 24, 3), ("dgp", 48, 24, 3), ("dgp", 96, 48, 4)])
def test_batch_step_equals_solo_steps(model, n, ns, count):
    """mgpu_batch_step (one launch per kernel over all members) gives every member the particles and
    device-resident means of mgpu_step on that member alone, bit for bit, over five steps, eagerly and as a
    per-step hipGraph; members start from different particles."""
    from helpers import lattice_particles
    box = float(n)
    kw = dict(model=M.DGP, mg_active=1, include_screening=1, filter=1) if model == "dgp" else {}
    mg = M.MG(a=0.6, coupling=0.4, mass2=0.0, phi_crit=0.0, dgp_fac=1.7, rsmooth=1.0, geff=1.0) if model == "dgp" \
        else None
    starts = []
    for r in range(count):
        pos0 = lattice_particles(ns, box, amp=1.0 + 0.2 * r, seed=11 + r)
        vel0 = (np.random.default_rng(5 + r).normal(size=pos0.shape) * 2.0).astype(np.float32)
        starts.append((pos0, vel0))
    coeffs = [(0.1 + 0.01 * s, 0.5 + 0.05 * s, 0.3, -0.2, 0.07 + 0.01 * s, 0.02, -0.01) for s in range(5)]
    solo = []
    for pos0, vel0 in starts:
        g = M.MGPU(n, ns, box, 0.3, **kw)
        g.set_particles(pos0, vel0)
        _member_start(g, mg)
        for c in coeffs:
            g.step(*c, mg=mg)
        solo.append(g.get_particles() + g.get_means())
        g.close()
    for graph in (False, True):
        members = M.MGPU.batch(count, n, ns, box, 0.3, **kw)
        for g, (pos0, vel0) in zip(members, starts):
            g.set_particles(pos0, vel0)
            _member_start(g, mg)
        for c in coeffs:
            M.MGPU.batch_step(members, *c, mg=mg, graph=graph)
        print(f"{model} n {n} count {count} graph {graph}", flush=True)
        for r, g in enumerate(members):
            p, v = g.get_particles()
            md, mv = g.get_means()
            np.testing.assert_array_equal(p, solo[r][0])
            np.testing.assert_array_equal(v, solo[r][1])
            np.testing.assert_array_equal(md, solo[r][2])
            np.testing.assert_array_equal(mv, solo[r][3])
        # a member is a whole context: the separate calls go on from the batch's state
        sD, _ = members[1].get_displacements(mg=mg)
        assert np.all(np.isfinite(sD))
        for g in members:
            g.close()
    assert not np.array_equal(solo[0][0], solo[1][0])


@pytest.mark.parametrize("n,ns,amp,sphere", [(32, 16, 1, 0), (48, 24, 0, 1)])
def test_batch_ic_equals_solo_ic(n, ns, amp, sphere):
    """mgpu_batch_ic_displacement_fields (one launch per kernel over the members, each with its own seed and
    inversion) gives every member the ZA and 2LPT lattice displacements of mgpu_ic_displacement_fields on it
    alone, bit for bit."""
    box = 2.0 * ns
    kf = 2 * np.pi / box
    n2 = np.arange(M.n2_table_len(n), dtype=np.float64)
    k = kf * np.sqrt(n2)
    pk = np.where(n2 > 0, 3e3 * k / (1 + (k / 0.05) ** 2.5), 0.0)
    seeds, inv = [777, 777, 901, 55], [0, 1, 0, 1]
    kw = dict(model=M.DGP, mg_active=1, include_screening=1)
    solo = []
    for sd, iv in zip(seeds, inv):
        g = M.MGPU(n, ns, box, 0.3, **kw)
        g.ic_displacement_fields(sd, pk, sphere_mode=sphere, amplitude_fixed=amp, inverted=iv)
        solo.append((g.get_lpt(0), g.get_lpt(1)))
        g.close()
    members = M.MGPU.batch(len(seeds), n, ns, box, 0.3, **kw)
    M.MGPU.batch_ic(members, seeds, pk, sphere_mode=sphere, amplitude_fixed=amp, inverted=inv)
    for r, g in enumerate(members):
        np.testing.assert_array_equal(g.get_lpt(0), solo[r][0])
        np.testing.assert_array_equal(g.get_lpt(1), solo[r][1])
    for g in members:
        g.close()
    assert not np.array_equal(solo[0][0], solo[2][0])
    np.testing.assert_array_equal(solo[1][0], -solo[0][0])


def test_batch_step_refuses_foreign_members():
    a = M.MGPU.batch(2, 32, 16, 32.0, 0.3)
    b = M.MGPU(32, 16, 32.0, 0.3)
    for g in a + [b]:
        g.set_particles(np.zeros((3, 16 ** 3), np.float32) + 1.0, np.zeros((3, 16 ** 3), np.float32))
        sD, _ = g.get_displacements()
        g.kick(0.1, 0.5, 0.0, 0.0, sD)
    for bad in ([a[1], a[0]], [a[0], b], [a[0]]):
        with pytest.raises(M.MGPUError):
            M.MGPU.batch_step(bad, 0.1, 0.5, 0.0, 0.0, 0.05, 0.0, 0.0)
    M.MGPU.batch_step(a, 0.1, 0.5, 0.0, 0.0, 0.05, 0.0, 0.0)
    for g in a + [b]:
        g.close()


def test_batched_runs_equal_solo_runs(tmp_path):
    """The ensemble as batches (ensemble.run_batch: one mgpu batch per size class, the members in lockstep, each
    step between outputs one launch per kernel over all members, eagerly or as one hipGraph per step, the
    batches one after the other or on two host threads) gives every realization the particles of its solo run
    with the separate calls, bit for bit."""
    sizes = (16, 16, 16, 16, 24, 24)
    hps = [params(tmp_path, f"s{i}", ns, 200 + i // 2, i % 2, nsteps=6) for i, ns in enumerate(sizes)]
    solo, _ = E.run_ensemble(hps, threads=1, keep_particles=True)
    for graph, threads in ((False, 1), (True, 1), (False, 2)):
        hps2 = [M.HostParams.from_buffer_copy(h) for h in hps]
        bat, _ = E.run_batch(hps2, keep_particles=True, graph=graph, threads=threads)
        print(f"graph {graph} threads {threads}: {len(bat)} realizations", flush=True)
        for a, b in zip(solo, bat):
            assert a["steps"] == b["steps"] == 7
            np.testing.assert_array_equal(a["pos"], b["pos"])
            np.testing.assert_array_equal(a["vel"], b["vel"])
    assert not np.array_equal(solo[0]["pos"], solo[2]["pos"])


def test_device_step_runs_through_the_host(tmp_path):
    """The C host with device_step = 1 (steps on the device) and 2 (each step one hipGraph) runs the
    ensemble's workload to the same particles as the separate calls, bit for bit."""
    res = []
    for mode in (0, 1, 2):
        hp = params(tmp_path, f"d{mode}", 32, 401, 1, nsteps=6)
        hp.device_step = mode
        r, _ = E.run_ensemble([hp], threads=1, keep_particles=True)
        res.append(r[0])
    for r in res[1:]:
        np.testing.assert_array_equal(r["pos"], res[0]["pos"])
        np.testing.assert_array_equal(r["vel"], res[0]["vel"])


def test_device_step_refuses_what_it_cannot_do():
    g = M.MGPU(32, 16, 32.0, 0.3, scale_dependent=1)
    with pytest.raises(M.MGPUError):
        g.step(0.1, 0.5, 0.0, 0.0, 0.05, 0.0, 0.0)
    g.close()
    g = M.MGPU(32, 16, 32.0, 0.3)
    with pytest.raises(M.MGPUError):  # no kick yet: no velocity sums on the device
        g.step(0.1, 0.5, 0.0, 0.0, 0.05, 0.0, 0.0)
    g.close()
