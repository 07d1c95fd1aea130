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
