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
