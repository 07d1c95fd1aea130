"""End-to-end GPU runs through the C host (libmgpicola -> libmgpu) against the
CPU oracle, and the initial conditions against the oracle's 2LPT restatement.

Tolerances: positions after the final step within 1e-4 of the box (FP32 GPU vs
FP64 oracle); initial displacements within 1e-5 of their maximum.
"""
import os

import numpy as np
import pytest

import mgpicola as M
import oracle as O
import workloads as W
from e2e import oracle_run, periodic_diff
from helpers import rel_err

pytestmark = pytest.mark.gpu


def make_params(tmp_path, name, nsteps, **over):
    model, vals, _ = W.values_for(name, str(tmp_path), **over)
    with open(vals["OutputRedshiftFile"], "w") as f:
        f.write(f"0, {nsteps}\n")
    pf = os.path.join(str(tmp_path), "params.txt")
    M.write_paramfile(pf, vals, model=model)
    return M.read_params(pf, model=model)


@pytest.mark.parametrize("case", [
    dict(name="lcdm_64_128", nsteps=10, over=dict(Nmesh=64, Nsample=32, Box=128.0)),
    dict(name="lcdm_64_128", nsteps=10, over=dict()),
    dict(name="lcdm_64_128", nsteps=5, over=dict(Nmesh=48, Nsample=24, Box=96.0, modified_gravity_active=1,
                                                 include_screening=1)),
])
def test_end_to_end_matches_oracle(tmp_path, case):
    hp = make_params(tmp_path, case["name"], case["nsteps"], **case["over"])
    print("case:", case, flush=True)
    sim = M.Sim(hp)
    sim.ic()
    while not sim.finished:
        sim.advance()
    gpos, gvel = sim.gpu.get_particles()
    sim.close()
    cosmo = M.Cosmo(hp)
    rpos, rvel, _ = oracle_run(hp, cosmo, case["nsteps"])
    d = np.abs(periodic_diff(gpos, rpos, hp.Box))
    print("max |dpos| / Box =", d.max() / hp.Box, flush=True)
    assert d.max() <= 1e-4 * hp.Box
    # output files of the reference's formats were written
    files = os.listdir(str(tmp_path))
    assert any(f.startswith("lcdm_64_128_z0p000") for f in files), files
    assert any(f.startswith("info_lcdm_64_128") for f in files), files


@pytest.mark.parametrize("n,ns,sd", [(32, 16, 0), (32, 32, 0), (24, 12, 1), (48, 24, 1)])
def test_ic_matches_oracle(n, ns, sd):
    box = 2.0 * n
    rng = np.random.default_rng(n)
    kf = 2 * np.pi / box
    n2 = np.arange(3 * (n // 2) ** 2 + 1)
    k = kf * np.sqrt(n2)
    pk = np.where(n2 > 0, 2e4 * k / (1 + (k / 0.05) ** 2.5), 0.0)
    g = M.MGPU(n, ns, box, 0.3, scale_dependent=sd)
    g.ic_displacement_fields(5001, pk)
    c = O.cfg(n, ns, box, 0.3)
    cd = O.ic_delta(c, 5001, pk)
    assert rel_err(g.get_mesh(M.MESH_CDELTA1), O.complex_view(cd, n)) < 1e-6
    if sd:
        (cd2,) = O.ic_displacements(c, cd, 1)
        assert rel_err(g.get_mesh(M.MESH_CDELTA2), O.complex_view(cd2, n)) < 5e-6
        g1 = rng.uniform(0.5, 1.5, size=n2.size)
        g2 = rng.uniform(-1e-4, 1e-4, size=n2.size)
        g.lpt_fields(0, g1, g2)
        ref = O.lpt_field(c, cd, g1).astype(np.float64) + O.lpt_field(c, cd2, g2).astype(np.float64)
        assert rel_err(g.get_lpt(0), ref) < 1e-5
    else:
        za, lp = O.ic_displacements(c, cd, 0)
        assert rel_err(g.get_lpt(0), za) < 1e-5
        assert rel_err(g.get_lpt(1), lp) < 2e-5
