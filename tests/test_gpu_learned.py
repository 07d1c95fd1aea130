"""The learned Green's-function multiplier end to end: trained on the GPU on
coarse (Nmesh = Nsample) / 3x-finer run pairs that the code generates, it
brings a held-out coarse run's P(k) at z = 0 closer to the plain fine-mesh run
than the plain coarse run gets, for LCDM and for f(R)."""
import math

import numpy as np
import pytest

import learned as LR
import mgpicola as M
import workloads as W

pytestmark = pytest.mark.gpu

NP, BOX, NSTEPS = 32, 64.0, 10


def pair(tmp_path, seed, fofr0):
    """(coarse, fine) HostParams: same seed and ICs, Nmesh = Np and 3 Np."""
    kny = math.pi * NP / BOX
    over = dict(Nsample=NP, Box=BOX, Seed=seed, pofk_compute_every_step=1, pofk_nbins=16, pofk_kmin=0.0,
                pofk_kmax=kny)
    if fofr0:
        over.update(modified_gravity_active=1, include_screening=1, fofr0=fofr0)
    out = []
    for nm in (NP, 3 * NP):
        d = tmp_path / f"s{seed}_f{fofr0}_n{nm}"
        d.mkdir()
        model, vals, _ = W.values_for("lcdm_64_128", str(d), Nmesh=nm, **over)
        with open(vals["OutputRedshiftFile"], "w") as f:
            f.write(f"0, {NSTEPS}\n")
        pf = str(d / "params.txt")
        M.write_paramfile(pf, vals, model=model)
        out.append(M.read_params(pf, model=model))
    return tuple(out)


def test_learned_multiplier_reduces_the_pofk_error(tmp_path):
    train = [pair(tmp_path, s, f) for s in (11, 12, 13) for f in (0.0, 1e-5)]
    mlp = M.MLP(seed=7)
    hist = LR.train(train, mlp, iterations=8, iters=1500, lr=3e-3, eta=0.6, eta_decay=0.93,
                    log=lambda s: print(s, flush=True))
    assert min(h[2] for h in hist) < 0.7 * hist[0][2], hist
    for f in (0.0, 1e-5):
        c, fine = pair(tmp_path, 21, f)
        fl = LR.run_logged(fine)
        plain = LR.pofk_error(fl, LR.run_logged(c))[0]
        lrn = LR.pofk_error(fl, LR.run_logged(c, mlp))[0]
        print(f"held-out seed, fR0 = {f}: max |P/P_fine - 1| for k < 1: plain {plain:.4f}, learned {lrn:.4f}",
              flush=True)
        assert lrn < 0.7 * plain, (plain, lrn)
    mlp.close()
