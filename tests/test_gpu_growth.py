"""Scale-dependent growth tables on the GPU (mgpu_growth_tables), which replace
calculate_scale_dependent_growth_factor (src/cosmo.c:798-981): the Einstein-de
Sitter closed form, and the tables of f(R), m-beta and f(R) with massive
neutrinos against the host's own ODE solver at and between the table nodes."""
import math

import numpy as np
import pytest

import mgpicola as M
import test_host as H

pytestmark = pytest.mark.gpu


def test_einstein_de_sitter_closed_form():
    """Omega = 1, no modified gravity: the initial data (1, 1, -3/7, -6/7) is the
    growing mode, so D1 = a/a_ini and D2 = -3/7 D1^2 exactly; the stored forms
    follow from Q = a^3 H with H = a^-1.5."""
    m = M.GrowthModel(model=0, mg_active=0, lcdm=0, omega=1.0)
    x0, na = math.log(1 / 20.0), 200
    dx = -x0 / (na - 1)
    lnk = np.log(np.geomspace(1e-3, 10.0, 37))
    t = M.growth_tables(m, x0, dx, na, lnk)
    x = x0 + dx * np.arange(na)
    a = np.exp(x)
    d1 = np.exp(x - x0)
    Q = a ** 3 * a ** -1.5
    want = [d1, d1 * Q / a, 1.5 * a * d1, -3 / 7 * d1 ** 2, -6 / 7 * d1 ** 2 * Q / a, 1.5 * a * (-3 / 7 - 1) * d1 ** 2]
    for f in range(6):
        np.testing.assert_allclose(t[f], np.broadcast_to(want[f], t[f].shape), rtol=1e-9, err_msg=f"field {f}")


def _pair(tmp_path, model, **over):
    hp, host = H.cosmo_for(tmp_path, model, **over) if model != "FOFRNU" else H.nu_cosmo(tmp_path, **over)
    hp.growth_on_gpu = 1
    gpu = M.Cosmo(hp)
    return hp, host, gpu


@pytest.mark.parametrize("model,over", [
    ("FOFR", dict(Nmesh=64, Box=200.0)),
    ("FOFR", dict(Nmesh=256, Box=1024.0, fofr0=1e-6, nfofr=2.0)),
    ("MBETA", dict(Nmesh=64, Box=200.0)),
    ("FOFRNU", dict(Nmesh=32, Box=512.0)),
])
def test_tables_match_host_solver(tmp_path, model, over):
    """Both solvers run Dormand-Prince 5(4) at tolerance 1e-11 on the same ODEs;
    the tables agree to 1e-8 relative (transcendental functions differ in the
    last bits, and the step sequence may differ with them).  The growth values
    are compared through mgp_growth_sd at the table nodes and halfway between."""
    hp, host, gpu = _pair(tmp_path, model, **over)
    print("case:", model, over, flush=True)
    kf = 2 * math.pi / hp.Box
    ks = np.geomspace(kf * 0.6, kf * hp.Nmesh * 1.7, 23)
    zi = hp.Init_Redshift
    As = np.geomspace(1 / (1 + zi), 1.0, 17)
    L = M.hlib()
    worst = 0.0
    for which in range(6):
        for k in ks:
            for A in As:
                h = L.mgp_growth_sd(host.h, which, k, A)
                g = L.mgp_growth_sd(gpu.h, which, k, A)
                worst = max(worst, abs(g - h) / max(abs(h), 1e-300))
    assert worst < 1e-8, worst
    if model == "FOFRNU":
        # the LCDM table with neutrinos, through the MG P(k) ratio that reads it
        for k in ks:
            for A in As:
                h, g = L.mgp_mg_pofk_ratio(host.h, k, A), L.mgp_mg_pofk_ratio(gpu.h, k, A)
                assert abs(g - h) <= 1e-8 * abs(h), (k, A, g, h)
    host.close()
    gpu.close()


def test_bad_model_is_rejected():
    m = M.GrowthModel(model=2, mg_active=1, lcdm=0, omega=0.3)  # DGP has no scale-dependent table
    with pytest.raises(M.MGPUError, match="scale-dependent models"):
        M.growth_tables(m, -3.0, 0.01, 10, np.zeros(4))
    m = M.GrowthModel(model=1, mg_active=1, lcdm=0, omega=0.3, fofr0=1e-5, nfofr=1.0, omega_nu=0.01)
    with pytest.raises(M.MGPUError, match="transfer splines"):
        M.growth_tables(m, -3.0, 0.01, 10, np.zeros(4))
