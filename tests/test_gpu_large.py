"""GPU parity at the mesh sizes the workloads run (192 ... 1536).

The FFT passes are compiled per size: the radix schedule, the rows per block
of the z passes and, for power-of-two sizes from 256 up, the two-tile
pipelined column pass are different code at every size.  These tests run each
size's code:

- r2c of a separable random field against its exact spectrum (product of FP64
  1D DFTs), with 8- and 16-column tiles where both fit;
- Forces (Green's x pass, dual y pass, z c2r with i n_z) on a density spectrum
  built so that the force component equals N^3 t(x) for a separable t, which
  checks every inverse pass and the +N/2 Nyquist convention of
  src/auxPM.c:468-476 in closed form;
- GetDisplacements with paired x passes (f(R) potential screening, nDGP
  density screening) against the CPU oracle at 256, 384, 512, 768 and 1024
  meshes: delta(x) to 1e-6 of 1 + delta, delta(k) to 2e-6, P(k) per bin to
  0.1 %, the force meshes and Disp to 3e-5 of their maximum.

Tolerances are relative to the maximum of the expected field (FP32 transforms
of FP64 references).
"""
import os

import numpy as np
import pytest

import mgpicola as M
import oracle as O
import workloads as W
from bigfield import Separable, compare, pmap
from helpers import lattice_particles, rel_err

pytestmark = pytest.mark.gpu

SIZES = [192, 256, 384, 512, 768, 1024, 1536]


def tile_widths(n):
    """16-column tiles where they fit a workgroup (160 KiB of LDS, 1024 threads at 12 elements per thread for
    the radix-3 sizes up to 768, 24 from 1536 on, 16 for the others)."""
    ept = (12 if n < 1536 else 24) if n % 3 == 0 else 16
    fits16 = n * 16 * 8 <= 160 * 1024 and n * 16 // ept <= 1024
    return (8, 16) if fits16 else (8,)


@pytest.mark.parametrize("n", SIZES)
def test_r2c_separable_field(n):
    s = Separable(n, rank=2, seed=n)
    field = s.real()
    g = M.MGPU(n, 8, float(n), 0.3)
    for t in tile_widths(n):
        g.set_fft_tile(0, t)
        assert g.fft_tile(0) == t
        g.set_mesh(M.MESH_DENSITY_X, field)
        g.pm_transform()
        got = g.get_mesh(M.MESH_DENSITY_K)
        emax, erms = compare(got, s.spectrum_slab, n)
        print(f"n {n} tile {t}: max err {emax:.2e} rms err {erms:.2e}", flush=True)
        assert emax < 2e-6 and erms < 5e-7, (t, emax, erms)
        del got
    g.close()


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("n", SIZES)
def test_forces_c2r_closed_form(n, axis):
    box = float(n)
    s = Separable(n, rank=2, seed=7 * n + axis, zero_mean_axes=(axis,))
    src = s.force_source(axis, box)
    g = M.MGPU(n, 8, box, 0.3)
    g.set_mesh(M.MESH_DENSITY_K, src)
    del src
    g.forces()
    got = g.get_mesh((M.MESH_FORCE_X, M.MESH_FORCE_Y, M.MESH_FORCE_Z)[axis])
    g.close()
    scale = float(n) ** 3
    emax, erms = compare(got, lambda a, b: s.real_slab(a, b, scale), n)
    print(f"n {n} axis {axis}: max err {emax:.2e} rms err {erms:.2e}", flush=True)
    assert emax < 3e-6 and erms < 1e-6, (emax, erms)


def host_params(tmp_path, name, **over):
    model, vals, nsteps = W.values_for(name, str(tmp_path), **over)
    with open(vals["OutputRedshiftFile"], "w") as f:
        f.write(f"0, {nsteps}\n")
    pf = os.path.join(str(tmp_path), "params.txt")
    M.write_paramfile(pf, vals, model=model)
    return M.read_params(pf, model=model)


DISP_CASES = {
    # f(R) f_R0 = 1e-5 with potential screening: the metric's kernels at 1024 (paired x passes
    # k_col_fused<1024, ...>, the two-row z screening pass), and the two-tile column passes at 256/512
    "fofr_256": dict(n=256, ns=128, box=256.0, model="fofr", a=0.5),
    "fofr_512": dict(n=512, ns=256, box=512.0, model="fofr", a=0.5),
    "fofr_1024": dict(n=1024, ns=256, box=1024.0, model="fofr", a=0.6),
    # nDGP rc H0 = 1 with density screening: radix-3 sizes of the ensemble's meshes
    "dgp_384": dict(n=384, ns=192, box=384.0, model="dgp", a=0.7),
    "dgp_768": dict(n=768, ns=256, box=768.0, model="dgp", a=0.7),
}


@pytest.mark.timeout(400)
@pytest.mark.parametrize("case", list(DISP_CASES))
def test_get_displacements_matches_oracle_large(tmp_path, case):
    cs = DISP_CASES[case]
    n, ns, box = cs["n"], cs["ns"], cs["box"]
    if cs["model"] == "fofr":
        hp = host_params(tmp_path, "lcdm_64_128", Nmesh=n, Nsample=ns, Box=box, modified_gravity_active=1,
                         include_screening=1)
        model, omodel, filt = M.FOFR, O.FOFR, O.FILTER_NONE
    else:
        hp = host_params(tmp_path, "dgp_96_192", Nmesh=n, Nsample=ns, Box=box)
        model, omodel, filt = M.DGP, O.DGP, O.FILTER_GAUSSIAN
    cosmo = M.Cosmo(hp)
    m = cosmo.mg(cs["a"])
    cosmo.close()
    nbins = n // 2
    pos = lattice_particles(ns, box, amp=1.5 * box / ns, seed=n + ns)
    print(f"case {case}: n {n} ns {ns}", flush=True)
    g = M.MGPU(n, ns, box, 0.3, model=model, mg_active=1, include_screening=1, filter=filt, pofk_nbins=nbins,
               pofk_subtract_shotnoise=0)
    g.set_particles(pos)
    s_first, (pk_first, _, n_first) = g.get_displacements(mg=m, pofk=True)  # unpaired: records mode counts
    s_gpu, (pk_gpu, k_gpu, n_gpu) = g.get_displacements(mg=m, pofk=True)  # paired x passes
    dk_gpu = g.get_mesh(M.MESH_DENSITY_K)
    f_gpu = [g.get_mesh(q) for q in (M.MESH_FORCE_X, M.MESH_FORCE_Y, M.MESH_FORCE_Z)]
    disp_gpu = g.get_disp()
    g.pm_deposit()  # delta(x) again: the step's force pass reuses its slot
    dx_gpu = g.get_mesh(M.MESH_DENSITY_X)
    g.close()
    print("GPU done", flush=True)

    c = O.cfg(n, ns, box, 0.3, model=omodel, mg_active=1, include_screening=1, filter=filt)
    d = O.deposit(c, pos)
    dx_ref = O.real_view(d, n)
    err = pmap(lambda a, b: float((np.abs(dx_gpu[a:b] - dx_ref[a:b]) / np.maximum(1.0, np.abs(1.0 + dx_ref[a:b])))
                                  .max()), n)
    print(f"delta(x) {max(err):.2e}", flush=True)
    assert max(err) < 1e-6
    del dx_gpu
    dxfull = d.copy()
    O.fft_r2c_3d(n, d)
    dk_ref = O.complex_view(d, n)
    e_dk = compare(dk_gpu, lambda a, b: dk_ref[a:b], n)[0]
    print(f"delta(k) {e_dk:.2e}", flush=True)
    assert e_dk < 2e-6
    del dk_gpu
    # the binned power without the shot-noise constant, so the per-bin error is the estimator's own
    nb, bt, sub, kmin, kmax = O.adjust_pofk(n, nbins, 0, 0, 0.0, 0.0)
    rp, rk, rn = O.pofk(c, d, nb, bt, kmin, kmax, sub)
    assert np.array_equal(n_gpu, rn) and np.array_equal(n_first, rn)
    sel = (rn > 0) & (rk > 0)  # the k = 0 bin holds the mean, zero by construction
    np.testing.assert_allclose(k_gpu[sel], rk[sel], rtol=1e-10)  # FP64 sums of up to 1e8 terms
    for pk in (pk_first, pk_gpu):
        e_pk = np.abs(pk[sel] - rp[sel]) / np.abs(rp[sel])
        print(f"P(k) max rel err per bin {e_pk.max():.2e}", flush=True)
        assert e_pk.max() < 1e-3
    phik = O.fifth_force(c, O.MG(m.a, m.coupling, m.mass2, m.phi_crit, m.dgp_fac, m.rsmooth, m.geff), d, dxfull)
    del dxfull
    f_ref = O.forces(c, d, phik)
    del d, dk_ref, phik
    for a in range(3):
        fr = O.real_view(f_ref[a], n)
        e = compare(f_gpu[a], lambda p, q: fr[p:q], n)[0]
        print(f"force {a} {e:.2e}", flush=True)
        assert e < 3e-5, a
    del f_gpu
    disp_ref, s_ref = O.gather(c, pos, f_ref)
    e_disp = rel_err(disp_gpu, disp_ref)
    print(f"Disp {e_disp:.2e}", flush=True)
    assert e_disp < 3e-5
    assert np.abs(s_gpu - s_ref).max() <= 3e-5 * np.abs(disp_ref).max()
    assert np.abs(s_first - s_gpu).max() <= 1e-6 * np.abs(disp_ref).max()
