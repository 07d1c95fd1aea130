"""The C ABI on the CPU: libmgpu.so and libmgpicola.so load, and every function
their headers declare is exported with C linkage (no C++ mangling).  Calls that
need no device are checked too: table sizes, kernel names, and the error
returned for invalid parameters before any device work.  No GPU is used."""
import ctypes as C
import os
import re
import subprocess

import pytest

import mgpicola as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set(re.findall(r"\b(" + prefix + r"[a-z0-9_]*)\s*\(", text))
    return names


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_mgpu_header_symbols_exported():
    names = declared(os.path.join(ROOT, "include", "mgpu.h"), "mgpu_")
    assert len(names) >= 30
    syms = exported(M.LIBMGPU)
    missing = sorted(n for n in names if n not in syms)
    assert not missing, missing
    assert set(M.MGPU_SYMBOLS) == names
    L = C.CDLL(M.LIBMGPU)
    for n in names:
        getattr(L, n)


def test_host_header_symbols_exported():
    hdr = os.path.join(ROOT, "mg-picola-public_amd", "host", "mgpicola.h")
    names = declared(hdr, "mgp_")
    syms = exported(M.LIBHOST)
    missing = sorted(n for n in names if n not in syms)
    assert not missing, missing


def test_no_mangled_abi_exports():
    syms = exported(M.LIBMGPU)
    public = [s for s in syms if s.startswith("mgpu_")]
    assert public and not [s for s in syms if s.startswith("_Z") and "mgpu_" in s and s.replace("_Z", "") in public]


def test_n2_table_len_and_kernel_names():
    L = M.lib()
    L.mgpu_n2_table_len.restype = C.c_size_t
    L.mgpu_n2_table_len.argtypes = [C.c_int]
    for n in (8, 64, 1024):
        assert L.mgpu_n2_table_len(n) >= 3 * (n // 2) ** 2 + 1
    M._kt_setup()
    names = [L.mgpu_kernel_name(i).decode() for i in range(len(M.KERNELS))]
    assert names == M.KERNELS


@pytest.mark.parametrize("bad", [dict(nmesh=0), dict(nsample=-1), dict(box=0.0), dict(nmesh=7)])
def test_create_rejects_invalid_parameters(bad):
    """Invalid sizes are refused with MGPU_E_INVALID before a device is touched."""
    L = M.lib()
    p = M.Params(nmesh=32, nsample=16, box=100.0, omega=0.3, model=0, mg_active=0, include_screening=0, filter=0,
                 scale_dependent=0, use_cola=1, pofk_nbins=0, pofk_bintype=0, pofk_subtract_shotnoise=0, pofk_kmin=0.0,
                 pofk_kmax=0.0)
    for k, v in bad.items():
        setattr(p, k, v)
    ctx = C.c_void_p()
    rc = L.mgpu_create(C.byref(p), C.byref(ctx))
    assert rc == -1  # MGPU_E_INVALID
    assert not ctx.value
    assert L.mgpu_last_error(None)


def test_no_device_fails_loudly():
    """Without a visible GPU the product path raises; it never falls back to a CPU path."""
    L = M.lib()
    p = M.Params(nmesh=32, nsample=16, box=100.0, omega=0.3, model=0, mg_active=0, include_screening=0, filter=0,
                 scale_dependent=0, use_cola=1, pofk_nbins=0, pofk_bintype=0, pofk_subtract_shotnoise=0, pofk_kmin=0.0,
                 pofk_kmax=0.0)
    ctx = C.c_void_p()
    rc = L.mgpu_create(C.byref(p), C.byref(ctx))
    if rc == 0:
        L.mgpu_destroy(ctx)
        pytest.skip("a GPU is visible")
    assert rc == -5  # MGPU_E_NODEVICE
    with pytest.raises(M.MGPUError):
        M.MGPU(32, 16, 100.0, 0.3)


def hp(n):
    """Complex entries per mesh row: n/2 + 1 rounded up to 16."""
    return (n // 2 + 1 + 15) // 16 * 16


@pytest.mark.parametrize("n,ns,sd", [(512, 256, 0), (1024, 1024, 1), (2048, 1024, 1), (384, 192, 1), (768, 512, 1)])
def test_memory_plan_by_part(n, ns, sd):
    """mgpu_memory_plan needs no device; its parts follow the buffers mgpu_create allocates."""
    parts, total = M.memory_plan(n, ns, scale_dependent=sd)
    np_ = ns ** 3
    sb = n * n * hp(n) * 8
    assert parts["particles"] == 60 * np_
    assert parts["meshes"] == 4 * sb
    assert parts["stored"] == (2 * sb if sd else 0)
    lat = 6 * ns * ns * hp(ns) * 8 if (sd and n % ns == 0) else 0
    assert parts["lattice"] == lat
    # tile bins plus a u32 rank and a float4 record per particle
    assert 20 * np_ <= parts["deposit"] <= 20 * np_ + 72 * (n // 8) ** 3 + (1 << 20)
    # per-block partial sums (3 doubles per 256 particles) plus the per-|n|^2 tables (26 bytes an entry)
    assert 0 < parts["tables"] < 3 * 8 * np_ // 256 + 26 * (3 * (n // 2) ** 2 + 1) + (16 << 20)
    assert parts["rsd"] == 16 * np_
    assert total == sum(v for k, v in parts.items() if k != "rsd")


def test_memory_plan_of_the_1024_learned_run_fits_one_gpu():
    """1024^3 particles on a 1024^3 mesh, scale-dependent growth, RSD: within 288 GB of HBM."""
    parts, total = M.memory_plan(1024, 1024, scale_dependent=1)
    assert total + parts["rsd"] < 288e9 * 0.9


@pytest.mark.parametrize("bad", [dict(nmesh=1000), dict(nsample=0), dict(nsample=4096)])
def test_memory_plan_rejects_bad_sizes(bad):
    L = M.lib()
    args = dict(nmesh=64, nsample=32)
    args.update(bad)
    p = M.make_params(args["nmesh"], args["nsample"], 100.0, 0.3)
    L.mgpu_memory_plan.argtypes = [C.POINTER(M.Params), C.c_void_p, C.c_void_p]
    assert L.mgpu_memory_plan(C.byref(p), None, None) == -1
