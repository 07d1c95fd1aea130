"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It wraps liboracle.so, the FP64 C restatement of the reference
PM/COLA step, with numpy arrays in the reference's padded FFTW layouts.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

LCDM, FOFR, DGP, BRANSDICKE, MBETA = 0, 1, 2, 3, 4
FILTER_NONE, FILTER_GAUSSIAN, FILTER_TOPHAT, FILTER_SHARPK = 0, 1, 2, 3


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


class Cfg(C.Structure):
    _fields_ = [("nmesh", C.c_int), ("nsample", C.c_int), ("box", C.c_double), ("omega", C.c_double),
                ("model", C.c_int), ("mg_active", C.c_int), ("include_screening", C.c_int), ("filter", C.c_int)]


class MG(C.Structure):
    _fields_ = [("a", C.c_double), ("coupling", C.c_double), ("mass2", C.c_double), ("phi_crit", C.c_double),
                ("dgp_fac", C.c_double), ("rsmooth", C.c_double), ("geff", C.c_double)]


class Ranlxd(C.Structure):
    _fields_ = [("x", C.c_double * 12), ("carry", C.c_double), ("ir", C.c_uint), ("jr", C.c_uint),
                ("ir_old", C.c_uint), ("pr", C.c_uint)]


_lib = None
_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_U = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_P3 = C.c_void_p * 3


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _load(LIB)
    return _lib


def build_planted(plant, outdir):
    """Compile a copy of the oracle with planted mistake number `plant` (see
    oracle.h) into outdir; returns the library path."""
    path = os.path.join(outdir, f"liboracle_plant{plant}.so")
    src = [os.path.join(HERE, f) for f in ("oracle_fft.c", "oracle_pm.c", "oracle_ic.c")]
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-fopenmp", "-std=gnu11", f"-DOR_PLANT={plant}", "-shared",
                           "-o", path] + src + ["-lm"])
    return path


class using:
    """Context manager: route every call of this module to another build of the
    oracle (a planted-mistake library) and restore the normal one afterwards."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _lib
        lib()
        self.saved = _lib
        _lib = _load(self.path)
        return _lib

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def _load(path):
    L = C.CDLL(path)
    L.or_num_threads.restype = C.c_int
    L.or_num_threads.argtypes = []
    L.or_real_len.restype = C.c_size_t
    L.or_real_len.argtypes = [C.c_int]
    L.or_fft_c2c_1d.argtypes = [C.c_int, _D, C.c_int]
    L.or_fft_r2c_3d.argtypes = [C.c_int, _D]
    L.or_fft_c2r_3d.argtypes = [C.c_int, _D]
    L.or_pm_deposit.argtypes = [C.POINTER(Cfg), C.c_size_t, _P3, _D]
    L.or_pofk.argtypes = [C.POINTER(Cfg), _D, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _D, _D, _D]
    L.or_adjust_pofk.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.or_pm_deposit_rsd.argtypes = [C.POINTER(Cfg), C.c_size_t, _P3, _P3, _P3, _P3, C.c_int, C.c_double, C.c_int,
                                    C.c_int, C.c_double, C.c_double, _D]
    L.or_rsd_multipoles.argtypes = [C.POINTER(Cfg), _D, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _D,
                                    _D, _D, _D, _D]
    L.or_mg_fifth_force.restype = C.c_int
    L.or_mg_fifth_force.argtypes = [C.POINTER(Cfg), C.POINTER(MG), _D, _D, _D]
    L.or_pm_forces.argtypes = [C.POINTER(Cfg), _D, C.c_void_p, _D, _D, _D]
    L.or_pm_gather.argtypes = [C.POINTER(Cfg), C.c_size_t, _P3, _D, _D, _D, _P3, C.POINTER(C.c_double)]
    L.or_kick.argtypes = [C.c_size_t, _P3, _P3, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int,
                          C.c_double, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double),
                          C.POINTER(C.c_double)]
    L.or_drift.argtypes = [C.c_size_t, _P3, _P3, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int,
                           C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double)]
    L.or_ranlxd_seed.argtypes = [C.POINTER(Ranlxd), C.c_ulong]
    L.or_ranlxd_uniform.restype = C.c_double
    L.or_ranlxd_uniform.argtypes = [C.POINTER(Ranlxd)]
    L.or_ranlxd_get.restype = C.c_ulong
    L.or_ranlxd_get.argtypes = [C.POINTER(Ranlxd)]
    L.or_seedtable.argtypes = [C.c_int, C.c_ulong, _U]
    L.or_ic_delta.argtypes = [C.POINTER(Cfg), C.c_ulong, C.c_int, C.c_int, C.c_int, _D, _D]
    L.or_ic_displacements.argtypes = [C.POINTER(Cfg), _D, C.c_int, _P3, _P3, C.c_void_p]
    L.or_lpt_field.argtypes = [C.POINTER(Cfg), _D, _D, _P3]
    return L


def num_threads():
    return lib().or_num_threads()


def cfg(nmesh, nsample, box, omega, model=LCDM, mg_active=0, include_screening=0, filter=FILTER_NONE):
    return Cfg(nmesh, nsample, box, omega, model, mg_active, include_screening, filter)


def real_len(n):
    return (n + 1) * n * 2 * (n // 2 + 1)


def new_mesh(n):
    return np.zeros(real_len(n), dtype=np.float64)


def _p3(arrs):
    return _P3(*[a.ctypes.data for a in arrs])


def _soa(x):
    """(3, np) float32 C-contiguous view list."""
    assert x.dtype == np.float32 and x.flags.c_contiguous and x.shape[0] == 3
    return [x[0], x[1], x[2]]


def real_view(m, n):
    """(n, n, n) view of the interior of a padded real mesh."""
    return m[: n * n * 2 * (n // 2 + 1)].reshape(n, n, 2 * (n // 2 + 1))[:, :, :n]


def complex_view(m, n):
    h = n // 2 + 1
    return m[: n * n * 2 * h].view(np.complex128).reshape(n, n, h)


def fft_r2c_3d(n, m):
    lib().or_fft_r2c_3d(n, m)


def fft_c2r_3d(n, m):
    lib().or_fft_c2r_3d(n, m)


def deposit(c, pos):
    d = new_mesh(c.nmesh)
    lib().or_pm_deposit(C.byref(c), pos.shape[1], _p3(_soa(pos)), d)
    return d


def adjust_pofk(nmesh, nbins, bintype, subtract, kmin, kmax):
    a = [C.c_int(nbins), C.c_int(bintype), C.c_int(subtract), C.c_double(kmin), C.c_double(kmax)]
    lib().or_adjust_pofk(nmesh, *[C.byref(x) for x in a])
    return a[0].value, a[1].value, a[2].value, a[3].value, a[4].value


def pofk(c, dk, nbins, bintype, kmin, kmax, subtract):
    p, k, n = np.zeros(nbins), np.zeros(nbins), np.zeros(nbins)
    lib().or_pofk(C.byref(c), dk, nbins, bintype, kmin, kmax, subtract, p, k, n)
    return p, k, n


def deposit_rsd(c, pos, vel, l0, l1, axis, vnorm, scale_dependent, use_cola, s1=0.0, s2=0.0):
    """PtoMesh_RSD: axis 1 = y, 2 = z; vnorm in mesh cells per velocity unit."""
    d = new_mesh(c.nmesh)
    arrs = [_soa(x) for x in (pos, vel, l0, l1)]
    lib().or_pm_deposit_rsd(C.byref(c), pos.shape[1], *[_p3(a) for a in arrs], axis, vnorm, scale_dependent,
                            use_cola, s1, s2, d)
    return d


def rsd_multipoles(c, dk, nbins, bintype, kmin, kmax, subtract):
    out = [np.zeros(nbins) for _ in range(5)]
    lib().or_rsd_multipoles(C.byref(c), dk, nbins, bintype, kmin, kmax, subtract, *out)
    return tuple(out)  # P0, P2, P4, k, n


def fifth_force(c, m, dk, dx):
    phik = new_mesh(c.nmesh)
    wrote = lib().or_mg_fifth_force(C.byref(c), C.byref(m), dk, dx, phik)
    return phik if wrote else None


def forces(c, dk, phik):
    f = [new_mesh(c.nmesh) for _ in range(3)]
    lib().or_pm_forces(C.byref(c), dk, None if phik is None else phik.ctypes.data, f[0], f[1], f[2])
    return f


def gather(c, pos, f):
    npart = pos.shape[1]
    disp = np.zeros((3, npart), dtype=np.float32)
    s = (C.c_double * 3)()
    lib().or_pm_gather(C.byref(c), npart, _p3(_soa(pos)), f[0], f[1], f[2], _p3(_soa(disp)), s)
    return disp, np.array(list(s))


def kick(vel, disp, d1, d2, s1, s2, sd, omega, dda, A, use_cola, sumD):
    sV = (C.c_double * 3)()
    sD = (C.c_double * 3)(*sumD)
    lib().or_kick(vel.shape[1], _p3(_soa(vel)), _p3(_soa(disp)),
                  None if d1 is None else C.cast(_p3(_soa(d1)), C.c_void_p),
                  None if d2 is None else C.cast(_p3(_soa(d2)), C.c_void_p),
                  s1, s2, sd, omega, dda, A, use_cola, sD, sV)
    return np.array(list(sV))


def drift(pos, vel, d1, d2, s1, s2, sd, box, dyyy, use_cola, sumV):
    sV = (C.c_double * 3)(*sumV)
    lib().or_drift(pos.shape[1], _p3(_soa(pos)), _p3(_soa(vel)),
                   None if d1 is None else C.cast(_p3(_soa(d1)), C.c_void_p),
                   None if d2 is None else C.cast(_p3(_soa(d2)), C.c_void_p),
                   s1, s2, sd, box, dyyy, use_cola, sV)


def seedtable(n, seed):
    t = np.zeros(n * n, dtype=np.uint32)
    lib().or_seedtable(n, seed, t)
    return t


def ranlxd_stream(seed, count):
    s = Ranlxd()
    lib().or_ranlxd_seed(C.byref(s), seed)
    return np.array([lib().or_ranlxd_uniform(C.byref(s)) for _ in range(count)])


def ranlxd_get_nth(seed, n):
    s = Ranlxd()
    lib().or_ranlxd_seed(C.byref(s), seed)
    v = 0
    for _ in range(n):
        v = lib().or_ranlxd_get(C.byref(s))
    return v


def ic_delta(c, seed, pk_n2, sphere_mode=0, amp_fixed=0, inverted=0):
    cd = new_mesh(c.nmesh)
    lib().or_ic_delta(C.byref(c), seed, sphere_mode, amp_fixed, inverted, np.ascontiguousarray(pk_n2, np.float64), cd)
    return cd


def ic_displacements(c, cd, scale_dependent):
    npart = c.nsample ** 3
    za = np.zeros((3, npart), dtype=np.float32)
    lp = np.zeros((3, npart), dtype=np.float32)
    cd2 = new_mesh(c.nmesh) if scale_dependent else None
    lib().or_ic_displacements(C.byref(c), cd, int(scale_dependent), _p3(_soa(za)), _p3(_soa(lp)),
                              None if cd2 is None else cd2.ctypes.data)
    return (cd2,) if scale_dependent else (za, lp)


def lpt_field(c, cd, g_n2):
    out = np.zeros((3, c.nsample ** 3), dtype=np.float32)
    lib().or_lpt_field(C.byref(c), cd, np.ascontiguousarray(g_n2, np.float64), _p3(_soa(out)))
    return out
