"""Python binding of libmgpu (include/mgpu.h) and libmgpicola (host/mgpicola.h).

Thin ctypes layer used by the tests, the benchmark and the smoke check.  The
compute path is the HIP library; there is no Python or CPU fallback: if
libmgpu.so is missing, or no GPU is visible, the calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBMGPU = os.path.join(HERE, "libmgpu.so")
LIBHOST = os.path.join(HERE, "libmgpicola.so")

LCDM, FOFR, DGP, BRANSDICKE, MBETA = 0, 1, 2, 3, 4
FILTER_NONE, FILTER_GAUSSIAN, FILTER_TOPHAT, FILTER_SHARPK = 0, 1, 2, 3
MESH_DENSITY_X, MESH_DENSITY_K, MESH_PHI_K, MESH_FORCE_X, MESH_FORCE_Y, MESH_FORCE_Z, MESH_CDELTA1, MESH_CDELTA2 = range(8)

# every symbol include/mgpu.h declares
MGPU_SYMBOLS = [
    "mgpu_n2_table_len", "mgpu_last_error", "mgpu_device_info", "mgpu_create", "mgpu_destroy", "mgpu_sync",
    "mgpu_set_particles", "mgpu_get_particles", "mgpu_set_lpt", "mgpu_get_lpt", "mgpu_get_disp",
    "mgpu_pm_density", "mgpu_pm_deposit", "mgpu_pm_transform", "mgpu_power_spectrum", "mgpu_rsd_power_spectrum",
    "mgpu_set_neutrinos", "mgpu_neutrino_response", "mgpu_growth_tables", "mgpu_mlp_create", "mgpu_mlp_destroy",
    "mgpu_mlp_get_params", "mgpu_mlp_set_params", "mgpu_mlp_forward", "mgpu_mlp_gradient", "mgpu_mlp_train",
    "mgpu_set_learned_multiplier", "mgpu_density_feature", "mgpu_kick_drift", "mgpu_set_fft_tile",
    "mgpu_get_fft_tile", "mgpu_autotune_fft", "mgpu_fifth_force", "mgpu_forces", "mgpu_mesh_to_particles",
    "mgpu_get_displacements", "mgpu_kick", "mgpu_drift", "mgpu_ic_displacement_fields", "mgpu_init_particles",
    "mgpu_lpt_fields", "mgpu_get_mesh", "mgpu_set_mesh", "mgpu_get_timers", "mgpu_enable_timers",
    "mgpu_kernel_name", "mgpu_enable_kernel_timing", "mgpu_kernel_timing", "mgpu_lpt_fields_slices",
    "mgpu_lpt_fields_slices2", "mgpu_ic_displacement_fields_noalloc", "mgpu_memory_plan", "mgpu_step",
    "mgpu_get_means", "mgpu_batch_create", "mgpu_batch_step", "mgpu_batch_ic_displacement_fields",
]


class Params(C.Structure):
    _fields_ = [("nmesh", C.c_int), ("nsample", C.c_int), ("box", C.c_double), ("omega", C.c_double),
                ("model", C.c_int), ("mg_active", C.c_int), ("include_screening", C.c_int), ("filter", C.c_int),
                ("scale_dependent", C.c_int), ("use_cola", C.c_int), ("pofk_nbins", C.c_int),
                ("pofk_bintype", C.c_int), ("pofk_subtract_shotnoise", C.c_int), ("pofk_kmin", C.c_double),
                ("pofk_kmax", C.c_double)]


class MG(C.Structure):
    _fields_ = [("a", C.c_double), ("coupling", C.c_double), ("mass2", C.c_double), ("phi_crit", C.c_double),
                ("dgp_fac", C.c_double), ("rsmooth", C.c_double), ("geff", C.c_double)]


class Pofk(C.Structure):
    _fields_ = [("pofk", C.POINTER(C.c_double)), ("kmean", C.POINTER(C.c_double)),
                ("nmodes", C.POINTER(C.c_double))]


class RsdPofk(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in ("P0", "P2", "P4", "k", "n")]


class Timers(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("ptomesh", "fft", "pofk", "modified_gravity", "forces", "mtoparticles",
                                          "kick", "drift", "lpt_fields", "ic")]


class MGPUError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libmgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBMGPU):
            raise MGPUError(f"{LIBMGPU} not built: run __graft_entry__.build() or make -C mg-picola-public_amd")
        L = C.CDLL(LIBMGPU)
        vp = C.c_void_p
        L.mgpu_n2_table_len.restype = C.c_size_t
        L.mgpu_n2_table_len.argtypes = [C.c_int]
        L.mgpu_last_error.restype = C.c_char_p
        L.mgpu_last_error.argtypes = [vp]
        L.mgpu_device_info.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(C.c_int),
                                       C.POINTER(C.c_size_t)]
        L.mgpu_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
        L.mgpu_destroy.argtypes = [vp]
        L.mgpu_destroy.restype = None
        for n in ("mgpu_sync", "mgpu_forces"):
            getattr(L, n).argtypes = [vp]
        L.mgpu_set_particles.argtypes = [vp, vp, vp]
        L.mgpu_get_particles.argtypes = [vp, vp, vp]
        L.mgpu_set_lpt.argtypes = [vp, C.c_int, vp]
        L.mgpu_get_lpt.argtypes = [vp, C.c_int, vp]
        L.mgpu_get_disp.argtypes = [vp, vp]
        L.mgpu_pm_density.argtypes = [vp, C.POINTER(Pofk)]
        L.mgpu_pm_deposit.argtypes = [vp]
        L.mgpu_pm_transform.argtypes = [vp, C.POINTER(Pofk)]
        L.mgpu_power_spectrum.argtypes = [vp, C.POINTER(Pofk)]
        L.mgpu_fifth_force.argtypes = [vp, C.POINTER(MG)]
        L.mgpu_mesh_to_particles.argtypes = [vp, C.POINTER(C.c_double)]
        L.mgpu_get_displacements.argtypes = [vp, C.POINTER(MG), C.POINTER(Pofk), C.POINTER(C.c_double)]
        L.mgpu_kick.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double),
                                C.POINTER(C.c_double)]
        L.mgpu_drift.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]
        L.mgpu_ic_displacement_fields.argtypes = [vp, C.c_ulong, C.c_int, C.c_int, C.c_int, vp]
        L.mgpu_init_particles.argtypes = [vp, C.c_double, C.c_double]
        L.mgpu_lpt_fields.argtypes = [vp, C.c_int, vp, vp]
        L.mgpu_get_mesh.argtypes = [vp, C.c_int, vp]
        L.mgpu_set_mesh.argtypes = [vp, C.c_int, vp]
        L.mgpu_get_timers.argtypes = [vp, C.POINTER(Timers), C.c_int]
        L.mgpu_enable_timers.argtypes = [vp, C.c_int]
        _lib = L
    return _lib


def n2_table_len(n):
    return 3 * (n // 2) ** 2 + 1


def device_info():
    name = C.create_string_buffer(256)
    mem, cus, lds = C.c_size_t(), C.c_int(), C.c_size_t()
    rc = lib().mgpu_device_info(name, 256, C.byref(mem), C.byref(cus), C.byref(lds))
    if rc:
        raise MGPUError(lib().mgpu_last_error(None).decode())
    return {"name": name.value.decode(), "global_mem": mem.value, "cu_count": cus.value, "lds_per_block": lds.value}


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f32(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float32)
    assert a.size == int(np.prod(shape)), (a.shape, shape)
    return a


MEM_PARTS = ("particles", "meshes", "deposit", "stored", "lattice", "tables", "rsd")


def make_params(nmesh, nsample, box, omega, model=LCDM, mg_active=0, include_screening=0, filter=FILTER_NONE,
                scale_dependent=0, use_cola=1, pofk_nbins=0, pofk_bintype=0, pofk_subtract_shotnoise=1,
                pofk_kmin=0.0, pofk_kmax=0.0):
    if pofk_nbins <= 0:
        pofk_nbins = nmesh
    if pofk_kmax <= pofk_kmin:
        pofk_kmin, pofk_kmax = (0.0 if pofk_bintype == 0 else 1.0), float(nmesh)
    return Params(nmesh, nsample, box, omega, model, mg_active, include_screening, filter, scale_dependent,
                  use_cola, pofk_nbins, pofk_bintype, pofk_subtract_shotnoise, pofk_kmin, pofk_kmax)


def memory_plan(nmesh, nsample, scale_dependent=0, pofk_nbins=0):
    """({part: bytes}, total bytes mgpu_create allocates) without a device."""
    p = make_params(nmesh, nsample, 1.0, 0.3, scale_dependent=scale_dependent, pofk_nbins=pofk_nbins)
    parts = (C.c_size_t * len(MEM_PARTS))()
    total = C.c_size_t()
    L = lib()
    L.mgpu_memory_plan.argtypes = [C.POINTER(Params), C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    if L.mgpu_memory_plan(C.byref(p), parts, C.byref(total)):
        raise MGPUError(f"mgpu_memory_plan failed: {L.mgpu_last_error(None).decode()}")
    return dict(zip(MEM_PARTS, list(parts))), total.value


class MGPU:
    """One simulation context on the GPU (owns particles, meshes, FFT tables)."""

    def __init__(self, nmesh, nsample, box, omega, model=LCDM, mg_active=0, include_screening=0,
                 filter=FILTER_NONE, scale_dependent=0, use_cola=1, pofk_nbins=0, pofk_bintype=0,
                 pofk_subtract_shotnoise=1, pofk_kmin=0.0, pofk_kmax=0.0):
        self.p = make_params(nmesh, nsample, box, omega, model, mg_active, include_screening, filter,
                             scale_dependent, use_cola, pofk_nbins, pofk_bintype, pofk_subtract_shotnoise, pofk_kmin,
                             pofk_kmax)
        self.n, self.ns, self.np = nmesh, nsample, nsample ** 3
        self.ctx = C.c_void_p()
        if getattr(self, "_no_create", False):
            return
        rc = lib().mgpu_create(C.byref(self.p), C.byref(self.ctx))
        if rc:
            raise MGPUError(f"mgpu_create failed ({rc}): {lib().mgpu_last_error(None).decode()}")

    @classmethod
    def batch(cls, count, *args, **kw):
        """count contexts of one configuration sharing one allocation and stream
        (mgpu_batch_create); step them together with batch_step."""
        members = []
        for _ in range(count):
            m = cls.__new__(cls)
            m._no_create = True
            cls.__init__(m, *args, **kw)
            members.append(m)
        L = lib()
        arr = (C.c_void_p * count)()
        L.mgpu_batch_create.argtypes = [C.POINTER(Params), C.c_int, C.c_void_p]
        rc = L.mgpu_batch_create(C.byref(members[0].p), count, arr)
        if rc:
            raise MGPUError(f"mgpu_batch_create failed ({rc}): {L.mgpu_last_error(None).decode()}")
        for m, h in zip(members, arr):
            m.ctx = C.c_void_p(h)
        return members

    def close(self):
        if self.ctx:
            lib().mgpu_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc:
            raise MGPUError(f"{what} failed ({rc}): {lib().mgpu_last_error(self.ctx).decode()}")

    # particles
    def set_particles(self, pos, vel=None):
        pos = _f32(pos, (3, self.np))
        vel = None if vel is None else _f32(vel, (3, self.np))
        self._ck(lib().mgpu_set_particles(self.ctx, _ptr(pos), _ptr(vel)), "set_particles")

    def get_particles(self):
        pos = np.empty((3, self.np), np.float32)
        vel = np.empty((3, self.np), np.float32)
        self._ck(lib().mgpu_get_particles(self.ctx, _ptr(pos), _ptr(vel)), "get_particles")
        return pos, vel

    def set_lpt(self, slot, d):
        d = _f32(d, (3, self.np))
        self._ck(lib().mgpu_set_lpt(self.ctx, slot, _ptr(d)), "set_lpt")

    def get_lpt(self, slot):
        d = np.empty((3, self.np), np.float32)
        self._ck(lib().mgpu_get_lpt(self.ctx, slot, _ptr(d)), "get_lpt")
        return d

    def get_disp(self):
        d = np.empty((3, self.np), np.float32)
        self._ck(lib().mgpu_get_disp(self.ctx, _ptr(d)), "get_disp")
        return d

    def _pofk_struct(self):
        nb = self.p.pofk_nbins
        bufs = [np.zeros(nb), np.zeros(nb), np.zeros(nb)]
        s = Pofk(*[b.ctypes.data_as(C.POINTER(C.c_double)) for b in bufs])
        return s, bufs

    # PM step
    def pm_density(self, pofk=False):
        if pofk:
            s, bufs = self._pofk_struct()
            self._ck(lib().mgpu_pm_density(self.ctx, C.byref(s)), "pm_density")
            return tuple(bufs)
        self._ck(lib().mgpu_pm_density(self.ctx, None), "pm_density")
        return None

    def pm_deposit(self):
        self._ck(lib().mgpu_pm_deposit(self.ctx), "pm_deposit")

    def pm_transform(self, pofk=False):
        if pofk:
            s, bufs = self._pofk_struct()
            self._ck(lib().mgpu_pm_transform(self.ctx, C.byref(s)), "pm_transform")
            return tuple(bufs)
        self._ck(lib().mgpu_pm_transform(self.ctx, None), "pm_transform")
        return None

    def power_spectrum(self):
        s, bufs = self._pofk_struct()
        self._ck(lib().mgpu_power_spectrum(self.ctx, C.byref(s)), "power_spectrum")
        return tuple(bufs)

    def rsd_power_spectrum(self, axis, vnorm, s1=0.0, s2=0.0):
        """P0, P2, P4, k, n of one line of sight (axis 1 = y, 2 = z)."""
        nb = self.p.pofk_nbins
        bufs = [np.zeros(nb) for _ in range(5)]
        r = RsdPofk(*[b.ctypes.data_as(C.POINTER(C.c_double)) for b in bufs])
        L = lib()
        L.mgpu_rsd_power_spectrum.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double,
                                              C.POINTER(RsdPofk)]
        self._ck(L.mgpu_rsd_power_spectrum(self.ctx, axis, vnorm, s1, s2, C.byref(r)), "rsd_power_spectrum")
        return tuple(bufs)

    def set_neutrinos(self, cdmfac, fac):
        """fac: a KSlice (norm num/den) of the neutrino factor, or None to turn the term off."""
        L = lib()
        L.mgpu_set_neutrinos.argtypes = [C.c_void_p, C.c_double, C.POINTER(KSlice)]
        self._ck(L.mgpu_set_neutrinos(self.ctx, cdmfac, None if fac is None else C.byref(fac)), "set_neutrinos")

    def neutrino_response(self):
        self._ck(lib().mgpu_neutrino_response(self.ctx), "neutrino_response")

    def set_learned_multiplier(self, mlp, a, theta):
        lib().mgpu_set_learned_multiplier.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double]
        self._ck(lib().mgpu_set_learned_multiplier(self.ctx, mlp.h if mlp is not None else None, a, theta),
                 "set_learned_multiplier")

    def density_feature(self):
        lib().mgpu_density_feature.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        s = C.c_double()
        self._ck(lib().mgpu_density_feature(self.ctx, C.byref(s)), "density_feature")
        return s.value

    def set_fft_tile(self, which, tcols):
        lib().mgpu_set_fft_tile.argtypes = [C.c_void_p, C.c_int, C.c_int]
        self._ck(lib().mgpu_set_fft_tile(self.ctx, which, tcols), "set_fft_tile")

    def fft_tile(self, which=0):
        lib().mgpu_get_fft_tile.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        t = C.c_int()
        self._ck(lib().mgpu_get_fft_tile(self.ctx, which, C.byref(t)), "get_fft_tile")
        return t.value

    def autotune_fft(self):
        lib().mgpu_autotune_fft.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        ms = (C.c_double * 4)()
        self._ck(lib().mgpu_autotune_fft(self.ctx, ms), "autotune_fft")
        return list(ms)

    def kick_drift(self, dda, A, s1k, s2k, sumD, dyyy, s1d, s2d, sumV):
        """Kick then Drift in one pass; sumV in: mean velocity before the kick; returns the mean after."""
        L = lib()
        d = C.c_double
        L.mgpu_kick_drift.argtypes = [C.c_void_p, d, d, d, d, C.POINTER(d), d, d, d, C.POINTER(d)]
        sd = (d * 3)(*sumD)
        sv = (d * 3)(*sumV)
        self._ck(L.mgpu_kick_drift(self.ctx, dda, A, s1k, s2k, sd, dyyy, s1d, s2d, sv), "kick_drift")
        return np.array(sv[:])

    def step(self, dda, A, s1k, s2k, dyyy, s1d, s2d, mg=None, graph=False):
        """One whole step with the means on the device (mgpu_step); nothing comes back."""
        L = lib()
        d = C.c_double
        L.mgpu_step.argtypes = [C.c_void_p, C.c_void_p, d, d, d, d, d, d, d, C.c_int]
        self._ck(L.mgpu_step(self.ctx, None if mg is None else C.byref(mg), dda, A, s1k, s2k, dyyy, s1d, s2d,
                             1 if graph else 0), "step")

    @staticmethod
    def batch_step(members, dda, A, s1k, s2k, dyyy, s1d, s2d, mg=None, graph=False):
        """mgpu_step for all members of one batch, one launch per kernel."""
        L = lib()
        d = C.c_double
        arr = (C.c_void_p * len(members))(*[m.ctx.value for m in members])
        L.mgpu_batch_step.argtypes = [C.c_void_p, C.c_int, C.c_void_p, d, d, d, d, d, d, d, C.c_int]
        rc = L.mgpu_batch_step(arr, len(members), None if mg is None else C.byref(mg), dda, A, s1k, s2k, dyyy, s1d,
                               s2d, 1 if graph else 0)
        if rc:
            raise MGPUError(f"batch_step failed ({rc}): {L.mgpu_last_error(members[0].ctx).decode()}")

    def get_means(self):
        """(mean Disp of the last gather, mean velocity of the last kick) from the device."""
        L = lib()
        sd, sv = (C.c_double * 3)(), (C.c_double * 3)()
        L.mgpu_get_means.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        self._ck(L.mgpu_get_means(self.ctx, sd, sv), "get_means")
        return np.array(sd[:]), np.array(sv[:])

    def fifth_force(self, mg):
        self._ck(lib().mgpu_fifth_force(self.ctx, C.byref(mg)), "fifth_force")

    def forces(self):
        self._ck(lib().mgpu_forces(self.ctx), "forces")

    def mesh_to_particles(self):
        s = (C.c_double * 3)()
        self._ck(lib().mgpu_mesh_to_particles(self.ctx, s), "mesh_to_particles")
        return np.array(list(s))

    def get_displacements(self, mg=None, pofk=False):
        s = (C.c_double * 3)()
        ps, bufs = self._pofk_struct() if pofk else (None, None)
        self._ck(lib().mgpu_get_displacements(self.ctx, None if mg is None else C.byref(mg),
                                              None if ps is None else C.byref(ps), s), "get_displacements")
        return np.array(list(s)), (tuple(bufs) if pofk else None)

    def kick(self, dda, A, s1, s2, sumD):
        sV = (C.c_double * 3)()
        self._ck(lib().mgpu_kick(self.ctx, dda, A, s1, s2, (C.c_double * 3)(*sumD), sV), "kick")
        return np.array(list(sV))

    def drift(self, dyyy, s1, s2, sumV):
        self._ck(lib().mgpu_drift(self.ctx, dyyy, s1, s2, (C.c_double * 3)(*sumV)), "drift")

    # initial conditions
    def ic_displacement_fields(self, seed, pk_n2, sphere_mode=0, amplitude_fixed=0, inverted=0):
        pk = np.ascontiguousarray(pk_n2, np.float64)
        assert pk.size == n2_table_len(self.n)
        self._ck(lib().mgpu_ic_displacement_fields(self.ctx, seed, sphere_mode, amplitude_fixed, inverted,
                                                   _ptr(pk)), "ic_displacement_fields")

    @staticmethod
    def batch_ic(members, seeds, pk_n2, sphere_mode=0, amplitude_fixed=0, inverted=None):
        """mgpu_ic_displacement_fields for all members of one batch, one launch per kernel."""
        L = lib()
        n = len(members)
        pk = np.ascontiguousarray(pk_n2, np.float64)
        assert pk.size == n2_table_len(members[0].n)
        arr = (C.c_void_p * n)(*[m.ctx.value for m in members])
        sd = (C.c_ulong * n)(*seeds)
        inv = (C.c_int * n)(*(inverted if inverted is not None else [0] * n))
        L.mgpu_batch_ic_displacement_fields.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                                        C.c_void_p, C.c_void_p]
        rc = L.mgpu_batch_ic_displacement_fields(arr, n, sd, sphere_mode, amplitude_fixed, inv, _ptr(pk))
        if rc:
            raise MGPUError(f"batch_ic failed ({rc}): {L.mgpu_last_error(members[0].ctx).decode()}")

    def init_particles(self, Di, Di2):
        self._ck(lib().mgpu_init_particles(self.ctx, Di, Di2), "init_particles")

    def lpt_fields(self, slot, g1_n2, g2_n2):
        g1 = np.ascontiguousarray(g1_n2, np.float64)
        g2 = np.ascontiguousarray(g2_n2, np.float64)
        self._ck(lib().mgpu_lpt_fields(self.ctx, slot, _ptr(g1), _ptr(g2)), "lpt_fields")

    # inspection
    def get_mesh(self, which):
        n = self.n
        real = which in (MESH_DENSITY_X, MESH_FORCE_X, MESH_FORCE_Y, MESH_FORCE_Z)
        out = np.empty(n * n * n if real else 2 * n * n * (n // 2 + 1), np.float64)
        self._ck(lib().mgpu_get_mesh(self.ctx, which, _ptr(out)), "get_mesh")
        if real:
            return out.reshape(n, n, n)
        return out.view(np.complex128).reshape(n, n, n // 2 + 1)

    def set_mesh(self, which, arr):
        a = np.asarray(arr)
        if np.iscomplexobj(a):
            a = np.ascontiguousarray(a, np.complex128).view(np.float64)
        a = np.ascontiguousarray(a, np.float64).ravel()
        self._ck(lib().mgpu_set_mesh(self.ctx, which, _ptr(a)), "set_mesh")

    def enable_timers(self, on=True):
        self._ck(lib().mgpu_enable_timers(self.ctx, 1 if on else 0), "enable_timers")

    def timers(self, reset=False):
        t = Timers()
        self._ck(lib().mgpu_get_timers(self.ctx, C.byref(t), 1 if reset else 0), "get_timers")
        return {n: getattr(t, n) for n, _ in Timers._fields_}

    def sync(self):
        self._ck(lib().mgpu_sync(self.ctx), "sync")


# ---------------------------------------------------------------------------
# host library (libmgpicola.so): paramfile interface, cosmology, the run
# ---------------------------------------------------------------------------
MODELS = {"FOFR": 0, "FOFR_LCDM": 1, "DGP": 2, "BRANSDICKE": 3, "MBETA": 4, "FOFRNU": 5}
G_D, G_dDdy, G_ddDddy, G_D2, G_dD2dy, G_ddD2ddy = range(6)
FIELD_D, FIELD_dDdy, FIELD_ddDddy, FIELD_deltaD = range(4)


class HostParams(C.Structure):
    _fields_ = (
        [(n, C.c_int) for n in ("model", "scale_dependent", "compute_pofk", "gadget_style",
                                "modified_gravity_active", "use_lcdm_growth_factors", "input_pofk_is_for_lcdm",
                                "input_sigma8_is_for_lcdm", "inverted_initial_condition",
                                "amplitude_fixed_initial_condition", "include_screening",
                                "pofk_compute_every_step", "pofk_compute_rsd_pofk", "pofk_nbins", "pofk_bintype",
                                "pofk_subtract_shotnoise")]
        + [(n, C.c_double) for n in ("pofk_kmin", "pofk_kmax", "fofr0", "nfofr", "range_symm", "assb_symm",
                                     "beta_symm", "Rsmooth", "rcH0_DGP", "wBD", "Omegah2", "Omegarh2", "Omegavh2")]
        + [(n, C.c_char * 500) for n in ("OutputDir", "FileBase", "OutputRedshiftFile")]
        + [("NumFilesWrittenInParallel", C.c_int), ("UseCOLA", C.c_int), ("Buffer", C.c_double),
           ("Nmesh", C.c_int), ("Nsample", C.c_int), ("Box", C.c_double), ("Init_Redshift", C.c_double),
           ("Seed", C.c_int), ("SphereMode", C.c_int), ("WhichSpectrum", C.c_int), ("WhichTransfer", C.c_int),
           ("FileWithInputSpectrum", C.c_char * 500), ("FileWithInputTransfer", C.c_char * 500)]
        + [(n, C.c_double) for n in ("Omega", "OmegaBaryon", "HubbleParam", "Sigma8", "PrimordialIndex",
                                     "UnitVelocity_in_cm_per_s", "UnitLength_in_cm", "UnitMass_in_g",
                                     "InputSpectrum_UnitLength_in_cm")]
        + [("Noutputs", C.c_int), ("out_z", C.c_double * 512), ("out_nsteps", C.c_int * 512)]
        + [("nu_include_massive_neutrinos", C.c_int), ("nu_SumMassNuEV", C.c_double),
           ("nu_FilenameTransferInfofile", C.c_char * 500), ("OmegaNu", C.c_double), ("OmegaCDM", C.c_double),
           ("growth_on_gpu", C.c_int), ("fuse_kick_drift", C.c_int), ("write_particles", C.c_int),
           ("device_step", C.c_int)]
    )


class RunStats(C.Structure):
    _fields_ = [("step_seconds", C.c_double), ("steps", C.c_long), ("final_A", C.c_double)]


_hlib = None


def hlib():
    """Load libmgpicola.so (the C host; it links libmgpu.so)."""
    global _hlib
    if _hlib is None:
        if not os.path.exists(LIBHOST):
            raise MGPUError(f"{LIBHOST} not built: run __graft_entry__.build()")
        L = C.CDLL(LIBHOST)
        vp, d = C.c_void_p, C.c_double
        P = C.POINTER(HostParams)
        L.mgp_params_defaults.argtypes = [P, C.c_int]
        L.mgp_params_defaults.restype = None
        L.mgp_read_paramfile.argtypes = [C.c_char_p, P, C.c_char_p, C.c_size_t]
        L.mgp_read_outputs.argtypes = [P, C.c_char_p, C.c_size_t]
        L.mgp_param_tag_count.argtypes = [P]
        L.mgp_param_tag_name.argtypes = [P, C.c_int]
        L.mgp_param_tag_name.restype = C.c_char_p
        L.mgp_cosmo_create.argtypes = [P, C.c_char_p, C.c_size_t]
        L.mgp_cosmo_create.restype = vp
        L.mgp_cosmo_free.argtypes = [vp]
        L.mgp_cosmo_free.restype = None
        for n in ("mgp_hubble", "mgp_dhubbleda", "mgp_coupling", "mgp_mass2", "mgp_beta", "mgp_beta_dgp",
                  "mgp_phi_crit", "mgp_power_spec", "mgp_mg_sigma8_enhancement"):
            getattr(L, n).argtypes = [vp, d]
            getattr(L, n).restype = d
        for n in ("mgp_growth", "mgp_growth_lcdm", "mgp_growth_lcdmfit"):
            getattr(L, n).argtypes = [vp, C.c_int, d]
            getattr(L, n).restype = d
        L.mgp_growth_sd.argtypes = [vp, C.c_int, d, d]
        L.mgp_growth_sd.restype = d
        for n in ("mgp_Sq", "mgp_Sphi"):
            getattr(L, n).argtypes = [vp, d, d, d]
            getattr(L, n).restype = d
        for n in ("mgp_T_nu", "mgp_T_cdm", "mgp_T_b", "mgp_T_tot", "mgp_T_cb", "mgp_munu_1lpt", "mgp_munu_2lpt"):
            getattr(L, n).argtypes = [vp, d, d]
            getattr(L, n).restype = d
        for n in ("mgp_pk_total", "mgp_pk_cb"):
            getattr(L, n).argtypes = [vp, d]
            getattr(L, n).restype = d
        for n in ("mgp_GeffoverG", "mgp_screening_potential", "mgp_screening_density", "mgp_mg_pofk_ratio"):
            getattr(L, n).argtypes = [vp, d, d]
            getattr(L, n).restype = d
        L.mgp_sigma8_norm.argtypes = [vp]
        L.mgp_sigma8_norm.restype = d
        L.mgp_mg_coeffs.argtypes = [vp, d, C.POINTER(MG)]
        L.mgp_mg_coeffs.restype = None
        L.mgp_ic_pk_table.argtypes = [vp, vp]
        L.mgp_ic_pk_table.restype = None
        L.mgp_lpt_tables.argtypes = [vp, C.c_int, d, d, vp, vp]
        L.mgp_lpt_tables.restype = None
        L.mgp_run.argtypes = [P, C.c_int, C.POINTER(RunStats), C.c_char_p, C.c_size_t]
        _hlib = L
    return _hlib


def write_paramfile(path, values, model="FOFR", compute_pofk=True):
    """Write a paramfile holding exactly the tags the configuration expects."""
    hp = HostParams()
    hlib().mgp_params_defaults(C.byref(hp), MODELS[model])
    hp.compute_pofk = 1 if compute_pofk else 0
    names = [hlib().mgp_param_tag_name(C.byref(hp), i).decode() for i in range(hlib().mgp_param_tag_count(C.byref(hp)))]
    missing = [n for n in names if n not in values]
    if missing:
        raise ValueError(f"paramfile values missing: {missing}")
    with open(path, "w") as f:
        f.write("% generated paramfile\n")
        for n in names:
            f.write(f"{n:<36s} {values[n]}\n")
    return names


def read_params(paramfile, model="FOFR", compute_pofk=True):
    hp = HostParams()
    hlib().mgp_params_defaults(C.byref(hp), MODELS[model])
    hp.compute_pofk = 1 if compute_pofk else 0
    err = C.create_string_buffer(4096)
    if hlib().mgp_read_paramfile(paramfile.encode(), C.byref(hp), err, 4096):
        raise ValueError(err.value.decode())
    if hlib().mgp_read_outputs(C.byref(hp), err, 4096):
        raise ValueError(err.value.decode())
    return hp


class Cosmo:
    """Growth factors, time integrals, hooks and P(k) of the host (libmgpicola)."""

    def __init__(self, hp):
        self.hp = hp
        err = C.create_string_buffer(4096)
        self.h = hlib().mgp_cosmo_create(C.byref(hp), err, 4096)
        if not self.h:
            raise MGPUError(err.value.decode())

    def close(self):
        if self.h:
            hlib().mgp_cosmo_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __getattr__(self, name):
        f = getattr(hlib(), "mgp_" + name)
        return lambda *a: f(self.h, *a)

    def mg(self, a):
        m = MG()
        hlib().mgp_mg_coeffs(self.h, a, C.byref(m))
        return m

    def ic_pk_table(self):
        out = np.zeros(n2_table_len(self.hp.Nmesh))
        hlib().mgp_ic_pk_table(self.h, out.ctypes.data)
        return out

    def lpt_tables(self, field, A, AFF):
        g1 = np.zeros(n2_table_len(self.hp.Nmesh))
        g2 = np.zeros_like(g1)
        hlib().mgp_lpt_tables(self.h, field, A, AFF, g1.ctypes.data, g2.ctypes.data)
        return g1, g2


def run(hp, verbose=False):
    st = RunStats()
    err = C.create_string_buffer(4096)
    if hlib().mgp_run(C.byref(hp), 1 if verbose else 0, C.byref(st), err, 4096):
        raise MGPUError(err.value.decode())
    return {"step_seconds": st.step_seconds, "steps": st.steps, "final_A": st.final_A}


KERNELS = ["deposit", "mean", "fft_r2c", "fft_c2r", "fft_col_fwd", "fft_col_inv", "pofk", "spectral", "real",
           "gather", "kick", "drift", "lpt", "readout", "fft_green", "fft_dual", "fft_screen", "fft_epi", "dep_sort",
           "fold", "fft_pot", "fft_pk", "fft_xdens", "fft_xforce"]


def _kt_setup():
    L = lib()
    if not getattr(L, "_kt", False):
        L.mgpu_enable_kernel_timing.argtypes = [C.c_void_p, C.c_int]
        L.mgpu_kernel_timing.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.c_int]
        L.mgpu_kernel_name.restype = C.c_char_p
        L.mgpu_kernel_name.argtypes = [C.c_int]
        L._kt = True
    return L


def enable_kernel_timing(g, on=True):
    g._ck(_kt_setup().mgpu_enable_kernel_timing(g.ctx, 1 if on else 0), "enable_kernel_timing")


def kernel_timing(g, reset=True):
    """{name: (launches, total_ms, total_algorithmic_bytes)} since the last reset."""
    L = _kt_setup()
    out = {}
    for k, name in enumerate(KERNELS):
        n, ms, b = C.c_int(), C.c_double(), C.c_double()
        g._ck(L.mgpu_kernel_timing(g.ctx, k, C.byref(n), C.byref(ms), C.byref(b), 1 if reset else 0), "kernel_timing")
        out[name] = (n.value, ms.value, b.value)
    return out


def _simlib():
    L = hlib()
    if not getattr(L, "_sim", False):
        L.mgp_sim_create.argtypes = [C.POINTER(HostParams), C.c_int, C.c_char_p, C.c_size_t]
        L.mgp_sim_create.restype = C.c_void_p
        for n in ("mgp_sim_ic", "mgp_sim_finished"):
            getattr(L, n).argtypes = [C.c_void_p]
        L.mgp_sim_advance.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.mgp_sim_free.argtypes = [C.c_void_p]
        L.mgp_sim_free.restype = None
        L.mgp_sim_gpu.argtypes = [C.c_void_p]
        L.mgp_sim_gpu.restype = C.c_void_p
        L.mgp_sim_cosmo.argtypes = [C.c_void_p]
        L.mgp_sim_cosmo.restype = C.c_void_p
        L.mgp_sim_error.argtypes = [C.c_void_p]
        L.mgp_sim_error.restype = C.c_char_p
        L.mgp_sim_a.argtypes = [C.c_void_p]
        L.mgp_sim_a.restype = C.c_double
        L.mgp_sim_create_batch.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_char_p, C.c_size_t]
        L._sim = True
    return L


class Sim:
    """The C host's simulation (libmgpicola mgp_sim_*): IC, then one time-loop iteration per advance()."""

    def __init__(self, hp, verbose=False, _handle=None):
        L = _simlib()
        self.hp = hp
        if _handle is not None:
            self.h = _handle
        else:
            err = C.create_string_buffer(4096)
            self.h = L.mgp_sim_create(C.byref(hp), 1 if verbose else 0, err, 4096)
            if not self.h:
                raise MGPUError(err.value.decode())
        # an MGPU view sharing the context (not owning it)
        g = MGPU.__new__(MGPU)
        g.p = None
        g.n, g.ns, g.np = hp.Nmesh, hp.Nsample, hp.Nsample ** 3
        g.ctx = C.c_void_p(L.mgp_sim_gpu(self.h))
        g.close = lambda: None
        self.gpu = g

    @classmethod
    def batch(cls, hps, verbose=False):
        """Runs of one configuration (HostParams differing in seed, pairing and output names) whose contexts
        form one mgpu batch (mgp_sim_create_batch); advance them in lockstep, member 0 first."""
        L = _simlib()
        n = len(hps)
        arr = (C.c_void_p * n)(*[C.cast(C.pointer(hp), C.c_void_p).value for hp in hps])
        out = (C.c_void_p * n)()
        err = C.create_string_buffer(4096)
        if L.mgp_sim_create_batch(arr, n, 1 if verbose else 0, out, err, 4096):
            raise MGPUError(err.value.decode())
        return [cls(hp, _handle=h) for hp, h in zip(hps, out)]

    def _ck(self, rc, what):
        if rc:
            raise MGPUError(f"{what} failed: {hlib().mgp_sim_error(self.h).decode()}")

    def ic(self):
        self._ck(hlib().mgp_sim_ic(self.h), "mgp_sim_ic")

    def advance(self):
        counted = C.c_int()
        self._ck(hlib().mgp_sim_advance(self.h, C.byref(counted)), "mgp_sim_advance")
        return bool(counted.value)

    def set_learned(self, mlp, track=True):
        """Apply the learned multiplier of mlp (an MLP, or None) from the next force
        evaluation on; track logs a, the density feature and the CDM P(k) per step."""
        L = hlib()
        L.mgp_sim_set_learned.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        self._mlp = mlp  # the context keeps the pointer
        self._ck(L.mgp_sim_set_learned(self.h, mlp.h if mlp is not None else None, 1 if track else 0),
                 "mgp_sim_set_learned")

    def log(self):
        """[(a, feature, P, k_mean, modes)] of the tracked force evaluations."""
        L = hlib()
        for n in ("mgp_sim_log_count", "mgp_sim_nbins"):
            getattr(L, n).argtypes = [C.c_void_p]
        L.mgp_sim_log_entry.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.c_void_p]
        nb = L.mgp_sim_nbins(self.h)
        out = []
        for i in range(L.mgp_sim_log_count(self.h)):
            a, f = C.c_double(), C.c_double()
            pk = np.zeros(3 * nb)
            L.mgp_sim_log_entry(self.h, i, C.byref(a), C.byref(f), pk.ctypes.data)
            out.append((a.value, f.value, pk[:nb], pk[nb:2 * nb], pk[2 * nb:]))
        return out

    @property
    def finished(self):
        return bool(hlib().mgp_sim_finished(self.h))

    @property
    def a(self):
        return hlib().mgp_sim_a(self.h)

    def close(self):
        if self.h:
            hlib().mgp_sim_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KSlice(C.Structure):
    _fields_ = [("lnk0", C.c_double), ("dlnk", C.c_double), ("nk", C.c_int),
                ("num", C.POINTER(C.c_double)), ("dnum", C.POINTER(C.c_double)),
                ("den", C.POINTER(C.c_double)), ("dden", C.POINTER(C.c_double)), ("norm", C.c_double)]


class MLP:
    """The learned multiplier's network on the GPU (libmgpu mgpu_mlp_*)."""
    NPARAMS = 8705

    def __init__(self, seed=1):
        L = lib()
        vp = C.c_void_p
        fp = C.POINTER(C.c_float)
        L.mgpu_mlp_create.argtypes = [C.c_ulonglong, C.POINTER(vp)]
        L.mgpu_mlp_destroy.argtypes = [vp]
        L.mgpu_mlp_destroy.restype = None
        L.mgpu_mlp_get_params.argtypes = [vp, fp]
        L.mgpu_mlp_set_params.argtypes = [vp, fp]
        L.mgpu_mlp_forward.argtypes = [vp, fp, C.c_int, fp]
        L.mgpu_mlp_gradient.argtypes = [vp, fp, fp, fp, C.c_int, fp, C.POINTER(C.c_double)]
        L.mgpu_mlp_train.argtypes = [vp, fp, fp, fp, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_double)]
        L.mgpu_set_learned_multiplier.argtypes = [vp, vp, C.c_double, C.c_double]
        L.mgpu_density_feature.argtypes = [vp, C.POINTER(C.c_double)]
        h = vp()
        if L.mgpu_mlp_create(seed, C.byref(h)):
            raise MGPUError(f"mgpu_mlp_create: {L.mgpu_last_error(None).decode()}")
        self.h = h

    @staticmethod
    def _f(a, shape=None):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if shape is not None:
            a = a.reshape(shape)
        return a, a.ctypes.data_as(C.POINTER(C.c_float))

    def _ck(self, rc, what):
        if rc:
            raise MGPUError(f"{what} failed ({rc}): {lib().mgpu_last_error(None).decode()}")

    @property
    def params(self):
        p, pp = self._f(np.zeros(self.NPARAMS))
        self._ck(lib().mgpu_mlp_get_params(self.h, pp), "mgpu_mlp_get_params")
        return p

    @params.setter
    def params(self, v):
        p, pp = self._f(v, (self.NPARAMS,))
        self._ck(lib().mgpu_mlp_set_params(self.h, pp), "mgpu_mlp_set_params")

    def forward(self, x):
        x, xp = self._f(x)
        y, yp = self._f(np.zeros(len(x)))
        self._ck(lib().mgpu_mlp_forward(self.h, xp, len(x), yp), "mgpu_mlp_forward")
        return y

    def gradient(self, x, t, w):
        x, xp = self._f(x)
        t, tp = self._f(t)
        w, wp = self._f(w)
        g, gp = self._f(np.zeros(self.NPARAMS))
        loss = C.c_double()
        self._ck(lib().mgpu_mlp_gradient(self.h, xp, tp, wp, len(x), gp, C.byref(loss)), "mgpu_mlp_gradient")
        return loss.value, g

    def train(self, x, t, w, iters, lr):
        x, xp = self._f(x)
        t, tp = self._f(t)
        w, wp = self._f(w)
        loss = C.c_double()
        self._ck(lib().mgpu_mlp_train(self.h, xp, tp, wp, len(x), iters, lr, C.byref(loss)), "mgpu_mlp_train")
        return loss.value

    def close(self):
        if getattr(self, "h", None):
            lib().mgpu_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Spline2D(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int)] + [(n, C.POINTER(C.c_double)) for n in
                                                     ("x", "y", "z", "zx", "zy", "zxy")]


class GrowthModel(C.Structure):
    _fields_ = [("model", C.c_int), ("mg_active", C.c_int), ("lcdm", C.c_int), ("omega", C.c_double),
                ("fofr0", C.c_double), ("nfofr", C.c_double), ("assb_symm", C.c_double), ("beta_symm", C.c_double),
                ("range_symm", C.c_double), ("omega_nu", C.c_double), ("omega_b", C.c_double),
                ("omega_cdm", C.c_double), ("t_nu", Spline2D), ("t_cdm", Spline2D), ("t_b", Spline2D)]


def growth_tables(model, x0, dx, na, lnk):
    """mgpu_growth_tables: (6, nk, na) array of D1, dD1/dy, 1LPT source, D2, dD2/dy, 2LPT source."""
    L = lib()
    L.mgpu_growth_tables.argtypes = [C.POINTER(GrowthModel), C.c_double, C.c_double, C.c_int,
                                     C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double)]
    lnk = np.ascontiguousarray(lnk, dtype=np.float64)
    out = np.zeros((6, lnk.size, na))
    rc = L.mgpu_growth_tables(C.byref(model), x0, dx, na, lnk.ctypes.data_as(C.POINTER(C.c_double)), lnk.size,
                              out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise MGPUError(f"mgpu_growth_tables failed ({rc}): {L.mgpu_last_error(None).decode()}")
    return out


def lpt_kslices(cosmo, field, A, AFF):
    """(KSlice order 1, KSlice order 2, backing buffer) from the host's scale-dependent tables."""
    L = hlib()
    L.mgp_lpt_kslices.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.POINTER(KSlice),
                                  C.POINTER(KSlice), C.c_void_p]
    buf = np.zeros(8 * 1000)
    o1, o2 = KSlice(), KSlice()
    if L.mgp_lpt_kslices(cosmo.h, field, A, AFF, C.byref(o1), C.byref(o2), buf.ctypes.data):
        raise MGPUError("no scale-dependent growth tables")
    return o1, o2, buf


def lpt_fields_slices(g, slot, o1, o2):
    L = lib()
    L.mgpu_lpt_fields_slices.argtypes = [C.c_void_p, C.c_int, C.POINTER(KSlice), C.POINTER(KSlice)]
    g._ck(L.mgpu_lpt_fields_slices(g.ctx, slot, C.byref(o1), C.byref(o2)), "lpt_fields_slices")


def lpt_fields_slices2(g, s0, s1):
    """Both LPT slots in one call: s0 = (order 1, order 2) slices of slot 0, s1 of slot 1."""
    L = lib()
    L.mgpu_lpt_fields_slices2.argtypes = [C.c_void_p] + [C.POINTER(KSlice)] * 4
    g._ck(L.mgpu_lpt_fields_slices2(g.ctx, C.byref(s0[0]), C.byref(s0[1]), C.byref(s1[0]), C.byref(s1[1])),
          "lpt_fields_slices2")
