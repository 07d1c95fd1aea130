"""mgpu_ic_displacement_fields_noalloc (the allocation-free IC entry point)
against mgpu_ic_displacement_fields through the C ABI: two calls in a row on
one GPU, the second with a smaller mesh (a smaller seed table in the shared
scratch), each bit-identical to the allocating entry point's result."""
import ctypes as C

import numpy as np
import pytest

import mgpicola as M

pytestmark = pytest.mark.gpu


def pk_table(n, box):
    kf = 2 * np.pi / box
    n2 = np.arange(M.n2_table_len(n))
    k = kf * np.sqrt(n2)
    return np.where(n2 > 0, 2e4 * k / (1 + (k / 0.05) ** 2.5), 0.0)


def outputs(g, sd):
    if sd:
        return [g.get_mesh(M.MESH_CDELTA1), g.get_mesh(M.MESH_CDELTA2)]
    return [g.get_lpt(0), g.get_lpt(1)]


def noalloc(g, seed, pk, sphere, fixed, inverted):
    L = M.lib()
    L.mgpu_ic_displacement_fields_noalloc.argtypes = [C.c_void_p, C.c_ulong, C.c_int, C.c_int, C.c_int, C.c_void_p]
    p = np.ascontiguousarray(pk, np.float64)
    rc = L.mgpu_ic_displacement_fields_noalloc(g.ctx, seed, sphere, fixed, inverted, p.ctypes.data)
    if rc:
        raise M.MGPUError(f"noalloc IC failed ({rc}): {L.mgpu_last_error(g.ctx).decode()}")


@pytest.mark.parametrize("sd", [0, 1])
def test_noalloc_ic_matches_allocating_entry_point(sd):
    # (Nmesh, Nsample, sphere, amplitude fixed, inverted): the second call is smaller
    calls = [(64, 32, 0, 0, 0), (48, 24, 1, 1, 1)] if sd else [(96, 48, 0, 1, 0), (32, 16, 1, 0, 1)]
    for n, ns, sphere, fixed, inv in calls:
        box = 2.0 * n
        pk = pk_table(n, box)
        print("call: n", n, "ns", ns, "sd", sd, flush=True)
        ref_ctx = M.MGPU(n, ns, box, 0.3, scale_dependent=sd)
        ref_ctx.ic_displacement_fields(4242, pk, sphere_mode=sphere, amplitude_fixed=fixed, inverted=inv)
        ref = outputs(ref_ctx, sd)
        ref_ctx.close()
        g = M.MGPU(n, ns, box, 0.3, scale_dependent=sd)
        noalloc(g, 4242, pk, sphere, fixed, inv)
        got = outputs(g, sd)
        g.close()
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
        assert np.abs(ref[0]).max() > 0


def test_noalloc_ic_rejects_bad_arguments():
    g = M.MGPU(16, 8, 32.0, 0.3)
    L = M.lib()
    L.mgpu_ic_displacement_fields_noalloc.argtypes = [C.c_void_p, C.c_ulong, C.c_int, C.c_int, C.c_int, C.c_void_p]
    assert L.mgpu_ic_displacement_fields_noalloc(g.ctx, 1, 0, 0, 0, None) != 0
    assert L.mgpu_ic_displacement_fields_noalloc(None, 1, 0, 0, 0, None) != 0
    g.close()
