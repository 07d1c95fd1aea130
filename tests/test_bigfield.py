"""The separable test fields of bigfield.py against numpy's FFT and the oracle's
Forces at small sizes (CPU): the exact half spectrum equals rfftn of the field,
and the Forces source built for each axis gives back N^3 t(x) in the oracle's
Green's function + c2r (src/auxPM.c:437-555), Nyquist planes included."""
import numpy as np
import pytest

import oracle as O
from bigfield import Separable, compare


@pytest.mark.parametrize("n", [8, 12, 16, 24])
def test_spectrum_is_rfftn(n):
    s = Separable(n, rank=3, seed=n)
    ref = np.fft.rfftn(s.real())
    np.testing.assert_allclose(s.spectrum(), ref, rtol=0, atol=1e-10 * np.abs(ref).max())
    assert compare(ref, s.spectrum_slab, n)[0] < 1e-12


@pytest.mark.parametrize("n", [8, 12, 16])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_force_source_gives_field_back_in_oracle(n, axis):
    box = 3.0 * n
    s = Separable(n, rank=2, seed=10 * n + axis, zero_mean_axes=(axis,))
    dk = O.new_mesh(n)
    O.complex_view(dk, n)[:] = s.force_source(axis, box)
    f = O.forces(O.cfg(n, n, box, 0.3), dk, None)
    got = O.real_view(f[axis], n)
    expect = s.real(float(n) ** 3)
    assert np.abs(got - expect).max() < 1e-9 * np.abs(expect).max()
    # the other components are not the field: the source is specific to its axis
    other = O.real_view(f[(axis + 1) % 3], n)
    assert np.abs(other - expect).max() > 1e-2 * np.abs(expect).max()
