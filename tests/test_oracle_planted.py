"""Each pin of the oracle must catch a plausible slip.  For every planted
mistake listed in oracle/oracle.h, a copy of the oracle is compiled with that
one line wrong (gcc -DOR_PLANT=k), the named pins of test_oracle_pins.py are
run against it, and at least one of them must fail; the same pins pass on the
normal build (test_oracle_pins.py itself).  CPU only."""
import numpy as np
import pytest

import oracle as O
import test_oracle_pins as P

# plant -> (what is wrong, pins that must catch it: (function, kwargs))
PLANTS = {
    1: ("force Nyquist index -N/2", [(P.test_force_nyquist_is_plus_half_n, {})]),
    2: ("2LPT factor +3/7", [(P.test_2lpt_two_orthogonal_waves, {})]),
    3: ("P(k) weight of kz > 0 modes", [(P.test_pofk_plane_wave, dict(m=(2, -1, 2))),
                                        (P.test_pofk_kmean_is_mode_average, {})]),
    4: ("ghost slab not folded", [(P.test_cic_ghost_slab_wraps_x, {})]),
    5: ("f(R) mass term without a^2", [(P.test_fofr_unscreened_multiplier_physical_units, {})]),
    6: ("IC delta ~ P", [(P.test_ic_delta_fixed_amplitude_cut_and_symmetry, dict(n=16, ns=8, sphere=0))]),
    7: ("IC: the skipped k = 0 mode draws nothing", [(P.test_ic_delta_phase_is_column_stream_draw,
                                                      dict(amp_fixed=0))]),
    8: ("Kick COLA term sign", [(P.test_kick_cola_identity_zero_residual, {}),
                                (P.test_zeldovich_cola_run_follows_growth, {})]),
    9: ("Drift 2LPT growth from order 1", [(P.test_drift_cola_moves_with_lpt_growth, {})]),
    10: ("LPT growth index not mirrored", [(P.test_lpt_field_plane_waves_with_scale_dependent_growth, {})]),
    11: ("seed table shell loops swapped", [(P.test_seedtable_draw_order_by_hand, {})]),
}


@pytest.fixture(scope="module")
def plant_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("planted"))


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_fails_a_pin(plant, plant_dir):
    what, pins = PLANTS[plant]
    path = O.build_planted(plant, plant_dir)
    caught = []
    with O.using(path):
        for fn, kw in pins:
            try:
                fn(**kw)
            except AssertionError:
                caught.append(fn.__name__)
    print(f"plant {plant} ({what}) caught by {caught}")
    assert caught, f"planted mistake {plant} ({what}) passes every pin"


def test_unplanted_build_passes_the_same_pins(plant_dir):
    """Control: the pins used above pass on a build with OR_PLANT=0 compiled the same way."""
    path = O.build_planted(0, plant_dir)
    with O.using(path):
        for _, pins in PLANTS.values():
            for fn, kw in pins:
                fn(**kw)
    assert np.isfinite(1.0)
