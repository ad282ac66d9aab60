"""The spreading tilings a grid admits (a host-side query of the library,
no GPU) and the autotuner's committed table."""
import json
import os

import pytest

from mpidopenmmplugin_amd import _capi, autotune


def test_tilings_of_the_benchmark_grids():
    # 128^3: 16 tiles of 8 points per axis, owned cores first (the built-in choice)
    assert _capi.spread_tilings((128, 128, 128)) == [(4, 1), (4, 2), (-1, 0)]
    # 40^3: 4 tiles of 10 points; their windows (18 points) would not fit in LDS
    assert _capi.spread_tilings((40, 40, 40)) == [(2, 1), (-1, 0)]
    # 24^3: tiles of 6 points take windows only; owned cores at 2 tiles of 12
    assert _capi.spread_tilings((24, 24, 24)) == [(2, 2), (1, 1), (-1, 0)]
    # an odd size has no tiling
    assert _capi.spread_tilings((15, 15, 15)) == [(-1, 0)]


def test_every_tiling_meets_the_kernels_limits():
    for g in (15, 16, 18, 20, 24, 25, 27, 30, 32, 36, 40, 48, 64, 96, 128):
        for level, mode in _capi.spread_tilings((g, g, g)):
            if level < 0:
                assert mode == 0
                continue
            T = g >> level
            assert g == T << level and T >= 6
            assert mode in (1, 2)
            if mode == 1:
                assert T >= 8  # a stencil meets two cores per axis at most


def test_grids_outside_the_pme_path_are_refused():
    with pytest.raises(ValueError):
        _capi.spread_tilings((7, 7, 7))
    with pytest.raises(ValueError):
        _capi.spread_tilings((4, 16, 16))


def test_options_carry_the_tiling():
    flat = {"nonbonded_method": 1, "polarization": 0, "cutoff": 0.8, "alpha": 3.0, "grid": (24, 24, 24),
            "default_thole": 0.39, "scale14": 1.0, "max_iter": 60, "target_eps": 1e-6}
    o = _capi.make_options(flat, 0, 0, (1, 1))
    assert (o.spread_level, o.spread_mode) == (1, 1)
    o = _capi.make_options(flat, 0, 0)
    assert (o.spread_level, o.spread_mode) == (0, 0)


def test_committed_table_is_consistent():
    if not os.path.exists(autotune.TABLE_PATH):
        pytest.skip("no tuned table (python -m mpidopenmmplugin_amd.autotune on the GPU writes it)")
    with open(autotune.TABLE_PATH) as fh:
        table = json.load(fh)
    assert table["entries"]
    for key, e in table["entries"].items():
        grid = tuple(int(x) for x in key.split(","))
        pick = (e["level"], e["mode"])
        assert pick in _capi.spread_tilings(grid), key
        times = e["spread_ms_per_evaluation"]
        assert f"{pick[0]},{pick[1]}" == min(times, key=times.get), key
        assert set(times) == {f"{lv},{md}" for lv, md in _capi.spread_tilings(grid)}, key
        assert set(e["ms_per_evaluation"]) == set(times), key
        assert autotune.lookup(grid) == pick
        assert e["energy_rel_spread"] < 1e-5, key


def test_lookup_of_an_untuned_grid_is_none():
    assert autotune.lookup((6, 6, 6)) is None
