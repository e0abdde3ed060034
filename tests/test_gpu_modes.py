"""The launch modes a run picks from its environment at mpic_create, each
held to the bars of the default mode:

  MPIC_UNTILED=1       the FP64 plain form's untiled push (markers pushed in
                       load order, no per-step sort; mpic_sort compacts the
                       array by hand) -- against the oracle over 20 steps;
  MPIC_ITEM_MARKERS=64 the split form's tiles cut into work items of at most
                       64 markers, so that the J gather sums several item
                       blocks per tile -- against the oracle with the split
                       form's bars;
  MPIC_TILE_WAVES=2,4  the split form's push with 2 or 4 wavefronts per
                       workgroup sharing one tile's staged fields and J block
                       (LDS atomics from several wavefronts) -- against the
                       oracle with the split form's bars, alone and with
                       64-marker items.

Both variables are read when a run is created, so the tests set them for
the creation only.
"""
import numpy as np
import pytest

import mandor3_amd as ma
from oracle_lib import load_fixture, oracle_from_fixture
from test_gpu_parity import by_id, cells, gpu_from_fixture, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture
def untiled(monkeypatch):
    monkeypatch.setenv("MPIC_UNTILED", "1")


@pytest.fixture
def small_items(monkeypatch):
    monkeypatch.setenv("MPIC_ITEM_MARKERS", "64")


@pytest.fixture(params=["2", "4"])
def tile_waves(request, monkeypatch):
    monkeypatch.setenv("MPIC_TILE_WAVES", request.param)
    return int(request.param)


@pytest.mark.parametrize("case", ["periodic", "mixed"])
def test_untiled_twenty_steps_parity(case, untiled):
    fx = load_fixture(case)
    g = gpu_from_fixture(fx)
    o = oracle_from_fixture(fx)
    for r in range(1, 21):
        dg = g.step(diag=True)
        do, _ = o.step()
        Mg, Mo = by_id(g.markers()), by_id(o.markers())
        assert np.array_equal(Mg["id"], Mo["id"]), (case, r)
        assert np.array_equal(cells(Mg, fx), cells(Mo, fx)), (case, r)
        Eg, Hg = g.fields()
        Eo, Ho = o.fields()
        assert rel_l2(Eg, Eo) <= 1e-10 and rel_l2(Hg, Ho) <= 1e-10, (case, r)
        np.testing.assert_allclose(dg[1:], do[1:], rtol=1e-10, atol=1e-300)


def _split_twenty_steps(case, what):
    from test_gpu_fp32 import FIELD_BAR, ENERGY_BAR, run32
    fx = load_fixture(case)
    g = run32(fx)
    o = oracle_from_fixture(fx)
    worst_f = worst_d = 0.0
    for _ in range(20):
        dg = g.step(diag=True)
        do, _ = o.step()
        Eg, Hg = g.fields()
        Eo, Ho = o.fields()
        worst_f = max(worst_f, rel_l2(Eg, Eo), rel_l2(Hg, Ho))
        worst_d = max(worst_d, max(abs(dg[q] - do[q]) / abs(do[q]) for q in range(1, 6)
                                   if abs(do[q]) > 1e-12))
    print("%s %s: fields %.2e, energies %.2e" % (what, case, worst_f, worst_d))
    assert worst_f <= FIELD_BAR[case] and worst_d <= ENERGY_BAR[case]
    # the mixed box's walls delete markers: the per-cell counts of a sort
    # still add up to the live markers
    counts = g.sort(with_counts=True)
    assert int(counts.sum()) == g.marker_count()
    assert abs(g.marker_count() - o.markers().size) <= 2


@pytest.mark.parametrize("case", ["periodic", "mixed"])
def test_small_items_twenty_steps(case, small_items):
    _split_twenty_steps(case, "64-marker items")


@pytest.mark.parametrize("case", ["periodic", "mixed", "laser_mur"])
def test_tile_waves_twenty_steps(case, tile_waves):
    _split_twenty_steps(case, "%d wavefronts per tile" % tile_waves)


@pytest.mark.parametrize("case", ["periodic", "mixed"])
def test_tile_waves_with_small_items(case, tile_waves, small_items):
    _split_twenty_steps(case, "%d wavefronts per 64-marker item" % tile_waves)


def _repeated_sorts(prec):
    from test_gpu_fp32 import test_repeated_sorts_after_wall_deletions_keep_the_live_count as t
    t(prec)


def test_untiled_repeated_sorts_keep_the_live_count(untiled):
    _repeated_sorts(ma.FP64)


def test_small_items_repeated_sorts_keep_the_live_count(small_items):
    _repeated_sorts(ma.FP32_SPLIT)
