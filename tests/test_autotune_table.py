"""The autotuner's table (phare_amd/autotune.py): entries keyed by dimension, order, precision
family and particles per cell; lookup picks the nearest particles per cell within a factor of 2."""
import json

from phare_amd import autotune
from phare_amd import workloads as W


def test_table_write_and_lookup(tmp_path):
    path = str(tmp_path / "table.json")
    wl = W.harris2d(ncells=(16, 8), ppc=100)
    best = dict(segment_shape=[4, 16, 1], rows_per_item=24, tile_row_pad=16, sort_interval=30,
                tile_copies=8)
    autotune.write_entry(wl, "fp64", best, 0.3, path=path)
    autotune.write_entry(wl, "fp32", dict(best, rows_per_item=48), 0.2, path=path)
    # a second write of the same key replaces the entry
    autotune.write_entry(wl, "fp64", dict(best, rows_per_item=16), 0.25, path=path)
    table = json.load(open(path))
    assert len(table) == 2
    assert autotune.lookup(2, 2, "fp64", 100, path=path)["rows_per_item"] == 16
    assert autotune.lookup(2, 2, "fp32", 70, path=path)["rows_per_item"] == 48
    assert autotune.lookup(2, 2, "fp64", 20, path=path) is None   # more than 2x away
    assert autotune.lookup(3, 2, "fp64", 100, path=path) is None
    assert autotune.lookup(2, 2, "fp64", 100, path=str(tmp_path / "none.json")) is None


def test_committed_table_is_well_formed():
    import os

    if not os.path.exists(autotune.TABLE):
        return
    for e in json.load(open(autotune.TABLE)):
        assert {"dim", "order", "precision", "ppc", "tuning"} <= set(e)
        t = e["tuning"]
        assert t["rows_per_item"] >= 1 and t["sort_interval"] >= 0
        assert t["tile_copies"] in (1, 2, 4, 8)
