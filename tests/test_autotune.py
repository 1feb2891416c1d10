"""The autotuned geometry table: the header cytom_create compiles in is the
one the committed sweep results generate, every row is a valid geometry,
and (GPU) a context with vertices_per_tile 0 takes its row's shape."""
import json
import os

import pytest

from cytom_amd import autotune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_is_generated_from_the_committed_sweep(tmp_path):
    hj, hh = tmp_path / "t.json", tmp_path / "t.h"
    autotune.install(autotune.TABLE_JSON, str(hj), str(hh))
    assert open(hh).read() == open(autotune.TABLE_H).read()
    assert json.load(open(hj)) == json.load(open(autotune.TABLE_JSON))


def test_rows_are_valid_and_winners_of_their_sweep():
    rows = json.load(open(autotune.TABLE_JSON))
    assert rows
    for r in rows:
        assert autotune.valid_vpt(r["scale"], r["vpt"])
        assert r["chunk_slots"] % 64 == 0 and r["chunk_slots"] <= 1024
        assert 1 <= r["source_bands"] <= 16
        best = min(r["candidates"], key=lambda c: c["device_ms"])
        assert (best["vpt"], best["chunk_slots"], best["source_bands"]) == \
            (r["vpt"], r["chunk_slots"], r["source_bands"])


@pytest.mark.gpu
def test_default_geometry_follows_the_table():
    import cytom_amd as C
    assert C.device_count() > 0, "no HIP device visible"
    rows = {r["scale"]: r for r in json.load(open(autotune.TABLE_JSON))
            if r["edge_factor"] == 16}
    scales = sorted(rows)
    for scale in (12, 16, 20, 22, 27):
        below = [s for s in scales if s <= scale]
        r = rows[below[-1] if below else scales[0]]
        g = C.Graph(1 << scale)
        vpt = min(r["vpt"], 1 << scale)
        while vpt < 16384 and ((1 << scale) // vpt) ** 2 > (1 << 26):
            vpt *= 2
        assert g.vpt == vpt, scale
        t = g.get_tuning()
        # source bands are at most the tile rows
        bands = min(r["source_bands"], (1 << scale) // vpt)
        assert t["chunk_slots"] == r["chunk_slots"] and t["source_bands"] == bands, scale
        g.close()
