"""The block-shape tuning table: lookups by size class, the block formula,
the defaults when tuning is off, and that the committed table was measured
for every knob."""
import pytest

from torchcor_amd import autotune


def test_committed_table_covers_every_knob():
    t = autotune.load_table()
    assert set(t) == set(autotune.CANDIDATES)
    for knob, rows in t.items():
        assert rows and rows[-1]["max_rows"] >= 2 ** 40       # last class open-ended
        assert all(a["max_rows"] < b["max_rows"] for a, b in zip(rows, rows[1:]))
        assert all(r["value"] in autotune.CANDIDATES[knob] for r in rows)


def test_lookup_and_defaults():
    t = autotune.load_table()
    first = t["serve_slices_per_block"][0]
    assert autotune.lookup("serve_slices_per_block", first["max_rows"]) == first["value"]
    autotune.set_enabled(False)
    try:
        for k, v in autotune.DEFAULTS.items():
            assert autotune.lookup(k, 12345) == v
    finally:
        autotune.set_enabled(True)


@pytest.mark.parametrize("n,spb,G", [(1030301, 8, 1024), (1030301, 16, 1000), (1030301, 32, 504),
                                     (531441, 16, 512), (9261, 8, 16), (9261, 32, 5),
                                     (729, 8, 2), (64, 8, 1)])
def test_pcg_blocks_formula(n, spb, G):
    assert autotune.pcg_blocks(n, spb) == G
