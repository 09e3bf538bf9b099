"""The autotuner's tuning table (vbt_tuning_* in the C-ABI): the JSON file
round trip and its errors on the CPU; on the GPU the key of an asset, a
timed first pick that later calls take from the table, and a table saved,
cleared and loaded again."""
import json
import os

import pytest

hip = pytest.importorskip("vbt_amd.hip")


def entries(path):
    with open(path) as fh:
        return {e["key"]: (e["tile"], e["ms"]) for e in json.load(fh)["entries"]}


def write_table(path, rows, fmt="vbt-tuning-1"):
    with open(path, "w") as fh:
        json.dump({"format": fmt, "arch": "gfx950",
                   "entries": [{"key": k, "tile": t, "ms": m} for k, (t, m) in rows.items()]}, fh, indent=1)


def test_table_file_round_trip(tmp_path):
    hip.tuning_clear()
    rows = {"p1 nt120 s19 r10.0.0 k2.0.0.3.4.1.0.0.0": (5, 0.1136), "p1 nt96 s22 r10.0.0 k0.0.0.1.4.3.1.0.0": (4, 0.25)}
    src = tmp_path / "in.json"
    write_table(src, rows)
    hip.tuning_load(str(src))
    assert hip.tuning_count() == 2
    out = tmp_path / "out.json"
    hip.tuning_save(str(out))
    assert entries(out) == rows
    # loading merges: a second file adds and overwrites
    write_table(src, {"p1 nt96 s22 r10.0.0 k0.0.0.1.4.3.1.0.0": (1, 0.2), "p1 nt150 s23 r7.0.3 k1.0.0.1.2.5.1.0.0": (5, 2.6)})
    hip.tuning_load(str(src))
    assert hip.tuning_count() == 3
    hip.tuning_save(str(out))
    assert entries(out)["p1 nt96 s22 r10.0.0 k0.0.0.1.4.3.1.0.0"] == (1, 0.2)
    assert not os.path.exists(str(out) + ".tmp")
    hip.tuning_clear()
    assert hip.tuning_count() == 0


def test_table_file_errors(tmp_path):
    hip.tuning_clear()
    with pytest.raises(hip.VbtError, match="cannot read tuning table"):
        hip.tuning_load(str(tmp_path / "missing.json"))
    bad = tmp_path / "bad.json"
    write_table(bad, {"k": (1, 0.1)}, fmt="something-else")
    with pytest.raises(hip.VbtError, match="not a tuning table"):
        hip.tuning_load(str(bad))
    write_table(bad, {"k": (hip.lib().vbt_cache_tile_count(), 0.1)})
    with pytest.raises(hip.VbtError, match="tile out of range"):
        hip.tuning_load(str(bad))
    assert hip.tuning_count() == 0


@pytest.mark.gpu
def test_autotune_through_the_table(tmp_path):
    import golden_io

    hip.tuning_clear()
    a = hip.Asset(golden_io.vbtp("smoke64"))
    b = hip.Asset(golden_io.vbtp("liquid_mini"))
    F = a.dims[3]
    key = a.tuning_key(F)
    assert key == hip.Asset(golden_io.vbtp("smoke64")).tuning_key(F)
    assert key.startswith(f"p1 nt{F} ") and key != b.tuning_key(b.dims[3])
    assert a.tuning_key(F) != a.tuning_key(F // 2)
    tile, measured = a.autotune_table(F, reps=1)
    assert measured and hip.tuning_count() == 1
    a.cache_tile = (tile + 1) % hip.lib().vbt_cache_tile_count()
    again, measured = a.autotune_table(F, reps=1)
    assert (again, measured) == (tile, False) and a.cache_tile == tile
    path = tmp_path / "t.json"
    hip.tuning_save(str(path))
    assert entries(path)[key][0] == tile
    hip.tuning_clear()
    hip.tuning_load(str(path))
    assert a.autotune_table(F, reps=1) == (tile, False)
    # scientific assets have one kernel shape: nothing is recorded
    s = hip.Asset(golden_io.vbtp("sci_mini"))
    s.autotune_table(s.dims[3], reps=1)
    assert hip.tuning_count() == 1
    hip.tuning_clear()
