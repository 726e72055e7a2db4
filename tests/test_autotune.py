"""Host logic of the launch-shape autotuner: table lookup and shape checks."""
import json
import os

import pytest

from fastrrt_amd import autotune


def test_valid_shape():
    assert autotune.valid_shape((4096, 2048, 2048))
    assert autotune.valid_shape((8, 1, 1))
    assert not autotune.valid_shape((12, 2048, 2048))  # not a multiple of 8
    assert not autotune.valid_shape((0, 1, 1))
    assert not autotune.valid_shape((8, 0, 1))
    assert not autotune.valid_shape((8, 1, 65536))
    assert autotune.valid_shape((4096, 2048, 2048, 64))
    assert autotune.valid_shape((4096, 2048, 2048, 192))
    assert not autotune.valid_shape((4096, 2048, 2048, 96))
    assert not autotune.valid_shape((4096, 2048, 2048, 512))
    assert not autotune.valid_shape((4096, 2048))


def test_lookup_takes_the_nearest_side_on_a_log_scale(tmp_path):
    tab = {"entries": [{"side": 128, "best": [1024, 512, 256]}, {"side": 256, "best": [2048, 1024, 1024]},
                       {"side": 512, "best": [8192, 4096, 2048]}]}
    p = tmp_path / "t.json"
    p.write_text(json.dumps(tab))
    assert autotune.lookup(128, str(p)) == (1024, 512, 256)
    assert autotune.lookup(100, str(p)) == (1024, 512, 256)
    assert autotune.lookup(181, str(p)) == (1024, 512, 256)   # below sqrt(128*256) = 181.02
    assert autotune.lookup(182, str(p)) == (2048, 1024, 1024)
    assert autotune.lookup(1024, str(p)) == (8192, 4096, 2048)
    assert autotune.lookup(256, str(tmp_path / "missing.json")) is None


def test_lookup_rejects_an_invalid_table_entry(tmp_path):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"entries": [{"side": 256, "best": [1001, 512, 256]}]}))
    with pytest.raises(ValueError):
        autotune.lookup(256, str(p))


def test_shipped_table_is_well_formed():
    if not os.path.exists(autotune.TABLE):
        pytest.skip("no tuned table shipped")
    with open(autotune.TABLE) as fh:
        tab = json.load(fh)
    assert tab["entries"]
    for e in tab["entries"]:
        assert autotune.valid_shape(e["best"])
        assert e["best_ms"] <= e["default_ms"]
        assert any(m["shape"] == e["best"] for m in e["measured"])
