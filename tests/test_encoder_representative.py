"""The synthetic asset encoder against the reference render encoder on the
same two 128^3 x 120 fields (a smoke256-like density and a liquid320-like
level set; tests/golden/encoder_compare.json, written by
tests/golden/compare_encoders.py from the reference encoder's own output):
the same route for every leaf count, the same packed controls per frame, and
keyframes, keys per control and payload within 10 %.  The synthetic encoder
is re-run here, so a change to it that moves it away from the reference
fails."""
import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import compare_encoders  # noqa: E402

with open(os.path.join(HERE, "golden", "encoder_compare.json")) as fh:
    RECORD = json.load(fh)


@pytest.mark.parametrize("name", sorted(compare_encoders.FIELDS))
def test_synthetic_encoder_matches_the_reference_encoder(name):
    ref = RECORD[name]["reference"]
    ours = compare_encoders.file_stats(compare_encoders.FIELDS[name].encode(threads=8))
    assert ours == RECORD[name]["synthetic"], "the synthetic encoder's output changed; rerun compare_encoders.py"
    assert ours["leaves_by_mode"] == ref["leaves_by_mode"]
    assert ours["controls_per_frame"] == ref["controls_per_frame"]
    assert ours["bytes"] == pytest.approx(ref["bytes"], rel=0.10)
    for stream in ("coarse", "fine"):
        assert ours["keyframes"][stream] == pytest.approx(ref["keyframes"][stream], rel=0.10)
        assert ours["keys_per_control"][stream]["mean_keys"] == pytest.approx(
            ref["keys_per_control"][stream]["mean_keys"], rel=0.10)
