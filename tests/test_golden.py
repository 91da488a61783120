"""The oracle's outputs on small fixed inputs against the committed digests
(tests/golden/oracle_digests.json, written by tests/golden/make_golden.py).

A mismatch means the oracle's arithmetic changed: every GPU parity result is
relative to the oracle, so such a change must be deliberate (regenerate the
fixtures in the same change and say why)."""
import importlib.util
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def test_oracle_matches_golden_digests():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    with open(os.path.join(HERE, "golden", "oracle_digests.json")) as fh:
        want = json.load(fh)
    got = mg.compute()
    assert sorted(got) == sorted(want)
    bad = [k for k in want if got[k] != want[k]]
    assert not bad, f"oracle outputs changed: {bad}"
