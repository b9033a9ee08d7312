"""Device extraction vs the exact oracle on the reference's fixtures
(tests/gpu/test_events.py:23-26 complete-oracle parity, test_peeler.py:93-136)."""

import numpy as np
import pytest

from peel_checks import accepted_fixtures, check_extraction, fixture_path, golden_layers

# Unresolved 3D crossings are outside the contour route's input contract
# (tests/gpu/test_events.py:19-20); they need the intersection resolver.
CROSSING = {"crossing", "perspective_crossing"} | {f"random_{i:02}" for i in range(12)}
CASES = [n for n in accepted_fixtures() if n not in CROSSING]


def expected_layers(name):
    layers = golden_layers()
    return layers.get(f"{name}-mesh", layers.get(name))


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_fixture_matches_oracle(name):
    import oracle
    from hip_peel import EventPeeler
    from hip_peel.io import load_scene

    print(f"[fixture] {name}", flush=True)
    arrays, labels, data = load_scene(fixture_path(name))
    host = oracle.load_scene(data)[:4]
    p = EventPeeler(*arrays)
    p.run()
    report = check_extraction(p, host, expected_layers=expected_layers(name))
    # replay resets every counter and reproduces the same labelled regions
    p.run()
    again = check_extraction(p, host, expected_layers=expected_layers(name))
    assert again["polygons"] == report["polygons"]
