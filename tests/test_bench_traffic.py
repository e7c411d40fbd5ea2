"""bench.py reports the counter-measured HBM traffic of the step kernel only
for the schedule the committed profile was taken with: the bytes per launch
depend on the tile size, the tile order and dry-tile skipping."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

KEY = "river_waf_2048000_rcm"


def test_traffic_only_for_the_profiled_schedule():
    with open(os.path.join(ROOT, "profiles", f"traffic_{KEY}.json")) as fh:
        prof = json.load(fh)
    opts = prof["schedule_options"]
    assert bench.traffic_from_profile(KEY, opts) == prof["hbm_bytes_per_launch"]
    assert abs(prof["hbm_bytes_per_launch"] - (prof["fetch_bytes"] + prof["write_bytes"])) < 1.0
    for other in (opts ^ 2, opts ^ 16, opts ^ 8):
        assert bench.traffic_from_profile(KEY, other) is None
    assert bench.traffic_from_profile("river_waf_2048000_random", opts) is None
