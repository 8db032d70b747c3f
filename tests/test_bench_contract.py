"""The bench's reporting rules that need no GPU: the profiled DRAM traffic
is reported only for the launch shapes it was profiled at, and a run on
more than one GPU is refused with a message."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

TUNED = {"slab_cols": 0, "blocks_per_cu": 4, "xcd_split": 0, "l2_budget_bytes": 0, "stream_nt": 0, "use_tiled": 0,
         "chunk_rows": 0, "chunk_period": 0, "row_lane_bytes": 0, "row_slab_cols": 0, "kubo_stack_mib": 0,
         "gamma_tile_rows": 0}


def test_traffic_reported_for_the_profiled_shapes_only():
    b0, src0 = bench.step_traffic(1024, 64, TUNED)
    b1, _ = bench.step_traffic(1024, 64, dict(TUNED, xcd_split=-1))
    assert b0 is not None and b1 is not None and b1 > b0 > 6.55e9
    assert "profiles/" in src0 and os.path.exists(os.path.join(ROOT, src0.split("(")[1].rstrip(")")))
    for other in (dict(TUNED, blocks_per_cu=8), dict(TUNED, use_tiled=1), dict(TUNED, chunk_rows=512)):
        b, why = bench.step_traffic(1024, 64, other)
        assert b is None and "not profiled" in why
    assert bench.step_traffic(512, 64, TUNED) == (None, None)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2 and "one GPU" in r.stderr
