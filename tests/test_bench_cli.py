"""bench.py on a host without a GPU: more than one GPU is refused with a message, before anything touches
a device; the roofline traffic it reports is looked up per job order."""
import os
import subprocess
import sys

from conftest import ROOT


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_traffic_is_recorded_for_the_fast_job_orders():
    """The bench line's roofline traffic comes from the committed counter summary, per job order: both
    orders the autotuner picks between on the cave batch have a figure, and an order without one gives
    None rather than another order's bytes."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    algorithmic = 116_634_192  # input 51,520,536 + output 65,113,656 bytes per launch
    for order in ("morton_cost", "morton_tailcost"):
        t = bench.measured_traffic(order)
        assert t is not None and algorithmic < t < 4 * algorithmic, order
    assert bench.measured_traffic("morton_tailcost") < bench.measured_traffic("morton_cost")
    assert bench.measured_traffic("linear") is None
