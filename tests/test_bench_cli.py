"""Host-side pieces of bench.py that run without a GPU: the one-GPU guard and
the traffic figures read from the committed rocprofv3 summaries."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one MI355X" in r.stderr and "--gpus 2" in r.stderr


def test_step_traffic_from_committed_profiles():
    """Bytes per step = the sum over the step's kernels of launches x
    (read + write) per launch, over the profiled steps (stage updates / 3)."""
    b = _bench()
    for prec in ("fp64", "mixed"):
        total, src = b.step_traffic("iso512", prec)
        assert src == "profiles/stage_update_iso512_%s.md" % prec
        # recompute by hand from the table
        rows = []
        for line in open(os.path.join(ROOT, src)):
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cols) >= 5 and cols[0].startswith("`k_") and not cols[0].startswith("`k_hyb_counts") \
                    and not cols[0].startswith("`k_copy"):
                rows.append((cols[0], int(cols[1]), float(cols[3]) + float(cols[4])))
        steps = sum(n for name, n, _ in rows if "visc_div_update" in name or "visc_fused" in name) / 3.0
        want = sum(n * mb for _, n, mb in rows) * 1e6 / steps
        assert abs(total - want) <= 1e-9 * want
        # the step moves several times the 360 B per cell-step it must
        assert 2.0 < total / (512 ** 3 * 360.0) < 12.0
    assert b.step_traffic("tgv64", "fp64") == (None, None)
