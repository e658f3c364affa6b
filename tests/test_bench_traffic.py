"""bench.py's roofline.traffic source: the committed rocprofv3 --pmc
summaries under profiles/ parse, and the one profiled at the run's own
launch parameters is chosen (else the first).  CPU only."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_traffic_follows_the_launch_parameters():
    b = _bench()
    paths = [os.path.join(ROOT, "profiles", f) for f in ("k_gather_c3_pmc.txt", "k_gather_c3_pmc_gwc24.txt")]
    base = dict(small_m=12, chunk_shells=64, coop_m=1 << 20, gather_waves_cu=16, post_width=1, lane_walk=0)
    seen = {}
    for gwc in (16, 24):
        tuned = dict(base, gather_waves_cu=gwc)
        traffic, tune = b.read_traffic(paths, "persistent", "k_gather", tuned)
        assert tune == tuned
        # between the algorithmic bytes (16.7 GB) and three times them
        assert 16.7e9 < traffic < 5e10
        seen[gwc] = traffic
    traffic, tune = b.read_traffic(paths, "persistent", "k_gather", dict(base, gather_waves_cu=32))
    assert tune["gather_waves_cu"] == 16 and traffic == seen[16]
    assert b.read_traffic(paths, "plain", "k_gather", base) == (None, None)
    assert b.read_traffic([os.path.join(ROOT, "profiles", "missing.txt")], "persistent", "k_gather", base) == (None, None)
