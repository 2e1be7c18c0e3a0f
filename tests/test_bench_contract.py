"""bench.py's reporting contract, checked without a GPU: the shipped tuning
table names a 2D kernel variant whose HBM traffic is in the committed PMC
summary, so the headline line always carries `roofline.traffic`."""
import importlib.util
import json
import math
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)   # stdlib imports only at module level
    return mod


def test_shipped_tuning_variant_has_profiled_traffic():
    b = _bench()
    table = json.load(open(b.TUNING_TABLE))
    key = f"bleq/single2d/{int(math.floor(math.log2(8192 * 8192)))}"
    assert key in table, "the shipped table covers the 8192^2 size class"
    tile, flatten = (int(v) for v in table[key][:2])
    traffic = b.traffic_from_profiles("bleq2d_8192_traffic.json", f"{tile},{flatten}")
    assert traffic is not None
    # at least the algorithmic bytes (1 B in + 4 B out per pixel), at most twice them
    assert 5 * 8192 * 8192 <= traffic <= 10 * 8192 * 8192


def test_traffic_is_none_for_unprofiled_variant():
    b = _bench()
    assert b.traffic_from_profiles("bleq2d_8192_traffic.json", "6,1") is None
    assert b.traffic_from_profiles("no_such_file.json", "3,0") is None


def test_roofline_fields():
    b = _bench()
    r = b.roofline(8192 * 8192, 130.0, 123, "x")
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert abs(r["achieved"] - 5 * 8192 * 8192 / 130e-6 / 1e9) < 0.1
    assert abs(r["frac"] - r["achieved"] / 8000.0) < 1e-3
