"""The tuning cache without a GPU: the planner takes a cached tile and K
split for a plan whose reduced problem has an entry (and asked for neither),
save / load round trips through a file at an explicit path, malformed files
are refused as a whole, entries a plan cannot run are ignored, and
tprims_product_autotune refuses bad arguments before touching the GPU."""
import ctypes

import pytest

import tprims_amd as tp
from harness import T, make_plan
from tprims_amd import _capi
from tprims_amd._capi import lib

HEADER = "# tprims tuning cache 1"


@pytest.fixture
def clean_cache():
    tp.tuning_cache_clear()
    yield
    tp.tuning_cache_clear()


def _f64_24():
    e = 24
    cm = tp.col_major([e] * 4)
    return (T([e] * 4, cm, [1, 2, 3, 4]), T([e] * 4, cm, [4, 5, 2, 6]), T([e] * 4, cm, [1, 3, 5, 6]))


def test_cached_entry_drives_the_plan(clean_cache, tmp_path):
    a, b, d = _f64_24()
    plan = make_plan(tp.DType.F64, a, b, d)
    r = plan.report()
    assert r["source"] == 0 and r["tile"] == 7  # the planner's model: the register-direct K-group tile, 32x48
    key = plan.tuning_key()
    assert key.startswith("dt=") and ";m=" in key and " " not in key
    f = tmp_path / "cache.txt"
    f.write_text("%s\n# comment\n\n4 3 %s\n" % (HEADER, key))
    tp.tuning_cache_load(str(f))
    assert tp.tuning_cache_size() == 1
    r = make_plan(tp.DType.F64, a, b, d).report()
    assert r["source"] == 2 and r["tile"] == 4 and r["kernel"].endswith(".splitk3"), r
    # An explicit configuration wins over the cache.
    r = make_plan(tp.DType.F64, a, b, d, config=tp.PlanConfig(tile=1)).report()
    assert r["source"] == 1 and r["tile"] == 1
    # Another precision or dtype of the same shapes has another key.
    assert make_plan(tp.DType.F32, a, b, d).tuning_key() != key
    # Save, clear, load: the same entry back.
    g = tmp_path / "saved.txt"
    tp.tuning_cache_save(str(g))
    lines = g.read_text().splitlines()
    assert lines[0] == HEADER and ("4 3 " + key) in lines
    tp.tuning_cache_clear()
    assert make_plan(tp.DType.F64, a, b, d).report()["source"] == 0
    tp.tuning_cache_load(str(g))
    assert make_plan(tp.DType.F64, a, b, d).report()["tile"] == 4


def test_malformed_files_are_refused_as_a_whole(clean_cache, tmp_path):
    a, b, d = _f64_24()
    key = make_plan(tp.DType.F64, a, b, d).tuning_key()
    bad_header = tmp_path / "h.txt"
    bad_header.write_text("tile split key\n4 1 %s\n" % key)
    with pytest.raises(tp.TprimsError) as e:
        tp.tuning_cache_load(str(bad_header))
    assert e.value.status == 1
    bad_line = tmp_path / "l.txt"
    bad_line.write_text("%s\n4 1 %s\n4 %s\n" % (HEADER, key, key))  # the third line lacks a field
    with pytest.raises(tp.TprimsError) as e:
        tp.tuning_cache_load(str(bad_line))
    assert e.value.status == 1 and ":3:" in str(e.value)
    assert tp.tuning_cache_size() == 0  # nothing merged from the refused file
    split_range = tmp_path / "s.txt"
    split_range.write_text("%s\n4 65 %s\n" % (HEADER, key))
    with pytest.raises(tp.TprimsError):
        tp.tuning_cache_load(str(split_range))
    with pytest.raises(tp.TprimsError):
        tp.tuning_cache_load(str(tmp_path / "missing.txt"))
    assert lib().tprims_tuning_cache_load(None) == 1
    assert lib().tprims_tuning_cache_save(None) == 1
    assert lib().tprims_tuning_cache_size(None) == 1


def test_entries_a_plan_cannot_run_are_ignored(clean_cache, tmp_path):
    """c64 has no 128x64 tile (id 3): a cache entry naming it is skipped and
    the planner's own choice stands."""
    a = T([300, 200], [1, 300], [0, 2])
    b = T([200, 250], [1, 200], [2, 1])
    d = T([300, 250], [1, 300], [0, 1])
    plan = make_plan(tp.DType.C64, a, b, d)
    want = plan.report()
    f = tmp_path / "c.txt"
    f.write_text("%s\n3 0 %s\n" % (HEADER, plan.tuning_key()))
    tp.tuning_cache_load(str(f))
    r = make_plan(tp.DType.C64, a, b, d).report()
    assert r["source"] == 0 and r["tile"] == want["tile"]


def test_autotune_refuses_bad_arguments_without_a_gpu_call(clean_cache):
    a, b, d = _f64_24()
    plan = make_plan(tp.DType.F64, a, b, d)
    one = ctypes.c_double(1.0)
    t, s, ms = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_double(0)
    args = (ctypes.byref(one), 16, 16, ctypes.byref(one), None, 16)
    assert lib().tprims_product_autotune(0, 0, *args, 3, ctypes.byref(t), ctypes.byref(s), ctypes.byref(ms)) == 1
    assert lib().tprims_product_autotune(plan.handle, 0, *args, 0, None, None, None) == 1
    assert lib().tprims_product_autotune(plan.handle, 0, *args, 1001, None, None, None) == 1
    # A null C with a nonzero beta is refused by the item validation.
    assert lib().tprims_product_autotune(plan.handle, 0, ctypes.byref(one), 16, 16, ctypes.byref(one), None, 16, 3,
                                         None, None, None) == 15
    n = ctypes.c_int64(0)
    assert lib().tprims_product_tuning_key(0, None, 0, ctypes.byref(n)) == 1
    assert lib().tprims_product_tuning_key(plan.handle, None, 4, ctypes.byref(n)) == 1
