"""The C ABI without a GPU: the library loads, exports every entry point
declared in include/*.h, validates configurations and reports errors as
status codes (the reference exits instead, cuda_util.cuh:19-27)."""
import ctypes
import re
from pathlib import Path

import pytest

from glint_pipeline_amd import runtime

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:[a-z_]+\s+\**)+\**(glint_\w+)\s*\(", text, flags=re.M))
    return names


def test_every_declared_symbol_is_exported_and_bound():
    lib = runtime.lib()
    names = declared_functions()
    assert len(names) > 30
    for n in sorted(names):
        assert hasattr(lib, n), f"{n} is declared but not exported"
    assert names <= set(runtime.SIGNATURES), sorted(names - set(runtime.SIGNATURES))


def test_status_strings():
    seen = set()
    for code in range(11):
        s = runtime.lib().glint_status_string(code)
        assert s and s not in seen
        seen.add(s)


def test_config_defaults_follow_the_reference():
    c = runtime.Config()
    runtime.lib().glint_config_default(ctypes.byref(c))
    # config.h:29,38,39,40,41,35
    assert (c.grid_capacity, c.schema_update_delay, c.min_meet_time, c.max_meet_time, c.dynamic_schema) == \
        (50, 1, 5, 4096, 1)
    assert c.reach_distance == 2.0


@pytest.mark.parametrize("field,value", [("num_objects", 0), ("grid_capacity", 0), ("reach_distance", -1.0),
                                         ("high_x", -5.0), ("candidate_mode", 7), ("precision", 9),
                                         ("dynamic_schema", 3)])
def test_invalid_configuration_is_rejected(field, value):
    c = runtime.make_config(100, (0, 0, 100, 100))
    setattr(c, field, value)
    h = ctypes.c_void_p()
    st = runtime.lib().glint_create(ctypes.byref(c), ctypes.byref(h))
    assert st == 1 and not h.value
    assert runtime.lib().glint_last_error_message(None)


def test_null_arguments():
    L = runtime.lib()
    assert L.glint_partition(None) == 1
    assert L.glint_destroy(None) == 1
    assert L.glint_set_positions(None, None, 0, 0) == 1


def test_reach_threshold_host_only():
    t = runtime.lib().glint_reach_threshold(2.0)
    assert 4.0 <= t < 4.0 + 1e-14
