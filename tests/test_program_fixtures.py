"""Call-program fixtures of reference scene files (tests/golden/program_*.json,
written by tests/golden/make_program_fixtures.py from the reference's own
scenelib): replaying a scene's recorded builder calls on this library's
scenelib yields exactly the reference's backend call program
(scenes/scenelib.py:528-679: sorted-name ids, NaN / None / -1 / -2
sentinels, step packing), and that program applied through the C ABI
describes the same scene as ours (CPU only)."""

import ctypes
import json
import math
import os

import pytest

import sundog_amd as sd
from sundog_amd import scenelib

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = ["smoke", "01-marble-run", "02-cornell-lume", "features", "06-spot-cascade", "09-ember-shore", "10-suncatcher", "14-toy-factory"]


def _load(name):
    with open(os.path.join(GOLDEN, "program_%s.json" % name)) as f:
        return json.load(f)


def _norm(x):
    if isinstance(x, (list, tuple)):
        return [_norm(v) for v in x]
    if isinstance(x, float) and math.isnan(x):
        return "nan"
    if isinstance(x, bool):
        return x
    if isinstance(x, int):
        return x
    return x


def _replay(calls):
    s = sd.Scene()
    for c in calls:
        getattr(s, c["method"])(*c["args"], **c["kwargs"])
    return s


@pytest.mark.parametrize("name", FIXTURES)
def test_program_equals_reference_program(name):
    fx = _load(name)
    ours = _replay(fx["builder_calls"]).program("scenes")
    assert _norm(ours) == _norm(fx["program"])


def _describe(lib, h):
    need = ctypes.c_size_t(0)
    assert lib.sundog_scene_describe(h, None, 0, ctypes.byref(need)) == 0
    buf = ctypes.create_string_buffer(need.value)
    assert lib.sundog_scene_describe(h, buf, need.value, None) == 0
    return json.loads(buf.value.decode())


@pytest.mark.parametrize("name", FIXTURES)
def test_reference_program_through_the_abi(lib, name):
    fx = _load(name)
    ref_prog = [tuple(c) for c in fx["program"]]
    h_ref = scenelib.apply_program(ref_prog, lib)
    h_ours = scenelib.apply_program(_replay(fx["builder_calls"]).program("scenes"), lib)
    try:
        d_ref, d_ours = _describe(lib, h_ref), _describe(lib, h_ours)
        assert d_ref == d_ours
        assert lib.sundog_scene_validate(h_ref) == 0
        calls = [c[0] for c in ref_prog]
        assert len(d_ref["objects"]) == calls.count("add_object")
        assert len(d_ref["materials"]) == sum(c.startswith("add_material") for c in calls)
    finally:
        lib.sundog_scene_destroy(h_ref)
        lib.sundog_scene_destroy(h_ours)
