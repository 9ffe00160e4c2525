"""Scene-construction ABI on the CPU (no GPU calls).

Behaviour specs restated from the reference's host tests
(tests/host/test_scene_build.cpp:59-734): defaults, call-for-call scene
equivalence, rejection paths, light-list order and phases, flames, water,
plastic, envmap, mesh lights, mirroring, the material cap and the camera.
State is read back through sundog_scene_describe (JSON).
"""

import ctypes
import json
import math
import os
import re

import pytest

import sundog_amd
from sundog_amd import _lib

NAN = float("nan")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def D3(x, y, z):
    return (ctypes.c_double * 3)(x, y, z)


def D2(x, y):
    return (ctypes.c_double * 2)(x, y)


def steps(*s):
    arr = (_lib.XformStep * max(len(s), 1))()
    for i, (k, a, b, c) in enumerate(s):
        arr[i] = _lib.XformStep(k, a, b, c)
    return arr, len(s)


SCALE, TRANSLATE, RX, RY, RZ = 0, 1, 2, 3, 4
SPHERE, RECT, DISK, CYL, PARA, MESH = 0, 1, 2, 3, 4, 5
NONE, DEFAULT = -1, -2


def describe(lib, h):
    need = ctypes.c_size_t(0)
    assert lib.sundog_scene_describe(h, None, 0, ctypes.byref(need)) == 0
    buf = ctypes.create_string_buffer(need.value)
    assert lib.sundog_scene_describe(h, buf, need.value, None) == 0
    return json.loads(buf.value.decode())


def add_obj(lib, h, kind, mat, back=DEFAULT, st=(), mesh=-1, cutout=-1, nee=-1):
    arr, n = steps(*st)
    return lib.sundog_add_object(h, kind, mesh, mat, back, cutout, arr, n, nee, None)


def expect_add_fail(lib, rc):
    assert rc == -1
    assert lib.sundog_last_error()


def expect_set_fail(lib, rc):
    assert rc != 0
    assert lib.sundog_last_error()


@pytest.fixture
def mini(lib):
    h = lib.sundog_scene_create(b"scenes")
    assert h
    assert lib.sundog_set_camera(h, D3(0, 1, 5), D3(0, 0, 0), None, NAN, NAN, NAN) == 0
    assert lib.sundog_add_material_lambert(h, D3(0.5, 0.5, 0.5), -1) == 0
    yield h
    lib.sundog_scene_destroy(h)


def test_every_declared_symbol_is_exported(lib):
    names = set()
    for hdr in ("sundog_api.h",):
        src = open(os.path.join(ROOT, "include", hdr)).read()
        names |= set(re.findall(r"SUNDOG_API\s+[^;(]*?\b(sundog_\w+)\s*\(", src))
    assert len(names) >= 40
    for n in sorted(names):
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, "ctypes binding missing for %s" % n


def test_minimal_defaults(lib, mini):
    assert add_obj(lib, mini, SPHERE, 0) == 0
    assert lib.sundog_scene_validate(mini) == 0
    d = describe(lib, mini)
    assert d["render"]["width"] == 1280 and d["render"]["height"] == 720
    assert d["render"]["spp"] == 64 and d["render"]["max_depth"] == 16
    assert d["render"]["clamp"] == 50 and d["render"]["seed"] == 7 and d["render"]["chunk"] == 16
    assert d["render"]["transparent_shadows"] is True
    assert len(d["objects"]) == 1 and d["lights"] == []
    assert d["base_dir"] == "scenes"
    h2 = lib.sundog_scene_create(None)
    assert lib.sundog_scene_validate(h2) != 0
    assert b"camera" in lib.sundog_last_error()
    assert describe(lib, h2)["base_dir"] == "."
    lib.sundog_scene_destroy(h2)
    h3 = lib.sundog_scene_create(None)
    assert lib.sundog_set_render(h3, -1, -1, -1, -1, NAN, -1, NAN, 0) == 0
    assert describe(lib, h3)["render"]["transparent_shadows"] is False
    lib.sundog_scene_destroy(h3)


def test_smoke_scene_equivalent(lib):
    h = lib.sundog_scene_create(b"scenes")
    assert lib.sundog_set_render(h, 256, 256, 16, 8, NAN, 7, NAN, -1) == 0
    assert lib.sundog_set_camera(h, D3(0, 1.5, 5), D3(0, 0.7, 0), None, 35.0, NAN, NAN) == 0
    assert lib.sundog_set_background_gradient(h, D3(1, 1, 1), D3(0.4, 0.6, 1.0)) == 0
    assert lib.sundog_add_texture_checker(h, D3(0.85, 0.85, 0.85), D3(0.15, 0.15, 0.15), D2(8, 8)) == 0
    ball = lib.sundog_add_material_lambert(h, D3(0.7, 0.3, 0.3), -1)
    ground = lib.sundog_add_material_lambert(h, None, 0)
    assert (ball, ground) == (0, 1)
    assert add_obj(lib, h, RECT, ground, st=[(SCALE, 4, 4, 4)]) == 0
    assert add_obj(lib, h, SPHERE, ball, st=[(SCALE, 0.7, 0.7, 0.7), (TRANSLATE, 0, 0.7, 0)]) == 1
    assert lib.sundog_add_point_light(h, D3(3, 4, 2), D3(40, 40, 40), 0.4) == 0
    assert lib.sundog_scene_validate(h) == 0
    d = describe(lib, h)
    r = d["render"]
    assert (r["width"], r["height"], r["spp"], r["max_depth"], r["seed"]) == (256, 256, 16, 8, 7)
    assert d["camera"]["lookfrom"][1] == pytest.approx(1.5, abs=1e-6)
    assert d["camera"]["vfov"] == pytest.approx(35.0)
    assert d["background"]["kind"] == 1
    assert d["background"]["a"][0] == pytest.approx(1.0) and d["background"]["b"][2] == pytest.approx(1.0)
    assert len(d["objects"]) == 2 and len(d["lights"]) == 1 and len(d["materials"]) == 2
    assert d["textures"][0]["kind"] == 2 and d["textures"][0]["scale"][0] == pytest.approx(8)
    assert d["materials"][1]["tex_id"] == 0 and d["materials"][0]["color"][0] == pytest.approx(0.7)
    rect, sph = d["objects"]
    assert rect["geom_kind"] == RECT and sph["geom_kind"] == SPHERE
    assert rect["xform"][0] == pytest.approx(4.0, abs=1e-5)   # x axis scaled by 4
    assert sph["xform"][7] == pytest.approx(0.7, abs=1e-5)    # translated to y = 0.7
    assert math.sqrt(sum(sph["xform"][i] ** 2 for i in (0, 4, 8))) == pytest.approx(0.7, abs=1e-5)
    assert rect["light_id"] == -1 and sph["light_id"] == -1
    l0 = d["lights"][0]
    assert l0["kind"] == 3 and l0["radius"] == pytest.approx(0.4) and l0["p"][0] == pytest.approx(3.0)
    assert l0["L"][0] == pytest.approx(40.0)
    assert rect["mat_front"] == rect["mat_back"] != 0xFFFF
    lib.sundog_scene_destroy(h)


def test_features_scene_equivalent(lib):
    h = lib.sundog_scene_create(b"scenes")
    assert lib.sundog_set_render(h, 1920, 1080, 64, 16, NAN, 7, NAN, -1) == 0
    assert lib.sundog_set_camera(h, D3(0, 3.2, 9), D3(0, 0.9, 0), None, 38.0, 0.02, NAN) == 0
    assert lib.sundog_add_texture_checker(h, D3(0.8, 0.8, 0.8), D3(0.25, 0.25, 0.3), D2(12, 12)) == 0
    ids = [lib.sundog_add_material_dielectric(h, 1.5, None, None, -1, NAN),
           lib.sundog_add_material_metal(h, D3(1.0, 0.78, 0.34), -1, 0.15),
           lib.sundog_add_material_lambert(h, None, 0),
           lib.sundog_add_material_emissive(h, D3(1.0, 0.95, 0.85), -1, 20.0, -1),
           lib.sundog_add_material_metal(h, D3(0.9, 0.9, 0.9), -1, 0.0),
           lib.sundog_add_material_lambert(h, D3(0.75, 0.25, 0.25), -1)]
    assert ids == list(range(6))
    glass, gold, ground, lamp, mirror, red = ids
    assert add_obj(lib, h, RECT, ground, st=[(SCALE, 8, 8, 8)]) == 0
    add_obj(lib, h, SPHERE, glass, st=[(SCALE, .8, .8, .8), (TRANSLATE, -3, .8, 1)])
    add_obj(lib, h, SPHERE, red, st=[(SCALE, .8, .8, .8), (TRANSLATE, -1, .8, 1)])
    add_obj(lib, h, CYL, gold, st=[(SCALE, .6, .8, .6), (TRANSLATE, 1, .8, 1)])
    assert add_obj(lib, h, PARA, mirror, back=NONE,
                   st=[(SCALE, 1.4, 1.4, 1.4), (RZ, 90, 0, 0), (TRANSLATE, 3.4, 1.4, 0)]) == 4
    add_obj(lib, h, DISK, mirror, st=[(SCALE, 1.6, 1.6, 1.6), (RX, 90, 0, 0), (TRANSLATE, 0, 1.6, -3.2)])
    assert add_obj(lib, h, RECT, lamp, st=[(RX, 180, 0, 0), (SCALE, 1.2, 1.2, 1.2), (TRANSLATE, 0, 6, 2)]) == 6
    assert lib.sundog_add_point_light(h, D3(5, 5, 4), D3(30, 30, 30), 0.3) >= 0
    assert lib.sundog_add_distant_light(h, D3(-0.5, -1, -0.3), D3(0.6, 0.6, 0.55)) >= 0
    assert lib.sundog_scene_validate(h) == 0
    d = describe(lib, h)
    assert len(d["objects"]) == 7 and len(d["materials"]) == 6
    kinds = [m["kind"] for m in d["materials"]]
    assert kinds.count(0) == 2 and kinds.count(1) == 2 and kinds.count(2) == 1 and kinds.count(3) == 1
    par = d["objects"][4]
    assert par["mat_back"] == 0xFFFF and d["materials"][par["mat_front"]]["roughness"] == 0.0
    lampo = d["objects"][6]
    assert lampo["light_id"] == 0
    L = d["lights"]
    assert len(L) == 3
    assert L[0]["kind"] == 0 and L[0]["area"] == pytest.approx(5.76, abs=1e-4)
    assert L[0]["n"][1] == pytest.approx(-1.0, abs=1e-5)
    assert L[0]["two_sided"] == 0 and L[0]["L"][0] == pytest.approx(20.0, abs=1e-4)
    assert L[0]["p"][1] == pytest.approx(6.0, abs=1e-5)
    assert L[1]["kind"] == 3 and L[1]["radius"] == pytest.approx(0.3)
    assert L[2]["kind"] == 4 and math.sqrt(sum(c * c for c in L[2]["dir"])) == pytest.approx(1, abs=1e-5)
    assert L[2]["dir"][1] < 0
    assert all(l["flame_id"] == -1 for l in L)
    assert d["camera"]["aperture"] == pytest.approx(0.02)
    lib.sundog_scene_destroy(h)


def test_error_paths_leave_no_side_effects(lib, mini):
    expect_add_fail(lib, lib.sundog_add_material_lambert(mini, None, 7))
    expect_add_fail(lib, lib.sundog_add_material_emissive(mini, D3(1, 1, 1), -1, -2.0, -1))
    expect_add_fail(lib, add_obj(lib, mini, 99, 0))
    expect_add_fail(lib, add_obj(lib, mini, MESH, 0, mesh=3))
    expect_add_fail(lib, add_obj(lib, mini, SPHERE, 42))
    expect_add_fail(lib, add_obj(lib, mini, SPHERE, NONE, back=NONE))
    expect_add_fail(lib, add_obj(lib, mini, SPHERE, 0, cutout=5))
    expect_add_fail(lib, add_obj(lib, mini, SPHERE, 0, st=[(99, 1, 1, 1)]))
    expect_add_fail(lib, lib.sundog_add_point_light(mini, D3(0, 5, 0), D3(10, -1, 10), 0.0))
    expect_add_fail(lib, lib.sundog_add_distant_light(mini, D3(0, -1, 0), D3(-0.1, 1, 1)))
    d = describe(lib, mini)
    assert d["objects"] == [] and d["lights"] == [] and len(d["materials"]) == 1

    h = lib.sundog_scene_create(None)
    g = lib.sundog_add_material_dielectric(h, NAN, D3(0.4, 0.05, 0.02), None, -1, NAN)
    p = lib.sundog_add_material_dielectric(h, NAN, None, None, -1, NAN)
    fr = lib.sundog_add_material_dielectric(h, NAN, None, None, -1, 0.35)
    assert (g, p, fr) == (0, 1, 2)
    m = describe(lib, h)["materials"]
    assert m[0]["absorb"][0] == pytest.approx(0.4) and m[1]["absorb"][0] == 0.0
    assert m[0]["roughness"] == 0.0 and m[2]["roughness"] == pytest.approx(0.35)
    assert m[0]["ior"] == pytest.approx(1.5)
    expect_add_fail(lib, lib.sundog_add_material_dielectric(h, NAN, D3(-0.1, 0, 0), None, -1, NAN))
    lib.sundog_scene_destroy(h)


def test_emissive_registration_rules(lib, mini):
    em = lib.sundog_add_material_emissive(mini, None, -1, 5.0, -1)
    se = lib.sundog_add_material_emissive(mini, D3(1, 1, 1), -1, 3.0, -1)
    tex = lib.sundog_add_texture_image(mini, b"textures/spot_texture.avif", -1)
    tem = lib.sundog_add_material_emissive(mini, None, tex, 2.0, -1)
    assert min(em, se, tex, tem) >= 0
    expect_add_fail(lib, add_obj(lib, mini, DISK, em, st=[(SCALE, 2, 1, 1)]))
    expect_add_fail(lib, add_obj(lib, mini, SPHERE, em, st=[(SCALE, 2, 1, 1)]))
    assert add_obj(lib, mini, DISK, em, st=[(SCALE, 2, 1, 1)], nee=0) >= 0
    d = describe(lib, mini)
    assert d["lights"] == [] and d["objects"][-1]["light_id"] == -1 and d["objects"][-1]["nee"] is False
    assert add_obj(lib, mini, SPHERE, se, st=[(SCALE, 2, 2, 2), (TRANSLATE, 1, 5, 0)]) >= 0
    d = describe(lib, mini)
    assert len(d["lights"]) == 1 and d["lights"][0]["kind"] == 2
    assert d["lights"][0]["radius"] == pytest.approx(2.0, abs=1e-5)
    assert d["lights"][0]["p"][1] == pytest.approx(5.0, abs=1e-5)
    assert d["lights"][0]["L"][0] == pytest.approx(3.0, abs=1e-5)
    assert d["objects"][-1]["light_id"] == 0
    expect_add_fail(lib, add_obj(lib, mini, SPHERE, tem, st=[(SCALE, 1, 1, 1)]))
    assert add_obj(lib, mini, PARA, NONE, back=0) >= 0
    d = describe(lib, mini)
    assert d["objects"][-1]["mat_front"] == 0xFFFF and d["objects"][-1]["mat_back"] == 0
    assert add_obj(lib, mini, RECT, 0, cutout=tex) >= 0
    d = describe(lib, mini)
    assert d["objects"][-1]["cutout_tex"] == tex and d["textures"][tex]["kind"] == 1


def test_flames(lib, mini):
    assert add_obj(lib, mini, RECT, 0, st=[(SCALE, 8, 8, 8)]) >= 0
    assert lib.sundog_add_flame(mini, D3(1, 0.2, -1), 1.5, 0.4, 25, 5, 2.5, 7, 16) == 0
    d = describe(lib, mini)
    f = d["flames"][0]
    assert f["base"][0] == pytest.approx(1.0) and f["height"] == pytest.approx(1.5)
    assert f["radius"] == pytest.approx(0.4) and f["intensity"] == pytest.approx(25)
    assert f["sigma"] == pytest.approx(5) and f["noise_scale"] == pytest.approx(2.5) and f["seed"] == 7
    L = d["lights"]
    assert len(L) == 2 and L[0]["kind"] == 3 and L[1]["kind"] == 3
    assert L[0]["p"][1] == pytest.approx(0.2 + 0.35 * 1.5, abs=1e-5)
    assert L[1]["p"][1] == pytest.approx(0.2 + 0.70 * 1.5, abs=1e-5)
    assert L[0]["radius"] == pytest.approx(0.12, abs=1e-6)
    assert L[0]["L"][0] > L[1]["L"][0]
    assert L[0]["flame_id"] == 0 and L[1]["flame_id"] == 0
    assert lib.sundog_add_flame(mini, D3(-2, 0, 1), 1.0, 0.3, NAN, NAN, NAN, -1, NAN) == 1
    d = describe(lib, mini)
    assert len(d["lights"]) == 4 and d["lights"][2]["flame_id"] == 1 and d["lights"][3]["flame_id"] == 1
    f1 = d["flames"][1]
    assert (f1["intensity"], f1["sigma"], f1["noise_scale"], f1["seed"]) == (20, 4, 3, 0)
    assert f1["octaves"] == 3 and f1["step"] == 0 and f1["bounds"] == 0   # reference march
    expect_add_fail(lib, lib.sundog_add_flame(mini, D3(0, 0, 0), 0, 0.5, NAN, NAN, NAN, -1, NAN))
    assert lib.sundog_set_flame_march(mini, 1, 5, 0.01, 1) == 0
    f1 = describe(lib, mini)["flames"][1]
    assert f1["octaves"] == 5 and f1["step"] == pytest.approx(0.01) and f1["bounds"] == 1
    expect_set_fail(lib, lib.sundog_set_flame_march(mini, 9, 5, 0.01, 1))


def test_water_and_plastic_defaults(lib):
    h = lib.sundog_scene_create(None)
    assert lib.sundog_set_camera(h, D3(0, 2, 8), D3(0, 0, 0), None, NAN, NAN, NAN) == 0
    lake = lib.sundog_add_material_water(h, NAN, None, NAN, NAN, None)
    pool = lib.sundog_add_material_water(h, 1.34, D3(0.6, 0.1, 0.05), 0.12, 3.5, None)
    toy = lib.sundog_add_material_plastic(h, D3(0.85, 0.2, 0.15), -1, NAN)
    matte = lib.sundog_add_material_plastic(h, None, -1, 0.6)
    assert (lake, pool, toy, matte) == (0, 1, 2, 3)
    assert add_obj(lib, h, RECT, lake, st=[(SCALE, 30, 30, 30)]) >= 0
    m = describe(lib, h)["materials"]
    lk = m[0]
    assert lk["kind"] == 4 and lk["ior"] == pytest.approx(1.33)
    assert lk["absorb"][0] == pytest.approx(0.45) and lk["absorb"][0] > lk["absorb"][1] > lk["absorb"][2]
    assert lk["wave_amp"] == pytest.approx(0.05) and lk["wave_freq"] == pytest.approx(2.0)
    assert lk["color"][0] == 1.0 and lk["wave_octaves"] == 3
    assert m[1]["ior"] == pytest.approx(1.34) and m[1]["absorb"][1] == pytest.approx(0.1)
    assert m[1]["wave_amp"] == pytest.approx(0.12) and m[1]["wave_freq"] == pytest.approx(3.5)
    assert m[2]["kind"] == 5 and m[2]["roughness"] == pytest.approx(0.15) and m[2]["ior"] == pytest.approx(1.5)
    assert m[3]["roughness"] == pytest.approx(0.6)
    assert describe(lib, h)["lights"] == []
    assert lib.sundog_set_water_octaves(h, lake, 4) == 0
    assert describe(lib, h)["materials"][0]["wave_octaves"] == 4
    expect_set_fail(lib, lib.sundog_set_water_octaves(h, toy, 4))
    expect_add_fail(lib, lib.sundog_add_material_water(h, NAN, None, -0.1, NAN, None))
    expect_add_fail(lib, lib.sundog_add_material_water(h, NAN, D3(-0.1, 0, 0), NAN, NAN, None))
    lib.sundog_scene_destroy(h)


def test_envmap_settings(lib, mini):
    assert lib.sundog_set_background_envmap(mini, b"../assets/sky.hdr", 90, 2.5, 0) == 0
    b = describe(lib, mini)["background"]
    assert b["kind"] == 2 and b["env_file"] == "../assets/sky.hdr"
    assert b["env_rotate_deg"] == pytest.approx(90) and b["env_intensity"] == pytest.approx(2.5)
    assert b["env_importance"] is False
    assert lib.sundog_set_background_envmap(mini, b"sky.hdr", NAN, NAN, -1) == 0
    b = describe(lib, mini)["background"]
    assert b["env_rotate_deg"] == 0 and b["env_intensity"] == 1 and b["env_importance"] is True
    expect_set_fail(lib, lib.sundog_set_background_envmap(mini, None, NAN, NAN, -1))
    expect_set_fail(lib, lib.sundog_set_background_envmap(mini, b"sky.hdr", NAN, -1.0, -1))


def test_light_order_and_phases(lib, mini):
    mesh = lib.sundog_add_mesh(mini, b"../assets/sparky.obj", -1, b"EmitYellow")
    assert mesh == 0
    em = lib.sundog_add_material_emissive(mini, D3(1, 1, 1), -1, 4.0, -1)
    assert add_obj(lib, mini, RECT, em, st=[(SCALE, 1, 1, 1)]) >= 0
    assert add_obj(lib, mini, MESH, em, st=[(SCALE, 1, 1, 1)], mesh=mesh) >= 0
    assert lib.sundog_add_flame(mini, D3(0, 0, 0), 1, 0.5, NAN, NAN, NAN, -1, NAN) == 0
    assert lib.sundog_add_point_light(mini, D3(0, 5, 0), D3(10, 10, 10), NAN) >= 0
    L = describe(lib, mini)["lights"]
    assert [l["kind"] for l in L] == [0, 5, 3, 3, 3]
    assert L[4]["p"][1] == pytest.approx(5.0)
    assert [l["flame_id"] for l in L] == [-1, -1, 0, 0, -1]
    expect_add_fail(lib, add_obj(lib, mini, SPHERE, 0))
    expect_add_fail(lib, lib.sundog_add_flame(mini, D3(0, 0, 0), 1, 0.5, NAN, NAN, NAN, -1, NAN))
    expect_set_fail(lib, lib.sundog_set_render(mini, 64, 64, -1, -1, NAN, -1, NAN, -1))
    expect_add_fail(lib, lib.sundog_add_material_lambert(mini, None, -1))


def test_mesh_light_registration(lib, mini):
    mesh = lib.sundog_add_mesh(mini, b"../assets/sparky.obj", -1, b"EmitYellow")
    tex = lib.sundog_add_texture_image(mini, b"textures/spot_texture.avif", -1)
    em = lib.sundog_add_material_emissive(mini, D3(1.0, 0.85, 0.2), -1, 6.0, -1)
    tem = lib.sundog_add_material_emissive(mini, None, tex, 3.0, -1)
    pose = [(SCALE, 2, 1, 3), (TRANSLATE, 0, 1, 0)]
    assert add_obj(lib, mini, MESH, em, st=pose, mesh=mesh) >= 0
    d = describe(lib, mini)
    assert len(d["lights"]) == 1 and d["lights"][0]["kind"] == 5
    assert d["objects"][-1]["light_id"] == 0 and d["lights"][0]["L"][0] == pytest.approx(6.0)
    assert d["lights"][0]["tex_id"] == -1 and d["lights"][0]["flame_id"] == -1
    assert d["lights"][0]["area"] == 0.0          # patched when the mesh loads
    assert add_obj(lib, mini, MESH, tem, st=pose, mesh=mesh) >= 0
    d = describe(lib, mini)
    assert d["lights"][1]["tex_id"] == tex and d["lights"][1]["L"][0] == pytest.approx(3.0)
    assert add_obj(lib, mini, MESH, em, st=pose, mesh=mesh, nee=0) >= 0
    assert describe(lib, mini)["objects"][-1]["light_id"] == -1
    expect_add_fail(lib, add_obj(lib, mini, MESH, em, st=pose, mesh=mesh, cutout=tex))
    expect_add_fail(lib, add_obj(lib, mini, RECT, em, st=pose, cutout=tex))
    assert len(describe(lib, mini)["lights"]) == 2
    assert add_obj(lib, mini, MESH, em, st=pose, mesh=mesh, cutout=tex, nee=0) >= 0


def test_mirrored_rect_light_keeps_front(lib, mini):
    em = lib.sundog_add_material_emissive(mini, D3(1, 1, 1), -1, 2.0, -1)
    assert add_obj(lib, mini, RECT, em, st=[(SCALE, -1, 1, 1)]) >= 0
    l = describe(lib, mini)["lights"][-1]
    assert l["n"][1] == pytest.approx(1.0, abs=1e-5)
    assert l["area"] == pytest.approx(4.0, abs=1e-5)


def test_material_cap(lib):
    h = lib.sundog_scene_create(None)
    for i in range(0xFFFF):
        assert lib.sundog_add_material_lambert(h, None, -1) == i
    expect_add_fail(lib, lib.sundog_add_material_lambert(h, None, -1))
    lib.sundog_scene_destroy(h)


def test_camera_90_degrees(lib):
    h = lib.sundog_scene_create(None)
    assert lib.sundog_set_render(h, 100, 100, -1, -1, NAN, -1, NAN, -1) == 0
    assert lib.sundog_set_camera(h, D3(0, 0, 5), D3(0, 0, 0), D3(0, 1, 0), 90.0, NAN, NAN) == 0
    c = describe(lib, h)["camera"]
    assert c["w"][2] == pytest.approx(1.0, abs=1e-5) and c["u"][0] == pytest.approx(1.0, abs=1e-5)
    assert c["v"][1] == pytest.approx(1.0, abs=1e-5) and c["origin"][2] == pytest.approx(5.0)
    assert c["horizontal"][0] == pytest.approx(10.0, abs=1e-4)
    assert c["vertical"][1] == pytest.approx(10.0, abs=1e-4)
    assert c["lower_left"] == pytest.approx([-5.0, -5.0, 0.0], abs=1e-4)
    assert c["lens_radius"] == 0.0
    lib.sundog_scene_destroy(h)


def test_physics_block_is_recorded(lib):
    h = lib.sundog_scene_create(None)
    assert lib.sundog_set_camera(h, D3(0, 4, 9), D3(0, 0, 0), None, NAN, NAN, NAN) == 0
    assert lib.sundog_set_physics(h, None, NAN, NAN, NAN, NAN, -1, -1, NAN, NAN) == 0
    p = describe(lib, h)["physics"]
    assert p["enabled"] and p["gravity"] == pytest.approx([0, -9.81, 0])
    assert p["timestep"] == pytest.approx(1 / 240) and p["max_time"] == 15 and p["pos_iters"] == 8
    expect_set_fail(lib, lib.sundog_set_physics(h, None, -1.0, NAN, NAN, NAN, -1, -1, NAN, NAN))
    expect_set_fail(lib, lib.sundog_set_physics(h, None, NAN, NAN, NAN, NAN, -1, -1, NAN, -2.0))
    lib.sundog_scene_destroy(h)


def test_render_requires_out_path(lib, mini):
    opt = _lib.default_options()
    assert lib.sundog_render(mini, ctypes.byref(opt)) == 1
    assert b"out_path" in lib.sundog_last_error()


def test_scene_program_ids_by_sorted_name(lib):
    s = sundog_amd.Scene()
    s.render(width=32, height=16, spp=4)
    s.camera(lookfrom=[0, 1, 5], lookat=[0, 0, 0])
    s.lambert("zeta", color=[0.1, 0.2, 0.3])
    s.metal("alpha", color=[0.9, 0.9, 0.9], roughness=0.2)
    s.add("sphere", "zeta")
    s.add("rect", "alpha", sundog_amd.scale(3))
    prog = s.program()
    mats = [c for c in prog if c[0].startswith("add_material")]
    assert [c[0] for c in mats] == ["add_material_metal", "add_material_lambert"]
    objs = [c for c in prog if c[0] == "add_object"]
    assert objs[0][3] == 1 and objs[1][3] == 0      # zeta -> 1, alpha -> 0
    assert objs[1][6] == [(0, 3.0, 3.0, 3.0)]       # uniform scale splat
    h = s.build()
    d = describe(lib, h.ptr)
    assert d["materials"][0]["kind"] == 1 and d["objects"][0]["mat_front"] == 1
    assert math.isnan(prog[1][5]) and prog[1][6] == -1   # clamp NaN, seed -1 sentinels


def test_ctypes_struct_layouts_match_the_header(tmp_path):
    """Every struct the ctypes bindings mirror has the header's size and
    field offsets (compiled with the host C compiler from include/): the
    library writes whole structs, so a shorter mirror would be overrun."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    pairs = [("sundog_xform_step", _lib.XformStep), ("sundog_physics_body", _lib.PhysicsBody),
             ("sundog_render_options", _lib.RenderOptions), ("sundog_device_info", _lib.DeviceInfo),
             ("sundog_session_stats", _lib.SessionStats), ("sundog_request_info", _lib.RequestInfo),
             ("sundog_server_stats", _lib.ServerStats), ("sundog_tune_result", _lib.TuneResult),
             ("sundog_bvh_layout", _lib.BvhLayout)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sundog_api.h"', "int main(void) {"]
    for cname, py in pairs:
        lines.append('  printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in py._fields_:
            lines.append('  printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        parts = line.split()
        if len(parts) == 3:
            got[(parts[0], parts[1])] = int(parts[2])
    for cname, py in pairs:
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
