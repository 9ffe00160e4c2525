"""Scene programs over the C ABI, plus resident render sessions.

`Scene` mirrors the reference's Python scene API (scenes/scenelib.py:161-717):
the same builder methods, transform-step constructors and call program —
texture / material / mesh ids allocated by sorted name, unset values sent as
NaN / NULL / -1 sentinels (scenes/scenelib.py:528-679) — replayed through
ctypes into libsundog.so (scenes/scenelib.py:790-880). Two additive builder
methods select the march / ripple variants of this build:
`flame(..., octaves=, step=, bounds=)` and `water(..., octaves=)`.

`Session` wraps the additive sundog_session_* entry points: the scene stays
resident in HBM and is rendered in progressive passes.
"""

import argparse
import ctypes
import os
import sys
import math

import numpy as np

from . import _lib

NAN = float("nan")
OMIT = object()

_SHAPES = {"sphere": 0, "rect": 1, "disk": 2, "cylinder": 3, "parabola": 4}
_MATERIAL_FIELDS = {
    "lambert": ("color", "texture"),
    "metal": ("color", "texture", "roughness"),
    "dielectric": ("ior", "absorb", "color", "texture", "roughness"),
    "emissive": ("color", "texture", "intensity", "two_sided"),
    "water": ("ior", "absorb", "wave_amp", "wave_freq", "color", "octaves"),
    "plastic": ("color", "texture", "roughness"),
}
_TEXTURE_FIELDS = {
    "solid": ("color",),
    "checker": ("a", "b", "scale"),
    "grid": ("a", "b", "scale", "width"),
    "image": ("file", "srgb"),
}


class SceneError(ValueError):
    pass


def scale(x, y=None, z=None):
    if y is None and z is None:
        return {"scale": list(x) if isinstance(x, (list, tuple)) else float(x)}
    return {"scale": [float(x), float(y), float(z)]}


def rotate_x(deg):
    return {"rotate_x": float(deg)}


def rotate_y(deg):
    return {"rotate_y": float(deg)}


def rotate_z(deg):
    return {"rotate_z": float(deg)}


def translate(x, y=None, z=None):
    if y is None and z is None:
        return {"translate": [float(c) for c in x]}
    return {"translate": [float(x), float(y), float(z)]}


def _put(d, k, v):
    if v is not OMIT:
        d[k] = v


# per-object physics bodies (recorded and passed to the ABI, which keeps them
# and renders baked poses: settling is not simulated)
def rigid_body(density=OMIT, velocity=OMIT, angular_velocity=OMIT, friction=OMIT, restitution=OMIT):
    b = {"dynamic": True}
    for k, v in (("density", density), ("velocity", velocity), ("angular_velocity", angular_velocity),
                 ("friction", friction), ("restitution", restitution)):
        _put(b, k, v)
    return b


def static_body(thickness=OMIT, friction=OMIT, restitution=OMIT):
    b = {}
    for k, v in (("thickness", thickness), ("friction", friction), ("restitution", restitution)):
        _put(b, k, v)
    return b


class Scene(object):
    def __init__(self):
        self._render = {}
        self._camera = None
        self._background = None
        self._textures = {}
        self._materials = {}
        self._meshes = {}
        self._objects = []
        self._lights = []
        self._flames = []
        self._physics = None

    # ---- blocks ---------------------------------------------------------------
    def physics(self, gravity=OMIT, timestep=OMIT, max_time=OMIT, friction=OMIT, restitution=OMIT,
                solver_iterations=OMIT, sleep_threshold=OMIT, stop_time=OMIT):
        if solver_iterations is not OMIT and (not isinstance(solver_iterations, (list, tuple))
                                              or len(solver_iterations) != 2):
            raise SceneError("physics: solver_iterations must be [posIters, velIters]")
        p = {}
        for k, v in (("gravity", gravity), ("timestep", timestep), ("max_time", max_time),
                     ("friction", friction), ("restitution", restitution),
                     ("solver_iterations", solver_iterations), ("sleep_threshold", sleep_threshold),
                     ("stop_time", stop_time)):
            _put(p, k, v)
        self._physics = p

    def render(self, width=OMIT, height=OMIT, spp=OMIT, max_depth=OMIT, clamp=OMIT,
               seed=OMIT, exposure=OMIT, transparent_shadows=OMIT):
        for k, v in (("width", width), ("height", height), ("spp", spp),
                     ("max_depth", max_depth), ("clamp", clamp), ("seed", seed),
                     ("exposure", exposure), ("transparent_shadows", transparent_shadows)):
            _put(self._render, k, v)

    def camera(self, lookfrom, lookat, up=OMIT, vfov=OMIT, aperture=OMIT, focus_dist=OMIT):
        c = {"lookfrom": list(lookfrom), "lookat": list(lookat)}
        for k, v in (("up", up), ("vfov", vfov), ("aperture", aperture),
                     ("focus_dist", focus_dist)):
            _put(c, k, v)
        self._camera = c

    def background_solid(self, color):
        self._background = {"type": "solid", "color": list(color)}

    def background_gradient(self, horizon, zenith):
        self._background = {"type": "gradient", "horizon": list(horizon), "zenith": list(zenith)}

    def background_envmap(self, file, rotate=OMIT, intensity=OMIT, importance=OMIT):
        b = {"type": "envmap", "file": file}
        for k, v in (("rotate", rotate), ("intensity", intensity), ("importance", importance)):
            _put(b, k, v)
        self._background = b

    # ---- registries -----------------------------------------------------------
    def texture(self, name, type, **fields):
        if type not in _TEXTURE_FIELDS:
            raise SceneError("texture %r: unknown type %r" % (name, type))
        for k in fields:
            if k not in _TEXTURE_FIELDS[type]:
                raise SceneError("texture %r: %r is not a field of %s textures" % (name, k, type))
        t = {"type": type}
        for k, v in fields.items():
            _put(t, k, v)
        self._textures[name] = t
        return name

    def material(self, name, type, **fields):
        if type not in _MATERIAL_FIELDS:
            raise SceneError("material %r: unknown type %r" % (name, type))
        for k in fields:
            if k not in _MATERIAL_FIELDS[type]:
                raise SceneError("material %r: %r is not a field of %s materials" % (name, k, type))
        m = {"type": type}
        for k, v in fields.items():
            _put(m, k, OMIT if v is None else v)
        self._materials[name] = m
        return name

    def lambert(self, name, color=OMIT, texture=OMIT):
        return self.material(name, "lambert", color=color, texture=texture)

    def metal(self, name, color=OMIT, texture=OMIT, roughness=OMIT):
        return self.material(name, "metal", color=color, texture=texture, roughness=roughness)

    def dielectric(self, name, ior=OMIT, absorb=OMIT, roughness=OMIT):
        return self.material(name, "dielectric", ior=ior, absorb=absorb, roughness=roughness)

    def emissive(self, name, color=OMIT, texture=OMIT, intensity=OMIT, two_sided=OMIT):
        return self.material(name, "emissive", color=color, texture=texture,
                             intensity=intensity, two_sided=two_sided)

    def water(self, name, ior=OMIT, absorb=OMIT, wave_amp=OMIT, wave_freq=OMIT, octaves=OMIT):
        return self.material(name, "water", ior=ior, absorb=absorb, wave_amp=wave_amp,
                             wave_freq=wave_freq, octaves=octaves)

    def plastic(self, name, color=OMIT, texture=OMIT, roughness=OMIT):
        return self.material(name, "plastic", color=color, texture=texture, roughness=roughness)

    def mesh(self, name, obj, normals=OMIT, usemtl=OMIT):
        m = {"obj": obj}
        _put(m, "normals", normals)
        _put(m, "usemtl", usemtl)
        self._meshes[name] = m
        return "mesh:" + name

    # ---- objects and lights ---------------------------------------------------
    def add(self, shape, material, *steps, **kw):
        unknown = set(kw) - {"material_back", "cutout", "nee", "physics"}
        if unknown:
            raise SceneError("add: unknown keyword(s) %s" % ", ".join(sorted(unknown)))
        if shape not in _SHAPES and not str(shape).startswith("mesh:"):
            raise SceneError("add: unknown shape %r" % (shape,))
        o = {"shape": shape, "material": material}
        if steps:
            o["transform"] = list(steps)
        for k in ("material_back", "cutout", "nee", "physics"):
            if k in kw and kw[k] is not OMIT:
                o[k] = kw[k]
        self._objects.append(o)
        return o

    def point_light(self, position, intensity, radius=OMIT):
        l = {"type": "point", "position": list(position), "intensity": list(intensity)}
        _put(l, "radius", radius)
        self._lights.append(l)

    def distant_light(self, direction, radiance):
        self._lights.append({"type": "distant", "direction": list(direction),
                             "radiance": list(radiance)})

    def flame(self, base, height, radius, intensity=OMIT, sigma=OMIT, noise_scale=OMIT,
              seed=OMIT, light_intensity=OMIT, octaves=OMIT, step=OMIT, bounds=OMIT):
        f = {"base": list(base), "height": float(height), "radius": float(radius)}
        for k, v in (("intensity", intensity), ("sigma", sigma), ("noise_scale", noise_scale),
                     ("seed", seed), ("light_intensity", light_intensity), ("octaves", octaves),
                     ("step", step), ("bounds", bounds)):
            _put(f, k, v)
        self._flames.append(f)

    # ---- call program ---------------------------------------------------------
    def program(self, base_dir="."):
        """Ordered backend calls (pure function of the builder calls)."""
        tex_ids = {n: i for i, n in enumerate(sorted(self._textures))}
        mat_ids = {n: i for i, n in enumerate(sorted(self._materials))}
        mesh_ids = {n: i for i, n in enumerate(sorted(self._meshes))}

        def fnum(d, k):
            return float(d[k]) if k in d else NAN

        def fint(d, k):
            return int(d[k]) if k in d else -1

        def ftri(d, k):
            return -1 if k not in d else (1 if d[k] else 0)

        def fvec(d, k):
            return [float(c) for c in d[k]] if k in d else None

        if self._camera is None:
            raise SceneError("scene has no camera")
        p = [("create", base_dir)]
        r = self._render
        p.append(("set_render", fint(r, "width"), fint(r, "height"), fint(r, "spp"),
                  fint(r, "max_depth"), fnum(r, "clamp"), fint(r, "seed"),
                  fnum(r, "exposure"), ftri(r, "transparent_shadows")))
        if self._physics is not None:
            ph = self._physics
            it = ph.get("solver_iterations")
            p.append(("set_physics", fvec(ph, "gravity"), fnum(ph, "timestep"), fnum(ph, "max_time"),
                      fnum(ph, "friction"), fnum(ph, "restitution"), int(it[0]) if it else -1,
                      int(it[1]) if it else -1, fnum(ph, "sleep_threshold"), fnum(ph, "stop_time")))
        c = self._camera
        p.append(("set_camera", fvec(c, "lookfrom"), fvec(c, "lookat"), fvec(c, "up"),
                  fnum(c, "vfov"), fnum(c, "aperture"), fnum(c, "focus_dist")))
        b = self._background
        if b is not None:
            if b["type"] == "solid":
                p.append(("set_background_solid", fvec(b, "color")))
            elif b["type"] == "gradient":
                p.append(("set_background_gradient", fvec(b, "horizon"), fvec(b, "zenith")))
            else:
                p.append(("set_background_envmap", b["file"], fnum(b, "rotate"),
                          fnum(b, "intensity"), ftri(b, "importance")))
        for name in sorted(self._textures):
            t = self._textures[name]
            k = t["type"]
            if k == "solid":
                p.append(("add_texture_solid", fvec(t, "color")))
            elif k == "checker":
                p.append(("add_texture_checker", fvec(t, "a"), fvec(t, "b"),
                          [float(v) for v in t["scale"]] if "scale" in t else None))
            elif k == "grid":
                p.append(("add_texture_grid", fvec(t, "a"), fvec(t, "b"),
                          [float(v) for v in t["scale"]] if "scale" in t else None,
                          fnum(t, "width")))
            else:
                p.append(("add_texture_image", t["file"], ftri(t, "srgb")))
        post = []
        for name in sorted(self._materials):
            m = self._materials[name]
            k = m["type"]
            tid = tex_ids[m["texture"]] if "texture" in m else -1
            if k == "lambert":
                p.append(("add_material_lambert", fvec(m, "color"), tid))
            elif k == "metal":
                p.append(("add_material_metal", fvec(m, "color"), tid, fnum(m, "roughness")))
            elif k == "dielectric":
                p.append(("add_material_dielectric", fnum(m, "ior"), fvec(m, "absorb"),
                          fvec(m, "color"), tid, fnum(m, "roughness")))
            elif k == "emissive":
                p.append(("add_material_emissive", fvec(m, "color"), tid,
                          fnum(m, "intensity"), ftri(m, "two_sided")))
            elif k == "plastic":
                p.append(("add_material_plastic", fvec(m, "color"), tid, fnum(m, "roughness")))
            else:
                p.append(("add_material_water", fnum(m, "ior"), fvec(m, "absorb"),
                          fnum(m, "wave_amp"), fnum(m, "wave_freq"), fvec(m, "color")))
                if "octaves" in m:
                    post.append(("set_water_octaves", mat_ids[name], int(m["octaves"])))
        p.extend(post)
        for name in sorted(self._meshes):
            ms = self._meshes[name]
            smooth = -1 if "normals" not in ms else (1 if ms["normals"] == "smooth" else 0)
            p.append(("add_mesh", ms["obj"], smooth, ms.get("usemtl")))
        for o in self._objects:
            shape = o["shape"]
            if shape.startswith("mesh:"):
                gk, mesh_id = 5, mesh_ids[shape[5:]]
            else:
                gk, mesh_id = _SHAPES[shape], -1
            front = -1 if o["material"] is None else mat_ids[o["material"]]
            if "material_back" not in o:
                back = -2
            elif o["material_back"] is None:
                back = -1
            else:
                back = mat_ids[o["material_back"]]
            cutout = tex_ids[o["cutout"]] if "cutout" in o else -1
            steps = []
            for st in o.get("transform", []):
                (k, v), = st.items()
                if k == "scale":
                    if isinstance(v, (int, float)):
                        steps.append((0, float(v), float(v), float(v)))
                    else:
                        steps.append((0, float(v[0]), float(v[1]), float(v[2])))
                elif k == "translate":
                    steps.append((1, float(v[0]), float(v[1]), float(v[2])))
                else:
                    steps.append(({"rotate_x": 2, "rotate_y": 3, "rotate_z": 4}[k],
                                  float(v), 0.0, 0.0))
            body = None
            if "physics" in o:
                pb = o["physics"]
                body = (1 if pb.get("dynamic") else 0, fnum(pb, "density"), 1 if "velocity" in pb else 0,
                        fvec(pb, "velocity") or [0.0, 0.0, 0.0], 1 if "angular_velocity" in pb else 0,
                        fvec(pb, "angular_velocity") or [0.0, 0.0, 0.0], fnum(pb, "friction"),
                        fnum(pb, "restitution"), fnum(pb, "thickness"))
            p.append(("add_object", gk, mesh_id, front, back, cutout, steps, ftri(o, "nee"), body))
        for i, fl in enumerate(self._flames):
            p.append(("add_flame", [float(v) for v in fl["base"]], float(fl["height"]),
                      float(fl["radius"]), fnum(fl, "intensity"), fnum(fl, "sigma"),
                      fnum(fl, "noise_scale"), fint(fl, "seed"), fnum(fl, "light_intensity")))
            if "octaves" in fl or "step" in fl or "bounds" in fl:
                p.append(("set_flame_march", i, int(fl.get("octaves", 3)),
                          float(fl.get("step", 0.0)),
                          {"cylinder": 0, "box": 1}[fl.get("bounds", "cylinder")]))
        for l in self._lights:
            if l["type"] == "point":
                p.append(("add_point_light", [float(v) for v in l["position"]],
                          [float(v) for v in l["intensity"]], fnum(l, "radius")))
            else:
                p.append(("add_distant_light", [float(v) for v in l["direction"]],
                          [float(v) for v in l["radiance"]]))
        return p

    def build(self, base_dir=".", lib=None):
        """Replay the call program; returns a SceneHandle."""
        return SceneHandle(apply_program(self.program(base_dir), lib), lib)

    def run(self, out, argv=None, base_dir=None):
        """Renders this scene from a scene file's __main__ block, with the
        reference's command line (scenes/scenelib.py:685-719, :898-929):
        --out, --spp, --size WxH, --seed, --denoise / --no-denoise,
        --opaque-shadows, --clamp, --physics-time, --stats, --aov-albedo,
        --aov-normal, --probe, --quiet. Exits 0 on success, 1 with
        'sundog: fatal: <message>' on stderr otherwise."""
        opts = parse_args(list(sys.argv[1:] if argv is None else argv), out)
        if opts.probe:
            sys.exit(print_probe())
        if base_dir is None:
            base_dir = os.path.dirname(os.path.abspath(sys.argv[0])) or "."
        sys.stdout.flush()
        try:
            h = self.build(base_dir)
            h.render_file(opts.out, **render_overrides(opts, os.path.basename(sys.argv[0]) or "(scene)"))
        except (SceneError, _lib.SundogError, OSError) as e:
            sys.stderr.write("sundog: fatal: %s\n" % e)
            sys.exit(1)
        sys.exit(0)


def _d3(v):
    return (ctypes.c_double * 3)(*v) if v is not None else None


def apply_program(program, lib=None):
    lib = lib or _lib.load()
    d = ctypes.c_double
    h = None
    try:
        for call in program:
            name, args = call[0], call[1:]
            if name == "create":
                h = lib.sundog_scene_create(args[0].encode())
                if not h:
                    raise SceneError(_lib.last_error(lib))
                continue
            if name == "set_render":
                w, ht, spp, md, cl, seed, ex, ts = args
                rc = lib.sundog_set_render(h, w, ht, spp, md, d(cl), seed, d(ex), ts)
            elif name == "set_physics":
                g, ts_, mt, fr, re, pi, vi, sl, st = args
                rc = lib.sundog_set_physics(h, _d3(g), d(ts_), d(mt), d(fr), d(re), pi, vi, d(sl), d(st))
            elif name == "set_camera":
                lf, la, up, vf, ap, fd = args
                rc = lib.sundog_set_camera(h, _d3(lf), _d3(la), _d3(up), vf, ap, fd)
            elif name == "set_background_solid":
                rc = lib.sundog_set_background_solid(h, _d3(args[0]))
            elif name == "set_background_gradient":
                rc = lib.sundog_set_background_gradient(h, _d3(args[0]), _d3(args[1]))
            elif name == "set_background_envmap":
                f, rot, inten, imp = args
                rc = lib.sundog_set_background_envmap(h, f.encode(), rot, inten, imp)
            elif name == "add_texture_solid":
                rc = lib.sundog_add_texture_solid(h, _d3(args[0]))
            elif name == "add_texture_checker":
                a, b, sc = args
                rc = lib.sundog_add_texture_checker(
                    h, _d3(a), _d3(b), (ctypes.c_double * 2)(*sc) if sc is not None else None)
            elif name == "add_texture_grid":
                a, b, sc, w = args
                rc = lib.sundog_add_texture_grid(
                    h, _d3(a), _d3(b), (ctypes.c_double * 2)(*sc) if sc is not None else None, w)
            elif name == "add_texture_image":
                rc = lib.sundog_add_texture_image(h, args[0].encode(), args[1])
            elif name == "add_material_lambert":
                rc = lib.sundog_add_material_lambert(h, _d3(args[0]), args[1])
            elif name == "add_material_metal":
                rc = lib.sundog_add_material_metal(h, _d3(args[0]), args[1], args[2])
            elif name == "add_material_dielectric":
                rc = lib.sundog_add_material_dielectric(h, args[0], _d3(args[1]), _d3(args[2]),
                                                        args[3], args[4])
            elif name == "add_material_emissive":
                rc = lib.sundog_add_material_emissive(h, _d3(args[0]), args[1], args[2], args[3])
            elif name == "add_material_water":
                rc = lib.sundog_add_material_water(h, args[0], _d3(args[1]), args[2], args[3],
                                                   _d3(args[4]))
            elif name == "add_material_plastic":
                rc = lib.sundog_add_material_plastic(h, _d3(args[0]), args[1], args[2])
            elif name == "set_water_octaves":
                rc = lib.sundog_set_water_octaves(h, args[0], args[1])
            elif name == "add_mesh":
                rc = lib.sundog_add_mesh(h, args[0].encode(), args[1],
                                         args[2].encode() if args[2] else None)
            elif name == "add_object":
                gk, mesh_id, front, back, cutout, steps, nee = args[:7]
                body = args[7] if len(args) > 7 else None
                arr = (_lib.XformStep * max(len(steps), 1))()
                for i, (k, a, b, c) in enumerate(steps):
                    arr[i] = _lib.XformStep(k, a, b, c)
                pb = None
                if body is not None:
                    dyn, dens, hv, vel, hav, avel, fr, re, th = body
                    pb = _lib.PhysicsBody(dyn, dens, hv, hav, (ctypes.c_double * 3)(*vel),
                                          (ctypes.c_double * 3)(*avel), fr, re, th)
                rc = lib.sundog_add_object(h, gk, mesh_id, front, back, cutout, arr, len(steps),
                                           nee, ctypes.byref(pb) if pb is not None else None)
            elif name == "add_flame":
                base, ht, rad, inten, sig, ns, seed, li = args
                rc = lib.sundog_add_flame(h, _d3(base), ht, rad, inten, sig, ns, seed, li)
            elif name == "set_flame_march":
                rc = lib.sundog_set_flame_march(h, args[0], args[1], args[2], args[3])
            elif name == "add_point_light":
                rc = lib.sundog_add_point_light(h, _d3(args[0]), _d3(args[1]), args[2])
            elif name == "add_distant_light":
                rc = lib.sundog_add_distant_light(h, _d3(args[0]), _d3(args[1]))
            else:
                raise SceneError("unknown program call %r" % name)
            if rc is None or rc < 0 or (name.startswith("set_") and rc != 0):
                raise SceneError(_lib.last_error(lib))
    except Exception:
        if h:
            lib.sundog_scene_destroy(h)
        raise
    return h


class SceneHandle(object):
    """Owns a sundog_scene*."""

    def __init__(self, ptr, lib=None):
        self.lib = lib or _lib.load()
        self.ptr = ptr

    def close(self):
        if self.ptr:
            self.lib.sundog_scene_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def flatten(self, **overrides):
        opt = _lib.default_options(**overrides)
        f = self.lib.sundog_scene_flatten(self.ptr, ctypes.byref(opt))
        if not f:
            raise SceneError(_lib.last_error(self.lib))
        return FlatScene(f, self.lib)

    def session(self, **overrides):
        return Session(self, **overrides)

    def render_file(self, out_path, **overrides):
        opt = _lib.default_options(out_path=out_path, **overrides)
        _lib.check(self.lib.sundog_render(self.ptr, ctypes.byref(opt)), self.lib)


class FlatScene(object):
    """Owns a sundog_flat_scene* (read-only view for the CPU oracle)."""

    def __init__(self, ptr, lib):
        self.ptr = ptr
        self.lib = lib

    def close(self):
        if self.ptr:
            self.lib.sundog_flat_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fptr(a):
    return a.ctypes.data_as(_lib.P_f)


def _iptr(a):
    return a.ctypes.data_as(_lib.P_i32)


class Session(object):
    """A resident render session on the GPU (sundog_session_*)."""

    def __init__(self, scene_handle, **overrides):
        self.lib = scene_handle.lib
        opt = _lib.default_options(**overrides)
        self.ptr = self.lib.sundog_session_create(scene_handle.ptr, ctypes.byref(opt))
        if not self.ptr:
            raise SceneError(_lib.last_error(self.lib))
        w, h, s = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self.lib.sundog_session_size(self.ptr, ctypes.byref(w), ctypes.byref(h), ctypes.byref(s))
        self.width, self.height, self.spp = w.value, h.value, s.value

    def close(self):
        if self.ptr:
            self.lib.sundog_session_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        _lib.check(self.lib.sundog_session_reset(self.ptr), self.lib)

    def count_rays(self, on=True):
        _lib.check(self.lib.sundog_session_set_count_rays(self.ptr, 1 if on else 0), self.lib)

    def tune(self, key, value):
        _lib.check(self.lib.sundog_session_set_tuning(self.ptr, key.encode(), int(value)), self.lib)

    def autotune(self, count=16):
        """Times the megakernel's occupancy variants on real passes, keeps
        the fastest, clears the film; returns the chosen waves per SIMD."""
        w = ctypes.c_int(0)
        _lib.check(self.lib.sundog_session_autotune(self.ptr, int(count), ctypes.byref(w)), self.lib)
        return w.value

    KERNELS = {"megakernel": 0, "wavefront": 1}

    def set_kernel(self, name):
        """'megakernel' (one work-item per pixel, inline traces) or
        'wavefront' (path state in HBM, queued traces, material-sorted
        shading); both give the same film bits."""
        _lib.check(self.lib.sundog_session_set_kernel(self.ptr, self.KERNELS[name]), self.lib)

    def count_work(self, on=True):
        _lib.check(self.lib.sundog_session_set_count_work(self.ptr, 1 if on else 0), self.lib)

    def render(self, offset, count):
        _lib.check(self.lib.sundog_session_render(self.ptr, offset, count), self.lib)

    def render_all(self, chunk=16):
        done = 0
        while done < self.spp:
            c = min(chunk, self.spp - done)
            self.render(done, c)
            done += c

    def sync(self):
        _lib.check(self.lib.sundog_session_sync(self.ptr), self.lib)

    def download(self):
        n = self.width * self.height * 4
        b = np.zeros(n, np.float32)
        a = np.zeros(n, np.float32)
        nn = np.zeros(n, np.float32)
        _lib.check(self.lib.sundog_session_download(self.ptr, _fptr(b), _fptr(a), _fptr(nn)), self.lib)
        shape = (self.height, self.width, 4)
        return b.reshape(shape), a.reshape(shape), nn.reshape(shape)

    def stats(self):
        st = _lib.SessionStats()
        _lib.check(self.lib.sundog_session_get_stats(self.ptr, ctypes.byref(st)), self.lib)
        return {k: getattr(st, k) for k, _ in st._fields_ if k != "pad"}

    def bvh(self):
        """The resident BVH: {"nodes": raw node bytes, "tris": (num_tris, 12)
        float32, "blas_roots": per instance (-1 for quadrics), "num_nodes",
        "tlas_nodes", "num_tris", "num_instances", "tlas_root"}."""
        lay = _lib.BvhLayout()
        _lib.check(self.lib.sundog_session_bvh(self.ptr, ctypes.byref(lay), None, 0, None, 0, None, 0), self.lib)
        nodes = np.zeros(lay.num_nodes * 32, np.float32)
        tris = np.zeros((lay.num_tris, 12), np.float32)
        roots = np.zeros(max(lay.num_instances, 1), np.int32)
        _lib.check(self.lib.sundog_session_bvh(self.ptr, ctypes.byref(lay), nodes.ctypes.data_as(ctypes.c_void_p),
                                               nodes.nbytes, _fptr(tris), tris.nbytes, _iptr(roots),
                                               lay.num_instances), self.lib)
        out = {k: getattr(lay, k) for k, _ in lay._fields_ if k != "pad"}
        out.update(nodes=nodes.tobytes(), tris=tris, blas_roots=roots[:lay.num_instances])
        return out

    def stream(self):
        return self.lib.sundog_session_stream(self.ptr)

    def trace(self, rays, brute=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        t = np.zeros(n, np.float32)
        inst = np.zeros(n, np.int32)
        prim = np.zeros(n, np.int32)
        _lib.check(self.lib.sundog_session_trace(self.ptr, _fptr(rays), n, 1 if brute else 0,
                                                 _fptr(t), _iptr(inst), _iptr(prim)), self.lib)
        return t, inst, prim

    def shadow(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        tr = np.zeros((n, 3), np.float32)
        _lib.check(self.lib.sundog_session_shadow(self.ptr, _fptr(rays), n, _fptr(tr)), self.lib)
        return tr

    def flame_field(self, flame, pts):
        """(sigma (n,), emission (n, 3)) of flame `flame` at (n, 3) points,
        evaluated on the device."""
        pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        n = pts.shape[0]
        sg = np.zeros(n, np.float32)
        em = np.zeros((n, 3), np.float32)
        _lib.check(self.lib.sundog_session_flame_field(self.ptr, int(flame), _fptr(pts), n, _fptr(sg), _fptr(em)),
                   self.lib)
        return sg, em

    def env_bins(self, u):
        u = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
        n = u.shape[0]
        row = np.zeros(n, np.int32)
        col = np.zeros(n, np.int32)
        _lib.check(self.lib.sundog_session_env_bins(self.ptr, _fptr(u), n, _iptr(row), _iptr(col)),
                   self.lib)
        return row, col

    def env_bins_mode(self, u, mode):
        """(row, col, kernel ms) of the environment bins by search `mode`:
        0 the renderer's per-lane 16-ary table, 1 the wave-cooperative
        search."""
        u = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
        n = u.shape[0]
        row = np.zeros(n, np.int32)
        col = np.zeros(n, np.int32)
        ms = ctypes.c_float(0.0)
        _lib.check(self.lib.sundog_session_env_bins_mode(self.ptr, _fptr(u), n, int(mode), _iptr(row), _iptr(col),
                                                         ctypes.byref(ms)), self.lib)
        return row, col, ms.value

    def light_samples(self, light, xs, seed0=1, seq=7):
        """Device light samples of light `light` from (n, 3) points, sample i
        from the stream init(seed0 + i, seq): the oracle's light_sample keys."""
        xs = np.ascontiguousarray(xs, np.float32).reshape(-1, 3)
        n = xs.shape[0]
        out = np.zeros((n, 12), np.float32)
        _lib.check(self.lib.sundog_session_light_samples(self.ptr, int(light), _fptr(xs), n, int(seed0), int(seq),
                                                         _fptr(out)), self.lib)
        flags = out[:, 8].astype(np.int32)
        return {"wi": out[:, 0:3].copy(), "dist": out[:, 3].copy(), "Li": out[:, 4:7].copy(), "pdf": out[:, 7].copy(),
                "valid": (flags & 1) != 0, "delta": (flags & 2) != 0, "draws": out[:, 9].astype(np.int32),
                "pdf_back": out[:, 10].copy()}


def device_pcg32(seed, seq, n, lib=None):
    lib = lib or _lib.load()
    out = np.zeros(n, np.uint32)
    _lib.check(lib.sundog_device_pcg32(seed, seq, n, out.ctypes.data_as(_lib.P_u32)), lib)
    return out


def parse_args(args, default_out):
    """The scene-file command line (the reference's flags and defaults:
    -1 / NaN / None leave the scene's own setting)."""
    ap = argparse.ArgumentParser(prog=os.path.basename(sys.argv[0]),
                                 description="sundog scene: running this file renders it", allow_abbrev=False)
    ap.add_argument("--out", default=default_out, metavar="FILE")
    ap.add_argument("--spp", type=int, default=-1, metavar="N")
    ap.add_argument("--size", default=None, metavar="WxH")
    ap.add_argument("--seed", type=int, default=-1, metavar="N")
    ap.add_argument("--denoise", dest="denoise", action="store_const", const=1, default=-1)
    ap.add_argument("--no-denoise", dest="denoise", action="store_const", const=0)
    ap.add_argument("--opaque-shadows", dest="transparent_shadows", action="store_const", const=0, default=-1)
    ap.add_argument("--clamp", type=float, default=NAN, metavar="F")
    ap.add_argument("--physics-time", dest="physics_time", type=float, default=NAN, metavar="F")
    ap.add_argument("--stats", default=None, metavar="FILE.json")
    ap.add_argument("--aov-albedo", dest="aov_albedo", default=None)
    ap.add_argument("--aov-normal", dest="aov_normal", default=None)
    ap.add_argument("--probe", action="store_true")
    ap.add_argument("--quiet", action="store_true")
    ns = ap.parse_args(args)
    ns.width = ns.height = -1
    if ns.size:
        w, sep, h = ns.size.partition("x")
        if not sep or not w.isdigit() or not h.isdigit():
            ap.error("--size expects WxH")
        ns.width, ns.height = int(w), int(h)
    return ns


def render_overrides(opts, scene_name):
    """sundog_render_options fields from parsed scene-file flags."""
    o = {"spp": opts.spp, "width": opts.width, "height": opts.height, "seed": opts.seed,
         "denoise": opts.denoise, "transparent_shadows": opts.transparent_shadows, "clamp": opts.clamp,
         "physics_time": opts.physics_time, "quiet": 1 if opts.quiet else 0, "scene_name": scene_name}
    if opts.stats:
        o["stats_path"] = opts.stats
    if opts.aov_albedo:
        o["aov_albedo_path"] = opts.aov_albedo
    if opts.aov_normal:
        o["aov_normal_path"] = opts.aov_normal
    return o


def print_probe(lib=None):
    """Prints the device report of --probe; returns the exit status."""
    lib = lib or _lib.load()
    di = _lib.DeviceInfo()
    if lib.sundog_probe(ctypes.byref(di)) != 0:
        sys.stderr.write("sundog: fatal: %s\n" % _lib.last_error(lib))
        return 1
    sys.stdout.write("GPU:            %s\n"
                     "Compute:        %d.%d\n"
                     "VRAM:           %d MB\n"
                     "HIP driver:     %d\n"
                     "HIP runtime:    %d\n"
                     % (di.name.decode(), di.cc_major, di.cc_minor, di.total_mem_mb, di.driver_version,
                        di.runtime_version))
    return 0


def run_scene_file(path, argv):
    """Runs a scene file written against the reference's scenelib (it does
    `from scenelib import ...`) with this library's builder in its place,
    as its __main__: `python -m sundog_amd.scenelib FILE.py [flags]`."""
    sys.modules["scenelib"] = sys.modules[__name__]
    sys.argv = [path] + list(argv)
    with open(path) as f:
        code = compile(f.read(), path, "exec")
    exec(code, {"__name__": "__main__", "__file__": path})


def probe(lib=None):
    lib = lib or _lib.load()
    di = _lib.DeviceInfo()
    _lib.check(lib.sundog_probe(ctypes.byref(di)), lib)
    return {"name": di.name.decode(), "cc": (di.cc_major, di.cc_minor),
            "total_mem_mb": di.total_mem_mb, "driver": di.driver_version,
            "runtime": di.runtime_version}


def read_pfm(path):
    """Reads a colour PFM ('PF', little-endian scale < 0) as (h, w, 3)
    float32, top row first."""
    with open(path, "rb") as f:
        if f.readline().strip() != b"PF":
            raise ValueError("%s: not a colour PFM" % path)
        w, h = (int(v) for v in f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(w * h * 12), "<f4" if scale < 0 else ">f4").astype(np.float32)
    return data.reshape(h, w, 3)[::-1].copy()


def rel_rmse(a, b):
    """||a - b||_2 / ||b||_2 over all pixels x RGB."""
    a = np.asarray(a, np.float64)[..., :3]
    b = np.asarray(b, np.float64)[..., :3]
    return math.sqrt(float(((a - b) ** 2).sum()) / max(float((b ** 2).sum()), 1e-300))


if __name__ == "__main__":
    if len(sys.argv) < 2:
        sys.stderr.write("usage: python -m sundog_amd.scenelib SCENE.py [scene flags]\n")
        sys.exit(2)
    run_scene_file(sys.argv[1], sys.argv[2:])
