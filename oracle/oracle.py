"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB points the checks at another build of the oracle (planted-mistake runs)
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")

P_f = ctypes.POINTER(ctypes.c_float)
P_i32 = ctypes.POINTER(ctypes.c_int32)


class OracleParams(ctypes.Structure):
    _fields_ = [("sample_offset", ctypes.c_int32), ("sample_count", ctypes.c_int32),
                ("tile_stride", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("brute", ctypes.c_int32)]


class OracleStats(ctypes.Structure):
    _fields_ = [("radiance_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
                ("paths", ctypes.c_uint64), ("seconds", ctypes.c_double),
                ("threads", ctypes.c_int32)]


class FlatMaterial(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("tex_id", ctypes.c_int32), ("color", ctypes.c_float * 3),
                ("roughness", ctypes.c_float), ("ior", ctypes.c_float), ("intensity", ctypes.c_float),
                ("two_sided", ctypes.c_int32), ("absorb", ctypes.c_float * 3),
                ("wave_amp", ctypes.c_float), ("wave_freq", ctypes.c_float),
                ("wave_octaves", ctypes.c_int32)]


_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.isfile(LIB_PATH):
        raise RuntimeError("oracle not built: run `make -C oracle`")
    lib = ctypes.CDLL(LIB_PATH)
    lib.oracle_render.restype = ctypes.c_int
    lib.oracle_render.argtypes = [ctypes.c_void_p, ctypes.POINTER(OracleParams), P_f, P_f, P_f,
                                  ctypes.POINTER(OracleStats)]
    lib.oracle_trace.restype = ctypes.c_int
    lib.oracle_trace.argtypes = [ctypes.c_void_p, P_f, ctypes.c_int, ctypes.c_int, P_f, P_i32, P_i32]
    lib.oracle_shadow.restype = ctypes.c_int
    lib.oracle_shadow.argtypes = [ctypes.c_void_p, P_f, ctypes.c_int, P_f]
    lib.oracle_env_bins.restype = ctypes.c_int
    lib.oracle_env_bins.argtypes = [ctypes.c_void_p, P_f, ctypes.c_int, P_i32, P_i32]
    lib.oracle_env_cdf.restype = ctypes.c_int
    lib.oracle_env_cdf.argtypes = [ctypes.c_void_p, P_f, P_f]
    lib.oracle_pcg32.restype = None
    lib.oracle_pcg32.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_uint32)]
    lib.oracle_bsdf.restype = ctypes.c_int
    lib.oracle_bsdf.argtypes = [ctypes.POINTER(FlatMaterial), P_f, P_f, P_f, ctypes.c_float, P_f, P_f]
    lib.oracle_bsdf_sample.restype = ctypes.c_int
    lib.oracle_bsdf_sample.argtypes = [ctypes.POINTER(FlatMaterial), P_f, P_f, ctypes.c_int,
                                       ctypes.c_float, ctypes.c_uint64, ctypes.c_uint64, P_f, P_f, P_f,
                                       P_i32, P_i32, P_i32]
    lib.oracle_flame_transmittance.restype = ctypes.c_float
    lib.oracle_flame_transmittance.argtypes = [ctypes.c_void_p, P_f, P_f, ctypes.c_float,
                                               ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, P_i32]
    lib.oracle_env_sample.restype = ctypes.c_int
    lib.oracle_env_sample.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, P_f, P_f, P_f]
    lib.oracle_march_flames.restype = ctypes.c_float
    lib.oracle_march_flames.argtypes = [ctypes.c_void_p, P_f, P_f, ctypes.c_float, ctypes.c_uint64,
                                        ctypes.c_uint64, P_f, P_f, P_i32]
    lib.oracle_texture.argtypes = [ctypes.c_void_p, ctypes.c_int, P_f, ctypes.c_int, P_f]
    lib.oracle_bilinear.argtypes = [P_f, ctypes.c_int, ctypes.c_int, P_f, ctypes.c_int, ctypes.c_int, P_f]
    lib.oracle_bilinear.restype = None
    lib.oracle_noise.argtypes = [P_f, ctypes.c_int, ctypes.c_uint32, ctypes.c_int, P_f]
    lib.oracle_noise.restype = None
    lib.oracle_water_normal.argtypes = [P_f, P_f, P_f, ctypes.c_float, ctypes.c_float, ctypes.c_int, P_f]
    lib.oracle_water_normal.restype = None
    lib.oracle_flame_field.argtypes = [ctypes.c_void_p, ctypes.c_int, P_f, ctypes.c_int, P_f, P_f]
    lib.oracle_mesh_light_sample.argtypes = [ctypes.c_void_p, ctypes.c_int, P_f, P_f, ctypes.c_int, P_f, P_f,
                                             P_f, P_i32, P_f]
    lib.oracle_light_sample.argtypes = [ctypes.c_void_p, ctypes.c_int, P_f, ctypes.c_int, ctypes.c_uint64,
                                        ctypes.c_uint64, P_f, P_f, P_f, P_f, P_i32, P_i32, P_f]
    lib.oracle_light_pdf.argtypes = [ctypes.c_void_p, ctypes.c_int, P_f, ctypes.c_int, P_f]
    lib.oracle_camera_jitter.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, P_f, P_f]
    lib.oracle_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _err(lib):
    return lib.oracle_last_error().decode()


def _fp(a):
    return a.ctypes.data_as(P_f)


def render(flat, width, height, sample_offset, sample_count, tile_stride=1, threads=0, brute=False,
           films=None):
    """Returns (beauty, albedo, normal) as (h, w, 4) float32 plus stats."""
    lib = load()
    if films is None:
        films = [np.zeros(width * height * 4, np.float32) for _ in range(3)]
    p = OracleParams(sample_offset, sample_count, tile_stride, threads, 1 if brute else 0)
    st = OracleStats()
    if lib.oracle_render(flat.ptr, ctypes.byref(p), _fp(films[0]), _fp(films[1]), _fp(films[2]),
                         ctypes.byref(st)) != 0:
        raise RuntimeError(_err(lib))
    shape = (height, width, 4)
    stats = {k: getattr(st, k) for k, _ in st._fields_}
    return films[0].reshape(shape), films[1].reshape(shape), films[2].reshape(shape), stats


def trace(flat, rays, brute=False):
    lib = load()
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = rays.shape[0]
    t = np.zeros(n, np.float32)
    inst = np.zeros(n, np.int32)
    prim = np.zeros(n, np.int32)
    if lib.oracle_trace(flat.ptr, _fp(rays), n, 1 if brute else 0, _fp(t),
                        inst.ctypes.data_as(P_i32), prim.ctypes.data_as(P_i32)) != 0:
        raise RuntimeError(_err(lib))
    return t, inst, prim


def shadow(flat, rays):
    lib = load()
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    tr = np.zeros((rays.shape[0], 3), np.float32)
    if lib.oracle_shadow(flat.ptr, _fp(rays), rays.shape[0], _fp(tr)) != 0:
        raise RuntimeError(_err(lib))
    return tr


def env_bins(flat, u):
    lib = load()
    u = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
    row = np.zeros(u.shape[0], np.int32)
    col = np.zeros(u.shape[0], np.int32)
    if lib.oracle_env_bins(flat.ptr, _fp(u), u.shape[0], row.ctypes.data_as(P_i32),
                           col.ctypes.data_as(P_i32)) != 0:
        raise RuntimeError(_err(lib))
    return row, col


def env_cdf(flat, w, h):
    lib = load()
    marg = np.zeros(h + 1, np.float32)
    cond = np.zeros(h * (w + 1), np.float32)
    if lib.oracle_env_cdf(flat.ptr, _fp(marg), _fp(cond)) != 0:
        raise RuntimeError(_err(lib))
    return marg, cond.reshape(h, w + 1)


def pcg32(seed, seq, n):
    lib = load()
    out = np.zeros(n, np.uint32)
    lib.oracle_pcg32(seed, seq, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def material(kind, color=(0.8, 0.8, 0.8), roughness=0.0, ior=1.5, tex_id=-1):
    m = FlatMaterial()
    m.kind = kind
    m.tex_id = tex_id
    m.color[:] = list(color)
    m.roughness = roughness
    m.ior = ior
    m.intensity = 1.0
    m.wave_octaves = 3
    return m


def bsdf_eval(m, albedo, wo, wi, eta=1.0):
    lib = load()
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    lib.oracle_bsdf(ctypes.byref(m), _fp(np.asarray(albedo, np.float32)), _fp(np.asarray(wo, np.float32)),
                    _fp(np.asarray(wi, np.float32)), eta, _fp(f), _fp(pdf))
    return f, float(pdf[0])


def bsdf_sample(m, albedo, raydir, frontface, eta_ext, seed, seq):
    lib = load()
    wi = np.zeros(3, np.float32)
    wt = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    delta = np.zeros(1, np.int32)
    valid = np.zeros(1, np.int32)
    draws = np.zeros(1, np.int32)
    lib.oracle_bsdf_sample(ctypes.byref(m), _fp(np.asarray(albedo, np.float32)),
                           _fp(np.asarray(raydir, np.float32)), 1 if frontface else 0, eta_ext, seed, seq,
                           _fp(wi), _fp(wt), _fp(pdf), delta.ctypes.data_as(P_i32),
                           valid.ctypes.data_as(P_i32), draws.ctypes.data_as(P_i32))
    return {"wi": wi, "weight": wt, "pdf": float(pdf[0]), "delta": bool(delta[0]),
            "valid": bool(valid[0]), "draws": int(draws[0])}


def flame_transmittance(flat, o, d, tend, seed, seq, skip=-1):
    lib = load()
    draws = np.zeros(1, np.int32)
    tr = lib.oracle_flame_transmittance(flat.ptr, _fp(np.asarray(o, np.float32)),
                                        _fp(np.asarray(d, np.float32)), tend, seed, seq, skip,
                                        draws.ctypes.data_as(P_i32))
    return float(tr), int(draws[0])


def env_sample(flat, u0, u1):
    lib = load()
    wi = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    back = np.zeros(1, np.float32)
    if lib.oracle_env_sample(flat.ptr, u0, u1, _fp(wi), _fp(pdf), _fp(back)) != 0:
        raise RuntimeError(_err(lib))
    return wi, float(pdf[0]), float(back[0])


def march_flames(flat, o, d, tend, seed, seq):
    lib = load()
    tr = np.zeros(3, np.float32)
    L = np.zeros(3, np.float32)
    draws = np.zeros(1, np.int32)
    lib.oracle_march_flames(flat.ptr, _fp(np.asarray(o, np.float32)), _fp(np.asarray(d, np.float32)), tend,
                            seed, seq, _fp(tr), _fp(L), draws.ctypes.data_as(P_i32))
    return tr, L, int(draws[0])


def texture(flat, tex, uv):
    """Texture `tex` of the flat scene at (n, 2) uv pairs -> (n, 4)."""
    lib = load()
    uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
    out = np.zeros((len(uv), 4), np.float32)
    if lib.oracle_texture(flat.ptr, tex, _fp(uv), len(uv), _fp(out)) != 0:
        raise RuntimeError(_err(lib))
    return out


def bilinear(img, uv, wrap_v):
    """The bilinear filter over an (h, w, 4) float32 image at (n, 2) uv."""
    lib = load()
    img = np.ascontiguousarray(img, np.float32)
    h, w = img.shape[:2]
    uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
    out = np.zeros((len(uv), 4), np.float32)
    lib.oracle_bilinear(_fp(img), w, h, _fp(uv), len(uv), int(bool(wrap_v)), _fp(out))
    return out


def noise(p, seed, octaves=0):
    """Value noise (octaves=0) or fbm at (n, 3) points."""
    lib = load()
    p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
    out = np.zeros(len(p), np.float32)
    lib.oracle_noise(_fp(p), len(p), seed, octaves, _fp(out))
    return out


def water_normal(x, ns, wo, amp, freq, octaves):
    lib = load()
    out = np.zeros(3, np.float32)
    lib.oracle_water_normal(_fp(np.asarray(x, np.float32)), _fp(np.asarray(ns, np.float32)),
                            _fp(np.asarray(wo, np.float32)), amp, freq, octaves, _fp(out))
    return out


def flame_field(flat, flame, p):
    """(sigma (n,), emission (n, 3)) of flame `flame` at (n, 3) points."""
    lib = load()
    p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
    sg = np.zeros(len(p), np.float32)
    em = np.zeros((len(p), 3), np.float32)
    if lib.oracle_flame_field(flat.ptr, flame, _fp(p), len(p), _fp(sg), _fp(em)) != 0:
        raise RuntimeError(_err(lib))
    return sg, em


def mesh_light_sample(flat, light, x, u):
    """Mesh-light samples for (n, 3) uniform triples seen from x: dict of
    wi (n, 3), dist, pdf, tri and the light's total area."""
    lib = load()
    u = np.ascontiguousarray(u, np.float32).reshape(-1, 3)
    n = len(u)
    wi = np.zeros((n, 3), np.float32)
    dist = np.zeros(n, np.float32)
    pdf = np.zeros(n, np.float32)
    tri = np.zeros(n, np.int32)
    area = np.zeros(1, np.float32)
    if lib.oracle_mesh_light_sample(flat.ptr, light, _fp(np.asarray(x, np.float32)), _fp(u), n, _fp(wi),
                                    _fp(dist), _fp(pdf), tri.ctypes.data_as(P_i32), _fp(area)) != 0:
        raise RuntimeError(_err(lib))
    return {"wi": wi, "dist": dist, "pdf": pdf, "tri": tri, "area": float(area[0])}


def light_sample(flat, light, xs, seed0=1, seq=7):
    """Samples of light `light` from (n, 3) points xs, sample i from the
    stream init(seed0 + i, seq): dict of wi (n, 3), dist, Li (n, 3), pdf,
    valid, delta, draws and pdf_back (the MIS-side pdf at the sampled point,
    NaN for mesh and delta lights)."""
    lib = load()
    xs = np.ascontiguousarray(xs, np.float32).reshape(-1, 3)
    n = len(xs)
    wi = np.zeros((n, 3), np.float32)
    dist = np.zeros(n, np.float32)
    li = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    flags = np.zeros(n, np.int32)
    draws = np.zeros(n, np.int32)
    back = np.zeros(n, np.float32)
    if lib.oracle_light_sample(flat.ptr, light, _fp(xs), n, seed0, seq, _fp(wi), _fp(dist), _fp(li), _fp(pdf),
                               flags.ctypes.data_as(P_i32), draws.ctypes.data_as(P_i32), _fp(back)) != 0:
        raise RuntimeError(_err(lib))
    return {"wi": wi, "dist": dist, "Li": li, "pdf": pdf, "valid": (flags & 1) != 0,
            "delta": (flags & 2) != 0, "draws": draws, "pdf_back": back}


def light_pdf(flat, light, frm, hit, nhit):
    """The MIS-side solid-angle pdf of light `light` for (n, 3) arrays of
    shading points, points on the light and the light's normal there."""
    lib = load()
    q = np.ascontiguousarray(np.concatenate([np.reshape(frm, (-1, 3)), np.reshape(hit, (-1, 3)),
                                             np.reshape(nhit, (-1, 3))], axis=1), np.float32)
    out = np.zeros(len(q), np.float32)
    if lib.oracle_light_pdf(flat.ptr, light, _fp(q), len(q), _fp(out)) != 0:
        raise RuntimeError(_err(lib))
    return out


def camera_jitter(flat, px, py, s):
    lib = load()
    jx = np.zeros(1, np.float32)
    jy = np.zeros(1, np.float32)
    if lib.oracle_camera_jitter(flat.ptr, px, py, s, _fp(jx), _fp(jy)) != 0:
        raise RuntimeError(_err(lib))
    return float(jx[0]), float(jy[0])
