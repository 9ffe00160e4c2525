"""Seed-generated benchmark scenes (the workloads the measurements use).

Every scene is a pure function of its parameters and seed: geometry
(icosphere OBJ), the environment image (Radiance .hdr) and all random
placements are generated here, written under `asset_dir`, and fed through
the unchanged scene-construction ABI.

  cornell(...)      Cornell-style box: 5 rect walls, lambert + glass spheres,
                    emissive ceiling rect (17, 12, 4); 256^2, 64 spp, depth 8.
  instanced(...)    4,096 instances of a displaced 20,480-triangle icosphere
                    on a 100 x 100 plane, 50% GGX metal / 50% lambert, 3 soft
                    point lights + a distant light; 1920x1080, 256 spp, depth 12.
  volumes(...)      HDR sky + 0.5 degree sun (2048x1024), box-bounded 5-octave
                    flame marched at step 0.01, rippled water (4 octaves, ior
                    1.33), rough glass panels at 0..0.6; 1280x720, 512 spp, depth 16.
"""

import math
import os
import struct

import numpy as np

from .scenelib import Scene, rotate_x, rotate_y, rotate_z, scale, translate


class Pcg32(object):
    """PCG-XSH-RR 64/32 (O'Neill 2014) for scene-layout draws."""

    MULT = 6364136223846793005
    MASK = (1 << 64) - 1

    def __init__(self, seed, seq=54):
        self.state = 0
        self.inc = ((seq << 1) | 1) & self.MASK
        self.next()
        self.state = (self.state + seed) & self.MASK
        self.next()

    def next(self):
        old = self.state
        self.state = (old * self.MULT + self.inc) & self.MASK
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xFFFFFFFF

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 8) * (1.0 / 16777216.0))


# ---------------------------------------------------------------------------
# geometry generators
# ---------------------------------------------------------------------------

def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint32) * np.uint32(73856093)) ^ (iy.astype(np.uint32) * np.uint32(19349663)) \
        ^ (iz.astype(np.uint32) * np.uint32(83492791)) ^ np.uint32(seed)
    h = h * np.uint32(747796405) + np.uint32(2891336453)
    h = ((h >> ((h >> np.uint32(28)) + np.uint32(4))) ^ h) * np.uint32(277803737)
    h = (h >> np.uint32(22)) ^ h
    return h.astype(np.float64) / 4294967296.0


def value_noise(p, seed):
    """Trilinear value noise in [0, 1) with smoothstep weights; p: (n, 3)."""
    f = np.floor(p)
    i = f.astype(np.int64)
    t = p - f
    t = t * t * (3.0 - 2.0 * t)
    acc = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = (t[:, 0] if dx else 1 - t[:, 0]) * (t[:, 1] if dy else 1 - t[:, 1]) * \
                    (t[:, 2] if dz else 1 - t[:, 2])
                acc += w * _hash3(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
    return acc


def icosphere(subdiv, amplitude=0.0, seed=1, freq=3.0):
    """Unit icosphere (20 * 4^subdiv triangles) with radial value-noise
    displacement r = 1 + amplitude * (2 n - 1)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
             (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(v, float) / np.linalg.norm(v) for v in verts]
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.array(verts)
    if amplitude:
        n = value_noise(v * freq + 17.0, seed)
        v = v * (1.0 + amplitude * (2.0 * n - 1.0))[:, None]
    return v.astype(np.float32), np.array(faces, np.int64)


def write_obj(path, verts, faces):
    with open(path, "w") as f:
        f.write("# seed-generated mesh\n")
        f.write("".join("v %.7g %.7g %.7g\n" % tuple(p) for p in verts))
        f.write("".join("f %d %d %d\n" % (a + 1, b + 1, c + 1) for a, b, c in faces))


def write_hdr(path, rgb):
    """Flat (uncompressed) Radiance RGBE writer; rgb: (h, w, 3) float."""
    h, w, _ = rgb.shape
    rgb = np.maximum(rgb.astype(np.float64), 0.0)
    m = rgb.max(axis=2)
    out = np.zeros((h, w, 4), np.uint8)
    nz = m > 1e-32
    e = np.zeros_like(m, dtype=np.int64)
    e[nz] = np.floor(np.log2(m[nz])).astype(np.int64) + 1
    # mantissas in [0, 256): value = mantissa * 2^(e - 8)
    scalev = np.zeros_like(m)
    scalev[nz] = np.ldexp(1.0, (8 - e[nz]).astype(np.int64))
    mant = np.floor(rgb * scalev[..., None])
    mant = np.clip(mant, 0, 255)
    out[..., :3] = mant.astype(np.uint8)
    out[..., 3] = np.where(nz, e + 128, 0).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n")
        f.write(("-Y %d +X %d\n" % (h, w)).encode())
        f.write(out.tobytes())


def sky_image(width=2048, height=1024, sun_elev_deg=32.0, sun_azim_deg=40.0, sun_diam_deg=0.5,
              sun_radiance=1e5):
    """Equirect sky: log-normal luminance profile over elevation, blue-tinted,
    a dim ground below the horizon, and a sun disc of the given angular size."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = v * math.pi                      # 0 = zenith (+Y)
    phi = 2.0 * math.pi * u - math.pi
    elev = 90.0 - np.degrees(theta)          # degrees above the horizon
    e = np.maximum(elev, 0.0) + 1.0
    # log-normal in elevation: brightest a few degrees above the horizon
    lum = 0.35 + 2.2 * np.exp(-(np.log(e) - math.log(6.0)) ** 2 / (2.0 * 0.9 ** 2))
    sky = np.stack([0.55 * lum, 0.75 * lum, 1.15 * lum], axis=-1)
    ground = np.array([0.18, 0.15, 0.12])
    img = np.where((elev >= 0.0)[:, None], 1.0, 0.0)[:, :, None] * sky[:, None, :] + \
        np.where((elev < 0.0)[:, None], 1.0, 0.0)[:, :, None] * ground[None, None, :]
    img = np.broadcast_to(img, (height, width, 3)).copy()
    # sun disc
    se, sa = math.radians(sun_elev_deg), math.radians(sun_azim_deg)
    sdir = np.array([math.cos(se) * math.cos(sa), math.sin(se), math.cos(se) * math.sin(sa)])
    th, ph = np.meshgrid(theta, phi, indexing="ij")
    d = np.stack([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)], axis=-1)
    cosang = d @ sdir
    disc = cosang >= math.cos(math.radians(sun_diam_deg * 0.5))
    if not disc.any():  # keep at least the closest texel lit
        disc = cosang >= cosang.max()
    img[disc] = sun_radiance
    return img.astype(np.float32)


# ---------------------------------------------------------------------------
# scenes
# ---------------------------------------------------------------------------

def cornell(width=256, height=256, spp=64, max_depth=8, seed=1, clamp=0.0):
    s = Scene()
    s.render(width=width, height=height, spp=spp, max_depth=max_depth, seed=seed, clamp=clamp)
    s.camera(lookfrom=[0, 0, 3.9], lookat=[0, 0, 0], vfov=40)
    s.background_solid(color=[0, 0, 0])
    s.lambert("grey", color=[0.73, 0.73, 0.73])
    s.lambert("red", color=[0.65, 0.05, 0.05])
    s.lambert("green", color=[0.12, 0.45, 0.15])
    s.lambert("ball", color=[0.73, 0.73, 0.73])
    s.dielectric("glass", ior=1.5)
    s.emissive("lamp", color=[17, 12, 4], intensity=1.0)
    s.add("rect", "grey", translate(0, -1, 0))                      # floor
    s.add("rect", "grey", rotate_x(180), translate(0, 1, 0))        # ceiling
    s.add("rect", "grey", rotate_x(90), translate(0, 0, -1))        # back wall
    s.add("rect", "red", rotate_z(-90), translate(-1, 0, 0))        # left wall
    s.add("rect", "green", rotate_z(90), translate(1, 0, 0))        # right wall
    s.add("sphere", "ball", scale(0.35), translate(-0.4, -0.65, -0.3))
    s.add("sphere", "glass", scale(0.35), translate(0.4, -0.65, 0.3))
    s.add("rect", "lamp", scale(0.25, 1, 0.2), rotate_x(180), translate(0, 0.999, 0))
    return s


def instanced(asset_dir, width=1920, height=1080, spp=256, max_depth=12, seed=1, instances=4096,
              subdiv=5, layout_seed=1, clamp=0.0):
    os.makedirs(asset_dir, exist_ok=True)
    obj = os.path.join(asset_dir, "icosphere_s%d.obj" % subdiv)
    if not os.path.exists(obj):
        v, f = icosphere(subdiv, amplitude=0.1, seed=7)
        tmp = obj + ".tmp%d" % os.getpid()
        write_obj(tmp, v, f)
        os.replace(tmp, obj)
    s = Scene()
    s.render(width=width, height=height, spp=spp, max_depth=max_depth, seed=seed, clamp=clamp)
    s.camera(lookfrom=[0, 22, 62], lookat=[0, 0, 6], vfov=48)
    s.background_gradient(horizon=[0.85, 0.9, 1.0], zenith=[0.25, 0.4, 0.75])
    s.mesh("ico", obj, normals="smooth")
    s.lambert("ground", color=[0.45, 0.43, 0.4])
    s.add("rect", "ground", scale(50))
    rng = Pcg32(layout_seed)
    for i in range(instances):
        x = rng.uniform(-50.0, 50.0)
        z = rng.uniform(-50.0, 50.0)
        yaw = rng.uniform(0.0, 360.0)
        sx, sy, sz = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0)
        name = "m%05d" % i
        if rng.uniform() < 0.5:
            c = rng.uniform(0.55, 0.95)
            s.metal(name, color=[c, c * 0.95, c * 0.9], roughness=rng.uniform(0.05, 0.6))
        else:
            a = rng.uniform(0.1, 0.9)
            s.lambert(name, color=[a, a * rng.uniform(0.6, 1.0), a * rng.uniform(0.6, 1.0)])
        s.add("mesh:ico", name, scale(sx, sy, sz), rotate_y(yaw), translate(x, sy, z))
    s.point_light(position=[-20, 12, 10], intensity=[1800, 1500, 1200], radius=0.2)
    s.point_light(position=[18, 10, -15], intensity=[900, 1100, 1500], radius=0.2)
    s.point_light(position=[0, 15, 30], intensity=[1400, 1400, 1400], radius=0.2)
    s.distant_light(direction=[-0.4, -1.0, -0.3], radiance=[2.2, 2.0, 1.8])
    return s


def volumes(asset_dir, width=1280, height=720, spp=512, max_depth=16, seed=1, env_w=2048, env_h=1024,
            clamp=0.0, flame_step=0.01, flame_octaves=5, water_octaves=4, flame_seed=3):
    os.makedirs(asset_dir, exist_ok=True)
    hdr = os.path.join(asset_dir, "sky_%dx%d.hdr" % (env_w, env_h))
    if not os.path.exists(hdr):
        tmp = hdr + ".tmp%d" % os.getpid()
        write_hdr(tmp, sky_image(env_w, env_h))
        os.replace(tmp, hdr)
    s = Scene()
    s.render(width=width, height=height, spp=spp, max_depth=max_depth, seed=seed, clamp=clamp,
             transparent_shadows=True)
    s.camera(lookfrom=[0, 1.6, 5.0], lookat=[0, 0.6, -1.0], vfov=45)
    s.background_envmap(hdr, rotate=0.0, intensity=1.0, importance=True)
    s.water("water", ior=1.33, absorb=[0.45, 0.06, 0.02], wave_amp=0.02, wave_freq=2.0,
            octaves=water_octaves)
    s.lambert("sand", color=[0.76, 0.68, 0.5])
    s.lambert("stone", color=[0.35, 0.33, 0.3])
    s.add("rect", "water", scale(20))                                   # water surface, front = up
    s.add("rect", "sand", scale(20), translate(0, -1.0, 0))             # pool floor
    s.add("rect", "stone", scale(0.7, 1, 0.7), translate(-1.8, 0.05, -1.2))  # flame plinth
    for i, r in enumerate([0.0, 0.15, 0.3, 0.45, 0.6]):
        g = "glass%d" % i
        s.dielectric(g, ior=1.5, roughness=r)
        x = -2.0 + i * 1.0
        # a thin closed slab: front face toward +z, back face toward -z
        s.add("rect", g, scale(0.4, 1, 0.6), rotate_x(90), translate(x, 0.8, -2.0 + 0.03))
        s.add("rect", g, scale(0.4, 1, 0.6), rotate_x(-90), translate(x, 0.8, -2.0 - 0.03))
    s.flame(base=[-1.8, 0.06, -1.2], height=2.0, radius=0.5, intensity=20.0, sigma=4.0,
            noise_scale=3.0, seed=flame_seed, light_intensity=12.0, octaves=flame_octaves,
            step=flame_step, bounds="box")
    return s


def scene_by_name(name, asset_dir, **kw):
    if name == "cornell":
        return cornell(**kw)
    if name == "instanced":
        return instanced(asset_dir, **kw)
    if name == "volumes":
        return volumes(asset_dir, **kw)
    raise ValueError("unknown scene %r" % name)
