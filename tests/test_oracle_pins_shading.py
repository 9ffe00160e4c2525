"""Pins of the oracle's shading helpers against sources outside the oracle:
the textbook bilinear filter (scipy.ndimage), the published hash and
interpolation that define value noise, the height-field normal computed by
an independent finite difference, invariants of the flame's heat ramp, and
the area and solid-angle closed forms of mesh-light sampling. CPU only.
"""

import math
import os

import numpy as np
import pytest
from scipy import ndimage

import sundog_amd as sd
from sundog_amd import configs
from sundog_amd.scenelib import scale, translate


def _write_pfm(path, rgb):
    """Colour PFM, little-endian; rgb is (h, w, 3), top row first."""
    h, w, _ = rgb.shape
    with open(path, "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (w, h))
        f.write(np.ascontiguousarray(rgb[::-1], "<f4").tobytes())


def _scipy_bilinear(img, uv, wrap_v):
    """Textbook bilinear filter with texel centres at (i + 0.5) / n, wrap in
    u and wrap or clamp-to-edge in v, by scipy.ndimage.map_coordinates over
    an image padded by one texel on each side."""
    u = uv[:, 0].astype(np.float64)
    v = uv[:, 1].astype(np.float64)
    u = u - np.floor(u)
    if wrap_v:
        v = v - np.floor(v)
    h, w = img.shape[:2]
    x = u * w - 0.5
    y = v * h - 0.5
    pad = np.pad(img.astype(np.float64), ((0, 0), (1, 1), (0, 0)), mode="wrap")
    pad = np.pad(pad, ((1, 1), (0, 0), (0, 0)), mode="wrap" if wrap_v else "edge")
    return np.stack([ndimage.map_coordinates(pad[..., c], [y + 1, x + 1], order=1, mode="nearest")
                     for c in range(img.shape[2])], axis=1)


# ---------------------------------------------------------------------------
# textures (device/texture_eval.cuh:11-46) and the environment lookup
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("wrap_v", [True, False])
def test_bilinear_equals_textbook_filter(orc, wrap_v):
    rng = np.random.default_rng(3 + wrap_v)
    img = np.ones((7, 11, 4), np.float32)
    img[..., :3] = rng.uniform(0, 4, (7, 11, 3))
    uv = rng.uniform(-1.5, 2.5, (3000, 2)).astype(np.float32)
    if not wrap_v:
        uv[:, 1] = rng.uniform(-0.2, 1.2, 3000)   # clamp-v is only asked for in [0, 1]
    got = orc.bilinear(img, uv, wrap_v)
    want = _scipy_bilinear(img, uv, wrap_v)
    assert np.allclose(got, want, rtol=2e-5, atol=2e-5), np.abs(got - want).max()
    # a texel centre returns that texel exactly
    cu = (np.arange(11) + 0.5) / 11
    cv = (np.arange(7) + 0.5) / 7
    g = np.stack(np.meshgrid(cu, cv), axis=-1).reshape(-1, 2).astype(np.float32)
    # (to float32 rounding of the texel-centre coordinates)
    assert np.allclose(orc.bilinear(img, g, wrap_v), img.reshape(-1, 4), rtol=0, atol=1e-5)


def test_image_texture_v_is_up_and_wraps(orc, asset_dir):
    # v = 0 is the image's bottom row (the OBJ / OpenGL texture convention),
    # and both coordinates wrap
    rng = np.random.default_rng(5)
    img = rng.uniform(0, 1, (6, 9, 3)).astype(np.float32)
    path = os.path.join(asset_dir, "pin_tex.pfm")
    _write_pfm(path, img)
    s = sd.Scene()
    s.render(width=8, height=8, spp=1)
    s.camera(lookfrom=[0, 0, 5], lookat=[0, 0, 0])
    s.texture("t", "image", file=path)
    s.lambert("m", texture="t")
    s.add("sphere", "m")
    flat = s.build().flatten()
    uv = rng.uniform(-2, 3, (2000, 2)).astype(np.float32)
    got = orc.texture(flat, 0, uv)
    rgba = np.concatenate([img, np.ones((6, 9, 1), np.float32)], axis=2)
    flipped = np.stack([uv[:, 0], 1.0 - uv[:, 1]], axis=1)
    want = _scipy_bilinear(rgba, flipped, True)
    assert np.allclose(got[:, :3], want[:, :3], rtol=2e-5, atol=2e-5)
    # the centre of the bottom-left texel
    bl = orc.texture(flat, 0, np.array([[0.5 / 9, 0.5 / 6]], np.float32))[0, :3]
    assert np.allclose(bl, img[5, 0], rtol=1e-6)


def test_two_colour_square_and_grid_textures(orc):
    s = sd.Scene()
    s.render(width=8, height=8, spp=1)
    s.camera(lookfrom=[0, 0, 5], lookat=[0, 0, 0])
    s.texture("sq", "checker", a=[1, 0, 0], b=[0, 0, 1], scale=[4, 4])
    s.texture("gr", "grid", a=[0, 1, 0], b=[1, 1, 1], scale=[2, 2], width=0.05)
    s.lambert("m0", texture="sq")
    s.lambert("m1", texture="gr")
    s.add("sphere", "m0")
    s.add("sphere", "m1", translate(3, 0, 0))
    flat = s.build().flatten()
    sq, gr = 1, 0                # textures are registered in name order
    a, b = [1, 0, 0], [0, 0, 1]
    # squares of side 1/4; floor() keeps the alternation across u < 0
    pts = {(0.1, 0.1): a, (0.3, 0.1): b, (0.3, 0.3): a, (-0.1, 0.1): b, (-0.1, -0.1): a, (0.6, 0.9): b}
    got = orc.texture(flat, sq, np.array(list(pts), np.float32))
    assert np.array_equal(got[:, :3], np.array(list(pts.values()), np.float32))
    # grid cells of side 1/2; lines of width 0.05 (in cell units) on every edge
    pts = {(0.25, 0.25): [0, 1, 0], (0.01, 0.25): [1, 1, 1], (0.25, 0.49): [1, 1, 1],
           (0.26, 0.74): [0, 1, 0], (0.76, 0.26): [0, 1, 0], (0.5 + 0.02, 0.3): [1, 1, 1]}
    got = orc.texture(flat, gr, np.array(list(pts), np.float32))
    assert np.array_equal(got[:, :3], np.array(list(pts.values()), np.float32))


# ---------------------------------------------------------------------------
# value noise and fbm (device/noise.cuh:13-79)
# ---------------------------------------------------------------------------

M32 = 0xFFFFFFFF


def _pcg_hash(x):
    """pcg_hash of Jarzynski and Olano, 'Hash Functions for GPU Rendering',
    JCGT 9(3), 2020: one PCG step, then the RXS-M-XS output permutation."""
    state = (x * 747796405 + 2891336453) & M32
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M32
    return (word >> 22) ^ word


def _lattice(ix, iy, iz, seed):
    # Teschner et al. 2003 spatial hash primes, then pcg_hash, as [0, 1)
    h = ((ix * 73856093) & M32) ^ ((iy * 19349663) & M32) ^ ((iz * 83492791) & M32) ^ seed
    return _pcg_hash(h) / 4294967296.0


def test_value_noise_at_lattice_points_is_the_published_hash(orc):
    rng = np.random.default_rng(11)
    pts = rng.integers(-40, 40, (500, 3))
    for seed in (0, 7, 1337):
        got = orc.noise(pts.astype(np.float32), seed)
        want = np.array([_lattice(int(a), int(b), int(c), seed) for a, b, c in pts])
        assert np.array_equal(got, want.astype(np.float32)), seed


def test_value_noise_is_smoothstep_trilinear_of_its_lattice(orc):
    # between lattice points: trilinear interpolation of the eight corner
    # values with the fade s(t) = 3t^2 - 2t^3 on each axis (C1 at the faces)
    rng = np.random.default_rng(12)
    p = rng.uniform(-20, 20, (400, 3)).astype(np.float32)
    got = orc.noise(p, 5)
    f = np.floor(p.astype(np.float64))
    t = p - f
    s = t * t * (3 - 2 * t)
    want = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                c = np.array([_lattice(int(a) + dx, int(b) + dy, int(cc) + dz, 5) for a, b, cc in f])
                w = ((s[:, 0] if dx else 1 - s[:, 0]) * (s[:, 1] if dy else 1 - s[:, 1])
                     * (s[:, 2] if dz else 1 - s[:, 2]))
                want += w * c
    assert np.allclose(got, want, rtol=0, atol=2e-6)
    # zero slope at the lattice points: the fade's derivative vanishes there
    q = rng.integers(-9, 9, (50, 3)).astype(np.float32)
    e = np.float32(1e-3)
    for ax in range(3):
        d = np.zeros(3, np.float32)
        d[ax] = e
        slope = (orc.noise(q + d, 5) - orc.noise(q - d, 5)) / (2 * e)
        assert np.abs(slope).max() < 5e-3


def test_fbm_first_octave_and_range(orc):
    rng = np.random.default_rng(13)
    p = rng.uniform(-30, 30, (1000, 3)).astype(np.float32)
    assert np.array_equal(orc.noise(p, 9, octaves=1), np.float32(0.5) * orc.noise(p, 9))
    for n in (2, 4, 5):
        v = orc.noise(p, 9, octaves=n)
        # amplitudes 1/2, 1/4, ... on values in [0, 1): the sum stays in [0, 1 - 2^-n)
        assert v.min() >= 0.0 and v.max() < 1.0 - 2.0 ** -n
        assert not np.array_equal(v, orc.noise(p, 9, octaves=n - 1))


# ---------------------------------------------------------------------------
# the water normal (device/bsdf.cuh water lobe; the fbm slope field)
# ---------------------------------------------------------------------------

def _height_normal(orc, x, z, amp, freq, octaves, d=5e-3):
    """Normal of the surface y = (amp / freq) * fbm(freq x, 0, freq z) by a
    central difference ten times finer than the material's own."""
    def H(a, b):
        return (amp / freq) * float(orc.noise(np.array([[freq * a, 0, freq * b]], np.float32), 1337,
                                              octaves)[0])
    hx = (H(x + d, z) - H(x - d, z)) / (2 * d)
    hz = (H(x, z + d) - H(x, z - d)) / (2 * d)
    n = np.array([-hx, 1.0, -hz])
    return n / np.linalg.norm(n)


@pytest.mark.parametrize("octaves,mean_tol", [(1, 2e-3), (2, 4e-3), (4, 3e-2)])
def test_water_normal_is_the_height_field_normal(orc, octaves, mean_tol):
    rng = np.random.default_rng(20 + octaves)
    amp, freq = 0.3, 2.0
    err, swapped = [], []
    for x, z in rng.uniform(-3, 3, (150, 2)):
        n = orc.water_normal([x, 0, z], [0, 1, 0], [0, 1, 0], amp, freq, octaves)
        ref = _height_normal(orc, x, z, amp, freq, octaves)
        assert abs(np.linalg.norm(n) - 1) < 1e-5
        err.append(np.linalg.norm(n - ref))
        swapped.append(np.linalg.norm(n - ref[[2, 1, 0]]))
    # the material's coarser difference (step 0.05 in noise space) stays
    # close; x/z transposed would be an order of magnitude further off
    assert np.mean(err) < mean_tol, np.mean(err)
    assert np.mean(swapped) > 3 * np.mean(err)


def test_water_normal_flat_sides_and_facing(orc):
    # no waves: the geometric up vector on the side the shading normal is on
    assert np.array_equal(orc.water_normal([0.3, 0, 0.7], [0, 1, 0], [0.2, 0.9, 0], 0.0, 2.0, 4), [0, 1, 0])
    assert np.array_equal(orc.water_normal([0.3, 0, 0.7], [0, -1, 0], [0, -1, 0], 0.0, 2.0, 4), [0, -1, 0])
    # from below the perturbed normal is the mirror image of the one from above
    up = orc.water_normal([1.1, 0, -0.4], [0, 1, 0], [0, 1, 0], 0.3, 2.0, 3)
    dn = orc.water_normal([1.1, 0, -0.4], [0, -1, 0], [0, -1, 0], 0.3, 2.0, 3)
    assert np.array_equal(dn, -up)
    # a perturbed normal facing away from wo falls back to the shading normal
    wo = -up + np.array([0, 0.05, 0], np.float32)   # grazing, behind the perturbed normal
    wo /= np.linalg.norm(wo)
    ns = np.array([0, 1, 0], np.float32)
    assert np.array_equal(orc.water_normal([1.1, 0, -0.4], ns, wo, 0.3, 2.0, 3), ns)


# ---------------------------------------------------------------------------
# the flame's density and heat ramp (device/volume.cuh:27-60)
# ---------------------------------------------------------------------------

def _flame_flat(intensity=20.0, sigma=4.0, seed=3):
    s = sd.Scene()
    s.render(width=8, height=8, spp=1)
    s.camera(lookfrom=[0, 0, 5], lookat=[0, 0, 0])
    s.lambert("a", color=[0.5] * 3)
    s.add("sphere", "a", scale(0.01), translate(0, -50, 0))
    s.flame(base=[0, 0, 0], height=2.0, radius=0.5, intensity=intensity, sigma=sigma, noise_scale=3.0,
            seed=seed)
    return s.build().flatten()


def _flame_points(rng, n):
    r = 0.7 * np.sqrt(rng.uniform(0, 1, n))
    a = rng.uniform(0, 2 * np.pi, n)
    y = rng.uniform(-0.2, 2.2, n)
    return np.stack([r * np.cos(a), y, r * np.sin(a)], axis=1).astype(np.float32)


def test_flame_heat_ramp_invariants(orc):
    rng = np.random.default_rng(30)
    p = _flame_points(rng, 20000)
    flat = _flame_flat()
    sg, em = orc.flame_field(flat, 0, p)
    # nothing outside the flame's height range
    out = (p[:, 1] <= 0.0) | (p[:, 1] >= 2.0)
    assert (sg[out] == 0).all() and (em[out] == 0).all()
    inside = sg > 0
    assert inside.sum() > 1000
    mask = sg[inside] / 4.0
    col = em[inside] / (20.0 * mask[:, None])
    # colour channels in [0, 1], ordered red >= green >= blue (a warm ramp)
    assert col.min() >= 0 and col.max() <= 1 + 1e-5
    assert (col[:, 0] >= col[:, 1] - 1e-6).all() and (col[:, 1] >= col[:, 2] - 1e-6).all()
    # a one-parameter monotone ramp: ordered by total brightness, no
    # channel ever falls
    c = col[np.argsort(col.sum(axis=1), kind="stable")]
    assert (np.diff(c, axis=0) >= -1e-5).all()
    # zero heat is black: a cold rim with density and no emission exists
    cold = (col.sum(axis=1) == 0)
    assert cold.any() and (~cold).any()
    assert (em[inside][cold] == 0).all()


def test_flame_emission_and_density_are_linear_in_their_scales(orc):
    rng = np.random.default_rng(31)
    p = _flame_points(rng, 4000)
    s1, e1 = orc.flame_field(_flame_flat(20.0, 4.0), 0, p)
    s2, e2 = orc.flame_field(_flame_flat(40.0, 4.0), 0, p)
    s3, e3 = orc.flame_field(_flame_flat(20.0, 8.0), 0, p)
    assert np.array_equal(s1, s2) and np.allclose(e2, 2 * e1, rtol=1e-6, atol=0)
    assert np.allclose(s3, 2 * s1, rtol=1e-6, atol=0) and np.array_equal(e1, e3)


def test_flame_profile_and_ramp_follow_the_published_closed_forms(orc):
    """The reference's report states the field in closed form
    (docs/report/12-volumes.md:62-76): radius r(h) = 3.2 sqrt(h) (1 - h)^1.55,
    widest at h* = 1 / (1 + 2 x 1.55) with r(h*) ~ 1; normalised distance
    d' = d / r(h) + (n - 0.45); heat = (0.88 - d')(1.25 - h) clamped to [0, 1];
    colour (min(2.6 heat, 1), min(2.1 heat^2, 1), min(1.7 heat^3, 1)). With
    noise_scale 0 every point samples the noise at the origin, so n is the
    pinned fbm's value there and each formula can be checked on its own."""
    H, R, SIG, INT, SEED, OCT = 2.0, 0.5, 4.0, 20.0, 3, 3
    s = sd.Scene()
    s.render(width=8, height=8, spp=1)
    s.camera(lookfrom=[0, 0, 5], lookat=[0, 0, 0])
    s.lambert("a", color=[0.5] * 3)
    s.add("sphere", "a", scale(0.01), translate(0, -50, 0))
    s.flame(base=[0, 0, 0], height=H, radius=R, intensity=INT, sigma=SIG, noise_scale=0.0, seed=SEED,
            octaves=OCT)
    flat = s.build().flatten()
    n0 = float(orc.noise(np.zeros((1, 3), np.float32), SEED, octaves=OCT)[0])

    hs = np.array([0.24390244])
    r64 = 3.2 * np.sqrt(hs) * (1 - hs) ** 1.55
    assert abs(r64[0] - 1.0) < 0.03          # the 3.2 normalises the peak radius to ~1 (1.025)

    rng = np.random.default_rng(41)
    h = rng.uniform(0.02, 0.98, 40000)
    dn = rng.uniform(0.0, 1.6, 40000)          # radial distance in flame radii
    a = rng.uniform(0, 2 * np.pi, 40000)
    p = np.stack([dn * R * np.cos(a), h * H, dn * R * np.sin(a)], axis=1).astype(np.float32)
    sg, em = orc.flame_field(flat, 0, p)
    # evaluate the closed forms in float64 at the float32 points' own h and d
    hp = p[:, 1].astype(np.float64) / H
    dp = np.hypot(p[:, 0].astype(np.float64), p[:, 2].astype(np.float64)) / R
    r = 3.2 * np.sqrt(hp) * (1 - hp) ** 1.55
    dd = dp / np.maximum(r, 1e-3) + (n0 - 0.45)
    # support: density exactly where d' < 1 (away from float rounding at the edge)
    clear = np.abs(dd - 1.0) > 1e-4
    assert ((sg > 0) == (dd < 1.0))[clear].all()
    inside = (sg > 0) & clear
    assert inside.sum() > 5000 and (~inside).sum() > 5000
    # colour = emission / (intensity x mask), mask = density / sigma
    col = em[inside] / (INT * (sg[inside] / SIG))[:, None]
    heat = np.clip((0.88 - dd[inside]) * (1.25 - hp[inside]), 0.0, 1.0)
    want = np.stack([np.minimum(2.6 * heat, 1), np.minimum(2.1 * heat ** 2, 1),
                     np.minimum(1.7 * heat ** 3, 1)], axis=1)
    assert np.abs(col - want).max() < 2e-4
    # the ramp is exercised below, through and above each channel's knee
    assert (heat < 0.2).any() and ((heat > 0.5) & (heat < 0.6)).any() and (heat > 0.9).any()


# ---------------------------------------------------------------------------
# mesh-light sampling (src/capi_render.cpp:119-172, device/light_sample.cuh:94-154)
# ---------------------------------------------------------------------------

# four triangles in the plane y = 0, all wound so that their normal is -y
TRIS = np.array([
    [[0, 0, 0], [1, 0, 0], [0, 0, 1]],            # area 0.5
    [[2, 0, 0], [4, 0, 0], [2, 0, 1]],            # area 1
    [[5, 0, 0], [8, 0, 0], [5, 0, 2]],            # area 3
    [[0, 0, 3], [0.5, 0, 3], [0, 0, 3.1]],        # area 0.025
], np.float64)


def _mesh_light_flat(asset_dir, steps, two_sided=False):
    path = os.path.join(asset_dir, "pin_tris.obj")
    configs.write_obj(path, TRIS.reshape(-1, 3), [(3 * i, 3 * i + 1, 3 * i + 2) for i in range(len(TRIS))])
    s = sd.Scene()
    s.render(width=8, height=8, spp=1)
    s.camera(lookfrom=[0, 5, 5], lookat=[0, 0, 0])
    s.emissive("e", color=[1, 1, 1], intensity=3.0, two_sided=two_sided)
    s.add(s.mesh("tris", path), "e", *steps)
    # NEE lights are numbered in object order: the mesh is light 0 (the
    # oracle's probe rejects an id that is not a mesh light)
    return s.build().flatten(), 0


def _world(steps_matrix, pts):
    return pts @ steps_matrix[:3, :3].T + steps_matrix[:3, 3]


def test_mesh_light_picks_triangles_by_world_area(orc, asset_dir):
    # a non-uniform scale changes the areas' ratios: the CDF is world-space
    flat, li = _mesh_light_flat(asset_dir, [scale(2.0, 1.0, 0.5), translate(0, 1, 0)])
    Mx = np.diag([2.0, 1.0, 0.5, 1.0])
    Mx[1, 3] = 1.0
    w = _world(Mx, TRIS.reshape(-1, 3)).reshape(-1, 3, 3)
    areas = 0.5 * np.linalg.norm(np.cross(w[:, 1] - w[:, 0], w[:, 2] - w[:, 0]), axis=1)
    N = 20000
    u = np.stack([(np.arange(N) + 0.5) / N, np.full(N, 0.5), np.full(N, 0.5)], axis=1)
    r = orc.mesh_light_sample(flat, li, [1.0, -5.0, 0.5], u)
    assert r["area"] == pytest.approx(areas.sum(), rel=1e-6)
    counts = np.bincount(r["tri"], minlength=len(TRIS))
    assert np.abs(counts - N * areas / areas.sum()).max() <= 1.0
    assert (np.diff(r["tri"]) >= 0).all()    # the pick is monotone in uc


def test_mesh_light_points_uniform_and_pdf_solid_angle(orc, asset_dir):
    flat, li = _mesh_light_flat(asset_dir, [translate(0, 1, 0)])
    x = np.array([3.0, -4.0, 1.0])
    total = 0.5 + 1 + 3 + 0.025
    # uc inside the big triangle's interval, (r0, r1) on a stratified grid
    g = (np.arange(60) + 0.5) / 60
    r0, r1 = np.meshgrid(g, g)
    uc = (0.5 + 1 + 1.5) / total
    u = np.stack([np.full(r0.size, uc), r0.ravel(), r1.ravel()], axis=1)
    r = orc.mesh_light_sample(flat, li, x, u)
    assert (r["tri"] == 2).all() and (r["pdf"] > 0).all()
    p = x + r["wi"] * r["dist"][:, None]
    a, b, c = TRIS[2] + [0, 1, 0]
    # barycentric coordinates of each point: all in the triangle, in its plane
    T = np.stack([b - a, c - a], axis=1)
    bc, res, _, _ = np.linalg.lstsq(T, (p - a).T, rcond=None)
    assert np.abs(p[:, 1] - 1.0).max() < 1e-5
    assert bc.min() > -1e-5 and (bc.sum(axis=0)).max() < 1 + 1e-5
    # uniform in area: the mean is the centroid and each corner's third of
    # the triangle (nearest-vertex regions) holds a third of the points
    assert np.allclose(p.mean(axis=0), (a + b + c) / 3, atol=5e-3)
    l0 = 1 - bc.sum(axis=0)
    near = np.argmax(np.stack([l0, bc[0], bc[1]]), axis=0)
    assert np.allclose(np.bincount(near, minlength=3) / len(near), 1 / 3, atol=0.02)
    # solid-angle pdf = d^2 / (|cos theta_l| * total area), normal from the winding
    ng = np.cross(b - a, c - a)
    ng /= np.linalg.norm(ng)
    cosl = np.abs(r["wi"] @ ng)
    assert np.allclose(r["pdf"], r["dist"] ** 2 / (cosl * total), rtol=2e-5)


def test_mesh_light_emits_on_its_normal_side_mirrored_or_not(orc, asset_dir):
    # the triangles' winding gives normal -y; a one-sided light is seen only
    # from below. A mirroring transform keeps the emitting side geometric.
    u = np.array([[0.7, 0.3, 0.6]])
    for steps in ([translate(0, 1, 0)], [scale(-1.0, 1.0, 1.0), translate(0, 1, 0)]):
        flat, li = _mesh_light_flat(asset_dir, steps)
        x0 = -6.0 if len(steps) == 2 else 6.0
        below = orc.mesh_light_sample(flat, li, [x0, -2.0, 1.0], u)
        above = orc.mesh_light_sample(flat, li, [x0, 4.0, 1.0], u)
        assert below["pdf"][0] > 0 and above["pdf"][0] == 0
    flat, li = _mesh_light_flat(asset_dir, [translate(0, 1, 0)], two_sided=True)
    assert orc.mesh_light_sample(flat, li, [6.0, 4.0, 1.0], u)["pdf"][0] > 0
