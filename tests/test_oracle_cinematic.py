"""Pins of the oracle's cinematic shading (oracle/frame.c; compute.wgsl:3499 `cinematic`,
get_color_supersample 4620-4648, get_color_single 4440-4523, frand 2454-2471, getSampleBiased
4425-4434, sunsky 520-640, LinearToSRGB 4600-4607) against what fixes them independently:

* frand's state is the LCG s -> 747796405 s + 1 (mod 2^32): checked against the closed-form
  jump s_n = a^n s_0 + (a^n - 1) / (a - 1) (mod 2^32); its output is PCG's RXS-M-XS permutation
  (shift by the top four bits + 4, multiply by 277803737, xorshift 22) mapped by f32(w) / 2^32;
* getSampleBiased draws the cosine-power lobe p(cos t) ~ cos^p t: unit vectors whose cosine to
  the axis has the distribution function c^(p+1) and whose azimuth is uniform (Kolmogorov-
  Smirnov bounds);
* sunsky against the Preetham / Hosek-style formulas of Fragmentarium's Sunsky.frag evaluated in
  float64 with textbook Rayleigh (3/16pi (1 + cos^2)) and Henyey-Greenstein phase functions, and
  below the horizon against its closed form (no extinction term: sky = E_sun * (bR PR + bM PM) /
  (bR + bM) * (1 - clamp((1 - s_z)^5)) / 100);
* LinearToSRGB against the sRGB transfer function of IEC 61966-2-1;
* the whole pixel (8 jittered samples, bounce, sun shadow ray, sky) on a floor that every
  primary ray hits and that no secondary ray can hit again, recomputed in float64 from those
  pieces and the oracle's own random stream; and frame invariants (alpha = the fraction of
  first rays that hit, nothing in view -> the reference's miss colour, repeatability).
"""
import math

import numpy as np
import pytest

import oracle
import temporalsetinversion_amd as T

L = oracle.lib()
A, C_INC = 747796405, 1
M32 = 1 << 32


def jump(s0: int, n: int) -> int:
    an = pow(A, n, (A - 1) * M32)
    geo = (an - 1) // (A - 1)           # 1 + a + ... + a^(n-1), exact before the reduction
    return (pow(A, n, M32) * s0 + C_INC * geo) % M32


def rxs_m_xs(states: np.ndarray) -> np.ndarray:
    s = states.astype(np.uint64)
    w = ((s >> ((s >> np.uint64(28)) + np.uint64(4))) ^ s) * np.uint64(277803737) % np.uint64(M32)
    w = (w >> np.uint64(22)) ^ w
    return w.astype(np.uint32)


@pytest.mark.parametrize("seed", [0, 1, 12345, 0xFFFFFFFF, 0x9E3779B9])
def test_frand_state_is_the_lcg_closed_form(seed):
    n = 300
    _, states = oracle.frand_seq(seed, n)
    for i in (0, 1, 2, 17, 128, n - 1):
        assert int(states[i]) == jump(seed, i + 1), i


def test_frand_output_permutation_and_ranges():
    vals01, states = oracle.frand_seq(777, 5000, unit=True)
    w = rxs_m_xs(states)
    want = (w.astype(np.float32) / np.float32(4294967296.0)).astype(np.float32)
    assert np.array_equal(vals01.view(np.uint32), want.view(np.uint32))
    vals, states2 = oracle.frand_seq(777, 5000)
    assert np.array_equal(states, states2)            # frand and frand0to1 advance alike
    assert np.array_equal(vals, (want * np.float32(2.0) - np.float32(1.0)).astype(np.float32))
    assert vals.min() >= -1.0 and vals.max() <= 1.0 and abs(vals.mean()) < 0.03
    # uniform on [0, 1]: Kolmogorov-Smirnov at 99.9 %
    u = np.sort(vals01.astype(np.float64))
    d = np.max(np.abs(u - (np.arange(1, u.size + 1) - 0.5) / u.size))
    assert d < 1.95 / math.sqrt(u.size)


def ks(u):
    u = np.sort(np.asarray(u, np.float64))
    return np.max(np.abs(u - (np.arange(1, u.size + 1) - 0.5) / u.size))


@pytest.mark.parametrize("d", [(0.0, 0.0, 1.0), (1.0, 1.0, 1.0), (-0.3, 2.0, 0.4), (5.0, -0.1, 0.2)])
def test_sample_biased_is_a_cosine_power_lobe(d):
    n, power = 6000, 3.0
    v = oracle.sample_biased(d, power, 4242, n).astype(np.float64)
    axis = np.asarray(d, np.float64) / np.linalg.norm(d)
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=2e-6)
    c = v @ axis
    assert c.min() >= -1e-6
    # P(cos <= c) = c^(p+1)
    assert ks(np.clip(c, 0, 1) ** (power + 1)) < 1.95 / math.sqrt(n)
    assert abs(c.mean() - (power + 1) / (power + 2)) < 4 * 0.15 / math.sqrt(n)
    # uniform azimuth around the axis
    e1 = np.cross(axis, [0.0, 0.0, 1.0] if abs(axis[2]) < 0.9 else [1.0, 0.0, 0.0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(axis, e1)
    phi = np.arctan2(v @ e2, v @ e1)
    assert ks((phi + math.pi) / (2 * math.pi)) < 1.95 / math.sqrt(n)


# ---- sunsky in float64 (Fragmentarium Sunsky.frag as in compute.wgsl:423-640)
BETA_R = np.array([5.176821105123963e-6, 1.2785348189936485e-5, 2.8530756026157178e-5])
LAMBDA = np.array([680e-9, 550e-9, 450e-9])
K_MIE = np.array([0.686, 0.678, 0.666])


def beta_mie(turbidity=1.0):
    c = 0.2 * turbidity * 1e-17
    return 0.434 * c * math.pi * (2 * math.pi / LAMBDA) ** 2 * K_MIE * 0.005


def sun_dir(sp):
    px, py = (sp[0] - 0.0) * 6.28, (sp[1] - 0.5) * 3.14
    v = np.array([math.cos(px) * math.sin(py), math.sin(px) * math.sin(py), math.cos(py)])
    return v / np.linalg.norm(v)


def rayleigh(c):
    return 3.0 / (16.0 * math.pi) * (1.0 + c * c)


def henyey_greenstein(c, g):
    return (1.0 - g * g) / (4.0 * math.pi * (1.0 + g * g - 2.0 * g * c) ** 1.5)


def sky64(view, sp=(-0.1, 0.7)):
    s = sun_dir(sp)
    v = np.asarray(view, np.float64)
    cvs, zc = float(v @ s), float(s[2])
    e_sun = 1000.0 * max(0.0, 1.0 - math.exp(-(1.6110731611233942 - math.acos(zc)) / 1.5))
    bm = beta_mie()
    z = max(0.0, float(v[2]))
    fex = np.exp(-(BETA_R * 8400.0 + bm * 1250.0) / z) if z > 0 else np.zeros(3)
    lin = e_sun * (BETA_R * rayleigh(cvs) + bm * henyey_greenstein(cvs, 0.8)) / (BETA_R + bm)
    f = min(max((1.0 - s[2]) ** 5, 0.0), 1.0)
    return 0.01 * lin * (1.0 - fex) * ((1.0 - f) + np.sqrt(lin * fex) * f)


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("view", [(0.2, 0.1, -0.97), (-0.6, 0.3, -0.2), (0.1, -0.99, -0.05)])
def test_sunsky_below_the_horizon_closed_form(view):
    v = unit(view)
    s = sun_dir((-0.1, 0.7))
    e_sun = 1000.0 * max(0.0, 1.0 - math.exp(-(1.6110731611233942 - math.acos(s[2])) / 1.5))
    bm = beta_mie()
    c = float(v @ s)
    want = 0.01 * e_sun * (BETA_R * rayleigh(c) + bm * henyey_greenstein(c, 0.8)) / (BETA_R + bm) \
        * (1.0 - min((1.0 - s[2]) ** 5, 1.0))
    got = oracle.sunsky(v.astype(np.float32))
    assert np.allclose(got, want, rtol=2e-5), (got, want)


@pytest.mark.parametrize("view", [(0.3, 0.2, 0.93), (0.0, 0.0, 1.0), (-0.7, 0.5, 0.1), (0.45, -0.35, 0.82),
                                  (0.9, 0.1, 0.02)])
def test_sunsky_above_the_horizon(view):
    v = unit(view).astype(np.float32)
    got = oracle.sunsky(v)
    want = sky64(v.astype(np.float64))
    assert np.all(got > 0) and np.allclose(got, want, rtol=5e-5), (got, want)


def test_linear_to_srgb_is_the_srgb_transfer_function():
    xs = np.concatenate([np.linspace(-0.1, 1.2, 3001), np.geomspace(1e-7, 1.0, 2000)]).astype(np.float32)
    got = np.array([L.orc_linear_to_srgb(float(x)) for x in xs], np.float64)
    x = np.clip(xs.astype(np.float64), 0.0, 1.0)
    want = np.where(x < 0.0031308, 12.92 * x, 1.055 * x ** (1 / 2.4) - 0.055)
    assert np.allclose(got, want, rtol=3e-6, atol=1e-7)


def test_new_det_math_accuracy():
    rng = np.random.default_rng(99)
    xs = rng.uniform(-60, 60, 3000).astype(np.float32)
    got = np.array([L.orc_det_cosf(float(x)) for x in xs])
    assert np.max(np.abs(got - np.cos(xs.astype(np.float64)))) < 3e-7
    a = rng.uniform(-1, 1, 3000).astype(np.float32)
    got = np.array([L.orc_det_acosf(float(x)) for x in a])
    assert np.max(np.abs(got - np.arccos(a.astype(np.float64)))) < 1e-6
    assert math.isnan(L.orc_det_acosf(1.5)) and L.orc_det_acosf(1.0) == 0.0
    b = rng.uniform(0, 4, 3000).astype(np.float32)
    e = rng.uniform(-3, 30, 3000).astype(np.float32)
    got = np.array([L.orc_det_powf(float(x), float(y)) for x, y in zip(b, e)])
    want = b.astype(np.float64) ** e.astype(np.float64)
    assert np.all(np.abs(got - want) <= 4e-6 * np.abs(want) + 1e-30)
    assert L.orc_det_powf(0.0, 0.25) == 0.0 and L.orc_det_powf(0.0, 0.0) == 1.0 and L.orc_det_powf(1.0, 7.5) == 1.0
    assert math.isnan(L.orc_det_powf(-0.5, 2.0))


# ---- whole frames

FLOOR_LL = ("define dso_local float @_Z1fffff(float %0, float %1, float %2, float %3) #0 {\n"
            "  %5 = fadd float %1, 1.000000e+00\n  ret float %5\n}\n")


def look_down(eye):
    """Camera at `eye` looking straight down (-y), image up = -z."""
    m = np.zeros((4, 4), np.float32)
    m[0, :3] = (1.0, 0.0, 0.0)      # side
    m[1, :3] = (0.0, 0.0, -1.0)     # up
    m[2, :3] = (0.0, 1.0, 0.0)      # back: the view ray is -back = -y
    m[3, :3] = eye
    m[3, 3] = 1.0
    return m.reshape(16)


def cinematic_frame(ll, L_, view, W, H, t=0.0, shading=T._lib.TSI_SHADE_CINEMATIC):
    fr = oracle.Frame(oracle.parse_ll(ll), localsubdiv=L_, voxel_capacity=1 << 21, shading=shading)
    fr.set_inversion(t)
    rgba, depth, visits = fr.trace(view, W, H, t, T.camera_zdist(H))
    fr.close()
    return rgba, depth, visits


def test_floor_pixels_recomputed_in_float64():
    """A floor (y = -1) under a camera looking down: every primary ray hits it; with gradient
    normals the hit normal is (0, 1, 0), so every bounce ray and sun ray leaves upwards and hits
    nothing.  Each sample is then lum * (3 sunlight + spec + sky(bounce)) with lum the fresh-voxel
    colour (0.7, 1e-4, 1e-4) (frame 0: movement colour, 4346-4379)."""
    W, H, eye = 12, 10, (0.3, 2.0, -0.2)
    view = look_down(eye)
    shading = T._lib.TSI_SHADE_CINEMATIC | T._lib.TSI_SHADE_GRADIENT
    rgba, depth, _ = cinematic_frame(FLOOR_LL, 5, view, W, H, shading=shading)
    assert np.all(rgba[..., 3] == 1.0)
    zdist = float(T.camera_zdist(H))
    m = view.reshape(4, 4).astype(np.float64)
    lum = np.array([np.float32(1) - np.float32(1 - np.float32(0.7)),
                    np.float32(1) - np.float32(1 - np.float32(0.0001)),
                    np.float32(1) - np.float32(1 - np.float32(0.0001))], np.float64)
    s = sun_dir((-0.1, 0.7))
    sundir = np.array([s[0], s[2], s[1]])
    n = np.array([0.0, 1.0, 0.0])
    o1, o2 = np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])   # ortho / cross of (0, 1, 0)
    srgb = lambda x: np.where(x < 0.0031308, 12.92 * x, 1.055 * np.clip(x, 0, 1) ** (1 / 2.4) - 0.055)
    worst = 0.0
    for y in range(H):
        for x in range(W):
            seed = (y * W + x + int(np.float32(12345.0) * (np.float32(L.orc_det_sinf(0.0)) * np.float32(0.5)
                                                           + np.float32(0.5)))) & 0xFFFFFFFF
            u, _ = oracle.frand_seq(seed, 8 * 7, unit=True)
            u = u.astype(np.float64)
            acc = np.zeros(3)
            for k in range(8):
                j = u[7 * k: 7 * k + 7]
                jx, jy = (2 * j[0] - 1) * 0.5, (2 * j[1] - 1) * 0.5
                d = unit([(x + jx) - W / 2 + 0.5, -((y + jy) - H / 2 + 0.5), -zdist])
                rd = m[0, :3] * d[0] + m[1, :3] * d[1] + m[2, :3] * d[2]
                r = j[3] ** 0.25
                om = math.sqrt(1 - r * r)
                bdir = math.cos(2 * math.pi * j[2]) * om * o1 + math.sin(2 * math.pi * j[2]) * om * o2 + r * n
                ssd = unit(sundir + 0.01 * (2 * j[4:7] - 1))
                half = unit(-rd + sundir)
                spec = 100.0 * max(float(half @ n), 0.0) ** 30
                bg = sky64(np.array([bdir[0], bdir[2], bdir[1]]))
                acc += lum * (3.0 * ssd[1] + spec) + lum * bg
            want = np.clip(srgb(np.clip(acc / 8, 0, 1)), 0, None) ** 0.45
            got = rgba[y, x, :3].astype(np.float64)
            worst = max(worst, float(np.max(np.abs(got - want) / np.maximum(want, 1e-3))))
    assert worst < 2e-4, worst


def test_nothing_in_view_gives_the_miss_colour_everywhere():
    from temporalsetinversion_amd import scenes
    view = T.look_at_inverse((3.0, 0.0, 0.0), target=(6.0, 0.0, 0.0))   # looking away
    rgba, depth, _ = cinematic_frame(scenes.sphere_unit(1.0), 5, view, 16, 12)
    miss = np.float32(L.orc_det_pow045(L.orc_linear_to_srgb(10.0)))
    assert np.all(rgba[..., 3] == 0.0) and np.all(rgba[..., :3] == miss) and np.all(depth == -1.0)


def test_alpha_counts_the_first_rays_that_hit():
    from temporalsetinversion_amd import scenes
    view = T.look_at_inverse((2.2, 1.7, 2.6))
    W, H = 40, 30
    rgba, depth, _ = cinematic_frame(scenes.SCENES["spheres"](), 6, view, W, H)
    a = rgba[..., 3]
    assert np.all(a * 8 == np.round(a * 8))
    assert np.mean(a == 1.0) > 0.01 and np.mean(a == 0.0) > 0.5 and np.any((a > 0) & (a < 1))
    # first-sample depth agrees with alpha: a hit there means some sample hit
    assert np.all((depth >= 0) <= (a > 0))
    again, _, _ = cinematic_frame(scenes.SCENES["spheres"](), 6, view, W, H)
    assert np.array_equal(rgba.view(np.uint32), again.view(np.uint32))
    later, _, _ = cinematic_frame(scenes.SCENES["spheres"](), 6, view, W, H, t=0.25)
    assert not np.array_equal(rgba, later)       # the seed follows the time (4667)
