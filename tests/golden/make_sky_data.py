"""Generates the FP64 golden vectors for the benchmark sky's bake (turbidity 3, solar elevation 30 deg).

Run only where the rtcamp8 tree is mounted: it calls the Hosek-Wilkie 2013 model rtcamp8 vendors
(ext/hosek_wilkie_2013, 3-clause BSD), compiled into oracle/_ref/libhosek_ref.so by oracle/ref/Makefile,
and reads the CIE 1931 2-degree colour-matching tables out of common/basic_types.cpp as text.

Outputs
  tests/golden/hosek_golden.npz -- FP64 sky radiance (RGB after the reference's colour pipeline,
      analytic_sky_kernels.cu:220-276) for a set of view directions including the solar disc and the
      lower hemisphere, evaluated with the vendored FP64 model.
"""
import ctypes as C
import math
import re
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
REF = Path("/root/reference")
TURBIDITY = 3.0
ELEVATION = math.radians(30.0)
SUN_AZIMUTH = math.radians(35.0)
NUM_BANDS = 16


def hosek_lib():
    so = ROOT / "oracle" / "_ref" / "libhosek_ref.so"
    if not so.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle" / "ref")], check=True)
    lib = C.CDLL(str(so))
    d = C.c_double
    lib.hosek_state_size.restype = C.c_int
    lib.hosek_state.argtypes = [d, d, d, C.c_void_p]
    lib.hosek_radiance.restype = d
    lib.hosek_radiance.argtypes = [d] * 6
    lib.hosek_solar_radiance.restype = d
    lib.hosek_solar_radiance.argtypes = [d] * 6
    lib.hosek_solar_table.argtypes = [C.c_int, C.c_void_p]
    lib.hosek_limb_table.argtypes = [C.c_int, C.c_void_p]
    return lib


def cie_tables():
    text = (REF / "common" / "basic_types.cpp").read_text(errors="replace")
    out = {}
    for name in ("xbar_CIE1931_2deg", "ybar_CIE1931_2deg", "zbar_CIE1931_2deg"):
        m = re.search(name + r"\s*\[\s*\]\s*=\s*\{([^}]*)\}", text)
        assert m, name
        nums = re.findall(r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?", m.group(1))
        out[name] = np.array([float(v) for v in nums], dtype=np.float32)
    return out["xbar_CIE1931_2deg"], out["ybar_CIE1931_2deg"], out["zbar_CIE1931_2deg"]


def discretized_cmfs(xbar, ybar, zbar, num_bands):
    """computeDiscretizedXyzCmfs (common/basic_types.cpp:195-248) in float32."""
    f = np.float32
    hi, lo = f(830.0), f(360.0)
    n = len(xbar)
    xs, ys, zs = np.zeros(num_bands, f), np.zeros(num_bands, f), np.zeros(num_bands, f)
    s_bin = 0
    next_p = f(s_bin + 1) / f(num_bands)
    x_sum = y_sum = z_sum = f(0)
    x_prev, y_prev, z_prev = xbar[0], ybar[0], zbar[0]
    interval = f(1)
    for i in range(1, n):
        cur_p = f(i) / f(n - 1)
        width = interval
        x_cur, y_cur, z_cur = xbar[i], ybar[i], zbar[i]
        if cur_p >= next_p:
            width = (cur_p - next_p) * (hi - lo)
            t = f(1) - width / interval
            x_in = x_prev * (f(1) - t) + x_cur * t
            y_in = y_prev * (f(1) - t) + y_cur * t
            z_in = z_prev * (f(1) - t) + z_cur * t
            x_sum += (x_prev + x_in) * (interval - width) * f(0.5)
            y_sum += (y_prev + y_in) * (interval - width) * f(0.5)
            z_sum += (z_prev + z_in) * (interval - width) * f(0.5)
            if s_bin < num_bands:
                xs[s_bin], ys[s_bin], zs[s_bin] = x_sum, y_sum, z_sum
            x_sum = y_sum = z_sum = f(0)
            x_prev, y_prev, z_prev = x_in, y_in, z_in
            s_bin += 1
            next_p = f(s_bin + 1) / f(num_bands)
        x_sum += (x_prev + x_cur) * width * f(0.5)
        y_sum += (y_prev + y_cur) * width * f(0.5)
        z_sum += (z_prev + z_cur) * width * f(0.5)
        x_prev, y_prev, z_prev = x_cur, y_cur, z_cur
    centers = np.array([lo + (f(i) + f(0.5)) * (hi - lo) / f(num_bands) for i in range(num_bands)], dtype=f)
    integral = f(0)
    for i in range(num_bands):
        integral += ys[i]
    return xs, ys, zs, centers, integral


def band_albedo(b):
    return 0.5 * (1.0 - (b + 0.5) / NUM_BANDS)


def sky_rgb_fp64(lib, view, sun, cmf, elevation=ELEVATION, turbidity=TURBIDITY):
    """FP64 restatement of generateArHosekSkyEnvironmentalTexture's per-texel pipeline
    (analytic_sky_kernels.cu:227-274) for one view direction given as (phi, theta)."""
    phi, theta = view
    xs, ys, zs, centers, integral = cmf
    lower = theta >= 0.5 * math.pi
    m_theta = math.pi - theta if lower else theta
    vv = np.array([-math.sin(phi) * math.sin(m_theta), math.cos(m_theta), math.cos(phi) * math.sin(m_theta)])
    gamma = math.acos(max(-1.0, min(1.0, float(vv @ sun))))
    xyz = np.zeros(3)
    for b in range(NUM_BANDS):
        wl = min(max(float(centers[b]), 320.0), 720.0)
        alb = band_albedo(b)
        if gamma < (0.51 * math.pi / 180.0) / 2.0 and not lower:
            v = lib.hosek_solar_radiance(elevation, turbidity, alb, m_theta, gamma, wl)
        else:
            v = lib.hosek_radiance(elevation, turbidity, alb, m_theta, gamma, wl)
        xyz += np.array([xs[b], ys[b], zs[b]], dtype=np.float64) * v
    xyz /= float(integral)
    m = np.array([[3.2404542, -1.5371385, -0.4985314], [-0.9692660, 1.8760108, 0.0415560],
                  [0.0556434, -0.2040259, 1.0572252]])
    rgb = m @ xyz
    max_v = rgb.max()
    rgb = rgb / max_v
    mn, mx = rgb.min(), rgb.max()
    h, s = 0.0, 0.0
    delta = mx - mn
    if not (delta < 1e-5 or mx == 0.0):
        s = delta / mx
        if rgb[0] >= mx:
            h = (rgb[1] - rgb[2]) / delta
        elif rgb[1] >= mx:
            h = 2.0 + (rgb[2] - rgb[0]) / delta
        else:
            h = 4.0 + (rgb[0] - rgb[1]) / delta
        h = math.fmod(h / 6 + 1.0, 1.0)
    diff_hue = math.fmod(h - 2.0 / 3 + 1.0, 1.0)
    t_sat = s * (0.5 * (1 + math.cos(2 * math.pi * diff_hue)))
    s = (1 - t_sat) * s + t_sat * 1.0  # lerp(s, 1, t) = (1 - t) s + t (common/common_shared.h:234-236)
    # HSVtoRGB(h, s, 1) (common/basic_types.h:3114-3137)
    if s == 0:
        out = np.array([1.0, 1.0, 1.0])
    else:
        hh = h - math.floor(h)
        hi = int(hh * 6)
        f = hh * 6 - hi
        mm, nn, kk = 1 - s, 1 - s * f, 1 - s * (1 - f)
        out = [np.array([1, kk, mm]), np.array([nn, 1, mm]), np.array([mm, 1, kk]), np.array([mm, nn, 1]),
               np.array([kk, mm, 1]), np.array([1, mm, nn])][hi] if hi < 6 else np.zeros(3)
    rgb = max_v * np.asarray(out, dtype=np.float64)
    if lower:
        sea = 0.5 * np.array([0.01, 0.01, (1 - sun[1]) * 0.01 + sun[1] * 0.1])
        t = math.cos(m_theta) ** (1 / 5.0)
        rgb = (1 - t) * rgb + t * sea
    return rgb


def main() -> int:
    lib = hosek_lib()
    cmf = discretized_cmfs(*cie_tables(), NUM_BANDS)
    # golden directions: a regular grid over the sphere plus points on and around the solar disc
    sun = np.array([math.cos(ELEVATION) * math.sin(SUN_AZIMUTH), math.sin(ELEVATION),
                    math.cos(ELEVATION) * math.cos(SUN_AZIMUTH)])
    sun32 = sun.astype(np.float32)
    sun64 = sun32.astype(np.float64)
    views = []
    for i in range(24):
        for j in range(48):
            views.append(((j + 0.37) / 48 * 2 * math.pi, (i + 0.61) / 24 * math.pi))
    sun_theta = math.acos(sun64[1])
    sun_phi = math.atan2(-sun64[0], sun64[2]) % (2 * math.pi)
    rng = np.random.default_rng(11)
    for k in range(64):  # inside / near the disc (radius 0.00445 rad)
        r = 0.01 if k < 24 else 0.003
        views.append((sun_phi + rng.uniform(-r, r) / math.sin(sun_theta), sun_theta + rng.uniform(-r, r)))
    views = np.array(views, dtype=np.float32).astype(np.float64)  # evaluate exactly at the FP32 inputs
    rgb = np.array([sky_rgb_fp64(lib, v, sun64, cmf) for v in views])
    np.savez_compressed(ROOT / "tests" / "golden" / "hosek_golden.npz", phi=views[:, 0].astype(np.float32),
                        theta=views[:, 1].astype(np.float32), sun=sun32, rgb=rgb)
    print(f"wrote hosek_golden.npz ({len(views)} directions), integral_cmf {cmf[4]}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
