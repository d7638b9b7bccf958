"""Writes the spectral Hosek-Wilkie 2013 model's tables for skies of any sun elevation and turbidity,
and golden vectors for the per-band states and the baked sky.

Run only where the rtcamp8 tree is mounted: it reads the tables through oracle/_ref/libhosek_ref.so,
built from the model rtcamp8 vendors (ext/hosek_wilkie_2013, 3-clause BSD) by oracle/ref/Makefile.

Outputs
  rtcamp8_amd/data/hosek_model.bin  -- little-endian, the layout rtc8x_hosek_model_create reads:
      char[8] "RTC8XHW1"; f64 configs[11][1080]; f64 radiances[11][120]; f64 solar[11][1800];
      f64 limb_darkening[11][6]; f32 cmf_x[16], cmf_y[16], cmf_z[16], center_wavelengths[16];
      f32 integral_cmf  (the CMF set as tests/golden/make_sky_data.py computes it)
  tests/golden/hosek_states.npz     -- for a grid of (solar elevation, turbidity) pairs given as float32,
      the 16 band states arhosekskymodelstate_alloc_init returns, cast to float32 field by field as
      AnalyticSkyEnvironment::computeDistribution does (renderer/renderer_scene.cpp:570-592); and the
      FP64 sky RGB of a second sun (elevation 70 deg, azimuth 200 deg, turbidity 6.5) at the
      directions of tests/golden/hosek_golden.npz.
"""
import ctypes as C
import math
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent.parent
sys.path.insert(0, str(Path(__file__).resolve().parent))

import make_sky_data as msd  # noqa: E402

PAIRS = [(0.0, 1.0), (0.05, 2.0), (math.radians(10.0), 3.0), (math.radians(30.0), 3.0), (math.radians(30.0), 4.5),
         (math.radians(45.0), 9.99), (math.radians(60.0), 10.0), (math.radians(70.0), 6.5),
         (math.radians(85.0), 1.25), (0.5 * math.pi, 7.0)]
SECOND_SUN = (math.radians(70.0), math.radians(200.0), 6.5)


def main() -> int:
    lib = msd.hosek_lib()
    lib.hosek_config_table.argtypes = [C.c_int, C.c_void_p]
    lib.hosek_radiance_table.argtypes = [C.c_int, C.c_void_p]
    configs = np.zeros((11, 1080), np.float64)
    radiances = np.zeros((11, 120), np.float64)
    solar = np.zeros((11, 1800), np.float64)
    limb = np.zeros((11, 6), np.float64)
    for wl in range(11):
        lib.hosek_config_table(wl, configs[wl].ctypes.data)
        lib.hosek_radiance_table(wl, radiances[wl].ctypes.data)
        lib.hosek_solar_table(wl, solar[wl].ctypes.data)
        lib.hosek_limb_table(wl, limb[wl].ctypes.data)
    cmf = msd.discretized_cmfs(*msd.cie_tables(), msd.NUM_BANDS)
    blob = b"RTC8XHW1" + configs.astype("<f8").tobytes() + radiances.astype("<f8").tobytes() + \
        solar.astype("<f8").tobytes() + limb.astype("<f8").tobytes()
    for a in cmf[:4]:
        blob += np.asarray(a, "<f4").tobytes()
    blob += np.float32(cmf[4]).astype("<f4").tobytes()
    (ROOT / "rtcamp8_amd" / "data" / "hosek_model.bin").write_bytes(blob)

    n_state = lib.hosek_state_size()
    states = np.zeros((len(PAIRS), msd.NUM_BANDS, n_state), np.float32)
    buf = np.zeros(n_state, np.float64)
    elev32 = np.array([p[0] for p in PAIRS], np.float32)
    turb32 = np.array([p[1] for p in PAIRS], np.float32)
    for k in range(len(PAIRS)):
        for b in range(msd.NUM_BANDS):
            albedo = np.float32(0.5) * (np.float32(1) - (np.float32(b) + np.float32(0.5)) / np.float32(msd.NUM_BANDS))
            lib.hosek_state(float(elev32[k]), float(turb32[k]), float(albedo), buf.ctypes.data)
            states[k, b] = buf.astype(np.float32)

    e, az, t = SECOND_SUN
    e32, t32 = float(np.float32(e)), float(np.float32(t))
    sun = np.array([math.cos(e) * math.sin(az), math.sin(e), math.cos(e) * math.cos(az)]).astype(np.float32)
    gold = np.load(ROOT / "tests" / "golden" / "hosek_golden.npz")
    views = np.stack([gold["phi"], gold["theta"]], axis=1).astype(np.float64)
    sun_theta = math.acos(float(sun[1]))
    sun_phi = math.atan2(-float(sun[0]), float(sun[2])) % (2 * math.pi)
    rng = np.random.default_rng(5)
    disc = [(sun_phi + rng.uniform(-0.003, 0.003) / math.sin(sun_theta), sun_theta + rng.uniform(-0.003, 0.003))
            for _ in range(32)]
    views = np.concatenate([views, np.array(disc, np.float32).astype(np.float64)])
    rgb = np.array([msd.sky_rgb_fp64(lib, v, sun.astype(np.float64), cmf, elevation=e32, turbidity=t32)
                    for v in views])
    np.savez_compressed(ROOT / "tests" / "golden" / "hosek_states.npz", elevation=elev32, turbidity=turb32,
                        states=states, sun2=sun, sun2_elevation=np.float32(e32), sun2_turbidity=np.float32(t32),
                        phi2=views[:, 0].astype(np.float32), theta2=views[:, 1].astype(np.float32), rgb2=rgb)
    print(f"wrote hosek_model.bin ({len(blob)} bytes) and hosek_states.npz ({len(PAIRS)} states, {len(views)} views)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
