"""Scene descriptions of the benchmark workloads and their device set-up.

The reference builds its scene from a text DSL and a .nvdb file (renderer/renderer_scene.cpp:1732-2471)
and fixes the medium parameters in runApp (renderer/renderer_main.cpp:1463-1474).  Here the two
workloads the path is measured on are described directly:

* ``dense128``  -- dense 128^3 fBm, sigma_t max 20, albedo 0.99, HG g = 0.8, constant white sky plus a
  directional sun, 256x256, 4 spp, at most 16 bounces.
* ``sparse512`` -- sparse 512^3 thresholded fBm cloud (25 % of the 8^3 leaf slots active; density =
  (g - threshold) / (g_max - threshold), contrast 1), sigma_t max 50, albedo 0.999, HG g = 0.85,
  Hosek-Wilkie sky (turbidity 3, sun elevation 30 deg) importance-sampled with alias tables, 1920x1080,
  16 spp, at most 256 bounces, Russian roulette after bounce 8.
* ``sparse512_saturated`` -- the same cloud with contrast 4: density = min(4 (g - threshold) /
  (g_max - threshold), 1), a saturated interior at the majorant (denser, more vertices per path).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import _abi
from ._lib import call
from .device import DeviceBuffer

# renderer/renderer_main.cpp:87: the per-pixel PCG states are drawn from mt19937_64 with this seed.
REFERENCE_RNG_SEED = 591842031321323413


@dataclass(frozen=True)
class FbmConfig:
    resolution: int
    seed: int
    octaves: int = 4
    base_frequency: float = 4.0
    lacunarity: float = 2.0
    gain: float = 0.5
    sparse: bool = False
    active_leaf_fraction: float = 0.25
    contrast: float = 1.0

    def desc(self) -> _abi.FbmDesc:
        return _abi.FbmDesc(self.resolution, self.octaves, self.seed, self.base_frequency, self.lacunarity, self.gain,
                            1 if self.sparse else 0, self.active_leaf_fraction, self.contrast)


@dataclass(frozen=True)
class SceneConfig:
    name: str
    fbm: FbmConfig
    density_coeff: float
    albedo: float
    phase: int
    forwardness: float
    width: int
    height: int
    spp: int
    max_path_length: int
    rr_after_path_length: int
    env_kind: str = "constant"          # "constant" | "hosek" | "none"
    env_radiance: tuple = (1.0, 1.0, 1.0)
    env_sampler: int = _abi.ENV_SAMPLER_ALIAS
    sky_width: int = 3200
    sky_height: int = 1600
    turbidity: float = 3.0
    sun_elevation_deg: float = 30.0
    sun_azimuth_deg: float = 35.0
    sun_enabled: bool = False
    sun_direction: tuple = (0.3, 0.85, 0.42)
    sun_irradiance: tuple = (3.0, 3.0, 3.0)
    sun_select_prob: float = 0.5
    camera_position: tuple = (1.45, 1.05, 2.35)
    camera_target: tuple = (0.5, 0.5, 0.5)
    fov_y_deg: float = 34.0
    tracking: int = _abi.TRACKING_GLOBAL_MAJORANT
    region_cells: int = 0               # local-majorant region size in 8^3 cells (0: the library's default)
    rng_seed: int = REFERENCE_RNG_SEED
    volume_file: str = ""               # a .nvdb file to render instead of the procedural volume
    volume_grid: int = 0                # which grid of the file (nanovdb::io::readGrid's n)

    def with_(self, **kw) -> "SceneConfig":
        return replace(self, **kw)


CONFIGS = {
    "dense128": SceneConfig(
        name="dense128",
        fbm=FbmConfig(resolution=128, seed=0x5EED0001, sparse=False),
        density_coeff=20.0, albedo=0.99, phase=_abi.PHASE_HENYEY_GREENSTEIN, forwardness=0.8,
        width=256, height=256, spp=4, max_path_length=16, rr_after_path_length=0,
        env_kind="constant", env_radiance=(1.0, 1.0, 1.0), sun_enabled=True),
    "sparse512": SceneConfig(
        name="sparse512",
        fbm=FbmConfig(resolution=512, seed=0x5EED0002, sparse=True, active_leaf_fraction=0.25, contrast=1.0),
        density_coeff=50.0, albedo=0.999, phase=_abi.PHASE_HENYEY_GREENSTEIN, forwardness=0.85,
        width=1920, height=1080, spp=16, max_path_length=256, rr_after_path_length=8,
        env_kind="hosek", env_sampler=_abi.ENV_SAMPLER_ALIAS, sun_enabled=False,
        camera_position=(1.27, 0.95, 2.0), fov_y_deg=32.0),
}
CONFIGS["sparse512_saturated"] = CONFIGS["sparse512"].with_(
    name="sparse512_saturated",
    fbm=FbmConfig(resolution=512, seed=0x5EED0002, sparse=True, active_leaf_fraction=0.25, contrast=4.0))


def look_at_quaternion(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """Unit quaternion (x, y, z, w) of a camera at `eye` looking at `target` (local -z forward, +y up),
    the orientation convention of PerspectiveCamera (common/common_renderer_types.h:1019-1034)."""
    eye, target, up = (np.asarray(v, dtype=np.float64) for v in (eye, target, up))
    back = eye - target
    back /= np.linalg.norm(back)
    right = np.cross(up, back)
    right /= np.linalg.norm(right)
    upv = np.cross(back, right)
    m = np.stack([right, upv, back], axis=1)  # columns
    t = m[0, 0] + m[1, 1] + m[2, 2]
    if t > 0:
        s = math.sqrt(t + 1.0) * 2
        w, x, y, z = 0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s
    elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s = math.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
        w, x, y, z = (m[2, 1] - m[1, 2]) / s, 0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s
    elif m[1, 1] > m[2, 2]:
        s = math.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
        w, x, y, z = (m[0, 2] - m[2, 0]) / s, (m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s
    else:
        s = math.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
        w, x, y, z = (m[1, 0] - m[0, 1]) / s, (m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s
    q = np.array([x, y, z, w])
    return (q / np.linalg.norm(q)).astype(np.float32)


def make_camera(cfg: SceneConfig) -> _abi.Camera:
    cam = _abi.Camera()
    pos = (C.c_float * 3)(*cfg.camera_position)
    q = (C.c_float * 4)(*look_at_quaternion(cfg.camera_position, cfg.camera_target).tolist())
    call("rtc8x_camera_setup", C.byref(cam), pos, q, C.c_float(math.radians(cfg.fov_y_deg)),
         C.c_float(cfg.width / cfg.height))
    return cam


def sun_direction(cfg: SceneConfig) -> np.ndarray:
    if cfg.env_kind == "hosek":
        el, az = math.radians(cfg.sun_elevation_deg), math.radians(cfg.sun_azimuth_deg)
        d = np.array([math.cos(el) * math.sin(az), math.sin(el), math.cos(el) * math.cos(az)])
    else:
        d = np.asarray(cfg.sun_direction, dtype=np.float64)
    return (d / np.linalg.norm(d)).astype(np.float32)


def world_dims(grid: _abi.Grid):
    """WorldDimInfo of a volume-only scene (compute_light_probs.cu:139-157): the aabb is the volume bbox,
    radius = |max - centre| in float32."""
    mn = np.array(grid.world_bbox_min[:], dtype=np.float32)
    mx = np.array(grid.world_bbox_max[:], dtype=np.float32)
    center = (mn + mx) * np.float32(0.5)
    d = mx - center
    radius = np.float32(np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])))
    return mn, mx, float(radius)


def fill_params(cfg: SceneConfig, grid: _abi.Grid, env: _abi.EnvLight, camera: _abi.Camera) -> _abi.FrameParams:
    """The parameter block shared by rtc8x_trace and (in the tests) the CPU oracle."""
    p = _abi.FrameParams()
    p.image_width, p.image_height = cfg.width, cfg.height
    p.camera = camera
    p.grid = grid
    p.density_coeff = cfg.density_coeff
    p.scattering_albedo = cfg.albedo
    p.scattering_forwardness = cfg.forwardness
    p.phase = cfg.phase
    p.env = env
    p.sun.enabled = 1 if cfg.sun_enabled else 0
    if cfg.sun_enabled:
        d = sun_direction(cfg)
        p.sun.direction[:] = d.tolist()
        p.sun.irradiance[:] = list(cfg.sun_irradiance)
        p.sun.select_prob = cfg.sun_select_prob
    mn, mx, radius = world_dims(grid)
    p.world_aabb_min[:] = mn.tolist()
    p.world_aabb_max[:] = mx.tolist()
    p.world_radius = radius
    p.rr_after_path_length = cfg.rr_after_path_length
    p.max_path_length = cfg.max_path_length
    p.tracking = cfg.tracking
    p.nrc.path_termination_factor = 0.01
    return p


class DeviceScene:
    """Grid, environment and camera of one workload, resident on the GPU."""

    def __init__(self, cfg: SceneConfig, stream=None):
        self.cfg = cfg
        if cfg.volume_file:
            from .nvdb import load_grid  # noqa: WPS433

            self.grid = load_grid(cfg.volume_file, cfg.volume_grid)
        else:
            self.grid = _abi.Grid()
            call("rtc8x_grid_build_fbm", C.byref(cfg.fbm.desc()), C.byref(self.grid), C.c_void_p(stream))
        if cfg.region_cells:
            call("rtc8x_grid_build_regions", C.byref(self.grid), cfg.region_cells)
        self.env = _abi.EnvLight()
        self._env_buffers = []
        self.env.power_coeff = 1.0
        if cfg.env_kind == "constant":
            self.env.kind = _abi.ENV_CONSTANT
            self.env.constant_radiance[:] = list(cfg.env_radiance)
        elif cfg.env_kind == "hosek":
            from .sky import bake_hosek_sky  # noqa: WPS433 (deferred: only the sky workload needs it)
            tex = bake_hosek_sky(cfg, sun_direction(cfg), stream)
            self._env_buffers.append(tex)
            self.env.kind = _abi.ENV_LATLONG
            self.env.sampler = cfg.env_sampler
            self.env.width, self.env.height = cfg.sky_width, cfg.sky_height
            self.env.texels = tex.ptr
            call("rtc8x_env_build_sampler", C.byref(self.env), C.c_void_p(stream))
        else:
            self.env.kind = _abi.ENV_NONE
        self.camera = make_camera(cfg)

    def params(self) -> _abi.FrameParams:
        return fill_params(self.cfg, self.grid, self.env, self.camera)

    def close(self) -> None:
        if self.grid.root_tiles:
            call("rtc8x_grid_destroy", C.byref(self.grid))
        if self.env.kind == _abi.ENV_LATLONG:
            call("rtc8x_env_release", C.byref(self.env))
        for b in self._env_buffers:
            b.free()
        self._env_buffers = []
