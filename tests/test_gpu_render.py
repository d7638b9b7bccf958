"""The batch renderer end to end on the GPU: its plain-tracer frames are the pixels the library gives for
the same frame rebuilt through the Python bindings (volume from the .nvdb, sky re-derived for the frame's
sun, camera at the frame's key interpolation, RNG carried from frame to frame, tone map and SDR
conversion), whose radiance sums equal the FP32 oracle's on the same inputs, and its NRC frames come out
as images of the same scene."""
import ctypes as C
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "rtcamp8_amd" / "rtc8x_render"

SCENE = """image 96 64
time 0 1 4

skyenv env
env-coeff env 2
skyenv-dir env 30 40
env-addkey env 0.0
skyenv-dir env 120 70
env-rotation env 90
env-addkey env 1.0

volume "cloud.nvdb"

camera cam
fovy cam 40
lookat cam 1.8 1.2 2.5 0.5 0.5 0.5 0 1 0
cam-addkey cam 0.0 false false
fovy cam 30
lookat cam 2.5 0.9 0.8 0.5 0.5 0.5 0 1 0
cam-addkey cam 1.0 false false
"""


@pytest.fixture(scope="module")
def scene_dir(tmp_path_factory):
    from rtcamp8_amd import nvdb
    from rtcamp8_amd.scene import FbmConfig
    from oracle_lib import oracle_build_grid

    d = tmp_path_factory.mktemp("render")
    hg = oracle_build_grid(FbmConfig(resolution=64, seed=5, sparse=True, active_leaf_fraction=0.4, contrast=2.0))
    (d / "cloud.nvdb").write_bytes(nvdb.write_grids([nvdb.encode_grid(hg.grid, "density")], "zip"))
    (d / "scene.txt").write_text(SCENE)
    return d


def render(scene_dir, out, *args):
    print(f"[render] {' '.join(map(str, args))}", flush=True)
    r = subprocess.run([str(EXE), "-scene", str(scene_dir / "scene.txt"), "-out", str(out), *map(str, args)],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


W, H = 96, 64


def _frame_params(grid):
    """The driver's frame parameters for the scene's volume (rtc8x_render.cpp's set-up)."""
    from rtcamp8_amd import _abi

    p = _abi.FrameParams()
    p.image_width, p.image_height = W, H
    p.grid = grid
    p.density_coeff, p.scattering_albedo, p.scattering_forwardness = 0.3, 0.99, 0.0
    p.phase = _abi.PHASE_SCHLICK
    mn = np.array(grid.world_bbox_min[:], np.float32)
    mx = np.array(grid.world_bbox_max[:], np.float32)
    center = (mn + mx) * np.float32(0.5)
    dd = mx - center
    r2 = np.float32(0)
    for a in range(3):
        r2 = np.float32(r2 + dd[a] * dd[a])
    p.world_aabb_min[:], p.world_aabb_max[:] = mn.tolist(), mx.tolist()
    p.world_radius = float(np.sqrt(r2))
    p.rr_after_path_length, p.max_path_length = 0, 0
    p.tracking = _abi.TRACKING_GLOBAL_MAJORANT
    p.nrc.path_termination_factor = 0.01
    return p


def _set_frame(p, f, tex):
    """Camera and sky of dumped frame f into p (the sky baked into tex); returns the sky to close and
    the environment to release after the frame."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.sky import HosekSky

    pos = (C.c_float * 3)(*f["position"])
    q = (C.c_float * 4)(*f["orientation"])
    call("rtc8x_camera_setup", C.byref(p.camera), pos, q, C.c_float(f["fov_y"]), C.c_float(np.float32(W) / np.float32(H)))
    sky = HosekSky(min(max(f["solar_elevation"], 0.0), 1.5707963), f["turbidity"])
    sd = (C.c_float * 3)(*f["sun"])
    call("rtc8x_sky_bake", None, C.byref(sky.struct), sd, 3200, 1600, C.c_void_p(tex.ptr))
    env = _abi.EnvLight()
    env.kind, env.sampler = _abi.ENV_LATLONG, _abi.ENV_SAMPLER_HIERARCHICAL
    env.width, env.height, env.texels = 3200, 1600, tex.ptr
    call("rtc8x_env_build_sampler", C.byref(env), None)
    env.power_coeff, env.rotation = f["env_coeff"], f["env_rotation"]
    p.env = env
    return sky, env


def library_frames(scene_dir, frames, spp, check_oracle=False):
    """The same plain-tracer frames through the Python bindings; with check_oracle, every frame's
    accumulation and carried PCG states also equal the FP32 oracle's on host copies of the same inputs
    (the .nvdb volume, the baked sky with its hierarchical importance map, the keyframed camera)."""
    from oracle_lib import download_env, download_grid, nthreads, oracle, ptr
    from rtcamp8_amd import _abi, image, nvdb
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.device import DeviceBuffer

    w, h = W, H
    grid = nvdb.load_grid(scene_dir / "cloud.nvdb")
    p = _frame_params(grid)
    rng, accum, px = DeviceBuffer(8 * w * h), DeviceBuffer(24 * w * h), DeviceBuffer(4 * w * h)
    tex = DeviceBuffer(16 * 3200 * 1600)
    call("rtc8x_seed_rng", C.c_void_p(rng.ptr), w, h, C.c_uint64(591842031321323413))
    p.rng_states, p.accum = rng.ptr, accum.ptr
    out = []
    hg = download_grid(grid) if check_oracle else None
    try:
        for f in frames:
            sky, env = _set_frame(p, f, tex)
            accum.fill(0)
            rng0 = rng.download(np.uint64, w * h) if check_oracle else None
            call("rtc8x_trace", None, C.byref(p), _abi.ENTRY_PATH_TRACE, spp)
            if check_oracle:
                he, keep = download_env(env)
                hp = _abi.FrameParams()
                C.memmove(C.byref(hp), C.byref(p), C.sizeof(hp))
                hp.grid, hp.env = hg.grid, he
                hrng, hacc = rng0.copy(), np.zeros(3 * w * h)
                hp.rng_states, hp.accum, hp.path_vertices, hp.stats = ptr(hrng), ptr(hacc), None, None
                assert oracle().oracle_trace_f32(C.byref(hp), _abi.ENTRY_PATH_TRACE, spp, 0, h, nthreads()) == 0
                assert np.array_equal(accum.download(np.float64, 3 * w * h).view(np.uint64), hacc.view(np.uint64))
                assert np.array_equal(rng.download(np.uint64, w * h), hrng)
            image.tone_map(accum.ptr, w, h, spp, None, px.ptr, image.sdr_config())
            out.append(px.download(np.uint32, w * h).reshape(h, w))
            call("rtc8x_env_release", C.byref(env))
            sky.close()
    finally:
        for b in (rng, accum, px, tex):
            b.free()
        call("rtc8x_grid_destroy", C.byref(grid))
    return out


class _DumpedScene:
    """What NrcRenderer reads of a scene: the image size, the PCG seed and the frame parameters (here
    the driver's, set per frame)."""

    def __init__(self, p):
        import types

        self.cfg = types.SimpleNamespace(width=W, height=H, rng_seed=591842031321323413)
        self.p = p

    def params(self):
        from rtcamp8_amd import _abi

        q = _abi.FrameParams()
        C.memmove(C.byref(q), C.byref(self.p), C.sizeof(q))
        return q


def library_nrc_frames(scene_dir, frames, spp):
    """The driver's NRC frames through NrcRenderer: the same network settings (two hidden layers, FP16,
    triangle waves, the default seed, loss scale and EMA), the same frame schedule (8 training iterations
    on frame 0, then 1; spp samples traced into the cache in one launch) and the accumulator cleared per
    frame while the PCG states carry over."""
    from rtcamp8_amd import _abi, image, nvdb
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.device import DeviceBuffer
    from rtcamp8_amd.nrc import NrcRenderer

    grid = nvdb.load_grid(scene_dir / "cloud.nvdb")
    p = _frame_params(grid)
    px, tex = DeviceBuffer(4 * W * H), DeviceBuffer(16 * 3200 * 1600)
    scene = _DumpedScene(p)
    r = NrcRenderer(scene, num_hidden_layers=2, precision=_abi.NRC_FP16, encoding=_abi.NRC_TRIANGLE_WAVE,
                    samples_per_launch=spp)
    out = []
    try:
        for f in frames:
            sky, env = _set_frame(p, f, tex)
            r.image()  # the renderer's earlier work is done before its accumulator is cleared
            r.accum.fill(0)
            r.samples = 0
            r.render_frame(spp=spp)
            r.image()
            image.tone_map(r.accum.ptr, W, H, spp, None, px.ptr, image.sdr_config())
            out.append(px.download(np.uint32, W * H).reshape(H, W))
            call("rtc8x_env_release", C.byref(env))
            sky.close()
    finally:
        r.close()
        for b in (px, tex):
            b.free()
        call("rtc8x_grid_destroy", C.byref(grid))
    return out


def test_plain_frames_match_the_library(scene_dir, tmp_path):
    from test_image import decode_png

    render(scene_dir, tmp_path, "-no-nrc", "-spp", 4, "-max-frames", 2)
    d = json.loads(subprocess.run([str(EXE), "-scene", str(scene_dir / "scene.txt"), "-dump", "-max-frames", "2"],
                                  capture_output=True, text=True, timeout=60, check=True).stdout)
    want = library_frames(scene_dir, d["frames"], 4, check_oracle=True)
    for k in range(2):
        got = decode_png((tmp_path / f"{k:03d}.png").read_bytes())
        assert np.array_equal(got, want[k]), k
        lum = (got & 255).astype(np.float64)
        assert 5 < lum.mean() < 250 and lum.std() > 1  # a lit cloud, not a flat image


def test_nrc_frames_match_the_library(scene_dir, tmp_path):
    """NRC frames with triangle-wave encoding (deterministic training): the driver's pixels equal those
    of NrcRenderer driven with the same frame parameters, frame after frame, the online-trained cache
    carried along."""
    from test_image import decode_png

    render(scene_dir, tmp_path, "-spp", 2, "-max-frames", 3, "-encoding", "triangle")
    d = json.loads(subprocess.run([str(EXE), "-scene", str(scene_dir / "scene.txt"), "-dump", "-max-frames", "3"],
                                  capture_output=True, text=True, timeout=60, check=True).stdout)
    want = library_nrc_frames(scene_dir, d["frames"], 2)
    for k in range(3):
        got = decode_png((tmp_path / f"{k:03d}.png").read_bytes())
        assert np.array_equal(got, want[k]), k
        lum = (got & 255).astype(np.float64)
        assert 5 < lum.mean() < 250 and lum.std() > 1


def test_nrc_frames(scene_dir, tmp_path):
    from test_image import decode_png

    r = render(scene_dir, tmp_path / "nrc", "-spp", 2, "-max-frames", 3)
    assert r.stdout.count("Frame ") == 3
    render(scene_dir, tmp_path / "plain", "-no-nrc", "-spp", 16, "-max-frames", 1)
    a = decode_png((tmp_path / "plain" / "000.png").read_bytes()).view(np.uint8).reshape(64, 96, 4)[..., :3]
    for k in range(3):
        img = decode_png((tmp_path / "nrc" / f"{k:03d}.png").read_bytes()).view(np.uint8).reshape(64, 96, 4)
        assert np.all(img[..., 3] == 255)
    b = decode_png((tmp_path / "nrc" / "000.png").read_bytes()).view(np.uint8).reshape(64, 96, 4)[..., :3]
    # the first NRC frame renders the same scene as the plain tracer: mean brightness within 25 %
    assert abs(b.astype(np.float64).mean() / a.astype(np.float64).mean() - 1) < 0.25
