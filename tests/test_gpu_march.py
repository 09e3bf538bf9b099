"""Device sampling and ray marching (vbt_sample_points, vbt_march_rays) over a
one-frame stride-576 cache, bit for bit against the oracle's restatement of
the viewer shader (VBTStudio/frontend/shaders/vbt_sample.comp)."""
import numpy as np
import pytest

import golden_io
import oracle_py
from vbt_amd import synth

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu


def bitexact(a, b, what):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
    assert len(bad) == 0, f"{what}: {len(bad)} of {a.size} differ, first at {bad[0]}: {a.flat[bad[0]]} vs {b.flat[bad[0]]}"


def device_cache(a, t):
    A = a.info.activeLeaves
    out = hip.DeviceArray((1, A * 576), np.float32, a.stream)
    a.frame_cache_into(out, t, 1, hip.CACHE_STRIDE576)
    return out


def rays_into_box(W, H, D, n, rng, tmax=None):
    """Rays from random points on a sphere around the volume towards random
    interior targets (origin, tmin, direction, tmax)."""
    c = np.array([W, H, D], np.float64) / 2
    r = np.linalg.norm(c) * 1.3
    u = rng.normal(size=(n, 3))
    org = c + r * u / np.linalg.norm(u, axis=1, keepdims=True)
    tgt = rng.random((n, 3)) * [W - 1, H - 1, D - 1]
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tmin = rng.random(n) * 2
    tm = np.full(n, tmax if tmax else 2.6 * r)
    return np.concatenate([org, tmin[:, None], d, tm[:, None]], 1).astype(np.float32)


ASSETS = {
    "smoke64": lambda: golden_io.vbtp("smoke64"),
    "liquid_mini": lambda: golden_io.vbtp("liquid_mini"),
    "levelset40": lambda: synth.Workload("ls", synth.LEVELSET, (40, 40, 32, 24), 8, (3.0, 9.0), 15).encode(threads=8),
}


@pytest.fixture(scope="module", params=sorted(ASSETS))
def asset(request):
    data = ASSETS[request.param]()
    a = hip.Asset(data)
    yield request.param, a, oracle_py.OracleFile(data)
    a.close()


def test_point_samples(asset):
    name, a, o = asset
    W, H, D, F = a.dims
    rng = np.random.default_rng(21)
    for t in (0, F // 2):
        dc = device_cache(a, t)
        hc = o.cache_576(t)
        bitexact(dc.host().reshape(-1, 576), hc, f"{name} cache t={t}")
        n = 30000
        xyz = (rng.random((n, 3)) * [W + 2, H + 2, D + 2] - 1).astype(np.float32)
        xyz[: n // 4] = np.round(xyz[: n // 4] * 2) / 2  # voxel centres and half-way points
        for mode in (0, 1):
            p = hip.march_params(frameIndex=t, renderMode=mode, backgroundValue=1.5, bboxMin=(0.0, 0.0, 0.0))
            for keep in (False, True):
                bitexact(a.sample_points(dc, p, xyz, keep), o.sample_points(hc, p, xyz, keep),
                         f"{name} samples t={t} mode={mode} keep={keep}")


@pytest.mark.parametrize("surface_model", [0, 1])
def test_density_rays(asset, surface_model):
    name, a, o = asset
    W, H, D, F = a.dims
    t = F // 3
    dc, hc = device_cache(a, t), o.cache_576(t)
    rays = rays_into_box(W, H, D, 2000, np.random.default_rng(5))
    p = hip.march_params(frameIndex=t, stepCount=96, surfaceModel=surface_model, bboxMin=(0.0, 0.0, 0.0))
    bitexact(a.march(dc, p, rays), o.march_rays(hc, p, rays), f"{name} density rays sm={surface_model}")


@pytest.mark.parametrize("surface_model,up", [(0, 1), (1, 1), (1, 2)])
def test_level_set_rays(asset, surface_model, up):
    name, a, o = asset
    W, H, D, F = a.dims
    t = F // 2
    dc, hc = device_cache(a, t), o.cache_576(t)
    rays = rays_into_box(W, H, D, 2000, np.random.default_rng(6 + up))
    p = hip.march_params(frameIndex=t, renderMode=1, backgroundValue=3.0, stepCount=160, maxStepVoxels=3.0,
                         surfaceModel=surface_model, floorHeight=1.0, shadowSteps=64, upAxis=up,
                         surfaceIso=0.05 if name == "smoke64" else 0.0)
    got, exp = a.march(dc, p, rays), o.march_rays(hc, p, rays)
    bitexact(got, exp, f"{name} level-set rays sm={surface_model} up={up}")
    if name != "smoke64":
        assert (exp[:, 3] >= 0).any()


def test_bbox_offset_and_invalid_calls(asset):
    name, a, o = asset
    W, H, D, F = a.dims
    dc, hc = device_cache(a, 1), o.cache_576(1)
    p = hip.march_params(frameIndex=1, renderMode=1, bboxMin=(-12.5, 3.0, 100.25), stepCount=64)
    rays = rays_into_box(W, H, D, 500, np.random.default_rng(2))
    rays[:, :3] += np.array([-12.5, 3.0, 100.25], np.float32)
    bitexact(a.march(dc, p, rays), o.march_rays(hc, p, rays), f"{name} offset rays")
    assert a.march(dc, p, rays[:0]).shape == (0, 8)
    with pytest.raises(hip.VbtError):
        hip.check(hip.lib().vbt_march_rays(a.h, None, hip.ctypes.byref(p), None, 1, None, None))


def test_scientific_assets_are_refused():
    data = golden_io.vbtp("sci_mini")
    a = hip.Asset(data)
    cache = hip.DeviceArray(576, np.float32)
    with pytest.raises(hip.VbtError):
        a.sample_points(cache, hip.march_params(), np.zeros((1, 3), np.float32))
    a.close()


def test_liquid320_level_set_image():
    """A 192x192 level-set image of the full-size liquid (320^3, 150 frames)
    at frame 75, with floor shadows; every ray bit for bit."""
    w = synth.WORKLOADS["liquid320"]
    data = w.encode(threads=16)
    a = hip.Asset(data)
    o = oracle_py.OracleFile(data)
    W, H, D, F = a.dims
    dc = device_cache(a, 75)
    hc = dc.host().reshape(-1, 576)
    # camera above and to the side, looking at the volume centre (y up)
    nx = ny = 192
    eye = np.array([W * 1.4, H * 1.1, -D * 0.6])
    look = np.array([W / 2, H * 0.35, D / 2])
    fwd = (look - eye) / np.linalg.norm(look - eye)
    right = np.cross(fwd, [0, 1, 0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    u, v = np.meshgrid(np.linspace(-0.5, 0.5, nx), np.linspace(-0.5, 0.5, ny))
    d = fwd[None] + u.reshape(-1, 1) * right[None] + v.reshape(-1, 1) * up[None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    n = len(d)
    rays = np.concatenate([np.repeat(eye[None], n, 0), np.zeros((n, 1)), d, np.full((n, 1), 1200.0)],
                          1).astype(np.float32)
    p = hip.march_params(frameIndex=75, renderMode=1, backgroundValue=4.0, stepCount=256, maxStepVoxels=4.0,
                         surfaceModel=1, floorHeight=0.5, shadowSteps=96, upAxis=1)
    got = a.march(dc, p, rays)
    exp = o.march_rays(hc, p, rays, threads=16)
    bitexact(got, exp, "liquid320 level-set image")
    assert (exp[:, 3] >= 0).mean() > 0.05
    a.close()
