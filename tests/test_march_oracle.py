"""The CPU restatement of the viewer's sampler and ray marcher
(oracle/vbt_oracle_march.c, VBTStudio/frontend/shaders/vbt_sample.comp),
pinned against what else fixes it:

- at voxel centres a sample equals the dense frame decode (the reference
  exporter's composition, pinned by the golden vectors) within 4 ulp of the
  field range: same controls, same lattice weights, only GLSL mix's rounding
  differs from the exporter's lerp;
- a density ray's integral equals the sum of point samples taken at the
  shader's step positions, recomputed here in float32 numpy;
- level-set tracing (hit t, normal, thickness, floor shadow) equals an
  independent float32 restatement of the shader's loops over the oracle's
  point samples, ray by ray;
- outside the volume a sample is the background of the render mode.
"""
import numpy as np
import pytest

import golden_io
import oracle_py
from vbt_amd.hip import march_params


def world_voxels(W, H, D, n, rng):
    return np.stack([rng.integers(0, W, n), rng.integers(0, H, n), rng.integers(0, D, n)], 1).astype(np.float32)


@pytest.mark.parametrize("name", ["smoke64", "liquid_mini"])
def test_samples_at_voxel_centres_equal_the_dense_decode(name):
    o = oracle_py.OracleFile(golden_io.vbtp(name))
    W, H, D, F = o.header.width, o.header.height, o.header.depth, o.header.frames
    rng = np.random.default_rng(3)
    for t in (0, F // 2, F - 1):
        cache = o.cache_576(t)
        for mode, bg in ((0, 0.0), (1, 2.5)):
            p = march_params(frameIndex=t, renderMode=mode, backgroundValue=bg)
            xyz = world_voxels(W, H, D, 4000, rng)
            got = o.sample_points(cache, p, xyz)
            # Mode2 outside the shell: the shader's background rule
            comp = oracle_py.compose(sign_rule=2, background=bg)
            dense = o.frame(t, comp)
            ix, iy, iz = xyz[:, 0].astype(int), xyz[:, 1].astype(int), xyz[:, 2].astype(int)
            exp = dense[iz, iy, ix]
            # empty leaves sample to the render mode's background
            compact, _ = o.page_table()
            lcx, lcy = (W + 7) // 8, (H + 7) // 8
            empty = compact[((iz // 8) * lcy + iy // 8) * lcx + ix // 8] == 0xFFFFFFFF
            exp = np.where(empty, np.float32(bg if mode == 1 else 0.0), exp)
            rng_field = float(np.abs(dense).max()) or 1.0
            tol = 4.0 * np.spacing(np.float32(rng_field))
            assert np.abs(got - exp).max() <= tol, (name, t, mode, np.abs(got - exp).max())


def test_samples_outside_the_volume_are_background():
    o = oracle_py.OracleFile(golden_io.vbtp("liquid_mini"))
    W, H, D = o.header.width, o.header.height, o.header.depth
    cache = o.cache_576(3)
    pts = np.array([[-0.01, 1, 1], [W - 1 + 0.01, 1, 1], [1, -1, 1], [1, H, 1], [1, 1, D - 0.5]], np.float32)
    for mode, bg in ((0, 7.0), (1, 7.0)):
        p = march_params(frameIndex=3, renderMode=mode, backgroundValue=bg)
        got = o.sample_points(cache, p, pts)
        assert (got == (bg if mode == 1 else 0.0)).all()


def test_bbox_offset_shifts_the_sampling_grid():
    o = oracle_py.OracleFile(golden_io.vbtp("smoke64"))
    cache = o.cache_576(9)
    rng = np.random.default_rng(5)
    xyz = (rng.integers(0, 60 * 64, (2000, 3)) / 64.0).astype(np.float32)  # exact under the offsets
    a = o.sample_points(cache, march_params(frameIndex=9), xyz)
    off = np.array([-3.0, 10.0, 0.5], np.float32)
    b = o.sample_points(cache, march_params(frameIndex=9, bboxMin=off), xyz + off)
    # world - bboxMin reproduces the local coordinate exactly for these values
    assert np.array_equal(((xyz + off) - off), xyz)
    assert np.array_equal(a, b)


def step_positions(ray, steps):
    o, tmin, d, tmax = ray[:3], ray[3], ray[4:7], ray[7]
    dt = np.float32(tmax) / np.float32(max(float(steps), 1.0))
    t = np.float32(tmin) + np.float32(0.5) * dt
    pts = []
    for _ in range(steps):
        pts.append(o + d * t)
        t = np.float32(t + dt)
    return np.array(pts, np.float32)


@pytest.mark.parametrize("surface_model", [0, 1])
def test_density_rays_sum_point_samples(surface_model):
    o = oracle_py.OracleFile(golden_io.vbtp("smoke64"))
    cache = o.cache_576(11)
    steps = 40
    p = march_params(frameIndex=11, stepCount=steps, surfaceModel=surface_model)
    rng = np.random.default_rng(8)
    rays = []
    for _ in range(24):
        org = np.array([-5.0, rng.random() * 64, rng.random() * 64], np.float32)
        d = rng.normal(size=3).astype(np.float32)
        d[0] = abs(d[0]) + 0.5
        d /= np.float32(np.linalg.norm(d))
        rays.append(np.concatenate([org, [np.float32(rng.random() * 4)], d, [np.float32(90.0)]]).astype(np.float32))
    rays = np.array(rays, np.float32)
    got = o.march_rays(cache, p, rays)
    for r, res in zip(rays, got):
        vals = o.sample_points(cache, p, step_positions(r, steps))
        if surface_model:
            bins = np.zeros(8, np.float32)
            for i, v in enumerate(vals):
                b = min(i * 8 // steps, 7)
                bins[b] = np.float32(bins[b] + v)
            assert np.array_equal(res, bins)
        else:
            acc = np.float32(0.0)
            for v in vals:
                acc = np.float32(acc + v)
            assert res[0] == acc and (res[1:] == 0).all()


f32 = np.float32


def gmin(x, y):
    return y if y < x else x


def gmax(x, y):
    return y if x < y else x


def gclamp(x, lo, hi):
    return gmin(gmax(x, lo), hi)


def trace_level_set_py(sample, ray, p):
    """traceLevelSet (vbt_sample.comp:476-551) in float32 numpy scalars, with
    `sample(points)` the oracle's point sampler: hit t and normal."""
    o, d, tmax = ray[:3], ray[4:7], f32(ray[7])
    if tmax <= 0:
        return f32(-1.0), None
    iso = f32(p.surfaceIso)
    eps = gmax(f32(p.surfaceEpsilonVoxels) * f32(p.voxelSize), f32(1e-5))
    phi = lambda t: sample(o + d * f32(t))[0] - iso
    prevT = f32(0.0)
    prev = sample(o)[0] - iso
    hitT = f32(-1.0)
    for _ in range(p.stepCount):
        step = gclamp(abs(prev) / gmax(f32(p.voxelSize), f32(1e-6)) * f32(0.8), f32(0.2),
                      gmax(f32(p.maxStepVoxels), f32(0.2)))
        curT = gmin(f32(prevT + step), tmax)
        cur = phi(curT)
        if (prev < 0) != (cur < 0):
            inside = cur < 0
            neg = cur if inside else prev
            supT = gmin(f32(curT + f32(0.35)), tmax) if inside else gmax(f32(prevT - f32(0.35)), f32(0.0))
            sup = phi(supT)
            if not (neg <= -eps * f32(0.5) or sup <= -eps * f32(0.25)):
                if curT >= tmax:
                    break
                prevT, prev = curT, gmax(cur, eps)
                continue
            low, high, lowPhi = prevT, curT, prev
            for _ in range(5):
                mid = f32(0.5) * f32(low + high)
                midPhi = phi(mid)
                if (lowPhi < 0) != (midPhi < 0):
                    high = mid
                else:
                    low, lowPhi = mid, midPhi
            hitT = f32(0.5) * f32(low + high)
            break
        if curT >= tmax:
            break
        prevT, prev = curT, cur
    if hitT < 0:
        return hitT, None
    h = o + d * hitT
    ns = gmax(f32(p.surfaceNormalStep), f32(0.25))
    g = []
    for k in range(3):
        e = np.zeros(3, np.float32)
        e[k] = ns
        g.append(sample(h + e)[0] - sample(h - e)[0])
    ln = np.sqrt(f32(f32(g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]))
    n = np.array(g, np.float32) / ln if ln > f32(1e-6) else -d
    return hitT, n.astype(np.float32)


def thickness_py(sample, ray, p, hitT):
    """estimateThickness (vbt_sample.comp:454-474), preserveInterior samples."""
    o, d, tmax = ray[:3], ray[4:7], f32(ray[7])
    iso = f32(p.surfaceIso)
    eps = gmax(f32(p.surfaceEpsilonVoxels) * f32(p.voxelSize), f32(1e-5))
    t = gmin(f32(hitT + f32(0.35)), tmax)
    prev = sample(o + d * t)[0] - iso
    entered = prev < f32(-0.25) * eps
    for _ in range(min(p.stepCount, 160)):
        step = gclamp(abs(prev) / gmax(f32(p.voxelSize), f32(1e-6)) * f32(0.7), f32(0.3), f32(3.0))
        nextT = gmin(f32(t + step), tmax)
        nxt = sample(o + d * nextT)[0] - iso
        entered = entered or nxt < f32(-0.25) * eps
        if entered and prev < 0 and nxt >= 0:
            return gmax(f32(f32(nextT - hitT) * f32(p.voxelSize)), f32(p.voxelSize) * f32(0.25))
        if nextT >= tmax:
            break
        t, prev = nextT, nxt
    return f32(p.voxelSize) * f32(6.0)


def shadow_py(sample, o, d, p, dims):
    """intersectWorldBox + traceShadow (vbt_sample.comp:417-452)."""
    mn = np.array(p.bboxMin[:], np.float32)
    mx = mn + np.array([dims[0] - 1, dims[1] - 1, dims[2] - 1], np.float32)
    inv = f32(1.0) / d
    a, b = (mn - o) * inv, (mx - o) * inv
    tmn = [gmin(a[k], b[k]) for k in range(3)]
    tmx = [gmax(a[k], b[k]) for k in range(3)]
    nearT = gmax(f32(0.0), gmax(tmn[0], gmax(tmn[1], tmn[2])))
    farT = gmin(tmx[0], gmin(tmx[1], tmx[2]))
    if not farT > nearT:
        return f32(0.0)
    iso = f32(p.surfaceIso)
    eps = gmax(f32(p.surfaceEpsilonVoxels) * f32(p.voxelSize), f32(1e-5))
    t = nearT
    prev = sample(o + d * t)[0] - iso
    for _ in range(p.shadowSteps):
        step = gclamp(abs(prev) / gmax(f32(p.voxelSize), f32(1e-6)) * f32(0.8), f32(0.35), f32(5.0))
        nextT = gmin(f32(t + step), farT)
        nxt = sample(o + d * nextT)[0] - iso
        if (prev < 0) != (nxt < 0):
            inside = nxt < 0
            neg = nxt if inside else prev
            supT = gmin(f32(nextT + f32(0.35)), farT) if inside else gmax(f32(t - f32(0.35)), nearT)
            sup = sample(o + d * supT)[0] - iso
            if neg <= -eps * f32(0.5) or sup <= -eps * f32(0.25):
                return f32(1.0)
        if nextT >= farT:
            break
        t = nextT
        prev = eps if abs(nxt) <= eps else nxt
    return f32(0.0)


@pytest.mark.parametrize("surface_model", [0, 1])
def test_level_set_trace_restated(surface_model):
    """Hit t and normal of every ray equal an independent float32 restatement
    of traceLevelSet over the oracle's point samples; hits are on sign
    changes of the sampled field, normals unit length."""
    o = oracle_py.OracleFile(golden_io.vbtp("liquid_mini"))
    W, H, D, F = o.header.width, o.header.height, o.header.depth, o.header.frames
    t = F // 2
    cache = o.cache_576(t)
    p = march_params(frameIndex=t, renderMode=1, backgroundValue=3.0, stepCount=200, maxStepVoxels=2.0,
                     surfaceModel=surface_model, floorHeight=-2.0, shadowSteps=48)
    rng = np.random.default_rng(9)
    n = 120
    org = np.stack([rng.random(n) * (W - 1), np.full(n, H - 1.0), rng.random(n) * (D - 1)], 1)
    tgt = np.stack([rng.random(n) * (W - 1), np.zeros(n), rng.random(n) * (D - 1)], 1)
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([org, np.zeros((n, 1)), d, np.full((n, 1), 40.0)], 1).astype(np.float32)
    res = o.march_rays(cache, p, rays)
    sample = lambda pts: o.sample_points(cache, p, np.asarray(pts, np.float32).reshape(-1, 3))
    keep = lambda pts: o.sample_points(cache, p, np.asarray(pts, np.float32).reshape(-1, 3), preserve_interior=True)
    hits = floors = 0
    kd = np.array([-0.55, 0.72, 0.42], np.float32)
    kd = kd / np.sqrt(f32(f32(kd[0] * kd[0] + kd[1] * kd[1]) + kd[2] * kd[2]))
    for r, got in zip(rays, res):
        hitT, nrm = trace_level_set_py(sample, r, p)
        assert got[3] == hitT
        if hitT >= 0:
            hits += 1
            assert np.array_equal(got[:3], nrm)
            assert abs(float(np.linalg.norm(nrm)) - 1.0) < 1e-5
            assert got[4] == (thickness_py(keep, r, p, hitT) if surface_model else 0.0)
            dot = f32(f32(nrm[0] * -r[4] + nrm[1] * -r[5]) + nrm[2] * -r[6])
            assert got[7] == abs(dot)
        elif surface_model and r[5] < f32(-1e-5):
            floorT = f32(f32(p.floorHeight) - r[1]) / r[5]
            if floorT > 0:
                floors += 1
                fp = r[:3] + r[4:7] * floorT
                assert got[5] == 1.0
                assert got[6] == shadow_py(sample, fp + kd * f32(0.5), kd, p, (W, H, D))
    assert 10 < hits < n
    assert floors > 0 or not surface_model
