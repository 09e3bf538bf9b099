"""The oracle's restatement of the reference's Cycles blob and samplers
(cycles_native_vbt/vbt_cycles_{loader,blob}.cpp, vbt_cycles_sampler.h;
blender_patch/native_cycles_vbt.patch:688-929), pinned against the
reference's own CPU probe: through the committed fixture
tests/golden/cycles_probe.npz, and live when the probe is built in
oracle/_ref.  CPU only."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

import golden_io
import oracle_py
from oracle_py import CYCLES_KERNEL, CYCLES_WORLD, cycles_blob, cycles_sample

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "oracle", "_ref", "vbt_sampler_cpu_probe")
FIXTURE = os.path.join(golden_io.GOLDEN, "cycles_probe.npz")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def fixture_cases():
    fx = dict(np.load(FIXTURE, allow_pickle=False))
    keys = sorted({k.rsplit("__", 1)[0] for k in fx})
    return [(k.split("__")[0], int(k.split("__")[1]), fx[k + "__points"], fx[k + "__density"],
             int(fx[k + "__blob_bytes"][0]), int(fx[k + "__frame_used"][0])) for k in keys]


@pytest.mark.parametrize("case", fixture_cases(), ids=lambda c: f"{c[0]}-t{c[1]}")
def test_standalone_sampler_matches_reference_probe_fixture(case):
    name, frame, pts, dens, size, used = case
    data = golden_io.vbtp(name)
    blob = cycles_blob(data, frame)
    assert len(blob) == size
    assert int(blob[44:48].view(np.uint32)[0]) == used  # frame_index after the clamp
    got = cycles_sample(blob, pts)
    bad = np.flatnonzero(bits(got) != bits(dens))
    assert len(bad) == 0, (name, frame, len(bad), pts[bad[:3]], got[bad[:3]], dens[bad[:3]])
    assert (dens != 0).mean() > 0.05


def run_probe(data: bytes, frame: int, pts: np.ndarray):
    with tempfile.NamedTemporaryFile(suffix=".vbtp", delete=False) as fh:
        fh.write(data)
        path = fh.name
    try:
        args = [PROBE, "--input-vbt", path, "--frame", str(frame)]
        for x, y, z in pts:
            args += ["--probe", f"{float(x):.9g},{float(y):.9g},{float(z):.9g}"]
        res = subprocess.run(args, capture_output=True, text=True, check=True)
    finally:
        os.unlink(path)
    rows = [line.split(",") for line in res.stdout.strip().splitlines()[1:]]
    return np.array([float(r[5]) for r in rows], np.float32), int(rows[0][4])


needs_probe = pytest.mark.skipif(not os.path.exists(PROBE), reason="reference probe not built (oracle/_ref)")


@needs_probe
@pytest.mark.parametrize("name", ["smoke64", "liquid_mini"])
def test_standalone_sampler_matches_reference_probe_live(name):
    data = golden_io.vbtp(name)
    W, H, D, F = np.frombuffer(data[12:28], np.uint32).tolist()
    rng = np.random.default_rng(77)
    for frame in (-3, 1, F // 2, F - 2):
        pts = (rng.random((1200, 3)) * [W + 2, H + 2, D + 2] - 1).astype(np.float32)
        pts[:200] = np.round(pts[:200] * 4) / 4
        exp, size = run_probe(data, frame, pts)
        blob = cycles_blob(data, frame)
        assert len(blob) == size
        got = cycles_sample(blob, pts, threads=4)
        assert np.array_equal(bits(got), bits(exp)), (name, frame)


def file_sections(data: bytes):
    w = np.frombuffer(data[8:52], np.uint32)
    L, nscale = int(w[8]), int(w[10])
    off = 52 + 4 * nscale
    return w, np.frombuffer(data[off: off + 4 * (L + 1)], np.uint32), data[off + 4 * (L + 1):]


@pytest.mark.parametrize("name", ["smoke64", "liquid_mini", "sci_mini"])
def test_blob_is_the_file_reframed(name):
    data = golden_io.vbtp(name)
    w, offsets, payload = file_sections(data)
    width, height, depth, frames, leaf = (int(x) for x in w[1:6])
    for frame, expect in ((-5, 0), (3, 3), (frames + 9, frames - 1)):
        blob = cycles_blob(data, frame)
        h = blob[:76].view(np.uint32)
        lc = [(d + 7) // 8 for d in (width, height, depth)]
        assert h[0] == 0x31544256 and bytes(blob[:4]) == b"VBT1" and h[1] == 1
        assert list(h[2:7]) == [width, height, depth, frames, leaf]
        assert list(h[7:10]) == lc and h[10] == int(w[8]) == lc[0] * lc[1] * lc[2]
        assert h[11] == expect
        assert h[12] == len(offsets) and h[13] * 4 == len(payload)
        assert h[14] == 76 and h[15] == 76 + 4 * len(offsets)
        assert not blob[64:76].any()  # bbox_min
        assert np.array_equal(blob[76: 76 + 4 * len(offsets)].view(np.uint32), offsets)
        assert bytes(blob[h[15]:]) == payload
        assert len(blob) == 76 + 4 * len(offsets) + len(payload)


def test_loader_checks_match_the_reference_messages():
    data = bytearray(golden_io.vbtp("smoke64"))
    w, offsets, _ = file_sections(bytes(data))
    cases = {
        "Unsupported VBT magic. Expected VBTPACK4.": lambda d: d.__setitem__(slice(0, 1), b"X"),
        "VBTPACK4 v4 requires leaf_size=8.": lambda d: d.__setitem__(slice(28, 32), (4).to_bytes(4, "little")),
        "Invalid VBTPACK4 profile type.": lambda d: d.__setitem__(slice(44, 48), (7).to_bytes(4, "little")),
        "VBTPACK4 payload is truncated.": lambda d: d.__delitem__(slice(len(d) - 4, len(d))),
        "VBTPACK4 file contains trailing bytes.": lambda d: d.extend(b"\0\0\0\0"),
    }
    off0 = 52 + 4 * int(w[10])

    def equal_offsets(d):
        d[off0 + 8: off0 + 12] = d[off0 + 4: off0 + 8]
    cases["VBTPACK4 offsets must be strictly increasing for every leaf."] = equal_offsets
    for msg, corrupt in cases.items():
        d = bytearray(data)
        corrupt(d)
        with pytest.raises(ValueError, match=msg.replace(".", r"\.")):
            cycles_blob(bytes(d), 0)


@pytest.mark.parametrize("name", ["smoke64", "liquid_mini"])
def test_kernel_form_at_voxels_and_between(name):
    """The kernel form is the standalone leaf sample at the eight voxel corners
    around the point, mixed trilinearly; Mode2 leaves give 0."""
    data = golden_io.vbtp(name)
    W, H, D, F = np.frombuffer(data[12:28], np.uint32).tolist()
    blob = cycles_blob(data, F // 2)
    o = oracle_py.OracleFile(data)
    rng = np.random.default_rng(3)
    vox = np.stack([rng.integers(0, W, 3000), rng.integers(0, H, 3000), rng.integers(0, D, 3000)], 1)
    mode = np.array([int(o.leaf_bytes(((z // 8) * ((H + 7) // 8) + y // 8) * ((W + 7) // 8) + x // 8)[0]) & 3
                     for x, y, z in vox])
    vf = vox.astype(np.float32)
    ker, std = cycles_sample(blob, vf, CYCLES_KERNEL), cycles_sample(blob, vf)
    assert (ker[mode == 2] == 0).all()
    keep = mode != 2
    assert np.array_equal(ker[keep], std[keep])  # every mix has t = 0 (+0 and -0 compare equal)
    if name == "liquid_mini":
        assert (mode == 2).any() and (std[mode == 2] != 0).any()
    # between voxels, from the standalone values at the corners
    p = (rng.random((2000, 3)) * [W - 1, H - 1, D - 1]).astype(np.float32)
    i0 = np.minimum(p.astype(np.uint32), np.array([W - 1, H - 1, D - 1], np.uint32))
    i1 = np.minimum(i0 + 1, np.array([W - 1, H - 1, D - 1], np.uint32))
    t = np.where(i1 > i0, p - i0.astype(np.float32), np.float32(0)).astype(np.float32)
    corner = [std_at(blob, np.where([(k >> a) & 1 for a in range(3)], i1, i0).astype(np.float32))
              for k in range(8)]
    mode_p = [leaf_modes(o, np.where([(k >> a) & 1 for a in range(3)], i1, i0), W, H) for k in range(8)]
    corner = [np.where(m == 2, np.float32(0), c) for c, m in zip(corner, mode_p)]
    one = np.float32(1)

    def mix(a, b, w):
        return (a * (one - w)).astype(np.float32) + (b * w).astype(np.float32)
    c00, c10 = mix(corner[0], corner[1], t[:, 0]), mix(corner[2], corner[3], t[:, 0])
    c01, c11 = mix(corner[4], corner[5], t[:, 0]), mix(corner[6], corner[7], t[:, 0])
    exp = mix(mix(c00, c10, t[:, 1]), mix(c01, c11, t[:, 1]), t[:, 2])
    assert np.array_equal(bits(cycles_sample(blob, p, CYCLES_KERNEL)), bits(exp))


def std_at(blob, pts):
    return cycles_sample(blob, pts)


def leaf_modes(o, vox, W, H):
    return np.array([int(o.leaf_bytes(((z // 8) * ((H + 7) // 8) + y // 8) * ((W + 7) // 8) + x // 8)[0]) & 3
                     for x, y, z in vox])


def test_world_coordinates_subtract_the_bbox():
    data = golden_io.vbtp("smoke64")
    blob = cycles_blob(data, 7)
    bb = np.array([-12.5, 3.25, 100.0], np.float32)
    blob[64:76] = np.frombuffer(bb.tobytes(), np.uint8)
    rng = np.random.default_rng(9)
    p = (rng.random((3000, 3)) * 66 - 1).astype(np.float32)
    for flags in (0, CYCLES_KERNEL):
        got = cycles_sample(blob, p + bb, flags | CYCLES_WORLD)
        exp = cycles_sample(blob, (p + bb) - bb, flags)
        assert np.array_equal(bits(got), bits(exp))


def test_bad_blob_header():
    data = golden_io.vbtp("smoke64")
    blob = cycles_blob(data, 0)
    p = np.array([[10.5, 20.5, 30.5]], np.float32)
    for mutate in (lambda b: b.__setitem__(0, 0), lambda b: b.__setitem__(4, 2)):
        b = blob.copy()
        mutate(b)
        with pytest.raises(ValueError, match="Invalid VBT blob header"):
            cycles_sample(b, p)
        assert cycles_sample(b, p, CYCLES_KERNEL)[0] == 0.0
    with pytest.raises(ValueError):
        cycles_sample(blob[:-4], p)


def test_product_validator_rejects_what_the_cycles_loader_rejects():
    """An asset that opens has passed every check of load_pack4, so packing it
    on the device needs none of its own beyond pack_blob's."""
    hip = pytest.importorskip("vbt_amd.hip")
    L = hip.lib()
    data = bytearray(golden_io.vbtp("smoke64"))
    w, _, _ = file_sections(bytes(data))
    off0 = 52 + 4 * int(w[10])
    bad = []
    for corrupt in (lambda d: d.__setitem__(slice(0, 1), b"X"), lambda d: d.extend(b"\0\0\0\0"),
                    lambda d: d.__setitem__(slice(off0 + 8, off0 + 12), d[off0 + 4: off0 + 8]),
                    lambda d: d.__setitem__(slice(28, 32), (4).to_bytes(4, "little"))):
        d = bytearray(data)
        corrupt(d)
        with pytest.raises(ValueError):
            cycles_blob(bytes(d), 0)
        buf = (ctypes.c_uint8 * len(d)).from_buffer(d)
        hdr = hip.FileHeader()
        bad.append(L.vbt_validate_file(buf, len(d), ctypes.byref(hdr)))
    assert all(rc != 0 for rc in bad)
