"""CPU tests: the oracle pinned against closed forms, the reference's own unit
tests restated, the reference decoder compiled in oracle/_ref (when built),
and the product library's host-side format functions and exported symbols."""
import ctypes
import math
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import golden_io
import oracle_py
import streams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
need_ref = pytest.mark.skipif(not oracle_py.have_ref(), reason="oracle/_ref not built (needs the reference sources)")


def O():
    return oracle_py.oracle()


# ------------------------------------------------------------ half floats --
def test_half_to_float_exhaustive_against_ieee():
    hb = np.arange(65536, dtype=np.uint32)
    got = np.array([O().vbto_half_bits_to_float(int(h)) for h in hb], np.float32)
    exp = hb.astype(np.uint16).view(np.float16).astype(np.float32)
    fin = np.isfinite(exp)
    np.testing.assert_array_equal(got[fin].view(np.uint32), exp[fin].view(np.uint32))
    assert np.isnan(got[np.isnan(exp)]).all() and np.array_equal(got[np.isinf(exp)], exp[np.isinf(exp)])


@need_ref
def test_half_conversions_against_reference():
    R = oracle_py.ref()
    for h in range(65536):
        a = np.float32(O().vbto_half_bits_to_float(h)).view(np.uint32)
        b = np.float32(R.vbtref_half_to_float(h)).view(np.uint32)
        assert a == b, h
    rng = np.random.default_rng(3)
    vals = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-9, 6, 20000),
                           np.array([0.0, -0.0, 65504.0, 65520.0, 1e-8, 6e-8, 2.0 ** -24, 2.0 ** -25, np.inf, -np.inf],
                                    np.float32)]).astype(np.float32)
    for v in vals:
        assert O().vbto_float_to_half_bits(float(v)) == R.vbtref_float_to_half(float(v)), v


def test_lerp_weight_identity_exhaustive(tmp_path):
    exe = str(tmp_path / "alpha_identity")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-o", exe,
                    os.path.join(ROOT, "tests", "native", "alpha_identity.c"), "-lm"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    checked, bad = map(int, r.stdout.split())
    assert r.returncode == 0 and bad == 0 and checked == 2147385345


# ---------------------------------------- reference unit tests, restated --
def oracle_stream_value(packed, s, frames, t, ci=0, limit=None, trace=None):
    buf = np.frombuffer(packed + b"\0" * 8, np.uint8)
    val = ctypes.c_float()
    rc = O().vbto_decode_control_value(buf.ctypes.data, s.codec, s.count, s.bins, 1, len(s.keys),
                                       len(packed) if limit is None else limit, 1, frames, ci, t, ctypes.byref(val),
                                       ctypes.byref(trace) if trace is not None else None)
    return rc, val.value


@pytest.mark.parametrize("codec", [0, 1])
@pytest.mark.parametrize("bins", [4, 8])
def test_temporal_decode_matches_lower_bound(codec, bins):
    """tests/temporal_decode_tests.cpp:103-108: frames 1..255, dense/sparse,
    hints {-1, 0, 255}; 2e-6 tolerance against the lower_bound lerp."""
    for frames in range(1, 256):
        for dense, hint in [(True, -1), (False, -1), (False, 0), (False, 255)]:
            s = streams.make_stream(frames, bins, codec, dense)
            if hint >= 0:
                s.bin_index = [hint] * len(s.bin_index)
            packed = s.pack()
            for t in range(-2, frames + 3):
                rc, v = oracle_stream_value(packed, s, frames, t)
                assert rc == 0
                assert abs(v - streams.lower_bound_reference(s.keys, frames, t)) <= 2e-6


def test_trace_bracket_against_linear_scan():
    """Integer selection: the (lo, hi, f0, f1) bracket equals a brute-force
    linear scan for increasing keys, whatever the hints."""
    rng = np.random.default_rng(11)
    for _ in range(300):
        frames = int(rng.integers(2, 256))
        n = int(rng.integers(1, min(frames, 40) + 1))
        kf = np.sort(rng.choice(frames, n, replace=False))
        s = streams.Stream(int(rng.integers(0, 2)), int(rng.choice([4, 8])), 255)
        s.keys = [(int(f), float(rng.standard_normal())) for f in kf]
        s.desc = [(0, n, 1)]
        s.bin_index = [int(x) for x in rng.integers(0, 256, s.bins)]
        packed = s.pack()
        for t in range(-1, frames + 1):
            tr = oracle_py.Trace()
            rc, _ = oracle_stream_value(packed, s, frames, t, trace=tr)
            assert rc == 0
            fr = min(max(t, 0), frames - 1)
            i = int(np.searchsorted(kf, fr, side="right")) - 1  # last key <= fr
            if fr <= kf[0]:
                assert tr.kind == 1 and tr.lo == 0
            elif fr >= kf[-1]:
                assert tr.kind in (1, 2) and max(tr.lo, tr.hi) == n - 1
            elif kf[i] == fr:
                assert (tr.kind == 1 and tr.lo == i) or (tr.kind == 2 and tr.hi == i)
            else:
                assert tr.kind == 3 and tr.lo == i and tr.hi == i + 1 and tr.f0 == kf[i] and tr.f1 == kf[i + 1]


def test_shell_test_restated():
    """tests/temporal_decode_tests.cpp:68-101 (Mode2 voxels {0,63,64,255,511})."""
    leaf, residual, voxels = streams.shell_test_leaf()
    buf = np.frombuffer(leaf, np.uint8)
    v = oracle_py.LeafView()
    assert O().vbto_parse_leaf(buf.ctypes.data, len(leaf), 1, ctypes.byref(v)) == 0
    val = ctypes.c_float()
    desc = ctypes.c_uint32()
    for t in range(-1, 102):
        ref = streams.lower_bound_reference(residual.keys, 101, t)
        for vox in range(512):
            assert O().vbto_decode_shell_voxel_value(buf.ctypes.data, ctypes.byref(v), 1, 101, vox, t,
                                                     ctypes.byref(val), ctypes.byref(desc)) == 0
            if vox in voxels:
                assert abs(val.value - ref) <= 2e-6 and desc.value == voxels.index(vox)
            else:
                assert val.value == 0.0 and desc.value == 0xFFFFFFFF


def make_valid_mode2_leaf():
    """makeValidMode2Leaf of tests/render_temporal_payload_tests.cpp:39-61."""
    hdr = streams.header_word(2, 0, 0, 1, 0)
    coarse = b"\0" * (4 + 4 + 0)  # 1 control: desc 4 + bins 8 (aligned) + no keys
    coarse = b"\0" * (1 * 4 + 8)
    occ = struct.pack("<8Q", 1, 0, 1 << 5, 0, 0, 0, 0, 0) + struct.pack("<8H", 0, 1, 1, 2, 2, 2, 2, 2)
    residual = b"\0" * (2 * 4 + 2 * 4 + 8)
    return streams.build_leaf(hdr, 0, 2, [coarse, occ, residual])


def parse_both(leaf: bytes):
    """(oracle rc, oracle msg, product rc, product msg)"""
    from vbt_amd import hip

    buf = np.frombuffer(leaf + b"\0\0\0\0", np.uint8)
    v = oracle_py.LeafView()
    rc = O().vbto_parse_leaf(buf.ctypes.data, len(leaf), 1, ctypes.byref(v))
    msg = oracle_py.last_error() if rc else ""
    L = hip.lib()
    pv = (ctypes.c_uint8 * 256)()
    L.vbt_parse_leaf.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    prc = L.vbt_parse_leaf(buf.ctypes.data, len(leaf), 1, pv)
    pmsg = L.vbt_last_error().decode() if prc else ""
    return rc, msg, prc, pmsg, v


def test_payload_tests_restated():
    """tests/render_temporal_payload_tests.cpp:65-114 on the oracle and the
    product's host parser."""
    assert (1 * 8 + 0) * 8 + 0 == 64  # renderTemporalLeafVoxelIndex(0,0,1) with the fixed 8 stride
    assert (1 * 8 + 2) * 8 + 2 == 82
    valid = make_valid_mode2_leaf()
    rc, msg, prc, pmsg, v = parse_both(valid)
    assert rc == 0 and prc == 0, (msg, pmsg)
    assert v.header.mode == 2 and v.shellActiveVoxelCount == 2
    assert v.shellResidualOffset + v.shellResidualBytes == len(valid)
    occ = v.layout.fine.descriptorOffset
    bad_prefix = bytearray(valid)
    bad_prefix[occ + 64 + 2: occ + 64 + 4] = b"\0\0"
    cases = [(bytes(bad_prefix), "group prefix"), (valid[:-4], "residual stream"),
             (valid + b"\0\0\0\0", "trailing bytes"),
             (valid[:occ] + b"\0" * 80 + valid[occ + 80:], "no active voxels")]
    for leaf, text in cases:
        rc, msg, prc, pmsg, _ = parse_both(leaf)
        assert rc != 0 and text in msg, (text, msg)
        assert prc != 0 and text in pmsg, (text, pmsg)


def vbtp_image(version=4, width=8, height=8, depth=8, frames=4, leaf_count=1, offsets=(0, 1), payload=(0,),
               trailing=False):
    """writeFile of tests/vbt_file_tests.cpp:38-68 as an in-memory image."""
    out = b"VBTPACK4" + struct.pack("<11I", version, width, height, depth, frames, 8, 4, 0, leaf_count, 1, 0)
    out += struct.pack(f"<{len(offsets)}I", *offsets) + struct.pack(f"<{len(payload)}I", *payload)
    return out + (b"\0" if trailing else b"")


def test_loader_tests_restated():
    """tests/vbt_file_tests.cpp:86-142 against the oracle loader and the
    product's vbt_validate_file."""
    from vbt_amd import hip

    L = hip.lib()
    hdr = hip.FileHeader()
    cases = [
        (vbtp_image(), None),
        (vbtp_image(version=3), "Unsupported VBTPACK4 version"),
        (vbtp_image(width=16), "leafCount does not match"),
        (vbtp_image(width=16, leaf_count=2, offsets=(0, 2, 1)), "strictly increasing"),
        (vbtp_image(offsets=(0, 0), payload=()), "strictly increasing"),
        (vbtp_image(trailing=True), "trailing bytes"),
        (vbtp_image(offsets=(0, 2)), "truncated"),
    ]
    for img, text in cases:
        buf = np.frombuffer(img + b"\0\0\0\0", np.uint8)
        vf = oracle_py.VFile()
        rc = O().vbto_load_file_memory(buf.ctypes.data, len(img), ctypes.byref(vf))
        prc = L.vbt_validate_file(buf.ctypes.data, len(img), ctypes.byref(hdr))
        if text is None:
            assert rc == 0 and prc == 0
            assert vf.header.version == 4 and vf.header.leafCount == 1
        else:
            assert rc != 0 and text in oracle_py.last_error(), (text, oracle_py.last_error())
            assert prc != 0 and text in L.vbt_last_error().decode()


# ---------------------------------------------------- closed-form pins --
def render_leaf_file(coarse_values_fn, res=4, frames=4):
    """A one-leaf 8^3 render asset whose coarse controls hold constant series
    given by coarse_values_fn(ix, iy, iz) (fp32 keyframes)."""
    C = res ** 3
    s = streams.Stream(1, 8, 255)
    for c in range(C):
        ix, iy, iz = c % res, (c // res) % res, c // (res * res)
        s.desc.append((len(s.keys), 1, 1))
        s.keys.append((0, float(np.float32(coarse_values_fn(ix, iy, iz)))))
        s.bin_index += [0] * 8
    leaf = streams.build_leaf(streams.header_word(1, 1, 0, res, 0), len(s.keys), 0, [s.pack()])
    words = len(leaf) // 4
    return b"VBTPACK4" + struct.pack("<11I", 4, 8, 8, 8, frames, 8, 4, 0, 1, 1, 0) + struct.pack("<2I", 0, words) + leaf


@pytest.mark.parametrize("res", [2, 3, 4, 5, 6, 7])
def test_trilinear_reproduces_linear_fields(res):
    sc = 7.0 / (res - 1)
    f = lambda x, y, z: 0.25 + 0.5 * x - 0.125 * y + 0.0625 * z  # noqa: E731
    img = render_leaf_file(lambda i, j, k: f(i * sc, j * sc, k * sc), res)
    of = oracle_py.OracleFile(img)
    got = of.frame(0)
    z, y, x = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    exp = f(x.astype(np.float64), y, z)
    assert np.max(np.abs(got - exp)) <= 1e-5
    # lattice end points reproduce the stored controls exactly
    assert got[0, 0, 0] == np.float32(f(0, 0, 0))
    assert got[7, 7, 7] == np.float32(f(7 * 1.0, 7, 7)) or abs(got[7, 7, 7] - f(7, 7, 7)) <= 1e-6


def test_trilinear_matches_independent_float64():
    rng = np.random.default_rng(5)
    for res in (2, 4, 6):
        vals = rng.standard_normal((res, res, res))
        img = render_leaf_file(lambda i, j, k: vals[k, j, i], res)
        of = oracle_py.OracleFile(img)
        got = of.frame(0).astype(np.float64)
        ctrl = vals.astype(np.float32).astype(np.float64)
        coords = np.arange(res) * (7.0 / (res - 1))

        def weights(p):
            i1 = int(np.clip(np.searchsorted(coords, p, side="left"), 1, res - 1))
            i0 = i1 - 1
            return i0, i1, (p - coords[i0]) / (coords[i1] - coords[i0])

        exp = np.zeros((8, 8, 8))
        for z in range(8):
            for y in range(8):
                for x in range(8):
                    (x0, x1, tx), (y0, y1, ty), (z0, z1, tz) = weights(x), weights(y), weights(z)
                    acc = 0.0
                    for cz, wz in ((z0, 1 - tz), (z1, tz)):
                        for cy, wy in ((y0, 1 - ty), (y1, ty)):
                            for cx, wx in ((x0, 1 - tx), (x1, tx)):
                                acc += wx * wy * wz * ctrl[cz, cy, cx]
                    exp[z, y, x] = acc
        assert np.max(np.abs(got - exp)) <= 1e-6 * max(1.0, np.abs(exp).max())


def test_dct_basis_is_orthonormal_dct2():
    N = 37
    B = np.array([[O().vbto_dct_basis(N, t, k) for k in range(N)] for t in range(N)], np.float64)
    np.testing.assert_allclose(B.T @ B, np.eye(N), atol=1e-6)
    t = np.arange(N)[:, None] + 0.5
    k = np.arange(N)[None, :]
    exact = np.sqrt(np.where(k == 0, 1.0 / N, 2.0 / N)) * np.cos(np.pi / N * t * k)
    np.testing.assert_array_equal(B.astype(np.float32), exact.astype(np.float32))


@pytest.mark.parametrize("name", ["sci_mini", "sci_dense_mini"])
def test_scientific_controls_match_float64_sum(name):
    of = oracle_py.OracleFile(golden_io.vbtp(name))
    F = of.header.frames
    scales = np.frombuffer(of.buf[52:52 + 4 * of.header.coarseAcScaleCount].tobytes(), np.float32)
    val = ctypes.c_float()
    for leaf in (0, of.header.leafCount // 2, of.header.leafCount - 1):
        lb = of.leaf_bytes(leaf)
        hw = int(lb[:4].view(np.uint32)[0])
        keep = ((hw >> 2) & 63) + 1
        base = int(of.offsets[leaf]) * 4
        for ci in (0, 21, 63):
            cb = 4 + ci * (keep + 1)
            dc = float(lb[cb:cb + 2].view(np.float16)[0])
            ac = lb[cb + 2: cb + 1 + keep].view(np.int8).astype(np.float32) * scales[:keep - 1]
            coef = np.concatenate([[dc], ac.astype(np.float64)])
            for t in (0, F // 2, F - 1):
                assert O().vbto_sci_coarse_control(ctypes.byref(of.vf), base, keep, ci, t, ctypes.byref(val)) == 0
                basis = np.sqrt(np.where(np.arange(keep) == 0, 1.0 / F, 2.0 / F)) * \
                    np.cos(np.pi / F * (t + 0.5) * np.arange(keep))
                assert abs(val.value - float(coef @ basis)) <= 1e-5 * max(1.0, abs(val.value))


def test_page_table_rank_invariant():
    of = oracle_py.OracleFile(golden_io.vbtp("smoke64"))
    compact, n = of.page_table()
    act = of.modes() != 0
    assert n == act.sum()
    np.testing.assert_array_equal(compact[act], np.arange(n))
    assert (compact[~act] == 0xFFFFFFFF).all()


# --------------------------------- oracle vs the reference's own decoder --
@need_ref
def test_random_streams_against_reference():
    """Fuzz: random descriptors, hints, codecs, unsorted / duplicate keyframe
    frames -- the oracle and the reference agree on value and on failure."""
    R = oracle_py.ref()
    rng = np.random.default_rng(99)
    val_r = ctypes.c_float()
    for it in range(400):
        codec = int(rng.integers(0, 2))
        bins = int(rng.choice([0, 1, 4, 8]))
        C = int(rng.integers(1, 9))
        K = int(rng.integers(0, 24))
        frames = int(rng.integers(1, 300))
        s = streams.Stream(codec, bins, 255)
        for _ in range(C):
            ks = int(rng.integers(0, K + 2))
            kc = int(rng.integers(0, 12))
            s.desc.append((ks, kc, int(rng.integers(0, 3))))
        s.bin_index = [int(x) for x in rng.integers(0, 20, C * bins)]
        fmax = 255 if codec == 0 else 400
        s.keys = [(int(f), float(v)) for f, v in zip(np.sort(rng.integers(0, fmax, K)) if rng.random() < 0.7
                                                      else rng.integers(0, fmax, K), rng.standard_normal(K))]
        packed = s.pack()
        buf = np.frombuffer(packed + b"\0" * 8, np.uint8)
        for _ in range(12):
            ci = int(rng.integers(0, C + 1))
            t = int(rng.integers(-3, frames + 3))
            rc, v = oracle_stream_value(packed, s, frames, t, ci)
            rr = R.vbtref_decode_control_value(buf.ctypes.data, codec, C, bins, K, len(packed), frames, ci, t,
                                               ctypes.byref(val_r))
            assert (rc == 0) == (rr == 0), (it, oracle_py.last_error(), R.vbtref_last_error())
            if rc == 0:
                assert np.float32(v).view(np.uint32) == np.float32(val_r.value).view(np.uint32), it


@need_ref
def test_leaf_corruptions_against_reference():
    R = oracle_py.ref()
    rng = np.random.default_rng(7)
    data = golden_io.vbtp("liquid_mini")
    of = oracle_py.OracleFile(data)
    out = (ctypes.c_uint32 * 24)()
    for it in range(600):
        leaf = int(rng.integers(0, of.header.leafCount))
        b = bytearray(of.leaf_bytes(leaf).tobytes())
        op = rng.integers(0, 3)
        if op == 0 and len(b) > 8:
            del b[int(rng.integers(4, len(b))) // 4 * 4:]
        elif op == 1:
            b += b"\0" * 4 * int(rng.integers(1, 3))
        else:
            i = int(rng.integers(0, min(len(b), 600)))
            b[i] ^= 1 << int(rng.integers(0, 8))
        buf = np.frombuffer(bytes(b) + b"\0" * 8, np.uint8)
        v = oracle_py.LeafView()
        rc = O().vbto_parse_leaf(buf.ctypes.data, len(b), 1, ctypes.byref(v))
        rr = R.vbtref_parse_leaf(buf.ctypes.data, len(b), out)
        assert (rc == 0) == (rr == 0), (it, oracle_py.last_error(), R.vbtref_last_error())
        if rc:
            assert oracle_py.last_error() == R.vbtref_last_error().decode()
        else:
            assert (v.shellActiveVoxelCount, v.shellResidualOffset, v.shellResidualBytes, v.layout.totalBytes) == \
                (out[14], out[15], out[16], out[13])


@need_ref
@pytest.mark.parametrize("name", ["sci_encoder_mix"])
def test_scientific_encoder_assets_against_reference(name, tmp_path):
    """This build's scientific encoder output (CoarseOnly, sparse, dense
    subtypes 0 and 1) decoded by the oracle and by the reference decoder."""
    from vbt_amd import synth

    w = synth.Workload(name, synth.SCIENTIFIC, (48, 40, 16, 30), 12, (0.0, 0.0), 0xABC123, spike_fraction=0.00005)
    data = w.encode(threads=4)
    of = oracle_py.OracleFile(data)
    hw = of.payload.view(np.uint32)[of.offsets[:-1]]
    modes = hw & 3
    sub = (hw >> 12) & 7
    assert (modes == 0).any() and (modes == 1).any() and ((modes >= 2) & (sub == 1)).any() and \
        ((modes >= 2) & (sub == 0)).any(), np.bincount(modes)
    path = str(tmp_path / "a.vbtp")
    with open(path, "wb") as fh:
        fh.write(data)
    R = oracle_py.ref()
    h = R.vbtref_load_file(path.encode())
    W, H, D, F = of.dims
    for t in (0, 13, F - 1):
        exp = np.empty((D, H, W), np.float32)
        assert R.vbtref_sci_frame(h, t, exp.ctypes.data) == 0
        np.testing.assert_array_equal(of.frame(t).view(np.uint32), exp.view(np.uint32))
    R.vbtref_free_file(h)


# ----------------------------------------------------- product library --
def test_library_exports_every_declared_symbol():
    from vbt_amd import hip

    L = hip.lib()
    with open(os.path.join(ROOT, "include", "vbt_hip.h")) as fh:
        text = fh.read()
    names = set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(vbt_[a-z_0-9]+)\s*\(", text, re.M))
    assert len(names) > 40
    for n in sorted(names):
        assert hasattr(L, n), n


def test_host_format_functions_agree_with_oracle():
    from vbt_amd import hip

    L = hip.lib()
    for h in range(0, 65536, 7):
        a = np.float32(L.vbt_half_bits_to_float(h)).view(np.uint32)
        b = np.float32(O().vbto_half_bits_to_float(h)).view(np.uint32)
        assert a == b
    rng = np.random.default_rng(1)
    for v in rng.standard_normal(5000).astype(np.float32) * 100:
        assert L.vbt_float_to_half_bits(float(v)) == O().vbto_float_to_half_bits(float(v))


def simulate_cache_windows(klo, khi, pending, block, window):
    """Host model of k_cache_render's LDS window passes (same arithmetic):
    returns the passes as (w0, nw, served lanes)."""
    klo = np.asarray(klo, np.int64)
    khi = np.asarray(khi, np.int64)
    pending = np.asarray(pending, bool).copy()
    n = len(klo)
    wave_fits = np.zeros(n, bool)
    wave_hi = np.zeros(n, np.int64)
    for w0 in range(0, n, 64):
        sl = slice(w0, min(n, w0 + 64))
        p = pending[sl]
        lo = klo[sl][p].min() if p.any() else 0xFFFFFFFF
        hi = khi[sl][p].max() if p.any() else 0
        wave_hi[sl] = hi
        wave_fits[sl] = lo != 0xFFFFFFFF and hi - (lo & ~3) <= window
    passes = []
    for _ in range(n + 2):
        if not pending.any():
            return passes
        lo = int(klo[pending].min())
        w0 = lo // 4
        limit = w0 * 4 + window
        mine = pending & (np.where(wave_fits, wave_hi <= limit, khi <= limit) | (klo == lo))
        hi = int(khi[mine].max()) if mine.any() else 0
        nw = (hi + 3) // 4 - w0 if hi > w0 * 4 else 0
        passes.append((w0, nw, np.nonzero(mine)[0]))
        pending &= ~mine
    raise AssertionError("window passes did not terminate")


@pytest.mark.parametrize("block,window", [(64, 8192), (64, 16384), (128, 16384), (128, 32768), (256, 32768),
                                          (256, 49152), (64, 4096), (128, 8192), (256, 16384), (64, 12288)])
def test_cache_window_passes_stay_in_bounds(block, window):
    rng = np.random.default_rng(block * 7 + window)
    for trial in range(400):
        n = block
        kb = int(rng.choice([3, 8]))
        counts = rng.integers(0, 256, n) if trial % 3 else rng.integers(0, 60, n)
        start = int(rng.integers(0, 4096))
        lo = start + np.concatenate([[0], np.cumsum(counts[:-1] * kb)]) + rng.integers(0, 4, n) * (trial % 2)
        hi = lo + counts * kb
        pending = (counts > 0) & (rng.random(n) < 0.95)
        served = np.zeros(n, int)
        for w0, nw, lanes in simulate_cache_windows(lo, hi, pending, block, window):
            assert 0 < nw <= window // 4
            assert (lo[lanes] >= w0 * 4).all() and (hi[lanes] <= (w0 + nw) * 4).all()
            served[lanes] += 1
        np.testing.assert_array_equal(served, pending.astype(int))


@need_ref
@pytest.mark.parametrize("name", ["smoke64", "liquid_mini"])
def test_packed_cache_equals_the_reference_per_value_decode(name):
    """Every slot of the oracle's packed frame cache against the reference's
    per-value decoders (oracle/_ref: decodeRenderTemporalLeafControlValue and
    decodeRenderTemporalLeafShellVoxelValue over every active leaf), bit for
    bit, at every frame."""
    data = golden_io.vbtp(name)
    of = oracle_py.OracleFile(data)
    rf = oracle_py.RefFile(data)
    offs = of.packed_offsets()
    got = np.zeros(int(offs[-1]), np.float32)
    for t in range(of.header.frames):
        rf.cache_frame_into(t, offs, got)
        exp = of.cache_packed(t)[0]
        assert (got.view(np.uint32) == exp.view(np.uint32)).all(), t
    rf.close()


@need_ref
@pytest.mark.parametrize("name", ["sci_mini", "sci_dense_mini"])
def test_scientific_queries_equal_the_reference_cpu_query_loop(name):
    """The oracle's scientific point queries against the reference's
    decodeScientificValueAtCpu in its serial query loop, bit for bit."""
    data = golden_io.vbtp(name)
    of = oracle_py.OracleFile(data)
    rf = oracle_py.RefFile(data)
    W, H, D, F = of.dims
    rng = np.random.default_rng(0x5EED0001)
    q = np.stack([rng.integers(0, W, 5000), rng.integers(0, H, 5000), rng.integers(0, D, 5000),
                  rng.integers(0, F, 5000)], axis=1).astype(np.uint32)
    assert (rf.sci_values(q).view(np.uint32) == of.query(q).view(np.uint32)).all()
    rf.close()


@need_ref
@pytest.mark.parametrize("pattern", ["random", "same-t", "same-xyz", "coherent"])
@pytest.mark.parametrize("dims", [(64, 64, 64, 32), (37, 21, 9, 5), (256, 256, 256, 120)])
def test_query_patterns_equal_the_reference(pattern, dims):
    """vbt_generate_queries / vbt_sort_queries_by_leaf against the reference's
    generateQueries and sortQueriesByLeaf (render/src/query_patterns.cpp,
    compiled into oracle/_ref) for several seeds, unsorted and leaf-sorted."""
    from vbt_amd import hip

    W, H, D, F = dims
    hdr = hip.FileHeader()
    hdr.width, hdr.height, hdr.depth, hdr.frames, hdr.leafSize = W, H, D, F, 8
    p = hip.QUERY_PATTERNS[pattern]
    for seed in (1, 0x5EED0001, 12345):
        for srt in (False, True):
            got = hip.generate_queries(hdr, pattern, 3000, seed, sort_by_leaf=srt)
            exp = np.zeros((3000, 4), np.uint32)
            assert oracle_py.ref().vbtref_generate_queries(W, H, D, F, p, 3000, seed, int(srt), exp.ctypes.data) == 0
            assert (got == exp).all(), (pattern, dims, seed, srt)
    # shapes: same-t fixes t, same-xyz fixes the voxel, coherent tiles stay in one frame
    q = hip.generate_queries(hdr, pattern, 512, 7)
    if pattern in ("same-t", "coherent"):
        assert len(set(q[:, 3])) == 1
    if pattern == "same-xyz":
        assert len({tuple(r) for r in q[:, :3]}) == 1
    assert (q[:, 0] < W).all() and (q[:, 1] < H).all() and (q[:, 2] < D).all() and (q[:, 3] < F).all()
    with pytest.raises(hip.VbtError):
        hip.generate_queries(hdr, "diagonal", 4, 1)
