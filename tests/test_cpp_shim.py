"""The C++ drop-in headers (include/vbt/*.h) used the way a reference consumer
uses the reference API (tests/native/shim_check.cpp), checked against the
oracle: container loading, header/prefix/layout/parse, the stream model round
trip, error texts (CPU); the per-value decoders and the scientific point /
frame decode through the GPU (gpu)."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

import golden_io
import oracle_py

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "tests", "native", "shim_check")
RENDER = ["smoke64", "liquid_mini"]
NAMES = RENDER + ["sci_mini", "sci_dense_mini"]
MASK = (1 << 64) - 1


def checksum(values) -> int:
    s = 0
    for b in np.ascontiguousarray(values, np.float32).view(np.uint32).tolist():
        s = (s * 1000003 + b) & MASK
    return s


def run(*args):
    if not os.path.exists(SHIM):
        pytest.fail(f"{SHIM} is not built (python -m vbt_amd.build shim)")
    out = subprocess.run([SHIM, *args], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    return [json.loads(line) for line in out.stdout.splitlines() if line.strip()]


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("vbtp")
    out = {}
    for n in NAMES:
        data = golden_io.vbtp(n)
        p = d / f"{n}.vbtp"
        p.write_bytes(data)
        out[n] = (str(p), oracle_py.OracleFile(data))
    return out


def test_shim_container_and_leaf_format(files):
    O = oracle_py.oracle()
    half = checksum([O.vbto_half_bits_to_float(h) for h in range(65536)
                     if (h & 0x7C00) != 0x7C00 or (h & 0x3FF) == 0])
    for name, (path, o) in files.items():
        lines = run("host", path)
        assert lines[0]["half_checksum"] == half
        h = o.header
        assert lines[1]["header"] == [h.version, h.width, h.height, h.depth, h.frames, h.leafCount, h.profileType]
        fh = o.vf.header
        for rec in (l["leaf_of"] for l in lines if "leaf_of" in l):
            x, y, z, leaf = rec
            assert leaf == O.vbto_leaf_index_for_voxel(ctypes.byref(fh), x, y, z), (name, rec)
        leaves = [l["leaf"] for l in lines if "leaf" in l]
        if h.profileType == 2:
            sci = [l["sci"] for l in lines if "sci" in l]
            assert len(sci) == h.leafCount
            assert [s[1] for s in sci] == list(o.modes())
            continue
        modes = o.modes()
        assert [l[0] for l in leaves] == [i for i in range(h.leafCount) if modes[i] != 0]
        for l in leaves:
            v = o.parse(l[0])
            exp = [l[0], v.header.mode, v.coarseKeyframeCount, v.fineKeyframeCount,
                   v.layout.coarse.totalBytes, v.layout.fine.totalBytes, v.layout.totalBytes, v.shellActiveVoxelCount,
                   v.shellResidualOffset, v.shellResidualBytes, v.layout.totalBytes]
            assert l == exp, (name, l, exp)
        assert lines[-1]["prefix_error"] == "render temporal prefix keyframe count exceeds uint16"


def test_shim_query_patterns(files):
    """generateQueries + fillLeafIndices + sortQueriesByLeaf through the C++
    drop-in equal the C-ABI's generator and leaf-major sort (which the CPU
    unit tests pin to the reference's query_patterns.cpp)."""
    from vbt_amd import hip

    O = oracle_py.oracle()
    for name, (path, o) in files.items():
        got = {l["queries"][0]: l["queries"][1] for l in run("host", path) if "queries" in l}
        hdr = hip.FileHeader()
        for f, _ in hdr._fields_:
            setattr(hdr, f, getattr(o.header, f))
        for pattern in ("random", "same-t", "same-xyz", "coherent"):
            q = hip.generate_queries(hdr, pattern, 1000, 77, sort_by_leaf=True)
            s = 0
            for x, y, z, t in q.tolist():
                leaf = O.vbto_leaf_index_for_voxel(ctypes.byref(o.vf.header), x, y, z)
                for v in (x, y, z, t, leaf):
                    s = (s * 1000003 + v) & MASK
            assert got[pattern] == s, (name, pattern)


def test_shim_load_error_text(tmp_path):
    data = bytearray(golden_io.vbtp("smoke64"))
    data[8:12] = (5).to_bytes(4, "little")  # version 5
    p = tmp_path / "bad.vbtp"
    p.write_bytes(bytes(data))
    lines = run("host", str(p))
    assert lines == [{"load_error": "Unsupported VBTPACK4 version."}] or \
        "Unsupported VBTPACK4 version" in lines[0]["load_error"]


@pytest.mark.gpu
def test_shim_per_value_decodes_match_the_oracle(files, tmp_path):
    O = oracle_py.oracle()
    rng = np.random.default_rng(77)
    for name in RENDER:
        path, o = files[name]
        F = o.header.frames
        modes = o.modes()
        leaves = [i for i in range(o.header.leafCount) if modes[i] != 0]
        reqs, exp = [], []
        for leaf in rng.choice(leaves, size=min(12, len(leaves)), replace=False):
            v = o.parse(int(leaf))
            lb = o.leaf_bytes(int(leaf))
            for _ in range(20):
                use_fine = int(v.header.mode == 3 and rng.random() < 0.5)
                count = (v.header.fineResolution ** 3) if use_fine else (v.header.coarseResolution ** 3)
                ci = int(rng.integers(0, count + 1))  # includes one past the end
                t = int(rng.integers(-2, F + 3))
                out = ctypes.c_float()
                rc = O.vbto_decode_leaf_control_value(lb.ctypes.data, ctypes.byref(v), use_fine, 1, F, ci, t,
                                                      ctypes.byref(out), None)
                reqs.append(f"c {leaf} {use_fine} {ci} {t}")
                exp.append({"error": oracle_py.last_error()} if rc else {"v": int(np.float32(out.value).view(np.uint32))})
            if v.header.mode == 2:
                for vox in range(0, 512, 3):
                    t = int(rng.integers(0, F))
                    out = ctypes.c_float()
                    rc = O.vbto_decode_shell_voxel_value(lb.ctypes.data, ctypes.byref(v), 1, F, vox, t,
                                                         ctypes.byref(out), None)
                    assert rc == 0
                    reqs.append(f"s {leaf} {vox} {t}")
                    exp.append({"v": int(np.float32(out.value).view(np.uint32))})
        rf = tmp_path / f"{name}.req"
        rf.write_text("\n".join(reqs) + "\n")
        got = run("gpu", path, str(rf))
        assert got == exp, name
    path, o = files["sci_mini"]
    W, H, D, F = o.dims
    q = np.stack([rng.integers(0, W, 40), rng.integers(0, H, 40), rng.integers(0, D, 40), rng.integers(0, F, 40)],
                 axis=1).astype(np.uint32)
    vals = o.query(q)
    reqs = [f"q {a} {b} {c} {d}" for a, b, c, d in q] + ["f 0", f"f {F - 1}"]
    rf = tmp_path / "sci.req"
    rf.write_text("\n".join(reqs) + "\n")
    got = run("gpu", path, str(rf))
    assert [g["v"] for g in got[:40]] == [int(x) for x in vals.view(np.uint32)]
    assert got[40] == {"frame": checksum(o.frame(0).ravel()), "n": W * H * D}
    assert got[41] == {"frame": checksum(o.frame(F - 1).ravel()), "n": W * H * D}
