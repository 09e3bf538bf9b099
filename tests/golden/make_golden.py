"""Regenerate the golden fixtures in tests/golden/ (run in a container that has
the reference sources at /root/reference; never on the GPU machine).

Inputs are this repository's synthetic fields (vbt_amd.synth).  They are
encoded by the REFERENCE encoders, built from the reference sources by
oracle/Makefile into oracle/_ref/:
  - vbt_render_temporal_kf_probe  (render temporal-first encoder)
  - vbt_spatialfirst_probe        (scientific spatial-first encoder)
and decoded by the REFERENCE decoder (oracle/_ref/libvbtref.so, a C-ABI
wrapper over render_temporal_decode.cpp / render_temporal_payload.cpp /
vbt_file.cpp / scientific_decode.cpp).  The fixtures are data only: the
encoded .vbtp bytes (gzip) and the reference decoder's outputs (npz).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes
import gzip
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from vbt_amd import synth  # noqa: E402
import oracle_py  # noqa: E402

REF_KF = os.path.join(ROOT, "oracle", "_ref", "vbt_render_temporal_kf_probe")
REF_SF = os.path.join(ROOT, "oracle", "_ref", "vbt_spatialfirst_probe")

SCI_FLAGS = ["--profile", "generic", "--dct-keep", "8", "--event-top-k", "128", "--event-threshold", "0.5",
             "--event-min-count", "3", "--split-scientific-render-modes", "--generic-adaptive-coarse-keep",
             "--generic-dense-crossover", "--generic-dense-grid-res", "3", "--generic-densegrid4-candidate",
             "--generic-densegrid4-dist-threshold", "0.3", "--generic-dense-bits", "4"]

FIXTURES = {
    # name: (workload, extra encoder flags, frames to record)
    "smoke64": ("smoke64", [], [0, 17, 31]),
    "liquid_mini": ("liquid_mini", [], [0, 7, 15]),
    "sci_mini": ("sci_mini", SCI_FLAGS, [0, 17, 47]),
    "sci_dense_mini": ("sci_dense_mini", SCI_FLAGS + ["--generic-norm-thr", "0.0005", "--generic-peak-thr", "0.0005"],
                       [0, 13, 39]),
    # every leaf dense with the tile / patch temporal basis (dense subtypes 2
    # and 3): the reference file decoder reads these blocks as per-frame BFP
    # data (subtype 2 decodes, subtype 3 reads past the pool and throws)
    "sci_tile_mini": ("sci_dense_mini", SCI_FLAGS + ["--generic-norm-thr", "0.0005", "--generic-peak-thr", "0.0005",
                                                     "--generic-tile-temporal-basis", "--generic-tile-temporal-force"],
                      [0, 13, 39]),
    "sci_patch_mini": ("sci_dense_mini", SCI_FLAGS + ["--generic-norm-thr", "0.0005", "--generic-peak-thr", "0.0005",
                                                      "--generic-patch-temporal-basis",
                                                      "--generic-patch-temporal-force"],
                       [0, 13, 39]),
    # tile size 4 (a DenseGrid4 leaf), 3-bit codes, six temporal terms
    "sci_tile4_mini": ("sci_dense_mini", SCI_FLAGS + ["--generic-norm-thr", "0.0005", "--generic-peak-thr", "0.0005",
                                                      "--generic-tile-temporal-basis", "--generic-tile-temporal-force",
                                                      "--generic-tile-size", "4", "--generic-tile-bits", "3",
                                                      "--generic-tile-temporal-keep", "6"],
                       [0, 13, 39]),
    # patch basis with 2-bit codes (the four-level table) and seven terms
    "sci_patch_b2_mini": ("sci_dense_mini", SCI_FLAGS + ["--generic-norm-thr", "0.0005", "--generic-peak-thr",
                                                         "0.0005", "--generic-patch-temporal-basis",
                                                         "--generic-patch-temporal-force", "--generic-multipatch-bits", "2",
                                                         "--generic-patch-temporal-keep", "7"],
                          [0, 13, 39]),
}


def encode_with_reference(w: synth.Workload, extra, tmp) -> bytes:
    f = w.field()
    W, H, D, F = w.dims
    raw = os.path.join(tmp, w.name + ".raw")
    with open(raw, "wb") as fh:
        for t in range(F):
            f.frame(t).tofile(fh)
    meta = {"width": W, "height": H, "depth": D, "frames": F, "grid_name": "density", "data_type": "float32",
            "voxel_size": 1.0}
    if w.kind == synth.LEVELSET:
        meta["conversion_mode"] = "levelset"
    mpath = os.path.join(tmp, w.name + ".metadata.json")
    with open(mpath, "w") as fh:
        json.dump(meta, fh)
    out = os.path.join(tmp, w.name + ".vbtp")
    if w.kind == synth.SCIENTIFIC:
        cmd = [REF_SF, "--input-raw", raw, "--metadata", mpath] + extra + [
            "--omp-threads", "8", "--save-vbt", out, "--report", os.path.join(tmp, "r.md")]
    else:
        cmd = [REF_KF, "--input-raw", raw, "--metadata", mpath, "--save-vbt", out, "--final-only", "--omp-threads",
               "8", "--report", os.path.join(tmp, "r.md")] + extra
    subprocess.run(cmd, check=True, capture_output=True)
    with open(out, "rb") as fh:
        return fh.read()


def render_vectors(data: bytes, frames):
    """Reference decoder outputs for every non-empty leaf at the given frames."""
    R = oracle_py.ref()
    of = oracle_py.OracleFile(data)
    F = of.header.frames
    rec = {"leaf": [], "frame": [], "coarse": [], "fine": [], "shell": []}
    val = ctypes.c_float()
    for leaf in range(of.header.leafCount):
        b = np.ascontiguousarray(of.leaf_bytes(leaf))
        v = of.parse(leaf)
        if v.header.mode == 0:
            continue
        cc = v.header.coarseResolution ** 3
        fc = v.header.fineResolution ** 3 if v.header.mode == 3 else 0
        for t in frames:
            co = np.zeros(64, np.float32)
            fi = np.zeros(216, np.float32)
            sh = np.zeros(512, np.float32)
            for i in range(cc):
                assert R.vbtref_leaf_control_value(b.ctypes.data, len(b), 0, F, i, t, ctypes.byref(val)) == 0
                co[i] = val.value
            for i in range(fc):
                assert R.vbtref_leaf_control_value(b.ctypes.data, len(b), 1, F, i, t, ctypes.byref(val)) == 0
                fi[i] = val.value
            if v.header.mode == 2:
                for vox in range(512):
                    assert R.vbtref_shell_voxel_value(b.ctypes.data, len(b), F, vox, t, ctypes.byref(val)) == 0
                    sh[vox] = val.value
            rec["leaf"].append(leaf)
            rec["frame"].append(t)
            rec["coarse"].append(co)
            rec["fine"].append(fi)
            rec["shell"].append(sh)
    return {k: np.array(v) for k, v in rec.items()}


def sci_vectors(path: str, data: bytes, frames, nq=2000, seed=7):
    R = oracle_py.ref()
    h = R.vbtref_load_file(path.encode())
    assert h, R.vbtref_last_error()
    of = oracle_py.OracleFile(data)
    W, H, D, F = of.dims
    out = {"frames": np.array(frames, np.int32)}
    dense = []
    for t in frames:
        fr = np.empty((D, H, W), np.float32)
        if R.vbtref_sci_frame(h, t, fr.ctypes.data) != 0:
            # the reference decoder throws at this frame: record its message
            # and the frames that decoded before it
            msg = R.vbtref_last_error().decode()
            R.vbtref_free_file(h)
            ok = frames[:len(dense)]
            return {"frames": np.array(ok, np.int32), "dense": np.stack(dense) if dense else np.zeros((0, D, H, W),
                    np.float32), "error": np.array(msg), "error_frame": np.array(t, np.int32)}
        dense.append(fr)
    out["dense"] = np.stack(dense)
    rng = np.random.default_rng(seed)
    q = np.stack([rng.integers(0, W, nq), rng.integers(0, H, nq), rng.integers(0, D, nq), rng.integers(0, F, nq)],
                 axis=1).astype(np.uint32)
    vals = np.empty(nq, np.float32)
    val = ctypes.c_float()
    for i, (x, y, z, t) in enumerate(q):
        assert R.vbtref_sci_value(h, int(x), int(y), int(z), int(t), ctypes.byref(val)) == 0
        vals[i] = val.value
    R.vbtref_free_file(h)
    out["queries"] = q
    out["values"] = vals
    return out


def main():
    """python tests/golden/make_golden.py [NAME ...]: regenerate all fixtures,
    or only the named ones (the manifest keeps the others)."""
    if not (os.path.exists(REF_KF) and os.path.exists(REF_SF) and oracle_py.have_ref()):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "-j4"], check=True)
    only = sys.argv[1:]
    manifest = {}
    mpath = os.path.join(HERE, "manifest.json")
    if only and os.path.exists(mpath):
        with open(mpath) as fh:
            manifest = json.load(fh)
    with tempfile.TemporaryDirectory() as tmp:
        for name, (wname, extra, frames) in FIXTURES.items():
            if only and name not in only:
                continue
            w = synth.WORKLOADS[wname]
            data = encode_with_reference(w, extra, tmp)
            path = os.path.join(tmp, name + ".vbtp")
            with open(path, "wb") as fh:
                fh.write(data)
            with open(os.path.join(HERE, name + ".vbtp.gz"), "wb") as raw_fh:
                with gzip.GzipFile(fileobj=raw_fh, mode="wb", compresslevel=9, mtime=0) as fh:
                    fh.write(data)
            of = oracle_py.OracleFile(data)
            if of.header.profileType == 2:
                vec = sci_vectors(path, data, frames)
            else:
                vec = render_vectors(data, frames)
            np.savez_compressed(os.path.join(HERE, name + ".expected.npz"), **vec)
            manifest[name] = {
                "workload": wname,
                "dims": list(w.dims),
                "encoder": "reference " + ("vbt_spatialfirst_probe" if w.kind == synth.SCIENTIFIC
                                           else "vbt_render_temporal_kf_probe"),
                "encoder_flags": extra,
                "frames": frames,
                "vbtp_sha256": hashlib.sha256(data).hexdigest(),
                "vbtp_bytes": len(data),
                "modes": np.bincount(of.modes(), minlength=4).tolist(),
            }
            print(name, manifest[name])
    with open(mpath, "w") as fh:
        json.dump(manifest, fh, indent=1)


if __name__ == "__main__":
    main()
