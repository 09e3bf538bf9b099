"""Expected outputs of the scientific fixtures whose leaves hold dense subtype
2 (tile temporal basis) or 3 (patch temporal basis) blocks, decoded as the
reference encoder reconstructs those blocks.

    python tests/golden/make_basis_golden.py      -> tests/golden/<name>.basis.npz

Needs oracle/_ref (make -C oracle ref).  For every fixture and recorded frame:
  value = coarse + residual, in float32, as decodeScientificValueAtCpu adds
  them (render/src/scientific_decode.cpp:335-359), where
  - coarse is the reference file decoder's value of the same file with every
    leaf's header word set to mode 0 (CoarseOnly: the residual is 0 and the
    coarse coefficients are read exactly as before), and
  - residual is the encoder's own decodeDenseTileTemporalBasisAt /
    decodeDensePatchTemporalBasisAt (src/spatial_first_encoder.cpp:2094-2176)
    through oracle/ref_sf_basis.cpp, which also checks that the encoder's
    append function writes each leaf's block back byte for byte.
The same 2000 seeded queries as the file-decoder fixtures are recorded too.
"""
from __future__ import annotations

import ctypes
import gzip
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_py  # noqa: E402

SF_SO = os.path.join(ROOT, "oracle", "_ref", "libvbtref_sf.so")
BASIS_FIXTURES = ["sci_tile_mini", "sci_tile4_mini", "sci_patch_mini", "sci_patch_b2_mini"]
FRAMES = [0, 13, 39]


def sf():
    L = ctypes.CDLL(SF_SO)
    L.vbtref_sf_basis_leaf.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
    return L


class Residuals:
    """The encoder's residual of one leaf at one frame, 512 voxels, cached."""

    def __init__(self, of: oracle_py.OracleFile):
        self.of, self.L, self.cache = of, sf(), {}

    def __call__(self, leaf: int, t: int) -> np.ndarray:
        key = (leaf, t)
        if key not in self.cache:
            of = self.of
            base = int(of.offsets[leaf]) * 4
            hw = int(of.payload[base:base + 4].view(np.uint32)[0])
            out = np.zeros(512, np.float32)
            if (hw & 3) >= 2 and ((hw >> 12) & 7) in (2, 3):
                keep = ((hw >> 2) & 63) + 1
                blk = np.ascontiguousarray(of.payload[base + 4 + 64 * (keep + 1):])
                used = ctypes.c_uint64()
                rc = self.L.vbtref_sf_basis_leaf(blk.ctypes.data, len(blk), hw, of.header.frames, t, out.ctypes.data,
                                                 ctypes.byref(used))
                if rc != 0:
                    raise RuntimeError(f"leaf {leaf}: encoder harness returned {rc}")
            self.cache[key] = out
        return self.cache[key]


def coarse_only(data: bytes) -> bytes:
    """The same image with every leaf's header word set to mode 0."""
    of = oracle_py.OracleFile(data)
    img = bytearray(data)
    pay = 52 + 4 * of.header.coarseAcScaleCount + 4 * (of.header.leafCount + 1)
    for leaf in range(of.header.leafCount):
        o = pay + 4 * int(of.offsets[leaf])
        hw = int.from_bytes(img[o:o + 4], "little")
        img[o:o + 4] = (hw & 0xFC).to_bytes(4, "little")
    return bytes(img)


def basis_vectors(name: str, tmp: str) -> dict:
    with gzip.open(os.path.join(HERE, name + ".vbtp.gz")) as fh:
        data = fh.read()
    of = oracle_py.OracleFile(data)
    W, H, D, F = of.dims
    path = os.path.join(tmp, name + ".coarse.vbtp")
    with open(path, "wb") as fh:
        fh.write(coarse_only(data))
    R = oracle_py.ref()
    h = R.vbtref_load_file(path.encode())
    assert h, R.vbtref_last_error()
    res = Residuals(of)
    lx, ly = (W + 7) // 8, (H + 7) // 8
    dense = []
    for t in FRAMES:
        fr = np.empty((D, H, W), np.float32)
        assert R.vbtref_sci_frame(h, t, fr.ctypes.data) == 0, R.vbtref_last_error()
        for z in range(0, D, 8):
            for y in range(0, H, 8):
                for x in range(0, W, 8):
                    leaf = ((z // 8) * ly + y // 8) * lx + x // 8
                    r = res(leaf, t).reshape(8, 8, 8)
                    blk = fr[z:z + 8, y:y + 8, x:x + 8]
                    blk[...] = blk + r[:blk.shape[0], :blk.shape[1], :blk.shape[2]]
        dense.append(fr)
    rng = np.random.default_rng(7)
    nq = 2000
    q = np.stack([rng.integers(0, W, nq), rng.integers(0, H, nq), rng.integers(0, D, nq), rng.integers(0, F, nq)],
                 axis=1).astype(np.uint32)
    vals = np.empty(nq, np.float32)
    val = ctypes.c_float()
    for i, (x, y, z, t) in enumerate(q):
        assert R.vbtref_sci_value(h, int(x), int(y), int(z), int(t), ctypes.byref(val)) == 0
        leaf = (int(z) // 8 * ly + int(y) // 8) * lx + int(x) // 8
        vals[i] = np.float32(val.value) + res(leaf, int(t))[((z % 8) * 8 + y % 8) * 8 + x % 8]
    R.vbtref_free_file(h)
    return {"frames": np.array(FRAMES, np.int32), "dense": np.stack(dense), "queries": q, "values": vals}


def main():
    if not (os.path.exists(SF_SO) and oracle_py.have_ref()):
        raise SystemExit("build oracle/_ref first (make -C oracle ref)")
    with tempfile.TemporaryDirectory() as tmp:
        for name in BASIS_FIXTURES:
            vec = basis_vectors(name, tmp)
            np.savez_compressed(os.path.join(HERE, name + ".basis.npz"), **vec)
            print(name, {k: v.shape for k, v in vec.items()})


if __name__ == "__main__":
    main()
