"""Writes tests/golden/cycles_probe.npz: density samples of the reference's
Cycles blob sampler, from its own CPU probe (cycles_native_vbt/
vbt_sampler_cpu_probe.cpp, compiled by `make -C oracle _ref/vbt_sampler_cpu_probe`
from the reference's sources), on the golden render assets.

For each asset and frame the fixture holds the probe points (index space),
the probe's densities (printed with 9 significant digits, so the float32
values are exact) and the blob size the probe reports.  Points: seeded
uniform points over the volume and half a voxel beyond it, voxel positions,
points just either side of leaf boundaries, and the volume's corners.

    python tests/golden/make_cycles_golden.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PROBE = os.path.join(ROOT, "oracle", "_ref", "vbt_sampler_cpu_probe")
sys.path.insert(0, os.path.dirname(HERE))

import golden_io  # noqa: E402

ASSETS = ("smoke64", "liquid_mini")


def probe_points(W: int, H: int, D: int, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    dim = np.array([W, H, D], np.float64)
    p = rng.random((n, 3)) * (dim + 1.0) - 0.5
    q = n // 5
    p[:q] = np.floor(p[:q])                                     # voxel positions
    e = np.round(p[q:2 * q] / 8.0) * 8.0                       # leaf boundaries
    p[q:2 * q] = e + rng.choice([-1e-3, 0.0, 1e-3], size=(q, 3))
    corners = np.array([[x, y, z] for x in (0, W - 1) for y in (0, H - 1) for z in (0, D - 1)], np.float64)
    p[2 * q:2 * q + 8] = corners
    return p.astype(np.float32)


def run_probe(path: str, frame: int, pts: np.ndarray):
    args = [PROBE, "--input-vbt", path, "--frame", str(frame)]
    for x, y, z in pts:
        args += ["--probe", f"{float(x):.9g},{float(y):.9g},{float(z):.9g}"]
    res = subprocess.run(args, capture_output=True, text=True, check=True)
    rows = [line.split(",") for line in res.stdout.strip().splitlines()[1:]]
    got = np.array([[float(r[0]), float(r[1]), float(r[2])] for r in rows], np.float32)
    assert np.array_equal(got, pts), "the probe did not read the points back exactly"
    dens = np.array([float(r[5]) for r in rows], np.float32)
    return dens, int(rows[0][4]), np.array([int(r[3]) for r in rows])


def main():
    if not os.path.exists(PROBE):
        sys.exit("build the probe first: make -C oracle _ref/vbt_sampler_cpu_probe")
    out = {}
    for ai, name in enumerate(ASSETS):
        data = golden_io.vbtp(name)
        W, H, D, F = np.frombuffer(data[12:28], np.uint32).tolist()
        with tempfile.NamedTemporaryFile(suffix=".vbtp", delete=False) as fh:
            fh.write(data)
            path = fh.name
        try:
            for fi, frame in enumerate((0, F // 3, F - 1, F + 4)):
                pts = probe_points(W, H, D, 600, 1000 * ai + fi)
                dens, size, used = run_probe(path, frame, pts)
                key = f"{name}__{frame}"
                out[key + "__points"] = pts
                out[key + "__density"] = dens
                out[key + "__blob_bytes"] = np.array([size], np.int64)
                out[key + "__frame_used"] = np.array([used[0]], np.int64)
        finally:
            os.unlink(path)
    np.savez_compressed(os.path.join(HERE, "cycles_probe.npz"), **out)
    print("wrote", len(out) // 4, "cases")


if __name__ == "__main__":
    main()
