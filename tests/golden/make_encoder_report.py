"""Record the reference encoder's own error report for the render golden
fixtures (tests/golden/encoder_reports.json).  Runs only where the
reference sources are present (this container), like make_golden.py.

The reference render encoder (oracle/_ref/vbt_render_temporal_kf_probe)
decodes the leaves it saves with its own packed-leaf sampler and reports the
RMSE against the source field at every sample-step-th voxel and frame
(render_temporal_kf_probe_main.cpp:1646-1775, 1835-1840; report table
:2507-2515).  This script re-encodes smoke64 and liquid_mini exactly as
make_golden.py does, checks that the bytes equal the committed fixtures, and
stores the reported RMSE of the saved variant (routed_empty_grid4_fine6_packed,
or levelset_surface_packed for level sets: the leaves --save-vbt writes,
:1457-1463), the coarse-only RMSE and the sample step.

    python tests/golden/make_encoder_report.py
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import make_golden  # noqa: E402
from make_golden import synth  # noqa: E402


def report_rows(text: str) -> dict:
    """{variant: sampled RMSE} from the report's variant table."""
    rows = {}
    for line in text.splitlines():
        m = re.match(r"\| `([A-Za-z0-9_]+)` \| `[^`]*` \| `([0-9.eE+-]+)` \|", line)
        if m:
            rows[m.group(1)] = float(m.group(2))
    return rows


def main():
    with open(os.path.join(HERE, "manifest.json")) as fh:
        manifest = json.load(fh)
    out = {}
    for name in ("smoke64", "liquid_mini"):
        wl, extra, _ = make_golden.FIXTURES[name]
        w = synth.WORKLOADS[wl]
        with tempfile.TemporaryDirectory() as tmp:
            data = make_golden.encode_with_reference(w, extra, tmp)
            with open(os.path.join(tmp, "r.md")) as fh:
                text = fh.read()
        if hashlib.sha256(data).hexdigest() != manifest[name]["vbtp_sha256"]:
            raise SystemExit(f"{name}: the reference encoder no longer reproduces the committed fixture")
        rows = report_rows(text)
        variant = "levelset_surface_packed" if w.kind == synth.LEVELSET else "routed_empty_grid4_fine6_packed"
        step = int(re.search(r"sample step: `(\d+)`", text).group(1))
        out[name] = {"workload": wl, "variant": variant, "sampled_rmse": rows[variant],
                     "coarse_only_rmse": rows["coarse_only"], "sample_step": step}
        print(name, out[name])
    with open(os.path.join(HERE, "encoder_reports.json"), "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main()
