"""The drop-in check: the reference's own unit tests (tests/vbt_file_tests.cpp,
tests/render_temporal_payload_tests.cpp, tests/temporal_decode_tests.cpp)
and one of its render tools (render/src/export_scientific_frame_main.cpp),
compiled from their sources with the reference's own headers and linked
against vbt_amd/libvbt_hip_cxx.so in place of the reference's decode .cpp
files (oracle/Makefile target `dropin`, built into oracle/_ref/ where the
reference sources exist).  Format-only programs run on the CPU; the decode
test and the tool decode on the GPU through libvbt_hip.so."""
import os
import subprocess

import numpy as np
import pytest

import golden_io
import oracle_py

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BIN = os.path.join(ROOT, "oracle", "_ref")


def binary(name):
    path = os.path.join(REF_BIN, name)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (make -C oracle dropin needs the reference sources)")
    return path


def run(args, timeout=120):
    out = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    return out.returncode, out.stdout + out.stderr


def test_reference_file_tests_pass_on_the_dropin():
    rc, text = run([binary("ref_vbt_file_tests")])
    assert rc == 0 and "PASS" in text, text[-2000:]


def test_reference_payload_tests_pass_on_the_dropin():
    rc, text = run([binary("ref_render_temporal_payload_tests")])
    assert rc == 0 and "PASS" in text, text[-2000:]


@pytest.mark.gpu
def test_reference_decode_tests_pass_on_the_dropin():
    """About 850k per-value decodes, each a device round trip."""
    rc, text = run([binary("ref_temporal_decode_tests")], timeout=110)
    assert rc == 0, text[-2000:]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["sci_mini", "sci_dense_mini"])
def test_reference_export_tool_on_the_dropin(name, tmp_path):
    """The tool checks 257 point decodes against its frame itself; its frame
    must equal the oracle's."""
    tool = binary("ref_export_scientific_frame")
    data = golden_io.vbtp(name)
    src = tmp_path / f"{name}.vbtp"
    src.write_bytes(data)
    o = oracle_py.OracleFile(data)
    F = o.header.frames
    for t in sorted({0, F // 2, F - 1}):
        raw = tmp_path / f"{name}_{t}.raw"
        rc, text = run([tool, "--input-vbt", str(src), "--frame", str(t), "--output-raw", str(raw)])
        assert rc == 0, text[-2000:]
        got = np.fromfile(raw, dtype=np.float32)
        exp = np.ascontiguousarray(o.frame(t), np.float32).ravel()
        assert got.shape == exp.shape
        assert (got.view(np.uint32) == exp.view(np.uint32)).all(), (name, t)
    bad = tmp_path / "missing.vbtp"
    rc, text = run([tool, "--input-vbt", str(bad), "--frame", "0", "--output-raw", str(tmp_path / "x.raw")])
    assert rc == 2 and "Failed to query file size" in text
