"""Loading the committed golden fixtures (tests/golden/)."""
from __future__ import annotations

import gzip
import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as fh:
        return json.load(fh)


def vbtp(name: str) -> bytes:
    with gzip.open(os.path.join(GOLDEN, name + ".vbtp.gz"), "rb") as fh:
        data = fh.read()
    assert hashlib.sha256(data).hexdigest() == manifest()[name]["vbtp_sha256"], name
    return data


def expected(name: str):
    return dict(np.load(os.path.join(GOLDEN, name + ".expected.npz"), allow_pickle=False))


def basis(name: str):
    """Expected outputs with dense subtypes 2 / 3 decoded as the encoder
    reconstructs them (tests/golden/make_basis_golden.py)."""
    return dict(np.load(os.path.join(GOLDEN, name + ".basis.npz"), allow_pickle=False))
