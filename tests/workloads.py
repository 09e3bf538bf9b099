"""The benchmark workloads at full size, encoded once per test process and
shared by the test files that use them (smoke256, the fire triple,
liquid320, sci192)."""
from __future__ import annotations

import functools
import os

from vbt_amd import synth

# the GPU machine's share of host cores is 16
THREADS = min(16, os.cpu_count() or 1)


@functools.lru_cache(maxsize=None)
def encoded(name: str) -> bytes:
    return synth.WORKLOADS[name].encode(threads=THREADS)
