"""ctypes front end of libvbt_loadgen.so: open-loop Poisson request traces
replayed against the decode server or the one-request-per-launch path."""
from __future__ import annotations

import ctypes
import os

from . import hip

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvbt_loadgen.so")

MODE_SERVER, MODE_PLAIN, MODE_PER_LAUNCH_BATCHED = 0, 1, 2


class Config(ctypes.Structure):
    _fields_ = [("nAssets", ctypes.c_uint32), ("serverIds", ctypes.POINTER(ctypes.c_int32)),
                ("assets", ctypes.POINTER(ctypes.c_void_p)), ("rate", ctypes.c_double),
                ("requests", ctypes.c_uint32), ("cacheFraction", ctypes.c_double),
                ("maxQueries", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("mode", ctypes.c_int32),
                ("workers", ctypes.c_uint32)]


class Result(ctypes.Structure):
    _fields_ = ([("seconds", ctypes.c_double)] +
                [(n, ctypes.c_uint64) for n in ("requests", "queryRequests", "cacheRequests", "queries",
                                                 "cacheValues", "failed")] +
                [(n, ctypes.c_double) for n in ("p50Us", "p99Us", "meanUs", "maxUs", "qp50Us", "qp99Us",
                                                 "cp50Us", "cp99Us", "queriesPerSec", "requestsPerSec",
                                                 "offeredPerSec")])

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        hip.lib()  # libvbt_hip.so first (the load generator links against it)
        L = ctypes.CDLL(LIB_PATH)
        L.vbtl_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(Config), ctypes.POINTER(Result)]
        _LIB = L
    return _LIB


def run(assets, mode: int, rate: float, requests: int, cache_fraction: float, max_queries: int = 65536,
        seed: int = 1, server: "hip.Server | None" = None, server_ids=None, workers: int = 1) -> dict:
    n = len(assets)
    ids = (ctypes.c_int32 * n)(*(server_ids or [0] * n))
    handles = (ctypes.c_void_p * n)(*[a.h.value for a in assets])
    c = Config(n, ids, handles, rate, requests, cache_fraction, max_queries, seed, mode, workers)
    r = Result()
    rc = lib().vbtl_run(server.h if server is not None else None, ctypes.byref(c), ctypes.byref(r))
    hip.check(rc)
    return r.as_dict()
