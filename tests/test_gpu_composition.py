"""The one-leaf assets whose composition the CPU tests pin by closed forms
(test_composition_pins.py: a Fine6 leaf of linear grids, a Mode2 leaf with a
sign-rule truth table), decoded on the GPU: dense frames under every sign
rule, point queries and packed caches, bit for bit against the oracle."""
import numpy as np
import pytest

import oracle_py
from test_composition_pins import BACKGROUND, SHELL_MAX, SHELL_SCALE, fine6_asset, mode2_asset

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu


def same(a, b, what):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape, what
    bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
    assert len(bad) == 0, f"{what}: {len(bad)} mismatches, first {bad[0]}"


@pytest.mark.parametrize("which", ["fine6", "mode2"])
def test_closed_form_leaves_on_the_gpu(which):
    data, F = (fine6_asset()[0], fine6_asset()[2]) if which == "fine6" else (mode2_asset()[0], mode2_asset()[3])
    a = hip.Asset(data)
    o = oracle_py.OracleFile(data)
    rules = [(hip.compose(0), oracle_py.compose(0)), (hip.compose(1), oracle_py.compose(1)),
             (hip.compose(2, background=BACKGROUND), oracle_py.compose(2, background=BACKGROUND)),
             (hip.compose(0, shell_empty_scale=SHELL_SCALE, shell_empty_max=SHELL_MAX),
              oracle_py.compose(0, shell_empty_scale=SHELL_SCALE, shell_empty_max=SHELL_MAX))]
    q = np.array([(x, y, z, t) for t in range(F) for z in range(8) for y in range(8) for x in range(8)], np.uint32)
    for gc, oc in rules:
        for t in range(F):
            same(a.dense(t, gc), o.frame(t, oc), f"{which} dense t={t} rule {oc.signRule}")
        same(a.query(q, gc), o.query(q, oc), f"{which} queries rule {oc.signRule}")
        same(a.query(q, gc, plain=True), o.query(q, oc), f"{which} plain queries rule {oc.signRule}")
    cache = a.frame_cache(0, F)
    for t in range(F):
        same(cache[t], o.cache_packed(t)[0], f"{which} packed cache t={t}")
    a.close()
