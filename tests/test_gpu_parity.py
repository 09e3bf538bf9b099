"""GPU parity: the HIP decode path (through the C-ABI) against the oracle and
the reference decoder's golden vectors.  Bit-exact on every value: indices,
control values, voxels and query answers (the FP32 bar is the reference's own
expression order; the stated tolerance ``1e-6 * field range`` is checked too,
but every assertion below is at 0 ulp)."""
import math

import numpy as np
import pytest

import golden_io
import oracle_py

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu

RENDER = ["smoke64", "liquid_mini"]
SCI = ["sci_mini", "sci_dense_mini", "sci_tile_mini"]
ALL = RENDER + SCI


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bitexact(got, exp, what):
    got = np.asarray(got, np.float32)
    exp = np.asarray(exp, np.float32)
    assert got.shape == exp.shape, (what, got.shape, exp.shape)
    bad = np.nonzero(bits(got) != bits(exp))[0] if got.ndim == 1 else np.argwhere(bits(got) != bits(exp))
    if len(bad):
        i = tuple(bad[0]) if got.ndim > 1 else bad[0]
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at {i}: got {got[i]!r} expected {exp[i]!r}")


@pytest.fixture(scope="module")
def assets():
    out = {}
    for n in ALL:
        data = golden_io.vbtp(n)
        out[n] = (hip.Asset(data), oracle_py.OracleFile(data))
    yield out
    for a, _ in out.values():
        a.close()


@pytest.mark.parametrize("name", ALL)
def test_page_table_is_rank_order(assets, name):
    a, o = assets[name]
    compact, n = o.page_table()
    got = a.page_table()
    np.testing.assert_array_equal(got, compact)
    assert a.info.activeLeaves == n
    # invariant: compact index is the rank of the leaf among active leaves
    act = np.nonzero(compact != 0xFFFFFFFF)[0]
    np.testing.assert_array_equal(compact[act], np.arange(len(act)))
    assert a.info.badLeaves == 0


@pytest.mark.parametrize("name", ALL)
def test_cache_offsets(assets, name):
    a, o = assets[name]
    _, offs = o.cache_packed(0)
    np.testing.assert_array_equal(a.cache_offsets(), offs)
    assert a.info.controlsPerFrame == offs[-1]


@pytest.mark.parametrize("name", ALL)
def test_frame_cache_packed_all_frames(assets, name):
    a, o = assets[name]
    F = a.dims[3]
    got = a.frame_cache(0, F)
    for t in range(F):
        exp, _ = o.cache_packed(t)
        assert_bitexact(got[t], exp, f"{name} cache frame {t}")


@pytest.mark.parametrize("name", RENDER)
def test_frame_cache_subranges(assets, name):
    a, o = assets[name]
    F = a.dims[3]
    for t0, nt in [(0, 1), (F - 1, 1), (F // 3, max(1, F // 2))]:
        got = a.frame_cache(t0, nt)
        for i in range(nt):
            exp, _ = o.cache_packed(t0 + i)
            assert_bitexact(got[i], exp, f"{name} cache [{t0},{t0 + nt}) frame {t0 + i}")


@pytest.mark.parametrize("name", RENDER)
def test_frame_cache_stride576(assets, name):
    a, o = assets[name]
    F = a.dims[3]
    for t in [0, F // 2, F - 1]:
        got = a.frame_cache(t, 1, hip.CACHE_STRIDE576).reshape(-1, 576)
        assert_bitexact(got, o.cache_576(t), f"{name} 576-cache frame {t}")


@pytest.mark.parametrize("name", RENDER)
def test_render_controls_equal_reference_vectors(assets, name):
    """Packed cache slots against the reference decoder's own outputs."""
    a, o = assets[name]
    exp = golden_io.expected(name)
    compact, _ = o.page_table()
    offs = a.cache_offsets()
    frames = sorted(set(int(t) for t in exp["frame"]))
    cache = {t: a.frame_cache(t, 1)[0] for t in frames}
    for row in range(len(exp["leaf"])):
        leaf, t = int(exp["leaf"][row]), int(exp["frame"][row])
        c = compact[leaf]
        v = o.parse(leaf)
        cc = v.header.coarseResolution ** 3
        slots = cache[t][offs[c]:offs[c + 1]]
        assert_bitexact(slots[:cc], exp["coarse"][row][:cc], f"{name} leaf {leaf} t {t} coarse")
        if v.header.mode == 3:
            fc = v.header.fineResolution ** 3
            assert_bitexact(slots[cc:cc + fc], exp["fine"][row][:fc], f"{name} leaf {leaf} t {t} fine")
        if v.header.mode == 2:
            sh = exp["shell"][row]
            occ = np.nonzero(sh != 0)[0]
            # shell slots are in descriptor order = ascending occupied voxel index
            got = slots[cc:]
            assert len(got) == v.shellActiveVoxelCount
            # every nonzero reference shell value appears at its descriptor index
            b = np.ascontiguousarray(o.leaf_bytes(leaf))
            occ_off = v.layout.fine.descriptorOffset
            mask = np.frombuffer(b[occ_off:occ_off + 64].tobytes(), np.uint64)
            vox = [g * 64 + i for g in range(8) for i in range(64) if (int(mask[g]) >> i) & 1]
            assert_bitexact(got, sh[vox], f"{name} leaf {leaf} t {t} shell")
            assert set(occ) <= set(vox)


@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("rule", [hip.SIGN_NONE, hip.SIGN_CLAMP_ZERO, hip.SIGN_BACKGROUND])
def test_dense_frame(assets, name, rule):
    a, o = assets[name]
    if rule != hip.SIGN_NONE and name in SCI:
        pytest.skip("sign rules apply to Mode2 leaves only")
    F = a.dims[3]
    for t in sorted({0, min(17, F - 1), F - 1}):
        comp_h = hip.compose(rule, background=-3.0)
        comp_o = oracle_py.compose(rule, background=-3.0)
        assert_bitexact(a.dense(t, comp_h), o.frame(t, comp_o), f"{name} dense t={t} rule={rule}")


def test_shell_empty_scale(assets):
    a, o = assets["liquid_mini"]
    comp_h = hip.compose(hip.SIGN_NONE, 0.0, 0.25, 0.5)
    comp_o = oracle_py.compose(0, 0.0, 0.25, 0.5)
    assert_bitexact(a.dense(5, comp_h), o.frame(5, comp_o), "shellEmptyScale")


def test_levelset_sign_identity(assets):
    """Mode2: decoded phi sign identical to the oracle at every |phi| < 2 voxel."""
    a, o = assets["liquid_mini"]
    for t in range(a.dims[3]):
        g = a.dense(t)
        e = o.frame(t)
        band = np.abs(e) < 2.0
        assert band.any()
        np.testing.assert_array_equal(np.signbit(g[band]), np.signbit(e[band]))


@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("plain", [False, True])
def test_point_queries(assets, name, plain):
    a, o = assets[name]
    W, H, D, F = a.dims
    rng = np.random.default_rng(0x5EED0001)
    n = 10000
    q = np.stack([rng.integers(0, W, n), rng.integers(0, H, n), rng.integers(0, D, n),
                  rng.integers(0, F + (0 if name in SCI else 5), n)], axis=1).astype(np.uint32)
    if name in SCI:
        q[:, 3] = np.minimum(q[:, 3], F - 1)
    got = a.query(q, plain=plain)
    exp = o.query(q)
    assert_bitexact(got, exp, f"{name} queries plain={plain}")


@pytest.mark.parametrize("name", SCI)
def test_scientific_against_reference_vectors(assets, name):
    a, _ = assets[name]
    exp = golden_io.expected(name)
    for i, t in enumerate(exp["frames"]):
        assert_bitexact(a.dense(int(t)), exp["dense"][i], f"{name} reference frame {t}")
    assert_bitexact(a.query(exp["queries"]), exp["values"], f"{name} reference queries")


def test_query_errors_are_per_query(assets):
    a, o = assets["sci_mini"]
    W, H, D, F = a.dims
    q = np.array([[0, 0, 0, 0], [W, 0, 0, 0], [1, 2, 3, F], [3, 3, 3, 1]], np.uint32)
    vals, st = a.query(q, with_status=True)
    assert st[0] == 0 and st[3] == 0
    assert st[1] == 32 and st[2] == 33
    assert math.isnan(vals[1]) and math.isnan(vals[2])
    exp = o.query(q[[0, 3]])
    assert_bitexact(vals[[0, 3]], exp, "valid queries beside invalid ones")


def test_fp32_tolerance_statement(assets):
    """The stated bar (max abs error <= 1e-6 * field range) -- implied by 0 ulp."""
    a, o = assets["smoke64"]
    g = a.dense(17)
    e = o.frame(17)
    rng = float(e.max() - e.min())
    assert np.max(np.abs(g - e)) <= 1e-6 * rng


def test_dense_patch_subtype_reads_past_the_pool_like_the_reference():
    """sci_patch_mini (dense subtype 3, read by the reference as per-frame
    BFP data): the leaves whose reads run past the payload pool at some frame
    are poisoned at open (for every frame: stricter than the reference, which
    throws only at the frames that read past it) and whole-frame decodes
    fail; point queries fail only in those leaves and elsewhere equal the
    oracle."""
    data = golden_io.vbtp("sci_patch_mini")
    a = hip.Asset(data)
    o = oracle_py.OracleFile(data)
    assert a.info.badLeaves > 0
    exp = golden_io.expected("sci_patch_mini")
    with pytest.raises(hip.VbtError):
        a.dense(int(exp["error_frame"]))
    W, H, D, F = a.dims
    rng = np.random.default_rng(4)
    q = np.stack([rng.integers(0, W, 5000), rng.integers(0, H, 5000), rng.integers(0, D, 5000),
                  rng.integers(0, F, 5000)], axis=1).astype(np.uint32)
    vals, st = a.query(q, with_status=True)
    exp = o.query(q, check=False)
    good = st == 0
    assert good.any() and (~good).any()
    assert np.isnan(vals[~good]).all()
    assert_bitexact(vals[good], exp[good], "patch-subtype queries in readable leaves")
    a.close()

