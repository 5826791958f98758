"""GPU Morton codes, radix sort and LBVH emission/refit against the oracle (bit-exact)."""
import numpy as np
import pytest

from conftest import golden, nodes_equal
from oracle import binding as ob
from scenes import GOLDEN_SCENES

import sequoia_amd as sq

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("bits", [10, 20, 21])
@pytest.mark.parametrize("name", ["ico1280", "clus3000", "ico69632"])
def test_morton_codes(ctx, name, bits):
    say("morton codes", name, bits)
    tris = GOLDEN_SCENES[name]()
    scene = ctx.upload(tris)
    got = ctx.morton_codes(scene, bits)
    assert np.array_equal(got, ob.morton_codes(tris, bits))


@pytest.mark.parametrize("key_bits,n", [(30, 1), (30, 4095), (30, 4097), (32, 100000), (63, 250000), (64, 70000)])
def test_radix_sort_is_stable(ctx, key_bits, n):
    say("radix sort", key_bits, n)
    rng = np.random.default_rng(key_bits * 1000 + n)
    hi = (1 << 12)
    keys = rng.integers(0, hi, n, dtype=np.uint64)
    keys = (keys * np.uint64(0x9E3779B97F4A7C15 >> (64 - key_bits)) if key_bits < 64 else keys * np.uint64(0x9E3779B97F4A7C15))
    keys &= np.uint64((1 << key_bits) - 1) if key_bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    vals = np.arange(n, dtype=np.uint32)
    k, v = ctx.radix_sort(keys, vals, key_bits)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order])
    assert np.array_equal(v, vals[order])


def _sort_case(name, n, key_bits, rng):
    full = (1 << key_bits) - 1
    if name == "uniform":
        keys = rng.integers(0, 1 << 62, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    elif name == "long_runs":
        # runs of equal top 16 bits far longer than one workgroup's LDS, low bits random
        top = rng.integers(0, 5, n, dtype=np.uint64) << np.uint64(key_bits - 16)
        keys = top | rng.integers(0, 1 << (key_bits - 16), n, dtype=np.uint64)
    elif name == "equal":
        keys = np.full(n, 0x123456789 & full, np.uint64)
    elif name == "low_bit":
        keys = rng.integers(0, 2, n, dtype=np.uint64)
    elif name == "two_digits":
        # only digits 0 and 5 vary (every other pass keeps the order)
        keys = (rng.integers(0, 256, n, dtype=np.uint64) | (rng.integers(0, 256, n, dtype=np.uint64) << np.uint64(40)))
    return keys & np.uint64(full)


@pytest.mark.parametrize("key_bits", [30, 63])
@pytest.mark.parametrize("name,n", [("uniform", 1_000_000), ("uniform", 8192), ("uniform", 5000),
                                    ("long_runs", 300_000), ("long_runs", 20_000), ("equal", 50_000),
                                    ("low_bit", 100_000), ("two_digits", 200_000)])
def test_radix_sort_spans(ctx, key_bits, name, n):
    """Sorts whose two top digits run globally and whose spans of equal top
    bits are finished per workgroup (in LDS, or in chunks through global
    memory for runs longer than the LDS), against a stable argsort."""
    say("radix sort spans", name, n, key_bits)
    rng = np.random.default_rng(n + key_bits)
    keys = _sort_case(name, n, key_bits, rng)
    vals = rng.permutation(n).astype(np.uint32)
    k, v = ctx.radix_sort(keys, vals, key_bits)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order])
    assert np.array_equal(v, vals[order])


def test_radix_sort_morton_codes_of_surfaces(ctx):
    """63-bit Morton codes of a 1.84M-triangle icosphere (runs of equal top
    bits of tens of thousands of keys) sort as a stable argsort does."""
    tris = sq.gen_icosphere(1837637, 6023, 0.1)
    scene = ctx.upload(tris)
    codes = ctx.morton_codes(scene, 21)
    k, v = ctx.radix_sort(codes, np.arange(len(codes), dtype=np.uint32), 63)
    order = np.argsort(codes, kind="stable")
    assert np.array_equal(v, order.astype(np.uint32))
    assert np.array_equal(k, codes[order])
    scene.free()


def test_radix_sort_matches_reference_fixture(ctx):
    g = golden("radix_sort32")
    k, v = ctx.radix_sort(g["keys"].astype(np.uint64), np.arange(len(g["keys"]), dtype=np.uint32), 32)
    assert np.array_equal(k.astype(np.uint32), g["sorted_keys"]) and np.array_equal(v, g["order"])


@pytest.mark.parametrize("digit", [4, 6, 8])
def test_radix_sort_digit_widths(ctx, digit):
    rng = np.random.default_rng(digit)
    keys = rng.integers(0, 1 << 30, 50000, dtype=np.uint64)
    k, v = ctx.radix_sort(keys, np.arange(len(keys), dtype=np.uint32), 30, digit)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(v, order.astype(np.uint32))


BUILD_CASES = [
    ("ico1280", "morton", 10), ("soup2000", "morton", 10), ("clus3000", "morton2", 20),
    ("ico69632", "morton", 10), ("ico69632", "morton63", 21), ("clus3000", "morton63", 21),
]


@pytest.mark.parametrize("layout", [sq.LAYOUT_REFERENCE, sq.LAYOUT_FIXED])
@pytest.mark.parametrize("threshold", [0, 64])
@pytest.mark.parametrize("name,kind,bits", BUILD_CASES)
def test_morton_build_bit_exact(ctx, name, kind, bits, layout, threshold):
    say("morton build", name, kind, layout, threshold)
    tris = GOLDEN_SCENES[name]()
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, kind, layout=layout, subtree_threshold=threshold)
    got = bvh.download()
    want, root, _, _ = ob.build_morton(tris, bits, layout=layout, refit=True)
    assert bvh.root == root
    assert len(nodes_equal(got, want)) == 0


@pytest.mark.parametrize("name", ["ico1280", "soup2000", "clus3000"])
def test_morton_topology_matches_reference_fixture(ctx, name):
    tris = GOLDEN_SCENES[name]()
    scene = ctx.upload(tris)
    g = golden(name)
    for kind in ("morton", "morton2"):
        got = ctx.build(scene, kind, refit=False).download()
        links = np.stack([got["count"].astype(np.uint32), got["flags"].astype(np.uint32), got["child"]], 1)
        assert np.array_equal(links, g["links_" + kind])
        assert not got["min"].any() and not got["max"].any()


def test_small_and_degenerate_scenes(ctx):
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 8, 9, 15, 64, 2047, 2048, 2049):
        tris = rng.random((n, 9), dtype=np.float32)
        say("tiny build", n)
        scene = ctx.upload(tris)
        for layout in (sq.LAYOUT_REFERENCE, sq.LAYOUT_FIXED):
            got = ctx.build(scene, "morton63", layout=layout).download()
            want, _, _, _ = ob.build_morton(tris, 21, layout=layout)
            assert len(nodes_equal(got, want)) == 0
    # all centroids equal -> midpoint splits
    tris = np.tile(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32), (5000, 1))
    scene = ctx.upload(tris)
    got = ctx.build(scene, "morton63").download()
    want, _, _, _ = ob.build_morton(tris, 21)
    assert len(nodes_equal(got, want)) == 0


def test_large_soup_morton63(ctx):
    say("1M soup morton63")
    tris = sq.gen_soup(1_000_000, 2, 0.01)
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, "morton63")
    got = bvh.download()
    want, _, _, _ = ob.build_morton(tris, 21)
    assert len(nodes_equal(got, want)) == 0
    t = ctx.timings()
    say("timings", {k: round(v, 3) if isinstance(v, float) else v for k, v in t.items()})


@pytest.mark.parametrize("kind,bits", [("morton63", 21), ("morton", 10)])
def test_large_icosphere_many_small_tasks(ctx, kind, bits):
    """A 1.84M-triangle icosphere whose Morton ranges of <= 2048 primitives
    split into more than T/4 + 2 tasks (517 in one range): the emission
    must still equal the oracle's tree for the default and smaller subtree
    thresholds."""
    say("large icosphere", kind)
    tris = sq.gen_icosphere(1837637, 6023, 0.1)
    want, _, _, _ = ob.build_morton(tris, bits)
    scene = ctx.upload(tris)
    for thr in (0, 1024):
        bvh = ctx.build(scene, kind, subtree_threshold=thr)
        got = bvh.download()
        assert len(nodes_equal(got, want)) == 0, (kind, thr, nodes_equal(got, want)[:5])
        bvh.free()


def test_phase_timing_opt_in(ctx):
    """Build phases are timed only on request; the total is always timed."""
    tris = sq.gen_soup(50_000, 7, 0.01)
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, "morton63")
    t = ctx.timings()
    assert t["total_build_ms"] > 0 and t["sort_ms"] == 0
    bvh.free()
    ctx.set_phase_timing(True)
    try:
        bvh = ctx.build(scene, "morton63")
        t = ctx.timings()
        assert t["bounds_ms"] > 0 and t["sort_ms"] > 0 and t["emit_ms"] > 0
        bvh.free()
    finally:
        ctx.set_phase_timing(False)


def test_top_phase_global_queue_and_search():
    """The top phase's lane-per-range kernel keeps a level's ranges in LDS up
    to 2048 of them, else in the scene's global queue, and takes a level in
    chunks of 1024 threads; splits come from the top-bits table or, when the
    range's codes agree on all of those bits, from the 8-ary search. A 1M soup
    with a top-phase threshold of 64 (up to ~15k ranges a level) and a dense
    cluster inside a wide scene (every top range within one top-bits cell)
    both equal the oracle's trees."""
    tris = sq.gen_soup(1_000_000, 2, 0.01)
    c = sq.Context(0)
    try:
        for kind, bits in (("morton63", 21), ("morton", 10)):
            say("1M soup, threshold 64,", kind)
            s = c.upload(tris)
            b = c.build(s, kind, subtree_threshold=64)
            assert c.timings()["top_levels"] > 10
            want, root, _, _ = ob.build_morton(tris, bits)
            assert b.root == root and len(nodes_equal(b.download(), want)) == 0
            b.free()
            s.free()
        rng = np.random.default_rng(41)
        dense = (0.5 + 1e-4 * rng.random((200_000, 9))).astype(np.float32)
        wide = np.array([[0, 0, 0, 1e-3, 0, 0, 0, 1e-3, 0], [1, 1, 1, 1, 1, 1 - 1e-3, 1, 1, 1 - 1e-3]], np.float32)
        clustered = np.concatenate([dense[:100_000], wide, dense[100_000:]])
        for threshold in (0, 64):
            say("dense cluster, threshold", threshold)
            s = c.upload(clustered)
            b = c.build(s, "morton63", subtree_threshold=threshold)
            want, root, _, _ = ob.build_morton(clustered, 21)
            assert b.root == root and len(nodes_equal(b.download(), want)) == 0
            b.free()
            s.free()
    finally:
        c.close()
