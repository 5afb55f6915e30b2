"""The oracle's mode with the reference's WebGPU caps enforced (FLAG_REF_CAPS), pinned against its
own lifted mode and the reference's rules:

  * no cap trips (a small scene, the default budget): every output equals the lifted mode's;
  * the reference budget (maxTotalTileRefs, uniform word 23, gpu_tile_splat_composite.wgsl:114):
    tile offsets are the lifted scan clamped to the budget, every stored entry sits below it, the
    stored entries are exactly the first entries of the lifted canonical lists, and a tile that
    keeps all its entries renders exactly as in the lifted mode; a tile past the budget is
    background;
  * the 65,535-entry large-splat list (gpu_classify_large_splats.wgsl:52,62): the large splats
    past it appear in no stored list, yet keep their slots (the offsets are the lifted ones up to
    the budget), and every tile keeps its listed entries;
  * the 1,024-chunk cap (gpu_tile_bucket_sort.wgsl:126,138): chunks 0..1023 are sorted, later
    chunks keep their stable bucket order (buckets non-decreasing, canonical order within a
    bucket), and each key holds its position.
"""
import numpy as np
import pytest

from hdsr import camera, compositor, scenes
from oracle import oracle
from scenes_for_tests import dense_scene

BG = np.array([0.02, 0.02, 0.04, 0.0], np.float32)


def frame(s, W, H, mult=None, f=300.0, eye=(0, 0, 3.0)):
    cam = camera.Camera(eye, (0, 0, 0), f, f, W, H)
    pin = {"viewportWidth": W, "viewportHeight": H, "tileSizePx": 16, "splatCount": s.count}
    if mult is not None:
        pin["tileEntryMultiplier"] = mult
    plan = compositor.planTileSplatCompositor(pin)
    srt = scenes.sort_back_to_front(s.positions, cam.view)
    return plan, srt, camera.frame_uniforms_for(cam, plan, s.sh_degree)


def big_scene(n=100_000, seed=5, sigma=0.3):
    """Splats of about 12 x 12 tiles at 512x384: every one is large (> 16 tiles)."""
    s = scenes.scene_config0(n, seed)
    rng = np.random.default_rng(seed)
    s.positions = (rng.uniform(-1, 1, (n, 3)) * np.array([1.2, 0.9, 0.3])).astype(np.float32)
    s.scales = (np.full((n, 3), np.log(sigma)) + rng.normal(0, 0.2, (n, 3))).astype(np.float32)
    return s


def test_no_cap_trips_equals_lifted():
    s = dense_scene(n=20_000, seed=4)
    plan, srt, u = frame(s, 320, 240)
    a = oracle.run(s, srt, u, 320, 240, 0)
    b = oracle.run(s, srt, u, 320, 240, oracle.FLAG_REF_CAPS)
    assert not (a.stats["cap_refs_tripped"] or a.stats["cap_large_splats_tripped"] or a.stats["cap_chunks_tripped"])
    for k in ("tile_counts", "tile_offsets", "entries_canon", "entries", "keys", "chunks", "small_tiles",
              "large_tiles", "rgba", "out_depth"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_reference_budget_drops_past_the_cap():
    s = dense_scene(n=20_000, seed=6, core_fraction=0.3)
    W, H = 320, 240
    plan, srt, u = frame(s, W, H, mult=2.0)
    cap = plan["maxTotalTileRefs"]
    a = oracle.run(s, srt, u, W, H, 0)
    b = oracle.run(s, srt, u, W, H, oracle.FLAG_REF_CAPS)
    assert a.stats["large_splats"] < 65535 and a.stats["refs"] > 2 * cap
    assert b.stats["cap_refs_tripped"] == 1 and b.stats["ref_cap"] == cap
    full_off, full_cnt = a["tile_offsets"].astype(np.int64), a["tile_counts"].astype(np.int64)
    np.testing.assert_array_equal(b["tile_offsets"], np.minimum(full_off, cap))
    np.testing.assert_array_equal(b["tile_counts"], np.minimum(full_off + full_cnt, cap) - np.minimum(full_off, cap))
    assert int(b["tile_counts"].sum()) == cap  # nothing unlisted here: the budget is filled exactly
    M = len(full_off)
    canon = a["entries_canon"]
    kept_whole = []
    for m in range(M):
        c = int(b["tile_counts"][m])
        if c == 0:
            continue
        got = b["entries"][b["tile_offsets"][m]:b["tile_offsets"][m] + c]
        np.testing.assert_array_equal(np.sort(got), np.sort(canon[full_off[m]:full_off[m] + c]))
        if c == full_cnt[m]:
            kept_whole.append(m)
            np.testing.assert_array_equal(got, a["entries"][full_off[m]:full_off[m] + c])
    # pixels of tiles that keep every entry are the lifted ones; tiles past the budget are background
    tx, ty = plan["tileColumns"], plan["tileRows"]
    ra, rb = a["rgba"].reshape(H, W, 4), b["rgba"].reshape(H, W, 4)
    whole, empty = 0, 0
    for y in range(ty):
        for x in range(tx):
            m = oracle.morton2(x, y)
            blk = (slice(y * 16, (y + 1) * 16), slice(x * 16, (x + 1) * 16))
            if m in kept_whole:
                np.testing.assert_array_equal(rb[blk], ra[blk])
                whole += 1
            elif b["tile_counts"][m] == 0:
                assert np.all(rb[blk] == BG)
                empty += 1
    assert whole > 10 and empty > 10


@pytest.fixture(scope="module")
def big():
    s = big_scene()
    W, H = 512, 384
    plan, srt, u = frame(s, W, H, mult=200.0)
    return s, plan, srt, u, oracle.run(s, srt, u, W, H, 0), oracle.run(s, srt, u, W, H, oracle.FLAG_REF_CAPS)


def test_large_splat_list_cap(big):
    s, plan, srt, u, a, b = big
    cap = plan["maxTotalTileRefs"]
    assert a.stats["large_splats"] > 65535 and b.stats["cap_large_splats_tripped"] == 1
    unlisted = set(a["large_list"][65535:].tolist())
    listed = a["large_list"][:65535]
    M = len(a["tile_offsets"])
    stored = np.concatenate([b["entries"][b["tile_offsets"][m]:b["tile_offsets"][m] + b["tile_counts"][m]]
                             for m in range(M)])
    assert not (set(stored.tolist()) & unlisted)
    # offsets keep the unlisted slots: the lifted scan, clamped to the budget
    np.testing.assert_array_equal(b["tile_offsets"], np.minimum(a["tile_offsets"].astype(np.int64), cap))
    # a tile below the budget keeps every listed splat that covers it: its count is the lifted
    # count minus the unlisted splats over it
    cache = a["cache"].reshape(-1, 14)
    unl = np.zeros(M, np.int64)
    for r in a["large_list"][65535:]:
        bd = int(cache[r, 7])
        x0, y0, x1, y1 = bd & 255, (bd >> 8) & 255, (bd >> 16) & 255, bd >> 24
        for yy in range(y0, y1 + 1):
            for xx in range(x0, x1 + 1):
                unl[oracle.morton2(xx, yy)] += 1
    full_end = a["tile_offsets"].astype(np.int64) + a["tile_counts"]
    below = full_end <= cap
    np.testing.assert_array_equal(b["tile_counts"][below], (a["tile_counts"].astype(np.int64) - unl)[below])
    assert below.sum() > 50 and len(listed) == 65535


def test_chunk_cap_keeps_bucket_order(big):
    s, plan, srt, u, a, b = big
    ch = b["chunks"].reshape(-1, 2)
    assert len(ch) > 1024 and b.stats["cap_chunks_tripped"] == 1
    keys = b["keys"]
    for i, (st, c) in enumerate(ch):
        k = keys[st:st + c].astype(np.int64)
        if i < 1024:
            assert np.all(np.diff(k) > 0), i  # sorted: (q20, local index) keys strictly ascending
        else:
            np.testing.assert_array_equal(k & 0xFFF, np.arange(c))  # each key holds its position
            assert np.all(np.diff((k >> 12) >> 13) >= 0), i  # bucket order (q20 >> 13 is the bucket)
    # across a large tile whose chunks are all past the cap, the whole list is in bucket order and
    # canonical (stable) order within each bucket
    offs = b["tile_offsets"]
    checked = 0
    for m in b["large_tiles"][::-1]:
        st, c = int(offs[m]), int(b["tile_counts"][m])
        first = np.nonzero(ch[:, 0] == st)[0]
        if not len(first) or first[0] < 1024:
            break
        k = keys[st:st + c].astype(np.int64)
        bucket = (k >> 12) >> 13
        assert np.all(np.diff(bucket) >= 0)
        ent = b["entries"][st:st + c]
        canon = a["entries_canon"][a["tile_offsets"][m]:a["tile_offsets"][m] + c]
        pos = {int(r): i for i, r in enumerate(canon)}
        for bk in np.unique(bucket):
            p = [pos[int(r)] for r in ent[bucket == bk]]
            assert p == sorted(p)
        checked += 1
    assert checked > 5
