"""Size-independent checks of one frame's intermediates (the stress workload's full size, where
the oracle would need minutes).  `get(name)` returns an intermediate under its debug-readback
name; the same checks run on the oracle's own arrays at a reduced size, which pins them."""
import numpy as np

from hdsr import scenes


def stress_camera(width, index=0):
    """The stress workload's orbit camera `index` at `width` x width*9/16 (focal scaled from 1500 at 3840)."""
    return scenes.cameras_orbit(16, width, width * 9 // 16, 1500.0 * width / 3840)[index]


def morton2(x, y):
    def spread(v):
        v = v.astype(np.uint32) & 0xFFFF
        v = (v | (v << 8)) & 0x00FF00FF
        v = (v | (v << 4)) & 0x0F0F0F0F
        v = (v | (v << 2)) & 0x33333333
        return (v | (v << 1)) & 0x55555555
    return spread(x) | (spread(y) << 1)


def invariants(get, st, plan, min_refs=1, min_max_tile=4097, min_chunks=1):
    """Tile counts / offsets from the tile boxes, the large-splat list, per-segment key order,
    small tiles as permutations of their canonical lists, large tiles' bucket order and splat
    multisets.  `st`: refs, visible, large_splats, small_tiles, large_tiles, chunks."""
    tx, ty = plan["tileColumns"], plan["tileRows"]
    # counts and offsets from the GPU's own tile boxes
    cache = get("projcache").reshape(-1, 14)
    box = cache[:, 7]
    vis = box != 0xFFFFFFFF
    x0, y0 = (box & 0xFF).astype(np.int64), ((box >> 8) & 0xFF).astype(np.int64)
    x1, y1 = ((box >> 16) & 0xFF).astype(np.int64), ((box >> 24) & 0xFF).astype(np.int64)
    area = (x1 - x0 + 1) * (y1 - y0 + 1)
    assert st["visible"] == int(vis.sum())
    np.testing.assert_array_equal(get("large_splats")[:st["large_splats"]],
                                  np.nonzero(vis & (area > 16))[0].astype(np.uint32))
    diff = np.zeros((ty + 1) * (tx + 1), np.int64)
    for xs, ys, sign in ((x0, y0, 1), (x1 + 1, y0, -1), (x0, y1 + 1, -1), (x1 + 1, y1 + 1, 1)):
        diff += sign * np.bincount((ys * (tx + 1) + xs)[vis], minlength=diff.size)
    grid = diff.reshape(ty + 1, tx + 1).cumsum(0).cumsum(1)[:ty, :tx]
    gy, gx = np.mgrid[0:ty, 0:tx]
    counts = get("tile_counts")
    offsets = get("tile_offsets")
    want = np.zeros_like(counts)
    want[morton2(gx.ravel(), gy.ravel())] = grid.ravel()
    np.testing.assert_array_equal(counts, want)
    refs = int(counts.astype(np.int64).sum())
    assert st["refs"] == refs and refs >= min_refs
    assert counts.max() >= min_max_tile
    np.testing.assert_array_equal(offsets.astype(np.int64), np.cumsum(counts, dtype=np.int64) - counts)

    entries = get("entries")[:refs]
    canon = get("entries_canon")[:refs]
    keys = get("keys20")[:refs]
    small = get("small_list")[:st["small_tiles"]]
    large = get("large_list")[:st["large_tiles"]]
    chunks = get("chunks")[:2 * st["chunks"]].reshape(-1, 2).astype(np.int64)
    assert st["chunks"] >= min_chunks

    # every segment's keys ascend; local indices stay inside their segment
    sstart = offsets[small].astype(np.int64)
    scount = counts[small].astype(np.int64)
    assert scount.max() <= 4096 and chunks[:, 1].max() <= 4096
    starts = np.concatenate([sstart, chunks[:, 0]])
    lens = np.concatenate([scount, chunks[:, 1]])
    first = np.zeros(refs, bool)
    first[starts] = True
    assert int(first.sum()) == starts.size and int(lens.sum()) == refs
    assert np.all((keys[1:] > keys[:-1]) | first[1:])
    seg_start = np.maximum.accumulate(np.where(first, np.arange(refs, dtype=np.int32), 0))
    lens_at = np.zeros(refs, np.int32)
    lens_at[starts] = lens
    assert np.all((keys & 0xFFF) < lens_at[seg_start])

    # small tiles: final entry = canonical entry at the key's local index
    pos = np.repeat(sstart, scount) + (np.arange(int(scount.sum())) - np.repeat(np.cumsum(scount) - scount, scount))
    np.testing.assert_array_equal(entries[pos], canon[seg_start[pos] + (keys[pos] & 0xFFF)])
    del pos, seg_start, lens_at

    # large tiles: the bucket (key >> 25) never decreases across a tile's chunks, and each tile
    # holds the same splats before and after its sort
    lstart = np.sort(offsets[large].astype(np.int64))
    lend = lstart + counts[large][np.argsort(offsets[large])].astype(np.int64)
    assert (lend - lstart).min() > 4096
    tile_first = np.zeros(refs, bool)
    tile_first[lstart] = True
    tile_first[sstart] = True
    bucket = keys >> 25
    assert np.all((bucket[1:] >= bucket[:-1]) | tile_first[1:])
    del bucket, tile_first
    idx = np.stack([lstart, lend], 1).ravel()
    pad = lambda v: np.concatenate([v, v[:1]])  # noqa: E731  (reduceat needs indices < len)
    e64, c64 = pad(entries.astype(np.uint64)), pad(canon.astype(np.uint64))
    for f in (lambda v: v, lambda v: v * v):
        np.testing.assert_array_equal(np.add.reduceat(f(e64), idx)[0::2], np.add.reduceat(f(c64), idx)[0::2])
    del e64, c64
    np.testing.assert_array_equal(np.bitwise_xor.reduceat(pad(entries), idx)[0::2],
                                  np.bitwise_xor.reduceat(pad(canon), idx)[0::2])

