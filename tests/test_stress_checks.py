"""The stress workload's size-independent checks (tests/stress_checks.py) run on the oracle's own
frame at a reduced size: they must hold there, and each must catch a planted mistake."""
import numpy as np
import pytest

from hdsr import scenes
from oracle import oracle
from scenes_for_tests import frame_inputs
from stress_checks import invariants, stress_camera

NAMES = {"projcache": "cache", "large_splats": "large_list", "tile_counts": "tile_counts",
         "tile_offsets": "tile_offsets", "entries": "entries", "entries_canon": "entries_canon", "keys20": "keys",
         "small_list": "small_tiles", "large_list": "large_tiles", "chunks": "chunks"}


@pytest.fixture(scope="module")
def frame():
    s = scenes.scene_config2(240_000)
    cam = stress_camera(960)
    plan, srt, u = frame_inputs(s, cam)
    orc = oracle.run(s, srt, u, cam.width, cam.height, oracle.FLAG_NO_COMPOSITE)
    arrays = {k: orc[v].copy() for k, v in NAMES.items()}
    st = {k: orc.stats[k] for k in ("refs", "visible", "large_splats", "small_tiles", "large_tiles", "chunks")}
    return arrays, st, plan


def check(arrays, st, plan):
    invariants(lambda n: arrays[n], st, plan, min_max_tile=40_000)


def test_invariants_hold_on_the_oracle_frame(frame):
    check(*frame)


def planted(frame, name, edit):
    arrays, st, plan = frame
    bad = dict(arrays)
    bad[name] = arrays[name].copy()
    edit(bad[name], arrays, st)
    with pytest.raises(AssertionError):
        check(bad, st, plan)


def first_small_tile(arrays):
    t = arrays["small_list"][np.argmax(arrays["tile_counts"][arrays["small_list"]])]
    return int(arrays["tile_offsets"][t])


def first_large_tile(arrays):
    t = arrays["large_list"][0]
    return int(arrays["tile_offsets"][t]), int(arrays["tile_counts"][t])


def test_each_check_catches_a_planted_mistake(frame):
    def swap_keys(k, a, st):        # two keys of a small tile out of order
        i = first_small_tile(a)
        k[i], k[i + 1] = k[i + 1], k[i]

    def swap_entries(e, a, st):     # two entries of a small tile exchanged, keys left alone
        i = first_small_tile(a)
        e[i], e[i + 1] = e[i + 1], e[i]

    def replace_entry(e, a, st):    # one splat of a large tile replaced by another
        i, _ = first_large_tile(a)
        e[i + 5000] = e[i + 5000] + 1

    def swap_chunks(k, a, st):      # two chunks of a large tile in the wrong bucket order
        i, c = first_large_tile(a)
        lo, hi = k[i:i + 4096].copy(), k[i + 4096:i + 8192].copy()
        if (lo >> 25).min() == (hi >> 25).max():
            pytest.skip("chunks share one bucket")
        k[i:i + 4096], k[i + 4096:i + 8192] = hi, lo

    def bump_count(c, a, st):       # a tile count off by one
        t = a["small_list"][0]
        c[t] += 1

    def drop_large(l, a, st):       # a large splat missing from the list
        l[3] = l[4]

    def bump_box(p, a, st):         # one projected tile box grown by a column
        r = np.nonzero(p.reshape(-1, 14)[:, 7] != 0xFFFFFFFF)[0][0]
        p[r * 14 + 7] += 1 << 16

    for name, edit in (("keys20", swap_keys), ("entries", swap_entries), ("entries", replace_entry),
                       ("keys20", swap_chunks), ("tile_counts", bump_count), ("large_splats", drop_large),
                       ("projcache", bump_box)):
        planted(frame, name, edit)
