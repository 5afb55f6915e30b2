"""Serving form: a request's frame is the same alone and inside a continuous batch, and it is the
oracle's frame.

Scenes of different sizes and SH degrees, requests at mixed resolutions and poses.  Every
request is rendered alone by the plain C-ABI path (identity sortedIndices, as the server uses),
checked against the oracle (pixels and depth within 1e-5), then submitted together to a batching
server (several slots, graph replay, eviction under a small byte budget); each result must equal
its alone image bit for bit, so every served frame is within 1e-5 of the oracle.  A serving-scale
case mixes a 1M-splat SH3 scene (the serving workload's largest kind) with a small SH0 scene at
1920x1080, 1280x720 and 640x480 in one batch.
"""
import numpy as np
import pytest

from hdsr import camera, compositor, scenes
from hdsr.serving import Server, poisson_run
from oracle import oracle
from scenes_for_tests import dense_scene

pytestmark = pytest.mark.gpu


def make_scenes():
    a = scenes.scene_config0(n=9000, seed=101)
    b = dense_scene(n=14000, seed=102, sh_degree=3, core_fraction=0.4)
    c = scenes.scene_config1(n=20000, seed=103)
    return {1: a, 2: b, 3: c}


def make_requests(sc, k=18, seed=5):
    rng = np.random.default_rng(seed)
    res = [(320, 240), (640, 480), (480, 272)]
    out = []
    for i in range(k):
        sid = [1, 2, 3][i % 3]
        w, h = res[int(rng.integers(0, 3))]
        cam = camera.Camera(camera.orbit_eye(float(rng.uniform(2.5, 3.5)), float(rng.uniform(0, 6.28)),
                                             float(rng.uniform(-0.3, 0.5))), (0, 0, 0), 500.0 * h / 480,
                            500.0 * h / 480, w, h)
        out.append((sid, cam))
    return out


def alone(scene, cam):
    plan = compositor.planTileSplatCompositor({"viewportWidth": cam.width, "viewportHeight": cam.height,
                                               "tileSizePx": 16, "splatCount": scene.count})
    res = compositor.createTileSplatCompositor(None, plan)
    compositor.bind_scene(res, scene, np.arange(scene.count, dtype=np.uint32))
    res.set_frame(camera.frame_uniforms_for(cam, plan, scene.sh_degree))
    compositor.encodeFullComputeCompositorPipeline(None, res)
    out = (res.read_rgba(), res.read_depth())
    res.destroy()
    return out


def check_oracle(scene, cam, rgba, depth):
    plan = compositor.planTileSplatCompositor({"viewportWidth": cam.width, "viewportHeight": cam.height,
                                               "tileSizePx": 16, "splatCount": scene.count})
    u = camera.frame_uniforms_for(cam, plan, scene.sh_degree)
    orc = oracle.run(scene, np.arange(scene.count, dtype=np.uint32), u, cam.width, cam.height)
    assert np.abs(rgba.reshape(-1) - orc["rgba"]).max() <= 1e-5
    assert np.abs(depth.reshape(-1) - orc["out_depth"]).max() <= 1e-5


@pytest.fixture(scope="module")
def workload(hip_device):
    sc = make_scenes()
    reqs = make_requests(sc)
    ref = [alone(sc[sid], cam) for sid, cam in reqs]
    for (sid, cam), (rgba, depth) in zip(reqs, ref):  # served == alone (below) == oracle
        check_oracle(sc[sid], cam, rgba, depth)
    return sc, reqs, ref


@pytest.mark.parametrize("graphs,budget", [(True, 0), (False, 0), (True, 6 << 20)])
def test_batched_equals_alone(workload, graphs, budget):
    sc, reqs, ref = workload
    srv = Server(hbm_budget_bytes=budget, max_views_per_batch=64, slots=2, use_graphs=graphs)
    for sid, s in sc.items():
        srv.add_scene(sid, s)
    for rep in range(2):  # the second round meets the first round's batch shapes again
        srv.hold()  # the burst queues whole, so batches form whatever the submit timing
        tickets = [srv.submit(sid, cam) for sid, cam in reqs]
        srv.drain()
        views = []
        for t, (rgba, depth) in zip(tickets, ref):
            info = srv.result(t)
            assert info["done"]
            views.append(info["batch_views"])
            np.testing.assert_array_equal(srv.read_rgba(t), rgba)
            np.testing.assert_array_equal(srv.read_depth(t), depth)
            srv.release(t)
        assert max(views) > 1, "requests were never batched"
    st = srv.stats()
    if budget:
        assert st["evictions"] > 0 and st["uploads"] > 3
    srv.close()


def test_graph_replay_equals_alone(workload):
    # batch shape depends on arrival timing; waiting on each request makes every batch one frame
    # of the same shape: the first runs eagerly, the second captures, every later one replays
    sc, reqs, ref = workload
    srv = Server(max_views_per_batch=64, slots=1, use_graphs=True)
    for sid, s in sc.items():
        srv.add_scene(sid, s)
    picks = [i for i, (sid, cam) in enumerate(reqs) if sid == reqs[0][0] and cam.width == reqs[0][1].width]
    for i in picks * 2:
        t = srv.submit(*reqs[i])
        srv.wait(t)
        np.testing.assert_array_equal(srv.read_rgba(t), ref[i][0])
        srv.release(t)
    assert srv.stats()["graph_replays"] >= len(picks) * 2 - 2
    srv.close()


def test_unbatched_server_equals_alone(workload):
    sc, reqs, ref = workload
    srv = Server(max_views_per_batch=1, slots=1, use_graphs=False)
    for sid, s in sc.items():
        srv.add_scene(sid, s)
    for (sid, cam), (rgba, _) in list(zip(reqs, ref))[:6]:
        t = srv.submit(sid, cam)
        srv.wait(t)
        assert srv.result(t)["batch_views"] == 1
        np.testing.assert_array_equal(srv.read_rgba(t), rgba)
        srv.release(t)
    srv.close()


def test_poisson_stream_completes(workload):
    sc, reqs, _ = workload
    srv = Server(max_views_per_batch=64, slots=2)
    for sid, s in sc.items():
        srv.add_scene(sid, s)
    pre = [(sid, cam.width, cam.height, srv.uniforms(sid, cam)) for sid, cam in reqs]
    st = poisson_run(srv, pre, rate_hz=400.0, duration_s=0.5)
    assert st["frames"] == st["submitted"] > 50
    assert st["p99_ms"] >= st["p50_ms"] > 0
    srv.close()


def test_serving_scale_mixed_resolutions(hip_device):
    """A 1M-splat SH3 clusters-and-sheets scene and a 50k SH0 scene, requests at 1920x1080,
    1280x720 and 640x480 batched together: each equals its alone frame bit for bit; one 1080p
    request of the big scene against the oracle at full size."""
    big = scenes.scene_config1(1_000_000, seed=301)
    small = scenes.scene_config0(50_000, seed=302)
    sc = {7: big, 8: small}
    reqs = []
    for i, (w, h) in enumerate([(1920, 1080), (1280, 720), (640, 480), (1280, 720), (1920, 1080), (640, 480)]):
        f = 1500.0 * h / 1080
        cam = camera.Camera(camera.orbit_eye(3.0, 0.8 * i, 0.2), (0, 0, 0), f, f, w, h)
        reqs.append((7 if i % 3 != 2 else 8, cam))
    ref = [alone(sc[sid], cam) for sid, cam in reqs]
    check_oracle(big, reqs[0][1], *ref[0])
    srv = Server(max_views_per_batch=64, slots=2, use_graphs=True)
    for sid, s in sc.items():
        srv.add_scene(sid, s)
    srv.hold()  # the six requests queue together, so they share launches
    tickets = [srv.submit(sid, cam) for sid, cam in reqs]
    srv.drain()
    batched = 0
    for t, (rgba, depth) in zip(tickets, ref):
        batched = max(batched, srv.result(t)["batch_views"])
        np.testing.assert_array_equal(srv.read_rgba(t), rgba)
        np.testing.assert_array_equal(srv.read_depth(t), depth)
        srv.release(t)
    assert batched > 1
    srv.close()


def test_batch_past_two_launch_classification(hip_device):
    """Nine 1920x1080 views of one scene in one batch hold 73,440 tiles, past the 65,536 that the
    two-launch tile classification takes, so the batch runs the scanned classification: each view
    equals its alone frame bit for bit."""
    s = scenes.scene_config0(50_000, seed=303)
    cams = [camera.Camera(camera.orbit_eye(3.0, 0.7 * i, 0.1), (0, 0, 0), 1500.0, 1500.0, 1920, 1080)
            for i in range(9)]
    ref = [alone(s, cam) for cam in cams]
    srv = Server(max_views_per_batch=64, slots=1, use_graphs=False)
    srv.add_scene(1, s)
    srv.hold()
    tickets = [srv.submit(1, cam) for cam in cams]
    srv.drain()
    views = [srv.result(t)["batch_views"] for t in tickets]
    for t, (rgba, depth) in zip(tickets, ref):
        np.testing.assert_array_equal(srv.read_rgba(t), rgba)
        np.testing.assert_array_equal(srv.read_depth(t), depth)
        srv.release(t)
    srv.close()
    assert max(views) == 9, views
