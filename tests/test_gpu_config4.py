"""The differentiation workload at full size: 500k SH3 splats (the 1M scene's recipe), 1280x720
orbit views, differentiable mode (logit opacities, fp32 rasterizer inputs), against the oracles.

Forward, views 0 and 8: the per-tile lists (Morton-indexed offsets and counts, entries) bit-exact
against the CPU oracle in differentiable mode, pixels and depth within 1e-5 of it.
Backward, view 0: dL/d(rgba) = 2 (rgba - target) with the target the seeded perturbation of the
scene (means + N(0, 0.01^2), log-scales + N(0, 0.1^2), logits + N(0, 0.3^2), SH + N(0, 0.02^2)),
kept on four tiles (lists of about 200, 1000, 3000 entries and a border tile) and zero elsewhere.
Every parameter group against the FP64 autodiff oracle evaluated over those tiles' lists:
relative 2-norm error <= 1e-3; splats in none of the four lists get exactly zero gradient.
"""
import numpy as np
import pytest
import torch

from hdsr import scenes
from oracle import diff64
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

N, SEED = 500_000, 4
NAMES = ("means", "log_scales", "quats", "logits", "sh")


def logit(p):
    p = np.asarray(p, np.float64)
    return np.log(p / (1.0 - p)).astype(np.float32)


@pytest.fixture(scope="module")
def setup(hip_device):
    from hdsr.diff import DiffRenderer
    sc = scenes.scene_config4(N, SEED)
    cams = scenes.cameras_config4()
    W, H = cams[0].width, cams[0].height
    sh = np.ascontiguousarray(sc.sh_data(), np.float32).reshape(N, -1)
    params = [np.ascontiguousarray(a, np.float32) for a in (sc.positions, sc.scales, sc.rotations,
                                                            logit(sc.opacities), sh)]
    r = DiffRenderer(N, W, H, 3)
    dev = [torch.from_numpy(a).to(hip_device) for a in params]
    yield sc, cams, params, r, dev
    r.close()


def oracle_frame(sc, params, u, W, H):
    means, log_scales, quats, logits, sh = params
    arrays = {"positions": means, "scales": log_scales, "rotations": quats, "opacities": logits,
              "materials": sc.materials, "normals": None, "sh_data": sh.reshape(-1),
              "sorted": np.arange(N, dtype=np.uint32), "sh_degree": 3}
    return orc.OracleFrame(arrays, u, W, H, orc.FLAG_DIFF)


@pytest.mark.parametrize("view", [0, 8])
def test_forward_full_size_matches_oracle(setup, view):
    sc, cams, params, r, dev = setup
    cam = cams[view]
    u = r.uniforms(cam)
    rgba, depth = r.render(u, *dev)
    torch.cuda.synchronize()
    fr = oracle_frame(sc, params, u, cam.width, cam.height)
    print(f"view {view}: refs {fr.stats['refs']}, large tiles {fr.stats['large_tiles']}")
    assert fr.stats["refs"] > 5_000_000
    np.testing.assert_array_equal(r.debug("tile_counts"), fr["tile_counts"])
    np.testing.assert_array_equal(r.debug("tile_offsets"), fr["tile_offsets"])
    np.testing.assert_array_equal(r.debug("entries"), fr["entries"])
    err = np.abs(rgba.cpu().numpy().reshape(-1) - fr["rgba"]).max()
    derr = np.abs(depth.cpu().numpy().reshape(-1) - fr["out_depth"]).max()
    print(f"view {view}: max |rgba - oracle| {err:.2e}, max |depth - oracle| {derr:.2e}")
    assert err <= 1e-5 and derr <= 1e-5


def pick_tiles(lists, tiles_x, tiles_y):
    """Tiles whose lists hold about 200, 1000 and 3000 entries, and the busiest bottom-row tile."""
    lens = np.array([len(x) for x in lists])
    out = []
    for want in (200, 1000, 3000):
        t = int(np.argmin(np.abs(lens - want)))
        if lens[t] > 0 and t not in out:
            out.append(t)
    bottom = np.arange((tiles_y - 1) * tiles_x, tiles_y * tiles_x)
    t = int(bottom[np.argmax(lens[bottom])])
    if lens[t] > 0 and t not in out:
        out.append(t)
    return out


def test_backward_full_size_matches_fp64(setup):
    sc, cams, params, r, dev = setup
    cam = cams[0]
    W, H = cam.width, cam.height
    u = r.uniforms(cam)
    rng = np.random.default_rng(SEED + 1)
    pert = [params[0] + rng.normal(0, 0.01, params[0].shape), params[1] + rng.normal(0, 0.1, params[1].shape),
            params[2], params[3] + rng.normal(0, 0.3, params[3].shape), params[4].copy()]
    pert[4][:, :-3] += rng.normal(0, 0.02, params[4][:, :-3].shape)
    target, _ = r.render(u, *[torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev[0].device)
                              for a in pert])
    target = target.cpu().numpy()
    rgba, _ = r.render(u, *dev)
    torch.cuda.synchronize()
    tiles_x, tiles_y = r.plan["tileColumns"], r.plan["tileRows"]
    lists = diff64.tile_lists(r.debug("entries"), r.debug("tile_offsets"), r.debug("tile_counts"),
                              np.arange(N, dtype=np.uint32), tiles_x, tiles_y)
    sel = pick_tiles(lists, tiles_x, tiles_y)
    print("tiles", sel, "list lengths", [len(lists[t]) for t in sel])
    assert len(sel) >= 3 and max(len(lists[t]) for t in sel) >= 1000
    mask = np.zeros((H, W, 1), np.float32)
    for t in sel:
        tx, ty = t % tiles_x, t // tiles_x
        mask[ty * 16:(ty + 1) * 16, tx * 16:(tx + 1) * 16] = 1.0
    G32 = (2.0 * (rgba.cpu().numpy() - target) * mask).astype(np.float32)
    g_hip = [t.cpu().numpy().astype(np.float64)
             for t in r.backward(torch.from_numpy(G32).to(dev[0].device), *dev)]
    g_hip = dict(zip(NAMES, g_hip))

    p64 = diff64.params64(*params)
    q = {k: v.clone().requires_grad_(True) for k, v in p64.items()}
    cam64 = diff64.parse_uniforms(u)
    P = diff64.project(q["means"], q["log_scales"], q["quats"], q["sh"], cam64, logits=q["logits"])
    img = diff64.render(P, lists, W, H, tiles_x, sel)
    (img * torch.as_tensor(G32.astype(np.float64))).sum().backward()
    touched = np.zeros(N, bool)
    for t in sel:
        touched[np.asarray(lists[t], np.int64)] = True
    for name in NAMES:
        ref = q[name].grad.numpy().reshape(g_hip[name].shape)
        rel = np.linalg.norm(g_hip[name] - ref) / np.linalg.norm(ref)
        print(f"{name}: |g_hip - g_fp64| / |g_fp64| = {rel:.2e}")
        assert rel <= 1e-3, (name, rel)
        assert np.count_nonzero(g_hip[name].reshape(N, -1)[~touched]) == 0, name
