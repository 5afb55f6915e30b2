"""Differentiable mode on the GPU against the oracles.

Forward: the HIP differentiable render (logit opacities, fp32 rasterizer inputs) must give the
same per-tile lists as the CPU oracle in differentiable mode (bit for bit) and pixels within 1e-5
of it; its distance to the FP64 twin is reported by the FP64 tests below.
Backward (dL/d(rgba) = 2 (rgba - target), the L2 loss of the differentiation workload):
  * against the FP64 autodiff oracle on the same lists: relative error of every parameter group
    (2-norm of the difference / 2-norm of the FP64 gradient) <= 1e-3;
  * against FP64 central differences at 256 seeded parameters (64 means, 48 log-scales,
    48 quaternion, 32 opacity-logit, 64 SH coefficients with non-negligible gradient; a sample
    that crosses a blend / cut-off decision is redrawn): per sample
    |g_hip - g_fd| <= 1e-3 * max(|g_fd|, 0.05 * rms of the group's FP64 gradient over its
    non-zero entries)  - a relative error of 1e-3, with a floor for the entries that are tiny
    against their group.
"""
import numpy as np
import pytest
import torch

from diff_cases import Case, sample_fd
from oracle import diff64

pytestmark = pytest.mark.gpu

NAMES = ("means", "log_scales", "quats", "logits", "sh")


@pytest.fixture(scope="module")
def setup(hip_device):
    from hdsr.diff import DiffRenderer
    case = Case()
    r = DiffRenderer(case.n, case.width, case.height, 3)
    dev = [torch.from_numpy(np.ascontiguousarray(a)).to(hip_device) for a in case.params()]
    rgba, depth = r.render(case.uniforms, *dev)
    torch.cuda.synchronize()
    lists_raw = (r.debug("entries"), r.debug("tile_offsets"), r.debug("tile_counts"))
    return case, r, dev, rgba, depth, lists_raw


def test_forward_matches_oracle(setup):
    case, r, dev, rgba, depth, (entries, offs, counts) = setup
    fr = case.oracle()
    np.testing.assert_array_equal(entries, fr["entries"])
    np.testing.assert_array_equal(offs, fr["tile_offsets"])
    np.testing.assert_array_equal(counts, fr["tile_counts"])
    err = np.abs(rgba.cpu().numpy().reshape(-1) - fr["rgba"]).max()
    assert err <= 1e-5, err
    derr = np.abs(depth.cpu().numpy().reshape(-1) - fr["out_depth"]).max()
    assert derr <= 1e-5, derr


def test_per_pixel_state(setup):
    from hdsr import _lib
    case, r, dev, rgba, depth, (entries, offs, counts) = setup
    tf = r.read_output(_lib.OUT_T_FINAL, np.float32)
    last = r.read_output(_lib.OUT_LAST_INDEX, np.uint32)
    np.testing.assert_array_equal(1.0 - tf, rgba.cpu().numpy()[..., 3])
    lists = case.lists(entries, offs, counts)
    tx = np.arange(case.width) // 16
    ty = np.arange(case.height) // 16
    cnt = np.array([len(l) for l in lists]).reshape(case.plan["tileRows"], case.plan["tileColumns"])
    assert np.all(last <= cnt[ty][:, tx])
    assert last.max() > 0


@pytest.fixture(scope="module")
def grads(setup):
    case, r, dev, rgba, depth, (entries, offs, counts) = setup
    target = case.target()
    G32 = (2.0 * (rgba.cpu().numpy() - target)).astype(np.float32)
    G = torch.from_numpy(G32).to(dev[0].device)
    r.render(case.uniforms, *dev)
    g_hip = [t.cpu().numpy().astype(np.float64) for t in r.backward(G, *dev)]
    lists = case.lists(entries, offs, counts)
    p64 = case.params64()
    g64 = diff64.vjp(p64, case.cam64, lists, case.width, case.height, G32.astype(np.float64))
    return case, dict(zip(NAMES, g_hip)), g64, p64, lists, G32.astype(np.float64)


def test_gradients_match_fp64_autodiff(grads):
    case, g_hip, g64, p64, lists, G = grads
    for name in NAMES:
        ref = g64[name].numpy().reshape(g_hip[name].shape)
        rel = np.linalg.norm(g_hip[name] - ref) / np.linalg.norm(ref)
        print(f"{name}: |g_hip - g_fp64| / |g_fp64| = {rel:.2e}")
        assert rel <= 1e-3, (name, rel)
    assert np.count_nonzero(g_hip["sh"][:, -3:]) == 0


def test_gradients_match_central_differences(grads):
    case, g_hip, g64, p64, lists, G = grads
    picks, replaced = sample_fd(case, p64, lists, G, g64, seed=17)
    total = 0
    worst = 0.0
    for name, items in picks.items():
        ref = g64[name].numpy()
        nz = ref[ref != 0]
        floor = 0.05 * float(np.sqrt(np.mean(nz * nz)))
        for idx, fd in items:
            got = float(g_hip[name].reshape(ref.shape)[idx])
            err = abs(got - fd) / max(abs(fd), floor)
            worst = max(worst, err)
            assert err <= 1e-3, (name, idx, got, fd, err)
            total += 1
    print(f"{total} samples, {replaced} redrawn, worst relative error {worst:.2e}")
    assert total >= 240


def test_autograd_function(setup):
    from hdsr.diff import RenderSplats
    case, r, dev, rgba, depth, _ = setup
    target = torch.from_numpy(case.target()).to(dev[0].device)
    leaves = [t.clone().requires_grad_(True) for t in dev]
    out, _ = RenderSplats.apply(r, case.uniforms, *leaves)
    loss = ((out - target) ** 2).sum()
    # a render of another view in between: backward must restore this view's state first
    cam2 = Case(azimuth=1.3).uniforms
    r.render(cam2, *dev)
    loss.backward()
    G = (2.0 * (out.detach() - target)).contiguous()
    r.render(case.uniforms, *dev)
    ref = r.backward(G, *dev)
    # float atomics add the per-tile partial sums in varying order: two runs agree to the rounding
    # of those partial sums (a few ulp of the largest gradients), not of each (possibly cancelled) total
    for leaf, g in zip(leaves, ref):
        gn = g.cpu().numpy()
        np.testing.assert_allclose(leaf.grad.cpu().numpy(), gn, rtol=1e-4, atol=1e-6 * float(np.abs(gn).max()))


def test_views_accumulate(setup):
    case, r, dev, rgba, depth, _ = setup
    u1, u2 = case.uniforms, Case(azimuth=2.1).uniforms
    gs = []
    for u in (u1, u2):
        img, _ = r.render(u, *dev)
        gs.append(r.backward(torch.ones_like(img), *dev))
    # one backward over two views equals the sum of the two
    from hdsr.diff import RenderSplats
    leaves = [t.clone().requires_grad_(True) for t in dev]
    tot = sum(RenderSplats.apply(r, u, *leaves)[0].sum() for u in (u1, u2))
    tot.backward()
    for k, leaf in enumerate(leaves):
        want = (gs[0][k] + gs[1][k]).cpu().numpy()
        np.testing.assert_allclose(leaf.grad.cpu().numpy(), want, rtol=1e-4,
                                   atol=max(1e-6, 1e-6 * float(np.abs(want).max())))
