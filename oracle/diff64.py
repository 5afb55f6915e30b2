"""FP64 autodiff oracle of the differentiable mode (torch float64 on the CPU).

TEST INFRASTRUCTURE ONLY: used by tests/ and the benchmark's checks, never by the product path.

A restatement, in float64 torch operations, of the forward that the differentiable mode
renders, so that torch's autograd gives the exact derivative of that function:
  * project(): project_splats (src/shaders/gpu_project_splats.wgsl:112-358) for the values
    the rasterizer uses - screen centre, conic (k0, k1, k2), opacity (sigmoid of the logit
    times the sqrt(det / det_blur) factor, clamped to [0, 0.99]) and the SH colour (all-positive
    basis, view direction camera - centre, :300-347) clamped to [0, 1];
  * composite(): the front-to-back walk of composite (gpu_tile_splat_composite.wgsl:123-370) over
    given per-tile lists: alpha = min(0.99, o exp(power)), an entry blends when power > -4,
    alpha > 1/255 and the transmittance before it is >= 0.001; rgba = (C + T (0.02, 0.02, 0.04),
    1 - T).
The lists, the visibility and the blend / cut-off decisions are held fixed (computed without
gradient), which is the function the backward differentiates.  Constants that the WGSL (and the
C++ oracle's FP64 twin, oracle/hdsr_oracle.cpp composite64) take as f32 literals keep their
f32 values here, so the two FP64 restatements agree to rounding.
"""
from __future__ import annotations

import numpy as np
import torch

F64 = torch.float64
# f32 literals of the projection, as float64
MIN_CLIP_W = float(np.float32(0.0001))
LOW_PASS = float(np.float32(0.3))
DET_EPS = float(np.float32(0.000001))
VZ_MAX = float(np.float32(-0.001))
OPAC_MAX = float(np.float32(0.99))
# SH basis constants (gpu_project_splats.wgsl:58-71), f32 literals there
SH_C1 = float(np.float32(0.4886025119029199))
SH_C2 = tuple(float(np.float32(v)) for v in (1.0925484305920792, 1.0925484305920792, 0.31539156525252005,
                                              1.0925484305920792, 0.5462742152960396))
SH_C3 = tuple(float(np.float32(v)) for v in (0.5900435899266435, 2.890611442640554, 0.4570457994644658,
                                              0.3731763325901154, 0.4570457994644658, 1.445305721320277,
                                              0.5900435899266435))
BG = (0.02, 0.02, 0.04)


def parse_uniforms(u) -> dict:
    """The 256-byte frame block (gpuTileSplatCompositor.ts:730-774) as float64 fields."""
    f = np.asarray(u, dtype=np.float32).reshape(64)
    w = f.view(np.uint32)
    d = lambda a: torch.tensor(np.asarray(a, dtype=np.float64), dtype=F64)  # noqa: E731
    return {"vp": d(f[0:16]), "vw": float(f[16]), "vh": float(f[17]), "tiles_x": int(w[20]), "tiles_y": int(w[21]),
            "splat_scale": float(f[24]), "sh_degree": int(w[25]), "cam": d(f[28:31]), "view": d(f[32:48]),
            "fx": float(f[48]), "fy": float(f[49])}


def _mat4_point(m, p):
    """Column-major mat4 times (p, 1): ((m[i] x + m[4+i] y) + m[8+i] z) + m[12+i]."""
    return [((m[i] * p[:, 0] + m[4 + i] * p[:, 1]) + m[8 + i] * p[:, 2]) + m[12 + i] for i in range(4)]


def _cross(a, b):
    return torch.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1], a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                        a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], dim=-1)


def sh_color(sh, deg: int, x, y, z):
    """SH colour (gpu_project_splats.wgsl:300-347): sh is (N, stride) packed as
    [dc rgb | K coefficients rgb-interleaved | emissive rgb]."""
    c = sh[:, 0:3]
    if deg >= 1:
        s1 = sh[:, 3:12].reshape(-1, 3, 3)
        c = c + SH_C1 * ((s1[:, 0] * y[:, None] + s1[:, 1] * z[:, None]) + s1[:, 2] * x[:, None])
    if deg >= 2:
        xx, yy, zz, xy, yz, xz = x * x, y * y, z * z, x * y, y * z, x * z
        b2 = [SH_C2[0] * xy, SH_C2[1] * yz, SH_C2[2] * ((2 * zz - xx) - yy), SH_C2[3] * xz, SH_C2[4] * (xx - yy)]
        s2 = sh[:, 12:27].reshape(-1, 5, 3)
        c = c + sum(s2[:, j] * b2[j][:, None] for j in range(5))
        if deg >= 3:
            b3 = [(SH_C3[0] * y) * (3 * xx - yy), ((SH_C3[1] * x) * y) * z, (SH_C3[2] * y) * ((4 * zz - xx) - yy),
                  (SH_C3[3] * z) * ((2 * zz - 3 * xx) - 3 * yy), (SH_C3[4] * x) * ((4 * zz - xx) - yy),
                  (SH_C3[5] * z) * (xx - yy), (SH_C3[6] * x) * (xx - 3 * yy)]
            s3 = sh[:, 27:48].reshape(-1, 7, 3)
            c = c + sum(s3[:, j] * b3[j][:, None] for j in range(7))
    return c


def project(means, log_scales, quats, sh, cam: dict, *, logits=None, opacities=None) -> dict:
    """Per-splat rasterizer inputs, float64.  Exactly one of logits / opacities is given."""
    vp, V = cam["vp"], cam["view"]
    clip = _mat4_point(vp, means)
    visible = (clip[3] > 0).detach()
    safe_w = torch.clamp(clip[3], min=MIN_CLIP_W)
    ndc_x, ndc_y = clip[0] / safe_w, clip[1] / safe_w
    px = (ndc_x * 0.5 + 0.5) * cam["vw"]
    py = (1.0 - (ndc_y * 0.5 + 0.5)) * cam["vh"]
    depth = clip[2] / safe_w

    qlen = torch.sqrt(((quats[:, 0] ** 2 + quats[:, 1] ** 2) + quats[:, 2] ** 2) + quats[:, 3] ** 2)
    qn = quats / torch.clamp(qlen, min=DET_EPS)[:, None]
    s = torch.exp(log_scales) * cam["splat_scale"]
    u = qn[:, 1:4]
    eye = torch.eye(3, dtype=F64)
    axes = []
    for i in range(3):
        e = eye[i].expand_as(u)
        inner = _cross(u, e) + qn[:, 0:1] * e
        axes.append((e + 2.0 * _cross(u, inner)) * s[:, i:i + 1])  # rotateAxis (wgsl:75-79) times scale
    w0 = torch.stack([V[0], V[4], V[8]])
    w1 = torch.stack([V[1], V[5], V[9]])
    w2 = torch.stack([V[2], V[6], V[10]])
    vc = _mat4_point(V, means)
    vz = torch.clamp(vc[2], max=VZ_MAX)
    j1x, j1y = cam["fx"] / vz, cam["fy"] / vz
    j2x = (-j1x / vz) * vc[0]
    j2y = (-j1y / vz) * vc[1]
    tt0 = j1x[:, None] * w0 + j2x[:, None] * w2
    tt1 = j1y[:, None] * w1 + j2y[:, None] * w2
    b0 = torch.stack([(a * tt0).sum(1) for a in axes], dim=1)
    b1 = torch.stack([(a * tt1).sum(1) for a in axes], dim=1)
    cxx_raw, cxy, cyy_raw = (b0 * b0).sum(1), (b0 * b1).sum(1), (b1 * b1).sum(1)
    det_orig = cxx_raw * cyy_raw - cxy * cxy
    cxx, cyy = cxx_raw + LOW_PASS, cyy_raw + LOW_PASS
    det_blur = cxx * cyy - cxy * cxy
    dbc = torch.clamp(det_blur, min=DET_EPS)
    aa = torch.sqrt(torch.clamp(det_orig / dbc, min=0.0))
    with torch.no_grad():
        mid = 0.5 * (cxx + cyy)
        lam = mid + torch.sqrt(torch.clamp(mid * mid - torch.clamp(det_blur, min=float(np.float32(0.1))),
                                           min=float(np.float32(0.1))))
        radius = torch.ceil(3.0 * torch.sqrt(lam))
        visible = visible & ~(radius <= 0)
    k0, k1, k2 = cyy / dbc, -cxy / dbc, cxx / dbc
    o_in = torch.sigmoid(logits) if logits is not None else opacities
    opac = torch.clamp(o_in * aa, 0.0, OPAC_MAX)
    v = cam["cam"][None, :] - means
    d = v / torch.sqrt((v * v).sum(1, keepdim=True))
    rgb = torch.clamp(sh_color(sh, cam["sh_degree"], d[:, 0], d[:, 1], d[:, 2]), 0.0, 1.0)
    return {"px": px, "py": py, "k0": k0, "k1": k1, "k2": k2, "opacity": opac, "rgb": rgb, "depth": depth,
            "visible": visible}


def _morton2(x: int, y: int) -> int:
    m = 0
    for b in range(16):
        m |= ((x >> b) & 1) << (2 * b) | ((y >> b) & 1) << (2 * b + 1)
    return m


def tile_lists(entries, tile_offsets, tile_counts, sorted_ids, tiles_x: int, tiles_y: int) -> list:
    """Per-tile splat ids, row-major tiles, from the Morton-indexed offsets / counts and the
    entries (sort ranks) of the binning, mapped to splat ids through sortedIndices."""
    out = []
    for ty in range(tiles_y):
        for tx in range(tiles_x):
            m = _morton2(tx, ty)
            s, c = int(tile_offsets[m]), int(tile_counts[m])
            out.append(np.asarray(sorted_ids, dtype=np.int64)[np.asarray(entries[s:s + c], dtype=np.int64)])
    return out


def composite_tile(P: dict, ids, X, Y) -> tuple:
    """One tile: pixels at (X + 0.5, Y + 0.5), entries 'ids' in order.  Returns (rgba (Q, 4),
    blend mask (Q, K), cut mask (Q, K)) - the masks are the discrete decisions."""
    idx = torch.as_tensor(ids, dtype=torch.long)
    qx = (X.to(F64) + 0.5)[:, None]
    qy = (Y.to(F64) + 0.5)[:, None]
    dx = qx - P["px"][idx][None, :]
    dy = qy - P["py"][idx][None, :]
    power = ((-0.5 * P["k0"][idx]) * dx * dx - P["k1"][idx] * dx * dy) - (0.5 * P["k2"][idx]) * dy * dy
    alpha = torch.clamp(P["opacity"][idx] * torch.exp(power), max=0.99)
    with torch.no_grad():
        ok = (power > -4.0) & (alpha > 1.0 / 255.0) & P["visible"][idx][None, :]
        a0 = torch.where(ok, alpha, torch.zeros_like(alpha))
        tex = torch.cumprod(torch.cat([torch.ones_like(a0[:, :1]), 1.0 - a0[:, :-1]], dim=1), dim=1)
        cut = tex >= 0.001
        blend = ok & cut
    a = torch.where(blend, alpha, torch.zeros_like(alpha))
    one = torch.ones_like(a[:, :1])
    T = torch.cumprod(torch.cat([one, 1.0 - a], dim=1), dim=1)
    w = a * T[:, :-1]
    Tf = T[:, -1]
    C = w @ P["rgb"][idx]
    bg = torch.tensor(BG, dtype=F64)
    rgba = torch.cat([C + Tf[:, None] * bg[None, :], (1.0 - Tf)[:, None]], dim=1)
    return rgba, blend, cut


def render(P: dict, lists: list, width: int, height: int, tiles_x: int, tiles: list | None = None):
    """(H, W, 4) float64 image over the given tile lists (row-major tile index -> ids); only the
    tiles in 'tiles' are evaluated when given (the rest stay 0)."""
    ly, lx = torch.meshgrid(torch.arange(16), torch.arange(16), indexing="ij")
    ys, xs, vals = [], [], []
    for t in (range(len(lists)) if tiles is None else tiles):
        ids = lists[t]
        tx, ty = t % tiles_x, t // tiles_x
        X, Y = (tx * 16 + lx).reshape(-1), (ty * 16 + ly).reshape(-1)
        keep = (X < width) & (Y < height)
        X, Y = X[keep], Y[keep]
        if len(ids) == 0:
            val = torch.tensor([BG[0], BG[1], BG[2], 0.0], dtype=F64).expand(len(X), 4)
        else:
            val, _, _ = composite_tile(P, ids, X, Y)
        ys.append(Y)
        xs.append(X)
        vals.append(val)
    img = torch.zeros(height, width, 4, dtype=F64)
    if vals:
        img = img.index_put((torch.cat(ys), torch.cat(xs)), torch.cat(vals))
    return img


def decisions(P: dict, lists: list, width: int, height: int, tiles_x: int, tiles: list) -> list:
    """The discrete blend / cut-off masks of the given tiles (for detecting a finite difference
    that crosses a discontinuity)."""
    ly, lx = torch.meshgrid(torch.arange(16), torch.arange(16), indexing="ij")
    out = []
    for t in tiles:
        ids = lists[t]
        tx, ty = t % tiles_x, t // tiles_x
        X, Y = (tx * 16 + lx).reshape(-1), (ty * 16 + ly).reshape(-1)
        keep = (X < width) & (Y < height)
        if len(ids) == 0:
            out.append(None)
            continue
        _, blend, cut = composite_tile(P, ids, X[keep], Y[keep])
        out.append((blend.clone(), cut.clone()))
    return out


PARAMS = ("means", "log_scales", "quats", "logits", "sh")


def params64(means, log_scales, quats, logits, sh) -> dict:
    """float64 leaf tensors of the five parameter groups (sh as (N, stride))."""
    n = np.asarray(means).reshape(-1, 3).shape[0]
    t = lambda a, shape: torch.tensor(np.asarray(a, np.float64).reshape(shape), dtype=F64)  # noqa: E731
    return {"means": t(means, (n, 3)), "log_scales": t(log_scales, (n, 3)), "quats": t(quats, (n, 4)),
            "logits": t(logits, (n,)), "sh": t(sh, (n, -1))}


def forward(p: dict, cam: dict, lists: list, width: int, height: int, tiles=None):
    P = project(p["means"], p["log_scales"], p["quats"], p["sh"], cam, logits=p["logits"])
    return render(P, lists, width, height, cam["tiles_x"], tiles), P


def vjp(p: dict, cam: dict, lists: list, width: int, height: int, G) -> dict:
    """d(sum(G * rgba))/d(parameters) by torch autograd in float64 (G: (H, W, 4))."""
    q = {k: v.clone().requires_grad_(True) for k, v in p.items()}
    img, _ = forward(q, cam, lists, width, height)
    (img * torch.as_tensor(G, dtype=F64)).sum().backward()
    return {k: v.grad.detach().clone() for k, v in q.items()}


def tiles_of(lists: list, splat: int) -> list:
    return [t for t, ids in enumerate(lists) if len(ids) and np.any(np.asarray(ids) == splat)]


def central_fd(p: dict, cam: dict, lists: list, width: int, height: int, G, name: str, index: tuple,
               h: float) -> tuple:
    """Central difference of sum(G * rgba) along one parameter, evaluated on the tiles whose lists
    hold that splat.  Returns (derivative, crosses) where crosses is True when a blend / cut-off
    decision differs between the two evaluations and the unperturbed one (the difference then
    straddles a discontinuity and is not a derivative)."""
    splat = index[0]
    tiles = tiles_of(lists, splat)
    if not tiles:
        return 0.0, False
    Gt = torch.as_tensor(G, dtype=F64)
    vals, decs = [], []
    with torch.no_grad():
        for sgn in (0.0, 1.0, -1.0):
            q = {k: v.clone() for k, v in p.items()}
            q[name][index] += sgn * h
            P = project(q["means"], q["log_scales"], q["quats"], q["sh"], cam, logits=q["logits"])
            img = render(P, lists, width, height, cam["tiles_x"], tiles)
            vals.append(float((img * Gt).sum()))
            decs.append(decisions(P, lists, width, height, cam["tiles_x"], tiles))
    crosses = False
    for a, b, c in zip(decs[0], decs[1], decs[2]):
        if a is None:
            continue
        if not (torch.equal(a[0], b[0]) and torch.equal(a[0], c[0]) and torch.equal(a[1], b[1])
                and torch.equal(a[1], c[1])):
            crosses = True
    return (vals[1] - vals[2]) / (2.0 * h), crosses
