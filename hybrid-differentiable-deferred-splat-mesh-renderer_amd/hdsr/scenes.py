"""Seeded synthetic scenes and cameras for the benchmark configurations, and the input packers.

Scene attributes follow the reference's SoA buffer layouts (src/splats.ts:376-500): positions 3N,
log-scales 3N, rotations 4N (w, x, y, z), activated opacities N, material words N
(pack2x16float(roughness, metalness) with the truncating f32ToF16 of src/splats.ts:431-442),
and shData N * (3 + 3K + 3) = [dc rgb | K SH coefficients, rgb-interleaved | emissive rgb].

Scenes (all fp32, numpy default_rng(seed)):
  scene_config0  50k Gaussians, means N(0, 0.5^2 I), log-scales N(-4, 0.5^2), uniform unit
                 quaternions (Shoemake), opacity sigmoid(N(0, 1.5^2)), SH0 colour U[0,1].
  scene_config1  1M Gaussians, SH3 (DC U[0,1], 45 higher coefficients N(0, 0.05^2)); means from 64
                 anisotropic Gaussian clusters plus 6 planar sheets (normal sigma 0.002).
  scene_config2  6M Gaussians as config 1, half inside a 0.2-unit ball in front of camera 0.
  scene_config4  500k Gaussians as config 1 (the differentiation workload).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .camera import Camera, orbit_eye

SHEET_FRACTION = 0.25


@dataclass
class Scene:
    positions: np.ndarray   # (N, 3) f32
    scales: np.ndarray      # (N, 3) f32, log-scale
    rotations: np.ndarray   # (N, 4) f32, w x y z
    opacities: np.ndarray   # (N,) f32 activated
    colors: np.ndarray      # (N, 3) f32 dc
    sh: np.ndarray | None   # (N, K, 3) f32 or None
    sh_degree: int
    materials: np.ndarray = field(default=None)   # (N,) u32
    normals: np.ndarray | None = None              # flat f32 [geo 3N | detail 3N] or None
    emissive: np.ndarray | None = None             # (N, 3) f32 or None
    info: dict = field(default_factory=dict)

    @property
    def count(self) -> int:
        return int(self.positions.shape[0])

    def sh_data(self) -> np.ndarray:
        return pack_sh_data(self.colors, self.sh, self.emissive)

    def subset(self, idx) -> "Scene":
        idx = np.asarray(idx)
        return Scene(self.positions[idx].copy(), self.scales[idx].copy(), self.rotations[idx].copy(),
                     self.opacities[idx].copy(), self.colors[idx].copy(),
                     None if self.sh is None else self.sh[idx].copy(), self.sh_degree,
                     self.materials[idx].copy(), None,
                     None if self.emissive is None else self.emissive[idx].copy(), dict(self.info))


def f32_to_f16_trunc(x: np.ndarray) -> np.ndarray:
    """The host material path's f32ToF16: truncating, flushes values below 2^-14 (splats.ts:431-442)."""
    bits = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.int64)
    sign = (bits >> 16) & 0x8000
    exp = ((bits >> 23) & 0xFF) - 127 + 15
    frac = bits & 0x7FFFFF
    out = sign | (exp << 10) | (frac >> 13)
    out = np.where(exp <= 0, sign, out)
    out = np.where(exp >= 31, sign | 0x7C00, out)
    return out.astype(np.uint32)


def pack_material(count: int, roughness=None, metalness=None) -> np.ndarray:
    """packMaterialBuffer (splats.ts:456-466); defaults roughness 0.75, metalness 0."""
    r = np.full(count, 0.75, np.float32) if roughness is None else np.asarray(roughness, np.float32)
    m = np.zeros(count, np.float32) if metalness is None else np.asarray(metalness, np.float32)
    return (f32_to_f16_trunc(r) | (f32_to_f16_trunc(m) << 16)).astype(np.uint32)


def pack_sh_data(colors, sh=None, emissive=None) -> np.ndarray:
    """packShDataBuffer (splats.ts:471-500)."""
    n = colors.shape[0]
    k = 0 if sh is None else sh.shape[1]
    out = np.zeros((n, 3 + 3 * k + 3), np.float32)
    out[:, 0:3] = colors
    if sh is not None:
        out[:, 3:3 + 3 * k] = sh.reshape(n, 3 * k)
    if emissive is not None:
        out[:, 3 + 3 * k:] = emissive
    return out.reshape(-1)


def sh_from_ply_f_rest(f_rest: np.ndarray) -> np.ndarray:
    """3DGS PLY f_rest_* (channel-major: all R coefficients, then G, then B) -> (N, K, 3)
    coefficient-major, RGB-interleaved (the splat_coeff_rgb layout of the SH sidecar)."""
    f = np.asarray(f_rest, np.float32)
    n, m = f.shape
    k = m // 3
    return np.ascontiguousarray(f.reshape(n, 3, k).transpose(0, 2, 1))


def sort_back_to_front(positions: np.ndarray, view16: np.ndarray) -> np.ndarray:
    """cpuDepthSort = sortSplatIdsBackToFront (src/splatSort.ts:32-55): ascending view-space z
    (farthest first), ties by id; depth evaluated in float64 as the JS does."""
    p = np.asarray(positions, dtype=np.float32).astype(np.float64)
    v = np.asarray(view16, dtype=np.float32).astype(np.float64)
    depth = ((v[2] * p[:, 0] + v[6] * p[:, 1]) + v[10] * p[:, 2]) + v[14]
    if not np.all(np.isfinite(depth)):
        raise ValueError("non-finite view-space depth")
    return np.argsort(depth, kind="stable").astype(np.uint32)


def _quaternions(rng, n):
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    x = np.sqrt(1 - u1) * np.sin(2 * np.pi * u2)
    y = np.sqrt(1 - u1) * np.cos(2 * np.pi * u2)
    z = np.sqrt(u1) * np.sin(2 * np.pi * u3)
    w = np.sqrt(u1) * np.cos(2 * np.pi * u3)
    return np.stack([w, x, y, z], axis=1).astype(np.float32)


def _common(rng, n, sh_degree):
    scales = rng.normal(-4.0, 0.5, (n, 3)).astype(np.float32)
    rot = _quaternions(rng, n)
    opac = (1.0 / (1.0 + np.exp(-rng.normal(0.0, 1.5, n)))).astype(np.float32)
    colors = rng.random((n, 3)).astype(np.float32)
    sh = None
    if sh_degree > 0:
        k = (sh_degree + 1) ** 2 - 1
        sh = rng.normal(0.0, 0.05, (n, k, 3)).astype(np.float32)
    return scales, rot, opac, colors, sh


def scene_config0(n=50_000, seed=0) -> Scene:
    rng = np.random.default_rng(seed)
    pos = rng.normal(0.0, 0.5, (n, 3)).astype(np.float32)
    scales, rot, opac, colors, sh = _common(rng, n, 0)
    return Scene(pos, scales, rot, opac, colors, sh, 0, pack_material(n), info={"config": 0, "seed": seed})


def _random_rotation(rng):
    q = _quaternions(rng, 1)[0].astype(np.float64)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def clustered_means(rng, n, clusters=64, sheets=6, sheet_fraction=SHEET_FRACTION, extent=0.6) -> np.ndarray:
    """Means from `clusters` anisotropic Gaussian clusters plus `sheets` planar sheets
    (normal sigma 0.002)."""
    n_sheet = int(round(n * sheet_fraction))
    n_cl = n - n_sheet
    out = np.empty((n, 3), np.float64)
    centers = rng.normal(0.0, extent, (clusters, 3))
    sig = np.exp(rng.normal(math.log(0.08), 0.5, (clusters, 3)))
    rots = [_random_rotation(rng) for _ in range(clusters)]
    lab = rng.integers(0, clusters, n_cl)
    local = rng.normal(0.0, 1.0, (n_cl, 3))
    for c in range(clusters):
        sel = lab == c
        out[:n_cl][sel] = centers[c] + (local[sel] * sig[c]) @ rots[c].T
    if n_sheet:
        slab = rng.integers(0, sheets, n_sheet)
        pts = np.empty((n_sheet, 3))
        for s in range(sheets):
            sel = slab == s
            k = int(sel.sum())
            R = _random_rotation(rng)
            o = rng.normal(0.0, extent * 0.5, 3)
            uv = rng.uniform(-0.7, 0.7, (k, 2))
            h = rng.normal(0.0, 0.002, k)
            pts[sel] = o + uv[:, 0:1] * R[:, 0] + uv[:, 1:2] * R[:, 1] + h[:, None] * R[:, 2]
        out[n_cl:] = pts
    perm = rng.permutation(n)
    return out[perm].astype(np.float32)


def scene_config1(n=1_000_000, seed=1, sh_degree=3) -> Scene:
    rng = np.random.default_rng(seed)
    pos = clustered_means(rng, n)
    scales, rot, opac, colors, sh = _common(rng, n, sh_degree)
    return Scene(pos, scales, rot, opac, colors, sh, sh_degree, pack_material(n),
                 info={"config": 1, "seed": seed, "sheet_fraction": SHEET_FRACTION, "clusters": 64, "sheets": 6})


def cameras_config0() -> list[Camera]:
    return [Camera(orbit_eye(3.0, 0.0, 0.0), (0, 0, 0), 525.0, 525.0, 640, 480)]


def cameras_orbit(count=8, width=1920, height=1080, f=1500.0, radius=3.0, elevation=0.25) -> list[Camera]:
    return [Camera(orbit_eye(radius, k * 2 * math.pi / count, elevation), (0, 0, 0), f, f, width, height)
            for k in range(count)]


def cameras_config1() -> list[Camera]:
    return cameras_orbit(8, 1920, 1080, 1500.0)


def cameras_config2() -> list[Camera]:
    return cameras_orbit(16, 3840, 2160, 1500.0)


def scene_config2(n=6_000_000, seed=2) -> Scene:
    rng = np.random.default_rng(seed)
    half = n // 2
    pos = np.empty((n, 3), np.float32)
    pos[:n - half] = clustered_means(rng, n - half)
    cam0 = cameras_config2()[0]
    fwd = (cam0.target - cam0.eye) / np.linalg.norm(cam0.target - cam0.eye)
    ball_c = cam0.eye + 1.0 * fwd
    d = rng.normal(0.0, 1.0, (half, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rad = 0.2 * rng.random(half) ** (1.0 / 3.0)
    pos[n - half:] = (ball_c + d * rad[:, None]).astype(np.float32)
    scales, rot, opac, colors, sh = _common(rng, n, 3)
    return Scene(pos, scales, rot, opac, colors, sh, 3, pack_material(n),
                 info={"config": 2, "seed": seed, "ball_center": ball_c.tolist()})


def scene_config4(n=500_000, seed=4) -> Scene:
    s = scene_config1(n, seed)
    s.info["config"] = 4
    return s


def cameras_config4() -> list[Camera]:
    return cameras_orbit(16, 1280, 720, 1500.0 * 720 / 1080)
