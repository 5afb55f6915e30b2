"""Seeded cases for the differentiable mode (the config-4 recipe at test sizes).

Scene: clustered means + sheets, SH3, as the 1M-splat scene; opacity logits = logit of the
scene's opacities.  Target: the same scene after the seeded perturbation of the differentiation
workload (means + N(0, 0.01^2), log-scales + N(0, 0.1^2), logits + N(0, 0.3^2), SH + N(0, 0.02^2)),
rendered by the CPU oracle in differentiable mode.  Loss: L2 per view, so dL/d(rgba) = 2 (rgba - target).
"""
from __future__ import annotations

import numpy as np

from hdsr import camera, compositor, scenes
from oracle import diff64
from oracle import oracle as orc


def logit(p):
    p = np.asarray(p, np.float64)
    return np.log(p / (1.0 - p)).astype(np.float32)


class Case:
    def __init__(self, n=3000, width=128, height=96, seed=11, f=260.0, radius=2.4, azimuth=0.4, elevation=0.2):
        sc = scenes.scene_config1(n, seed=seed, sh_degree=3)
        self.scene = sc
        self.n, self.width, self.height = n, width, height
        self.means = np.ascontiguousarray(sc.positions, np.float32)
        self.log_scales = np.ascontiguousarray(sc.scales, np.float32)
        self.quats = np.ascontiguousarray(sc.rotations, np.float32)
        self.logits = logit(sc.opacities)
        self.sh = np.ascontiguousarray(sc.sh_data(), np.float32).reshape(n, -1)
        rng = np.random.default_rng(seed + 1)
        self.target_params = (
            (self.means + rng.normal(0, 0.01, self.means.shape)).astype(np.float32),
            (self.log_scales + rng.normal(0, 0.1, self.log_scales.shape)).astype(np.float32),
            self.quats.copy(),
            (self.logits + rng.normal(0, 0.3, self.logits.shape)).astype(np.float32),
            self._perturb_sh(rng),
        )
        self.cam = camera.Camera(camera.orbit_eye(radius, azimuth, elevation), (0, 0, 0), f, f, width, height)
        self.plan = compositor.planTileSplatCompositor({"viewportWidth": width, "viewportHeight": height,
                                                        "tileSizePx": 16, "splatCount": n})
        self.uniforms = camera.frame_uniforms_for(self.cam, self.plan, 3)
        self.identity = np.arange(n, dtype=np.uint32)
        self.cam64 = diff64.parse_uniforms(self.uniforms)

    def _perturb_sh(self, rng):
        sh = self.sh.copy()
        sh[:, :-3] += rng.normal(0, 0.02, sh[:, :-3].shape).astype(np.float32)  # emissive tail unchanged
        return sh

    def params(self):
        return (self.means, self.log_scales, self.quats, self.logits, self.sh)

    def arrays(self, params):
        means, log_scales, quats, logits, sh = params
        return {"positions": means, "scales": log_scales, "rotations": quats, "opacities": logits,
                "materials": self.scene.materials, "normals": None, "sh_data": np.ascontiguousarray(sh).reshape(-1),
                "sorted": self.identity, "sh_degree": 3}

    def oracle(self, params=None, flags=0):
        """The CPU oracle in differentiable mode (logit opacities, fp32 rasterizer inputs)."""
        return orc.OracleFrame(self.arrays(params or self.params()), self.uniforms, self.width, self.height,
                               orc.FLAG_DIFF | flags)

    def target(self):
        return self.oracle(self.target_params)["rgba"].reshape(self.height, self.width, 4)

    def lists(self, entries, tile_offsets, tile_counts):
        return diff64.tile_lists(entries, tile_offsets, tile_counts, self.identity, self.plan["tileColumns"],
                                 self.plan["tileRows"])

    def params64(self):
        return diff64.params64(*self.params())


# sampled parameters per group for the finite-difference checks (256 in all)
FD_SAMPLES = {"means": 64, "log_scales": 48, "quats": 48, "logits": 32, "sh": 64}
FD_STEP = 1e-6


def sample_fd(case: Case, p64: dict, lists: list, G, g64: dict, counts=None, seed=0, max_tries=4):
    """Central differences (float64) of sum(G * rgba) at seeded parameters with a non-negligible
    autodiff gradient (|g| > 1e-4 of the group's largest).  A sample whose difference crosses a
    blend / cut-off decision is replaced by another draw (up to max_tries per sample).
    Returns {group: [(index, fd value), ...]} and the number of replaced samples."""
    rng = np.random.default_rng(seed)
    out, replaced = {}, 0
    for name, k in (counts or FD_SAMPLES).items():
        g = g64[name]
        cand = np.argwhere(np.abs(g.numpy()) > 1e-4 * float(np.abs(g.numpy()).max()))
        if name == "sh":
            cand = cand[cand[:, 1] < p64["sh"].shape[1] - 3]
        picks = []
        for _ in range(k):
            for _t in range(max_tries):
                idx = tuple(int(v) for v in cand[rng.integers(len(cand))])
                fd, crosses = diff64.central_fd(p64, case.cam64, lists, case.width, case.height, G, name, idx,
                                                FD_STEP)
                if not crosses:
                    picks.append((idx, fd))
                    break
                replaced += 1
        out[name] = picks
    return out, replaced
