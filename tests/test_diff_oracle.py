"""The FP64 autodiff oracle of the differentiable mode, pinned on the CPU.

* Its forward (torch float64) equals the C++ oracle's independent FP64 twin of the
  differentiable forward (loops over tiles / pixels / entries, oracle/hdsr_oracle.cpp composite64
  without f16 rounding) to 1e-12: same projection, same decisions, same compositing.
* Its gradients (torch autograd) equal float64 central differences at seeded parameters of every
  group to 1e-5 relative, on the tiles each parameter reaches; a difference that crosses a
  blend / cut-off decision is detected (decisions compared bit for bit) and redrawn.
* The fp32 differentiable-mode oracle (what the HIP path must match) is within 1e-3 of FP64.
"""
import numpy as np
import pytest
import torch

from diff_cases import Case, sample_fd
from oracle import diff64
from oracle import oracle as orc


@pytest.fixture(scope="module")
def case():
    return Case(n=1500, width=96, height=64, seed=5)


@pytest.fixture(scope="module")
def frame(case):
    return case.oracle(flags=orc.FLAG_FP64_NOQUANT)


def test_fp64_forward_equals_cpp_twin(case, frame):
    lists = case.lists(frame["entries"], frame["tile_offsets"], frame["tile_counts"])
    img, _ = diff64.forward(case.params64(), case.cam64, lists, case.width, case.height)
    ref = frame["rgba64"].reshape(case.height, case.width, 4)
    assert np.abs(img.numpy() - ref).max() < 1e-12


def test_fp32_oracle_close_to_fp64(case, frame):
    err = np.abs(frame["rgba"].reshape(-1) - frame["rgba64"]).max()
    assert err < 1e-3, err


def test_autodiff_equals_central_differences(case, frame):
    lists = case.lists(frame["entries"], frame["tile_offsets"], frame["tile_counts"])
    p = case.params64()
    img, _ = diff64.forward(p, case.cam64, lists, case.width, case.height)
    G = 2.0 * (img.detach().numpy() - case.target())
    g = diff64.vjp(p, case.cam64, lists, case.width, case.height, G)
    picks, replaced = sample_fd(case, p, lists, G, g, counts={"means": 8, "log_scales": 6, "quats": 6, "logits": 6,
                                                              "sh": 8}, seed=3)
    for name, items in picks.items():
        assert len(items) >= 4, name
        for idx, fd in items:
            ad = float(g[name][idx])
            assert abs(ad - fd) <= 1e-5 * abs(fd) + 1e-9, (name, idx, ad, fd)
    assert replaced <= 8


def test_emissive_tail_has_no_gradient(case, frame):
    lists = case.lists(frame["entries"], frame["tile_offsets"], frame["tile_counts"])
    p = case.params64()
    G = np.ones((case.height, case.width, 4))
    g = diff64.vjp(p, case.cam64, lists, case.width, case.height, G)
    assert torch.count_nonzero(g["sh"][:, -3:]) == 0
    assert torch.count_nonzero(g["sh"][:, :3]) > 0
