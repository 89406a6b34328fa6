"""s2dw_hip_stat_moments_device_noalloc (moments without allocation: the
chunk records in the library's static device scratch) against the entry point
it varies, s2dw_hip_stat_moments_device (records in the plan), through the
C ABI: bit-identical rows on two calls in a row, the second smaller, with and
without a mask, and a clean refusal above the scratch's size."""
import sys

import numpy as np
import pytest

import oracle as O
from conftest import seeded_flm
from s2dw_amd import core

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def analysed(B, N, alpha, seed):
    import torch

    st = O.Setup(B, N, alpha)
    plan = core.Plan(B, N, st.J, alpha, st.K, st.Phi2, st.S)
    d_f = torch.from_numpy(np.ascontiguousarray(seeded_flm(B, seed, red=True).T)).cuda()
    d_w, d_sc = plan.empty_wav(), plan.empty_scoeff()
    plan.analysis(d_f, d_w, d_sc)
    plan.sync()
    return st, plan, d_w


def test_noalloc_matches_plan_moments_twice():
    import torch

    for B, N, alpha, seed in [(128, 5, 2.0, 11), (64, 3, 3.0, 12)]:  # the second call is smaller
        say(f"moments without allocation B={B} N={N} alpha={alpha}")
        st, plan, d_w = analysed(B, N, alpha, seed)
        ref = plan.stat_moments(d_w.clone())
        got = torch.full((8, st.J + 1), -7.0, dtype=torch.float64, device=d_w.device)
        core.stat_moments_device_noalloc(d_w.clone(), st.J, B, N, alpha, got)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.cpu().numpy(), ref.cpu().numpy())
        # masked: the same mask through both entry points (each masks its own copy)
        g = torch.Generator(device="cpu").manual_seed(seed)
        mask = (torch.rand(d_w.numel(), generator=g, dtype=torch.float64) < 0.7).to(torch.float64).cuda()
        w1, w2 = d_w.clone(), d_w.clone()
        ref_m = plan.stat_moments(w1, d_mask=mask)
        got_m = torch.full_like(got, -7.0)
        core.stat_moments_device_noalloc(w2, st.J, B, N, alpha, got_m, d_mask=mask)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got_m.cpu().numpy(), ref_m.cpu().numpy())
        np.testing.assert_array_equal(w2.cpu().numpy(), w1.cpu().numpy())
        plan.close()


def test_noalloc_refuses_beyond_scratch():
    import torch

    d = torch.zeros(16, dtype=torch.float64, device="cuda")
    out = torch.zeros((8, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(core.S2dwError) as e:
        core.stat_moments_device_noalloc(d, core.s2dw_core_comp_Jmax(2048, 2.0), 2048, 8, 2.0, out)
    assert e.value.code == core.S2DW_ERROR_SIZE_INVALID
