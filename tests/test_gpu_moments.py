"""GPU parity of the per-scale moments (s2dw_stat_moments,
src/mod/s2dw_stat_mod.f90:81-275) computed by libs2dw_hip.so.

- the host entry point against the reference routine's own outputs
  (tests/golden/ref_moments.npz, written by the reference's Fortran compiled
  from its sources), including the in-place masking, bit for bit;
- the device entry point on coefficients produced by a device analysis,
  against the oracle on the same coefficients, with and without a mask, up to
  the benchmark size L=512, N=8, J=7;
- the flags of the device entry point (zero variance, nj < 2) and run-to-run
  determinism.

Tolerances as in test_moments_oracle.py (1e-12 relative, scaled by the
condition number |mean| / standard deviation).  On coefficients from a device
analysis (up to 8.4e6 values per scale) the oracle's sums run in array order,
whose rounding error is bounded by n u (u = 2^-53) relative to the sum of
non-negative terms; there the GPU result is held to 1e-12 against the same
formulas evaluated with numpy's pairwise sums (error ~ u log2 n), and to the
oracle within max(1e-12, n u).
"""
import sys

import numpy as np
import pytest

import oracle as O
from conftest import seeded_flm
from s2dw_amd import core
from test_moments_oracle import CASES, assert_moments, fixture_case, split

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", CASES)
def test_host_moments_match_reference_routine(golden, name):
    say(f"moments vs reference routine: {name}")
    B, N, J, alpha, x, m, want, after = fixture_case(golden, name)
    wavs = split(x, B, N, alpha, J)
    masks = split(m, B, N, alpha, J) if m is not None else None
    if name == "const":
        with pytest.warns(RuntimeWarning, match="variance zero"):
            got = core.s2dw_stat_moments(wavs, J, B, N, alpha, masks)
    else:
        got = core.s2dw_stat_moments(wavs, J, B, N, alpha, masks)
    assert_moments(got, want, name)
    flat = np.concatenate([w.ravel(order="F") for w in wavs])
    assert np.array_equal(flat, after if after is not None else x)


def test_host_moments_halt_on_too_few_samples():
    B, N, alpha = 16, 2, 2.0
    J = O.comp_Jmax(B, alpha)
    his, _ = O.band_limits(J, B, alpha)
    rng = np.random.default_rng(3)
    wavs = [np.asfortranarray(rng.standard_normal((2 * h - 1, 2 * h, N))) for h in his]
    masks = [np.asfortranarray(np.ones_like(w)) for w in wavs]
    masks[2][:] = 0.0
    masks[2].flat[5] = 1.0  # nj = 1
    with pytest.raises(core.S2dwError) as e:
        core.s2dw_stat_moments(wavs, J, B, N, alpha, masks)
    assert e.value.code == core.S2DW_ERROR_SIZE_INVALID


def device_analysis(B, N, alpha, J, seed):
    import torch

    st = O.Setup(B, N, alpha, J=J)
    f = seeded_flm(B, seed, red=True)
    plan = core.Plan(B, N, st.J, alpha, st.K, st.Phi2, st.S)
    d_f = torch.from_numpy(np.ascontiguousarray(f.T)).cuda()
    d_w, d_sc = plan.empty_wav(), plan.empty_scoeff()
    plan.analysis(d_f, d_w, d_sc)
    plan.sync()
    return st, plan, d_w


def pairwise_moments(wavs, masks=None):
    """The finishing formulas of s2dw_stat_mod.f90:143-268 with numpy's
    pairwise sums (the test's accurate reference at large n)."""
    out = np.zeros((6, len(wavs)))
    for j, w in enumerate(wavs):
        x = w.ravel(order="F")
        if masks is None:
            v, nj = x, x.size
        else:
            m = masks[j].ravel(order="F")
            v, nj = (x * m)[np.abs(m - 1.0) < np.float64(np.float32(1e-12))], int(np.rint(m.sum()))
        mu = np.sum(v) / nj
        s = v - mu
        ep, s2 = np.sum(s), s * s
        T = {2: np.sum(s2), 3: np.sum(s2 * s), 4: np.sum(s2 * s2), 5: np.sum(s2 * s2 * s), 6: np.sum(s2 * s2 * s2)}
        var = (T[2] - ep * ep / nj) / (nj - 1)
        sd = np.sqrt(var)
        out[:, j] = [mu, var, T[3] / (nj * sd ** 3), T[4] / (nj * var ** 2) - 3, T[5] / (nj * sd ** 5),
                     T[6] / (nj * var ** 3) - 15]
    return out


def host_scales(plan, d_w):
    # Fortran (a, b, g) arrays of each scale
    return [np.asfortranarray(np.transpose(plan.scale_view(d_w, j).cpu().numpy(), (2, 1, 0)))
            for j in range(plan.J + 1)]


@pytest.mark.parametrize("B,N,alpha,J", [(64, 3, 2.0, None), (81, 5, 3.0, None), (128, 8, 2.0, 5),
                                         (512, 8, 2.0, 7)])
def test_device_moments_after_analysis(B, N, alpha, J):
    import torch

    say(f"device moments after analysis B={B} N={N} alpha={alpha} J={J}")
    st, plan, d_w = device_analysis(B, N, alpha, J, 40 + B)
    wavs = host_scales(plan, d_w)
    d_out = plan.stat_moments(d_w)
    plan.sync()
    out = d_out.cpu().numpy()
    nmax = max(w.size for w in wavs)
    rtol_oracle = max(1e-12, nmax * 2.0 ** -53)
    assert_moments(list(out[:6]), list(pairwise_moments(wavs)), "unmasked vs pairwise")
    code, *want = O.stat_moments([w.copy(order="F") for w in wavs], plan.J, B, N, alpha)
    assert code == 0
    assert_moments(list(out[:6]), want, "unmasked", rtol=rtol_oracle)
    assert np.all(out[7] == 0)
    np.testing.assert_array_equal(out[6], [w.size for w in wavs])
    # masked: about a quarter of the entries dropped, coefficients masked in place
    g = torch.Generator(device="cuda").manual_seed(B)
    d_mask = (torch.rand(d_w.shape, generator=g, device="cuda", dtype=torch.float64) < 0.75).double()
    d_w2 = d_w.clone()
    d_out2 = plan.stat_moments(d_w2, d_mask)
    plan.sync()
    masks = host_scales(plan, d_mask)
    out2 = d_out2.cpu().numpy()
    assert_moments(list(out2[:6]), list(pairwise_moments(wavs, masks)), "masked vs pairwise")
    code, *want2 = O.stat_moments(wavs, plan.J, B, N, alpha, masks)  # masks wavs in place
    assert code == 0
    assert_moments(list(out2[:6]), want2, "masked", rtol=rtol_oracle)
    np.testing.assert_array_equal(out2[6], [float(np.rint(m.sum())) for m in masks])
    for a, b in zip(host_scales(plan, d_w2), wavs):
        np.testing.assert_array_equal(a, b)
    plan.close()


def test_device_moments_flags_and_determinism():
    import torch

    B, N, alpha = 32, 2, 2.0
    say("device moments flags B=32 N=2")
    st, plan, d_w = device_analysis(B, N, alpha, None, 7)
    a = plan.stat_moments(d_w).clone()
    b = plan.stat_moments(d_w).clone()
    plan.sync()
    assert torch.equal(a, b)  # fixed merge order: identical results
    plan.scale_view(d_w, 0).fill_(0.5)  # zero variance on scale 0
    d_mask = torch.ones_like(d_w)
    plan.scale_view(d_mask, 1).zero_()  # nj = 0 on scale 1
    d_out = plan.stat_moments(d_w.clone())
    plan.sync()
    out = d_out.cpu().numpy()
    assert out[7, 0] == 1 and out[0, 0] == 0.5 and np.all(out[2:6, 0] == 0)
    d_out = plan.stat_moments(d_w.clone(), d_mask)
    plan.sync()
    out = d_out.cpu().numpy()
    assert out[7, 1] == 2 and out[6, 1] == 0 and np.all(np.isnan(out[:6, 1]))
    assert np.all(out[7, 2:] == 0)
    plan.close()


@pytest.mark.parametrize("B,N,alpha,J", [(64, 3, 2.0, None), (81, 5, 3.0, None), (27, 4, 3.0, None),
                                         (128, 8, 2.0, 5), (40, 5, 2.0, 3), (512, 8, 2.0, 7)])
def test_analysis_moments_without_storing(B, N, alpha, J):
    """The analysis that reduces to moments in its alpha stage against the
    moments of the stored coefficients (same transform, different chunking
    of the sums) and against pairwise sums over them."""
    import torch

    say(f"analysis reducing to moments B={B} N={N} alpha={alpha} J={J}")
    st, plan, d_w = device_analysis(B, N, alpha, J, 60 + B)
    assert plan.moments_fusable
    wavs = host_scales(plan, d_w)
    d_ref = plan.stat_moments(d_w)
    f = seeded_flm(B, 60 + B, red=True)
    d_f = torch.from_numpy(np.ascontiguousarray(f.T)).cuda()
    d_sc, d_sc2 = plan.empty_scoeff(), plan.empty_scoeff()
    plan.analysis(d_f, plan.empty_wav(), d_sc)
    d_out = plan.analysis_moments(d_f, d_sc2)
    plan.sync()
    ref, out = d_ref.cpu().numpy(), d_out.cpu().numpy()
    assert torch.equal(d_sc, d_sc2)
    np.testing.assert_array_equal(out[6:], ref[6:])  # nj and flags
    assert_moments(list(out[:6]), list(ref[:6]), "fused vs stored")
    assert_moments(list(out[:6]), list(pairwise_moments(wavs)), "fused vs pairwise")
    plan.close()


def test_analysis_moments_graph_replay():
    import torch

    B, N, alpha = 64, 5, 2.0
    say("analysis reducing to moments, graph replay B=64 N=5")
    st = O.Setup(B, N, alpha)
    plan = core.Plan(B, N, st.J, alpha, st.K, st.Phi2, st.S)
    d_f = torch.from_numpy(np.ascontiguousarray(seeded_flm(B, 5, red=True).T)).cuda()
    d_sc = plan.empty_scoeff()
    eager = plan.analysis_moments(d_f, d_sc).clone()
    plan.sync()
    plan.set_graphs(True)
    d_out = torch.empty_like(eager)
    for r in range(3):
        d_out.fill_(-1.0)
        plan.analysis_moments(d_f, d_sc, d_out)
        plan.sync()
        assert torch.equal(d_out, eager), r
        torch.cuda.synchronize()
        _ = (d_f * 2).sum().item()  # eager work between replays
    plan.close()


def test_host_moments_many_scales():
    """alpha = 1.1 at B = 64: 45 scales, more than one launch batch of the
    moments kernels (32 scales), with a mask."""
    B, N, alpha = 64, 2, 1.1
    J = O.comp_Jmax(B, alpha)
    assert J + 1 > 32
    say(f"moments with {J + 1} scales")
    his, _ = O.band_limits(J, B, alpha)
    rng = np.random.default_rng(45)
    wavs = [np.asfortranarray(rng.standard_normal((2 * h - 1, 2 * h, N)) * (1 + j)) for j, h in enumerate(his)]
    masks = [np.asfortranarray((rng.random(w.shape) < 0.8).astype(np.float64)) for w in wavs]
    got = core.s2dw_stat_moments([w.copy(order="F") for w in wavs], J, B, N, alpha)
    code, *want = O.stat_moments([w.copy(order="F") for w in wavs], J, B, N, alpha)
    assert code == 0
    assert_moments(got, want, "many scales")
    wg = [w.copy(order="F") for w in wavs]
    got = core.s2dw_stat_moments(wg, J, B, N, alpha, masks)
    wo = [w.copy(order="F") for w in wavs]
    code, *want = O.stat_moments(wo, J, B, N, alpha, masks)
    assert code == 0
    assert_moments(got, want, "many scales, masked")
    for a, b in zip(wg, wo):
        np.testing.assert_array_equal(a, b)


def test_analysis_moments_refused_when_not_fusable():
    """A plan whose FFTs run outside the grouped launches (here the wave FFT
    engine, chosen by S2DW_HIP_FFT at plan creation) cannot reduce the
    analysis to moments; the entry point says so instead of computing
    something else, and the unfused path still works on that plan."""
    import os

    import torch

    B, N, alpha = 64, 3, 2.0
    say("moments fusion refused on a wave-engine plan")
    st = O.Setup(B, N, alpha)
    old = os.environ.get("S2DW_HIP_FFT")
    os.environ["S2DW_HIP_FFT"] = "wave"
    try:
        plan = core.Plan(B, N, st.J, alpha, st.K, st.Phi2, st.S)
    finally:
        if old is None:
            del os.environ["S2DW_HIP_FFT"]
        else:
            os.environ["S2DW_HIP_FFT"] = old
    assert not plan.moments_fusable
    f = seeded_flm(B, 9, red=True)
    d_f = torch.from_numpy(np.ascontiguousarray(f.T)).cuda()
    d_sc = plan.empty_scoeff()
    with pytest.raises(core.S2dwError) as e:
        plan.analysis_moments(d_f, d_sc)
    assert e.value.code == core.S2DW_ERROR_ARG_INVALID
    d_w = plan.empty_wav()
    plan.analysis(d_f, d_w, d_sc)
    d_out = plan.stat_moments(d_w)
    plan.sync()
    code, *want = O.stat_moments(host_scales(plan, d_w), plan.J, B, N, alpha)
    assert code == 0
    assert_moments(list(d_out.cpu().numpy()[:6]), want, "wave engine")
    plan.close()
