"""The request loop (s2dw_hip_server_*): analyses and syntheses of mixed
shapes submitted together, some synthesising from coefficients an earlier
analysis left resident, give the same results as each request alone on a
device plan of the same form, bit for bit (tolerance 0: every request runs the same kernels on
buffers of its own).  A small HBM budget forces queueing and plan eviction
without changing any result."""
import sys

import numpy as np
import pytest

from conftest import seeded_flm
from s2dw_amd import core
from s2dw_amd.serve import Server

pytestmark = pytest.mark.gpu

# (B, N, alpha, J): J None = J_max
SHAPES = [(32, 3, 2.0, None), (64, 5, 2.0, None), (27, 4, 3.0, None), (128, 8, 2.0, 5), (64, 1, 2.0, None),
          (100, 3, 2.0, None), (256, 5, 2.0, None)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(B, N, alpha, J):
    if J is None:
        J = core.s2dw_core_comp_Jmax(B, alpha)
    K, Phi2, bl = core.s2dw_core_init_kernels(J, B, alpha)
    S = core.s2dw_core_init_directionality(B, N)
    return J, K, Phi2, S, bl


def alone(B, N, alpha, J, flm):
    """The same request alone: one device plan of the server's form (the
    default, streamed), analysis then synthesis of its output."""
    import torch

    J, K, Phi2, S, bl = setup(B, N, alpha, J)
    plan = core.Plan(B, N, J, alpha, K, Phi2, S)
    d_f = torch.from_numpy(np.ascontiguousarray(flm.T)).cuda()
    d_w, d_sc, d_g = plan.empty_wav(), plan.empty_scoeff(), plan.empty_flm()
    plan.analysis(d_f, d_w, d_sc)
    plan.synthesis(d_w, d_sc, d_g)
    plan.sync()
    w = [np.asfortranarray(np.transpose(plan.scale_view(d_w, j).cpu().numpy(), (2, 1, 0))) for j in range(J + 1)]
    sc = np.asfortranarray(d_sc.cpu().numpy().T)
    g = np.asfortranarray(d_g.cpu().numpy().T)
    plan.close()
    return J, w, sc, g


def run_mix(srv, cases):
    """Submit every analysis, then syntheses (alternately from the resident
    result and from host copies) while the analyses may still run."""
    ta = [srv.submit_analysis(B, N, J, alpha, f, keep_resident=True) for (B, N, alpha, J, f) in cases]
    ts = []
    for i, (B, N, alpha, J, f) in enumerate(cases):
        if i % 2 == 0:
            ts.append(srv.submit_synthesis(B, N, J, alpha, wav_from=ta[i]))
        else:
            w, sc = srv.result(ta[i])
            ts.append(srv.submit_synthesis(B, N, J, alpha, wavs=w, scoeff=sc))
    out = []
    for i in range(len(cases)):
        w, sc = srv.result(ta[i])
        g = srv.result(ts[i])
        out.append((w, sc, g))
    for t in ta:
        srv.release(t)
    return out, ta, ts


@pytest.mark.parametrize("budget_gb,inflight", [(64, 8), (1, 2)])
def test_mix_equals_alone(budget_gb, inflight):
    say(f"server mix: budget {budget_gb} GB, {inflight} in flight")
    cases, ref = [], []
    for k, (B, N, alpha, J) in enumerate(SHAPES):
        f = seeded_flm(B, 500 + k, red=True)
        Jr, w, sc, g = alone(B, N, alpha, J, f)
        cases.append((B, N, alpha, Jr, f))
        ref.append((w, sc, g))
    with Server(budget_gb << 30, inflight) as srv:
        out, ta, ts = run_mix(srv, cases)
        st = srv.stats()
        say("  stats", st, "latencies (ms)", [round(t.latency_ms, 2) for t in ta + ts])
    for (B, N, alpha, J, f), (w0, sc0, g0), (w1, sc1, g1) in zip(cases, ref, out):
        for a, b in zip(w0, w1):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(sc0, sc1)
        np.testing.assert_array_equal(g0, g1)
        assert np.abs(g1 - f).max() <= 1e-12


def test_bad_requests_fail_cleanly():
    say("server error paths")
    with Server(8 << 30, 4) as srv:
        with pytest.raises(core.S2dwError):
            srv.submit_synthesis(32, 3, 4, 2.0, wav_from=type("T", (), {"id": 12345})())
        with pytest.raises(core.S2dwError):
            srv.submit_analysis(32, 40, 4, 2.0, seeded_flm(32, 1))  # N > B
        f = seeded_flm(32, 2)
        J = core.s2dw_core_comp_Jmax(32, 2.0)
        ta = srv.submit_analysis(32, 3, J, 2.0, f, keep_resident=True)
        ts = srv.submit_synthesis(32, 3, J, 2.0, wav_from=ta)
        g = srv.result(ts)
        assert np.abs(g - f).max() <= 1e-12
        srv.release(ta)
        with pytest.raises(core.S2dwError):
            srv.release(ta)  # already released
        # a budget too small for any plan fails the request instead of hanging
    with Server(1 << 16, 1) as tiny:
        t = tiny.submit_analysis(64, 3, core.s2dw_core_comp_Jmax(64, 2.0), 2.0, seeded_flm(64, 3))
        with pytest.raises(core.S2dwError) as e:
            tiny.wait(t)
        assert e.value.code == core.S2DW_ERROR_MEM_ALLOC_FAIL
