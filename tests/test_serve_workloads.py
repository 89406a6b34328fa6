"""The batch (64 mixed transforms) and serving (Poisson stream) workloads
through the request loop, on the GPU: every result bit for bit against the
same request alone on a device plan of the server's precision, and a subset
against the FP64 oracle.

Batch: 32 analyses and 32 syntheses of their resident outputs, L from
{64, 128, 256, 512}, N from {1, 3, 5, 8}, alpha = 2, J = J_max, seeded
real-signal f_lm, split-precision contraction, all submitted at once; oracle
bound 1e-10 (split form).  Stream: Poisson arrivals, B log-uniform in
[32, 1024], N uniform in 1..8, alpha in {2, 3}, FP64; oracle bound 1e-12
(red spectrum)."""
import sys
import time

import numpy as np
import pytest

import oracle as O
from conftest import seeded_flm
from s2dw_amd import core
from s2dw_amd.serve import Server

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def alone(B, N, alpha, J, flm, precision):
    import torch

    K, Phi2, bl = core.s2dw_core_init_kernels(J, B, alpha)
    S = core.s2dw_core_init_directionality(B, N)
    plan = core.Plan(B, N, J, alpha, K, Phi2, S, precision=precision)
    d_f = torch.from_numpy(np.ascontiguousarray(flm.T)).cuda()
    d_w, d_sc, d_g = plan.empty_wav(), plan.empty_scoeff(), plan.empty_flm()
    plan.analysis(d_f, d_w, d_sc)
    plan.synthesis(d_w, d_sc, d_g)
    plan.sync()
    w = [np.asfortranarray(np.transpose(plan.scale_view(d_w, j).cpu().numpy(), (2, 1, 0))) for j in range(J + 1)]
    g = np.asfortranarray(d_g.cpu().numpy().T)
    plan.close()
    return w, g


def check_oracle(B, N, alpha, J, f, w, g, tol):
    st = O.Setup(B, N, alpha, J=J)
    w_ref, _ = st.analysis(f)
    e_w = max(float(np.abs(a - b).max()) for a, b in zip(w, w_ref))
    e_g = float(np.abs(g - f).max())
    say(f"  oracle B={B} N={N} alpha={alpha}: wav {e_w:.1e}, round trip {e_g:.1e}")
    assert e_w <= tol and e_g <= tol


def test_batch_64_mixed_split():
    rng = np.random.default_rng(4)
    reqs = []
    for _ in range(32):
        B = int(rng.choice([64, 128, 256, 512]))
        N = int(rng.choice([1, 3, 5, 8]))
        J = core.s2dw_core_comp_Jmax(B, 2.0)
        reqs.append((B, N, 2.0, J, seeded_flm(B, int(rng.integers(1 << 30)))))
    say("batch: 32 analyses + 32 syntheses, split precision, all submitted at once")
    with Server(64 << 30, 16, core.PREC_SPLIT) as srv:
        t0 = time.perf_counter()
        ta = [srv.submit_analysis(B, N, J, a, f, keep_resident=True) for (B, N, a, J, f) in reqs]
        ts = [srv.submit_synthesis(B, N, J, a, wav_from=t) for (B, N, a, J, f), t in zip(reqs, ta)]
        res = [(srv.result(a)[0], srv.result(s)) for a, s in zip(ta, ts)]
        say(f"  64 transforms in {1e3 * (time.perf_counter() - t0):.0f} ms (plan builds included)")
        for t in ta:
            srv.release(t)
    checked = 0
    for (B, N, a, J, f), (w, g) in zip(reqs, res):
        w0, g0 = alone(B, N, a, J, f, core.PREC_SPLIT)
        for x, y in zip(w, w0):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(g, g0)
        if B <= 128 and checked < 4:
            check_oracle(B, N, a, J, f, w, g, 1e-10)
            checked += 1
    assert checked >= 1


def test_serving_stream_fp64():
    rng = np.random.default_rng(5)
    n, rate = 40, 200.0
    reqs = []
    for _ in range(n):
        B = int(round(float(np.exp(rng.uniform(np.log(32), np.log(1024))))))
        N = min(int(rng.integers(1, 9)), B)
        a = float(rng.choice([2.0, 3.0]))
        J = core.s2dw_core_comp_Jmax(B, a)
        reqs.append((B, N, a, J, seeded_flm(B, int(rng.integers(1 << 30)), red=True)))
    gaps = rng.exponential(1.0 / rate, n)
    say(f"serving: {n} requests, Poisson arrivals at {rate}/s, B log-uniform in [32, 1024]")
    with Server(64 << 30, 16) as srv:
        tickets = []
        t_next = time.perf_counter()
        for (B, N, a, J, f), gap in zip(reqs, gaps):
            t_next += gap
            while time.perf_counter() < t_next:
                time.sleep(1e-4)
            ta = srv.submit_analysis(B, N, J, a, f, keep_resident=True)
            tickets.append((ta, srv.submit_synthesis(B, N, J, a, wav_from=ta)))
        res = [(srv.result(ta)[0], srv.result(ts)) for ta, ts in tickets]
        lat = [t.latency_ms for pair in tickets for t in pair]
        say(f"  latency p50 {np.percentile(lat, 50):.2f} ms, p99 {np.percentile(lat, 99):.2f} ms")
        for ta, _ in tickets:
            srv.release(ta)
    checked = 0
    for (B, N, a, J, f), (w, g) in zip(reqs, res):
        w0, g0 = alone(B, N, a, J, f, core.PREC_FP64)
        for x, y in zip(w, w0):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(g, g0)
        if B <= 128 and checked < 4:
            check_oracle(B, N, a, J, f, w, g, 1e-12)
            checked += 1
    assert checked >= 1
