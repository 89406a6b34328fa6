"""Benchmark of the S2DW directional wavelet transform on one MI355X.

Workload: L=512, N=8, alpha=2, J_min=2 (reference J = J_max - 2 = 7),
seeded real-signal f_lm ~ N(0,1) scaled by (l+1)^-1.  One step = one analysis
plus one synthesis (two transforms) on a device-resident f_lm; inputs are in
HBM before the timed region.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--precision fp64|split]
                    [--no-cpu-baseline]

The timed region runs the steps back to back (value, per-transform latency).
A second pass over the same steps records HIP events around every stage
launch on the plan's stream; the dominant stage's average launch time gives
the roofline figures.  HBM traffic per launch comes from rocprofv3 counter
runs, reduced by

    python bench.py --traffic-from FETCH_DIR WRITE_DIR

into profiles/traffic.json (FETCH_SIZE / WRITE_SIZE per dispatch, scaled by
the same counters read on a plain device copy of known size, which
``python bench.py --calib-copy`` runs).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "transforms/s at L=512,N=8 (split-prec vs FP64, err≤1e-10); p50/p99 latency"
L, N, ALPHA, J_MIN = 512, 8, 2.0, 2
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def seeded_flm(B, seed, red=True):
    rng = np.random.default_rng(seed)
    f = np.zeros((B, B), dtype=np.complex128, order="F")
    for l in range(B):
        f[l, 0] = rng.standard_normal()
        if l:
            f[l, 1:l + 1] = rng.standard_normal(l) + 1j * rng.standard_normal(l)
        if red:
            f[l, :] /= (l + 1)
    return f


def contraction_cmacs(hi, lo, N):
    """Useful complex MACs of the l-contraction, per transform:
    analysis sum_j sum_{l=lo_j+1}^{hi_j-1} n_a(l) (l+1)^2 with n_a(l) = #{n: |n| <= min(N-1,l), N+n odd};
    synthesis the same with n >= 0 only."""
    ana = syn = 0
    for h, lo_ in zip(hi, lo):
        for l in range(lo_ + 1, h):
            nmax = min(N - 1, l)
            na = sum(1 for n in range(-nmax, nmax + 1) if (N + n) % 2)
            ns = sum(1 for n in range(0, nmax + 1) if (N + n) % 2)
            ana += na * (l + 1) ** 2
            syn += ns * (l + 1) ** 2
    return ana, syn


def stage_bytes(hi, lo, N, ncs, grouped=True):
    """Compulsory HBM bytes per stage for one transform: reading each stage's
    input once and writing its output once.  With grouped launches the FFT
    stages 'ana_beta_fft' ... 'syn_beta_fft' cover scales 0 and 1 and the
    'small' stages cover scales 2..J; otherwise the FFT stages cover all."""
    def beta_a(h): return N * h * h * 16 + N * 2 * h * h * 16
    def alpha_a(h): return N * 2 * h * h * 16 + N * 2 * h * (2 * h - 1) * 8
    def alpha_s(h): return N * 2 * h * (2 * h - 1) * 8 + N * 2 * h * h * 16
    def beta_s(h): return N * 2 * h * h * 16 + ncs * h * h * 16
    sc = [j for j, (h, l_) in enumerate(zip(hi, lo)) if l_ + 1 <= h - 1]
    big = [hi[j] for j in sc if j < 2 or not grouped]
    small = [hi[j] for j in sc if j >= 2 and grouped]
    b = {"ana_beta_fft": sum(map(beta_a, big)), "ana_alpha_c2r": sum(map(alpha_a, big)),
         "syn_alpha_r2c": sum(map(alpha_s, big)), "syn_beta_fft": sum(map(beta_s, big))}
    if small:
        b["ana_small_ffts"] = sum(beta_a(h) + alpha_a(h) for h in small)
        b["syn_small_ffts"] = sum(alpha_s(h) + beta_s(h) for h in small)
    return b


def host_cpu():
    """nproc, the cores this process may run on, sockets and the CPU model."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    try:
        with open("/proc/cpuinfo") as fh:
            txt = fh.read()
        models = [ln.split(":", 1)[1].strip() for ln in txt.splitlines() if ln.startswith("model name")]
        phys = {ln.split(":", 1)[1].strip() for ln in txt.splitlines() if ln.startswith("physical id")}
        info["model"] = models[0] if models else "unknown"
        info["sockets"] = max(1, len(phys))
    except OSError:
        info["model"], info["sockets"] = "unknown", None
    return info


def cpu_baseline(K, Phi2, S, J, bl, flm, rounds=3):
    """The CPU oracle (a C/OpenMP restatement of the reference's algorithm,
    loop for loop) on the host's cores, as the reference's timing is taken
    (s2dw_test.f90:114-135: wall clock around each call, kernels built
    before): one warm-up round trip, then the median of `rounds` analyses and
    of `rounds` syntheses.  Threads: every core this process may use, capped
    by OMP_NUM_THREADS when that is set (the job's CPU share).  Returns the
    record and the oracle's coefficients (the FP64 reference for the
    per-scale errors)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    hw = host_cpu()
    cores = hw["affinity"]
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    O.set_threads(cores)
    w, sc = O.analysis_dynamic(flm, K, S, J, L, N, bl, ALPHA)  # warm-up
    O.synthesis_dynamic(w, sc, K, Phi2, S, J, L, N, bl, ALPHA)
    ta, ts = [], []
    for _ in range(rounds):
        t0 = time.perf_counter()
        w, sc = O.analysis_dynamic(flm, K, S, J, L, N, bl, ALPHA)
        t1 = time.perf_counter()
        g = O.synthesis_dynamic(w, sc, K, Phi2, S, J, L, N, bl, ALPHA)
        t2 = time.perf_counter()
        ta.append(t1 - t0)
        ts.append(t2 - t1)
        print(f"bench: cpu baseline round {len(ta)}: analysis {ta[-1]:.2f} s, synthesis {ts[-1]:.2f} s",
              file=sys.stderr, flush=True)
    err = float(np.abs(g - flm).max())
    a_med, s_med = statistics.median(ta), statistics.median(ts)
    rec = {"value": 2.0 / (a_med + s_med), "unit": "transforms/s", "cores": cores, "kind": "port",
           "sample": f"L=512 N=8 J=7 round trips by oracle/s2dw_oracle.c (the reference's loop structure, "
                     f"OpenMP, {cores} threads): one warm-up, then the median of {rounds} analyses "
                     f"({a_med:.2f} s) and of {rounds} syntheses ({s_med:.2f} s); round-trip err {err:.1e}",
           "analysis_s": a_med, "synthesis_s": s_med, "analysis_runs_s": ta, "synthesis_runs_s": ts,
           "host": hw, "planning_in_timed_region": "no FFT planning (the oracle's FFT has none)"}
    return rec, w


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "traffic.json")
CALIB_BYTES = 256 << 20  # bytes read and written by the calibration copy
CALIB_REPS = 4


def calib_copy():
    """A plain device-to-device copy of known size, for scaling the HBM
    byte counters (run under rocprofv3 --pmc)."""
    import torch

    n = CALIB_BYTES // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(CALIB_REPS):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calib_copy_bytes": CALIB_BYTES, "reps": CALIB_REPS}))


def _counter_rows(path):
    """All rows of every *counter_collection.csv under path (or of path)."""
    import csv

    files = [path]
    if os.path.isdir(path):
        files = [os.path.join(r, f) for r, _, fs in os.walk(path) for f in fs if f.endswith("counter_collection.csv")]
    rows = []
    for f in files:
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    return rows


def traffic_from(fetch_path, write_path):
    """Per-dispatch HBM bytes of every kernel from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE, in KiB), each scaled by its reading on the
    calibration copy (the largest copyBuffer dispatch of the run)."""
    out = {}
    for counter, path in (("FETCH_SIZE", fetch_path), ("WRITE_SIZE", write_path)):
        rows = [r for r in _counter_rows(path) if r["Counter_Name"] == counter]
        per = {}
        for r in rows:
            per.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]) * 1024.0)
        copies = [v for name, vals in per.items() if "copyBuffer" in name for v in vals]
        scale = CALIB_BYTES / max(copies) if copies else 1.0
        for name, vals in per.items():
            if "s2dw" not in name:
                continue
            key = name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").replace("s2dw::", "")
            d = out.setdefault(key, {})
            d[counter.lower() + "_bytes"] = scale * sum(vals) / len(vals)
            d[counter.lower() + "_scale"] = scale
            d["dispatches"] = len(vals)
    for d in out.values():
        d["hbm_bytes"] = d.get("fetch_size_bytes", 0.0) + d.get("write_size_bytes", 0.0)
    os.makedirs(os.path.dirname(TRAFFIC_JSON), exist_ok=True)
    with open(TRAFFIC_JSON, "w") as fh:
        json.dump({"source": "rocprofv3 --pmc FETCH_SIZE, then --pmc WRITE_SIZE, each over "
                             "'python3 bench.py --steps 2 --warmup 1 --no-cpu-baseline --no-autotune --skip-forms "
                             "--skip-serving --no-pipeline' and "
                             "'python3 bench.py --calib-copy'; per-dispatch KiB scaled by the counter's reading "
                             f"on the {CALIB_BYTES >> 20} MiB calibration copy", "kernels": out}, fh, indent=1)
    print(json.dumps(out, indent=1))


def _mix_requests(rng, count, Ls, Ns, alphas, log_uniform=False):
    """Seeded request shapes (B, N, alpha, J=J_max) and f_lm."""
    from s2dw_amd import core

    reqs = []
    for _ in range(count):
        if log_uniform:
            B = int(round(float(np.exp(rng.uniform(np.log(Ls[0]), np.log(Ls[1]))))))
        else:
            B = int(rng.choice(Ls))
        N = int(rng.choice(Ns)) if not log_uniform else int(rng.integers(Ns[0], Ns[1] + 1))
        N = min(N, B)
        alpha = float(rng.choice(alphas))
        J = core.s2dw_core_comp_Jmax(B, alpha)
        reqs.append((B, N, alpha, J, seeded_flm(B, int(rng.integers(1 << 30)))))
    return reqs


def bench_batch(args, precision=0):
    """64 transforms (32 analyses, 32 syntheses of their resident outputs),
    B from {64,128,256,512}, N from {1,3,5,8}, alpha = 2, J = J_max, all
    submitted at once to the request loop.  Two warm passes (the first builds
    the plans, the second still grows the pinned staging pool), then the
    median of three timed passes.  Plan builds are timed separately (the
    first warm pass)."""
    from s2dw_amd.serve import Server

    rng = np.random.default_rng(4)
    reqs = _mix_requests(rng, 32, [64, 128, 256, 512], [1, 3, 5, 8], [2.0])
    with Server(200 << 30, 16, precision) as srv:
        def one_pass():
            t0 = time.perf_counter()
            ta = [srv.submit_analysis(B, N, J, a, f, keep_resident=True, want_wav=False) for (B, N, a, J, f) in reqs]
            ts = [srv.submit_synthesis(B, N, J, a, wav_from=t) for (B, N, a, J, f), t in zip(reqs, ta)]
            for t in ta + ts:
                srv.wait(t)
            t1 = time.perf_counter()
            err = max(float(np.abs(srv.result(t) - r[4]).max()) for t, r in zip(ts, reqs))
            for t in ta:
                srv.release(t)
            return t1 - t0, err, [t.latency_ms for t in ta + ts]
        warm = [one_pass()[0] for _ in range(2)]
        runs = sorted((one_pass() for _ in range(3)), key=lambda r: r[0])
        wall, err, lat = runs[1]
        err = max(r[1] for r in runs)
    out = {"workload": "batch of 64 transforms (32 analysis + 32 synthesis), L in {64,128,256,512}, "
                       "N in {1,3,5,8}, alpha=2, J=J_max, seeded f_lm",
           "transforms_per_s": 64 / wall, "wall_ms": 1e3 * wall, "max_round_trip_err": err,
           "timed_passes_ms": [1e3 * r[0] for r in runs], "warm_passes_ms": [1e3 * w for w in warm],
           "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
           "first_pass_ms_with_plan_builds": 1e3 * warm[0],
           "precision": "split" if precision & 0xF else "fp64", "dtype": "f64", "n_gpus": 1}
    return out


def bench_serve(args, precision=0):
    """Request loop under Poisson arrivals: B log-uniform in [32, 1024]
    (rounded), N uniform in 1..8, alpha in {2, 3}, J = J_max; every request
    is an analysis whose coefficients stay resident, half of them followed by
    a synthesis of those coefficients.  A first pass builds the plans; the
    second replays the arrivals in real time."""
    import threading

    from s2dw_amd.serve import Server

    rng = np.random.default_rng(5)
    n_req, rate = args.serve_requests, args.serve_rate
    reqs = _mix_requests(rng, n_req, [32, 1024], [1, 8], [2.0, 3.0], log_uniform=True)
    follow = rng.random(n_req) < 0.5
    gaps = rng.exponential(1.0 / rate, n_req)
    with Server(200 << 30, 16, precision) as srv:
        def one_pass(realtime):
            tickets, pairs, lock = [], [], threading.Lock()
            total = n_req + int(follow.sum())

            def waiter():
                i = 0
                while i < total:
                    with lock:
                        t = tickets[i] if i < len(tickets) else None
                    if t is None:
                        time.sleep(1e-4)
                        continue
                    srv.wait(t)
                    i += 1

            th = threading.Thread(target=waiter)
            th.start()
            t0 = time.perf_counter()
            tnext = t0
            for i, (B, N, a, J, f) in enumerate(reqs):
                if realtime:
                    tnext += gaps[i]
                    while time.perf_counter() < tnext:
                        time.sleep(max(0.0, min(1e-3, tnext - time.perf_counter())))
                ta = srv.submit_analysis(B, N, J, a, f, keep_resident=True, want_wav=False)
                with lock:
                    tickets.append(ta)
                if follow[i]:
                    ts = srv.submit_synthesis(B, N, J, a, wav_from=ta)
                    pairs.append((ts, f))
                    with lock:
                        tickets.append(ts)
            th.join()
            t1 = time.perf_counter()
            for t in tickets:
                if t.kind == "analysis":
                    srv.release(t)
            err = max(float(np.abs(ts.out - f).max()) for ts, f in pairs) if pairs else 0.0
            return t1 - t0, [t.latency_ms for t in tickets], tickets, err
        tb0 = time.perf_counter()
        one_pass(False)  # builds every plan of the stream
        t_build = time.perf_counter() - tb0
        wall, lat, tickets, err = one_pass(True)
    ntr = len(tickets)
    out = {"workload": f"serving loop: {n_req} requests, Poisson arrivals at {rate}/s, B log-uniform in [32,1024], "
                       "N uniform in 1..8, alpha in {2,3}, J=J_max; each an analysis kept resident, half followed "
                       "by a synthesis of it",
           "transforms": ntr, "transforms_per_s": ntr / wall, "wall_s": wall,
           "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "max": float(np.max(lat))},
           "max_round_trip_err": err, "warm_pass_s_with_plan_builds": t_build,
           "precision": "split" if precision & 0xF else "fp64", "dtype": "f64", "n_gpus": 1}
    return out


# stage name -> kernel whose launches dominate it
STAGE_KERNEL = {"ana_contract": "k_ana_contract_mfma_w", "syn_contract": "k_syn_contract_mfma",
                "ana_beta_fft": "k_beta_ana_h2", "ana_alpha_c2r": "k_alpha_c2r_o2",
                "syn_alpha_r2c": "k_alpha_r2c_o2", "syn_beta_fft": "k_beta_syn_h2",
                "ana_small_ffts": "k_alpha_c2r_g", "syn_small_ffts": "k_alpha_r2c_g"}


def load_traffic(stage):
    try:
        with open(TRAFFIC_JSON) as fh:
            data = json.load(fh)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    kern = STAGE_KERNEL.get(stage)
    hits = [v for k, v in data.items() if kern and k.startswith(kern)]
    if not hits:
        return None
    return sum(h["hbm_bytes"] * h["dispatches"] for h in hits) / sum(h["dispatches"] for h in hits)


def bench_pipelined(args, L, N, J, K, Phi2, S, prec, d_f, flm):
    """Round trips with the analysis of step i+1 overlapping the synthesis of
    step i: one plan per direction, each on its own stream, wavelet and
    scaling coefficients double-buffered, events for the two dependencies
    (synthesis i reads what analysis i wrote; analysis i+2 overwrites what
    synthesis i read).  Returns transforms/s over the same number of steps."""
    import torch

    from s2dw_amd import core

    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    pa = core.Plan(L, N, J, ALPHA, K, Phi2, S, precision=prec, stream=sa)
    ps = core.Plan(L, N, J, ALPHA, K, Phi2, S, precision=prec, stream=sb)
    for p in (pa, ps):
        if not args.no_autotune:
            p.autotune(2)
        p.set_graphs(not args.eager)
    d_w = [pa.empty_wav(), pa.empty_wav()]
    d_sc = [pa.empty_scoeff(), pa.empty_scoeff()]
    d_g = ps.empty_flm()
    e_ana = [torch.cuda.Event(), torch.cuda.Event()]
    e_syn = [torch.cuda.Event(), torch.cuda.Event()]

    def run(n):
        for i in range(n):
            k = i & 1
            sa.wait_event(e_syn[k]) if i >= 2 else None
            pa.analysis(d_f, d_w[k], d_sc[k])
            e_ana[k].record(sa)
            sb.wait_event(e_ana[k])
            ps.synthesis(d_w[k], d_sc[k], d_g)
            e_syn[k].record(sb)

    torch.cuda.synchronize()
    run(max(2, args.warmup))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    err = float(np.abs(d_g.cpu().numpy().T - flm).max())
    pa.close()
    ps.close()
    return {"transforms_per_s": 2 * args.steps / wall, "ms_per_step": 1e3 * wall / args.steps,
            "round_trip_max_err": err,
            "note": "analysis of step i+1 overlapping synthesis of step i (two plans, two streams)"}


def time_form(args, J, K, Phi2, S, prec, d_f, flm, steps):
    """Round trips of one precision / rounding form of the same workload:
    transforms/s over `steps` back-to-back round trips (graph replay unless the
    form runs eagerly), stage times of its contractions, round-trip error,
    and its wavelet coefficients (host copy, packed)."""
    import torch

    from s2dw_amd import core

    stream = torch.cuda.current_stream()
    t_b0 = time.perf_counter()
    plan = core.Plan(L, N, J, ALPHA, K, Phi2, S, precision=prec, stream=stream)
    build_ms = 1e3 * (time.perf_counter() - t_b0)
    tuned = plan.autotune(2) if not args.no_autotune else "not run"
    d_w, d_sc, d_g = plan.empty_wav(), plan.empty_scoeff(), plan.empty_flm()
    plan.set_graphs(not args.eager)
    for _ in range(max(1, args.warmup)):
        plan.analysis(d_f, d_w, d_sc)
        plan.synthesis(d_w, d_sc, d_g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.analysis(d_f, d_w, d_sc)
        plan.synthesis(d_w, d_sc, d_g)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    plan.set_timing(True)
    for _ in range(3):
        plan.analysis(d_f, d_w, d_sc)
        plan.synthesis(d_w, d_sc, d_g)
    torch.cuda.synchronize()
    plan.set_timing(False)
    st = plan.stage_times()
    rt = float(np.abs(d_g.cpu().numpy().T - flm).max())
    wav = d_w.cpu().numpy()
    rec = {"transforms_per_s": 2.0 * steps / wall, "ms_per_step": 1e3 * wall / steps,
           "round_trip_max_err": rt, "plan_device_bytes": plan.device_bytes, "plan_build_ms": build_ms,
           "wigner_form": {0: "streamed", 16: "table", 48: "reference"}.get(plan.wigner_form, "?"),
           "contraction_ms": {k: st[k][0] / 3 for k in ("ana_contract", "syn_contract") if k in st},
           "autotune": tuned}
    offs, his = plan.wav_offsets, plan.bl_hi
    plan.close()
    return rec, wav, offs, his


def per_scale_err(wav, ref_scales, offs, his):
    """Max abs difference per scale between packed coefficients and per-scale
    (a, b, g) arrays."""
    out = []
    for j, h in enumerate(his):
        n = (2 * h - 1) * 2 * h * N
        x = wav[offs[j]:offs[j] + n].reshape((N, 2 * h, 2 * h - 1))
        out.append(float(np.abs(np.transpose(x, (2, 1, 0)) - ref_scales[j]).max()))
    return out


def host_api_ms(J, K, Phi2, S, bl, flm):
    """The Fortran drop-in's cost: the host-buffer entry points at this
    workload (reference-rounding form, PCIe copies of f_lm and of the 156 MB
    of coefficients each way), median of three after one warm call."""
    from s2dw_amd import core

    ta, ts = [], []
    for i in range(4):
        t0 = time.perf_counter()
        w, sc = core.s2dw_core_analysis_flm2wav_dynamic(flm, K, S, J, L, N, bl, ALPHA)
        t1 = time.perf_counter()
        core.s2dw_core_synthesis_wav2flm_dynamic(w, sc, K, Phi2, S, J, L, N, bl, ALPHA)
        t2 = time.perf_counter()
        if i:
            ta.append(1e3 * (t1 - t0))
            ts.append(1e3 * (t2 - t1))
    del w
    # the static layout wav(0:J, 0:2B-2, 0:2B-1, 0:N-1): (J+1) x 1023 x 1024 x 8
    # doubles (536 MB at this workload) each way, packed and unpacked on the device
    sa, ss = [], []
    for i in range(3):
        t0 = time.perf_counter()
        w, sc = core.s2dw_core_analysis_flm2wav(flm, K, S, J, L, N, bl, ALPHA)
        t1 = time.perf_counter()
        core.s2dw_core_synthesis_wav2flm(w, sc, K, Phi2, S, J, L, N, bl, ALPHA)
        t2 = time.perf_counter()
        if i:
            sa.append(1e3 * (t1 - t0))
            ss.append(1e3 * (t2 - t1))
        del w
    return {"analysis_ms": statistics.median(ta), "synthesis_ms": statistics.median(ts),
            "static_analysis_ms": statistics.median(sa), "static_synthesis_ms": statistics.median(ss),
            "static_bytes": 8 * (J + 1) * (2 * L - 1) * (2 * L) * N,
            "note": "host-buffer entry points (the Fortran drop-in): reference-rounding form, f_lm and the "
                    "wavelet coefficients copied over PCIe each way; dynamic layout median of 3 after a warm "
                    "call, static layout (jj fastest, packed on the device) median of 2 after a warm call"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the GPU's clocks settle over tens of milliseconds of sustained work: the
    # defaults time ~0.1 s of round trips after ~30 ms of warm-up
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "split"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calib-copy", action="store_true", help="run the byte-counter calibration copy and exit")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH", "WRITE"),
                    help="reduce two rocprofv3 counter runs into profiles/traffic.json and exit")
    ap.add_argument("--workload", default="c3", choices=["c3", "batch", "serve"],
                    help="c3: the L=512 N=8 round trip (default); batch: 64 mixed transforms; "
                         "serve: the request loop under Poisson arrivals")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=100.0, help="requests per second")
    ap.add_argument("--eager", action="store_true", help="launch kernels one by one instead of replaying "
                                                         "captured HIP graphs")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default kernel paths")
    ap.add_argument("--no-pipeline", action="store_true", help="skip the round trips with the analysis of step i+1 "
                                                               "overlapping the synthesis of step i")
    ap.add_argument("--skip-forms", action="store_true", help="skip the split, table and reference-rounding forms")
    ap.add_argument("--skip-serving", action="store_true", help="skip the batch (64 transforms) and serving "
                                                                "(Poisson stream) workloads")
    args = ap.parse_args()
    if args.traffic_from:
        traffic_from(*args.traffic_from)
        return
    if args.calib_copy:
        calib_copy()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(independent transforms shard across GPUs with no communication: run one process per GPU)", file=sys.stderr)
        sys.exit(2)

    import torch

    from s2dw_amd import core

    if args.workload == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return
    dev = torch.device("cuda:0")
    J = core.s2dw_core_comp_Jmax(L, ALPHA) - J_MIN
    K, Phi2, bl = core.s2dw_core_init_kernels(J, L, ALPHA)
    S = core.s2dw_core_init_directionality(L, N)
    flm = seeded_flm(L, 1234)
    prec = core.PREC_FP64 if args.precision == "fp64" else core.PREC_SPLIT
    stream = torch.cuda.current_stream(dev)
    plan = core.Plan(L, N, J, ALPHA, K, Phi2, S, precision=prec, stream=stream)
    d_f = torch.from_numpy(np.ascontiguousarray(flm.T)).to(dev)
    d_w, d_sc, d_g = plan.empty_wav(dev), plan.empty_scoeff(dev), plan.empty_flm(dev)

    tuned = plan.autotune(3) if not args.no_autotune else "not run"
    # graph replay (captured at the first warmup step) unless --eager
    plan.set_graphs(not args.eager)
    for _ in range(max(1, args.warmup)):
        plan.analysis(d_f, d_w, d_sc)
        plan.synthesis(d_w, d_sc, d_g)
    torch.cuda.synchronize()
    plan.stage_times()  # discard anything recorded so far

    steps = args.steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        plan.analysis(d_f, d_w, d_sc)
        plan.synthesis(d_w, d_sc, d_g)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    # per-transform latencies in a pass of their own: timing events between
    # the transforms add a few microseconds to each, so they stay out of the
    # throughput loop above
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    for i in range(steps):
        ev[i][0].record(stream)
        plan.analysis(d_f, d_w, d_sc)
        ev[i][1].record(stream)
        plan.synthesis(d_w, d_sc, d_g)
        ev[i][2].record(stream)
    torch.cuda.synchronize()
    pipelined = None if args.no_pipeline else bench_pipelined(args, L, N, J, K, Phi2, S, prec, d_f, flm)
    # second pass: HIP events around every stage launch, on the plan's stream
    plan.set_timing(True)
    for i in range(steps):
        plan.analysis(d_f, d_w, d_sc)
        plan.synthesis(d_w, d_sc, d_g)
    torch.cuda.synchronize()
    plan.set_timing(False)
    st = plan.stage_times()
    lat_a = [ev[i][0].elapsed_time(ev[i][1]) for i in range(steps)]
    lat_s = [ev[i][1].elapsed_time(ev[i][2]) for i in range(steps)]
    lat_all = sorted(lat_a + lat_s)

    g = d_g.cpu().numpy().T
    rt_err = float(np.abs(g - flm).max())

    # per-scale moments of the analysis output (s2dw_stat_moments) on the
    # device: one read of the packed coefficients, HIP events on the plan's stream
    plan.analysis(d_f, d_w, d_sc)
    d_mo = torch.empty((8, J + 1), dtype=torch.float64, device=dev)
    for _ in range(2):
        plan.stat_moments(d_w, d_out=d_mo)
    m0, m1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    m0.record(stream)
    for _ in range(steps):
        plan.stat_moments(d_w, d_out=d_mo)
    m1.record(stream)
    torch.cuda.synchronize()
    mo_ms = m0.elapsed_time(m1) / steps
    mo_bytes = 8.0 * plan.wav_doubles
    moments = {"ms": mo_ms, "GBps": mo_bytes / (mo_ms * 1e-3) / 1e9,
               "frac_hbm_peak": mo_bytes / (mo_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "bytes": mo_bytes,
               "kernels": ["k_mom_partial", "k_mom_final"], "mean": d_mo[0].cpu().tolist(),
               "var": d_mo[1].cpu().tolist()}

    # the statistics pipeline (analysis, then the moments of every scale):
    # stored coefficients read back, against the analysis whose alpha stage
    # reduces to moments records without storing them
    def per_call_ms(fn):
        for _ in range(2):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    if plan.moments_fusable:
        t_ana = per_call_ms(lambda: plan.analysis(d_f, d_w, d_sc))
        t_two = per_call_ms(lambda: (plan.analysis(d_f, d_w, d_sc), plan.stat_moments(d_w, d_out=d_mo)))
        d_mf = torch.empty_like(d_mo)
        t_fused = per_call_ms(lambda: plan.analysis_moments(d_f, d_sc, d_mf))
        moments.update({"analysis_ms": t_ana, "analysis_then_moments_ms": t_two,
                        "analysis_reducing_to_moments_ms": t_fused,
                        "pipeline_speedup": t_two / t_fused,
                        "max_rel_diff_fused_vs_stored": float(
                            ((d_mf[:6] - d_mo[:6]).abs() / d_mo[:6].abs().clamp(min=1.0)).max())})

    hi, lo = plan.bl_hi, plan.bl_lo
    out_offs = list(plan.wav_offsets)
    ana_cmac, syn_cmac = contraction_cmacs(hi, lo, N)
    grouped = st.get("ana_small_ffts", (0.0, 0))[1] > 0
    sb = stage_bytes(hi, lo, N, ncs=sum(1 for n in range(N) if (N + n) % 2), grouped=grouped)
    stages = {}
    for name, (ms, cnt) in st.items():
        if cnt:
            stages[name] = {"ms_per_step": ms / steps, "launches_per_step": cnt / steps}
    # dominant stage (largest time per step) and the roofline of its kernel:
    # algorithmic work per launch over the average launch time from the events
    dom = max(stages, key=lambda k: stages[k]["ms_per_step"])
    nl = max(1.0, stages[dom]["launches_per_step"])
    launch_ms = stages[dom]["ms_per_step"] / nl
    traffic = load_traffic(dom)
    if dom in ("ana_contract", "syn_contract"):
        cm = ana_cmac if dom == "ana_contract" else syn_cmac
        work = 8.0 * cm / nl  # 8 real flops per useful complex MAC
        achieved = work / (launch_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                "kernel": STAGE_KERNEL.get(dom, dom), "stage": dom, "avg_launch_ms": launch_ms,
                "work_per_launch": work, "work_unit": "flop (8 per useful complex MAC)"}
    else:
        work = sb.get(dom, 0) / nl
        achieved = work / (launch_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": STAGE_KERNEL.get(dom, dom), "stage": dom, "avg_launch_ms": launch_ms,
                "work_per_launch": work, "work_unit": "byte (compulsory)"}
    contraction = {}
    for name, cm in (("ana_contract", ana_cmac), ("syn_contract", syn_cmac)):
        if name in stages:
            tf = 8.0 * cm / (stages[name]["ms_per_step"] * 1e-3) / 1e12
            contraction[name] = {"tflops": tf, "frac_fp64_peak": tf / FP64_PEAK_TFLOPS}
    ffts = {}
    for name, byts in sb.items():
        if name in stages:
            gbs = byts / (stages[name]["ms_per_step"] * 1e-3) / 1e9
            ffts[name] = {"GBps": gbs, "frac_hbm_peak": gbs / HBM_PEAK_GBS}

    out = {
        "metric": METRIC,
        "value": 2.0 * steps / wall,
        "unit": "transforms/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 (contraction split on MFMA)",
        "data": "synthetic: seeded real-signal flm, Re/Im ~ N(0,1) scaled by (l+1)^-1",
        "config": {"workload": "analysis+synthesis round trip, L=512 N=8 alpha=2 J_min=2 (J=7)", "L": L, "N": N,
                   "alpha": ALPHA, "J": J, "bl_scoeff": bl, "precision": args.precision,
                   "transforms_per_step": 2, "launch": "eager" if args.eager else "hip graph replay",
                   "autotune": tuned},
        "roofline": roof,
        "latency_ms": {"analysis_p50": float(np.percentile(lat_a, 50)),
                       "analysis_p99": float(np.percentile(lat_a, 99)),
                       "synthesis_p50": float(np.percentile(lat_s, 50)),
                       "synthesis_p99": float(np.percentile(lat_s, 99)),
                       "p50": float(np.percentile(lat_all, 50)), "p99": float(np.percentile(lat_all, 99))},
        "round_trip_max_err": rt_err,
        "pipelined": pipelined,
        "plan_device_bytes": plan.device_bytes,
        "stages": stages,
        "contraction": contraction,
        "fft_stages": ffts,
        "moments": moments,
    }
    wav64 = d_w.cpu().numpy()
    plan.close()
    print("bench: fp64 round trips done", file=sys.stderr, flush=True)
    # the split-precision form (analysis contraction on int8 MFMA) and the
    # table / reference-rounding forms of the plain transform, same workload
    forms = {}
    wavs = {}
    for name, prec in (("split", core.PREC_SPLIT), ("table", core.WIGNER_TABLE),
                       ("reference", core.ROUNDING_REFERENCE)):
        if args.skip_forms:
            break
        forms[name], wavs[name], offs, his = time_form(args, J, K, Phi2, S, prec, d_f, flm, steps)
        print(f"bench: form {name} done", file=sys.stderr, flush=True)
    if "split" in forms:
        sp = forms["split"]
        sp["ratio_vs_fp64"] = sp["transforms_per_s"] / out["value"]
        sp["max_err_per_scale_vs_gpu_fp64"] = per_scale_err(
            wavs["split"], [np.transpose(wav64[o:o + (2 * h - 1) * 2 * h * N].reshape((N, 2 * h, 2 * h - 1)), (2, 1, 0))
                            for o, h in zip(offs, his)], offs, his)
    out["forms"] = forms
    if not args.skip_forms:
        out["host_api"] = host_api_ms(J, K, Phi2, S, bl, flm)
        print("bench: host api done", file=sys.stderr, flush=True)
    if not args.skip_serving:
        out["batch"] = {"fp64": bench_batch(args, core.PREC_FP64), "split": bench_batch(args, core.PREC_SPLIT)}
        print("bench: batch done", file=sys.stderr, flush=True)
        out["serve"] = {"fp64": bench_serve(args, core.PREC_FP64), "split": bench_serve(args, core.PREC_SPLIT)}
        print("bench: serve done", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        rec, w_orc = cpu_baseline(K, Phi2, S, J, bl, flm)
        out["cpu_baseline"] = rec
        # the errors per scale against the FP64 oracle (the reference's algorithm)
        offs64 = [int(o) for o in out_offs]
        out["max_err_per_scale_vs_oracle"] = {"fp64": per_scale_err(wav64, w_orc, offs64, hi)}
        for name in wavs:
            out["max_err_per_scale_vs_oracle"][name] = per_scale_err(wavs[name], w_orc, offs64, hi)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
