#!/usr/bin/env python3
"""Particle-updates/s of the constant-velocity bootstrap PF step on one
MI355X: N = 2^24 particles, FP32 SoA state (x, y, vx, vy), FP64 log-weights,
systematic resampling when ESS < N/2, observations simulated from a seed
(synthetic data).

A step is one pass of the hot path over the particle set, as the Sampler loop
runs it on the device (vsmch_cv_iterate): the fused move + log-likelihood +
weight pass (reading every particle through the previous event's ancestors),
the finalize (ESS, log-evidence, decision, histories, monitor), and, when the
device decides so, the exact-count resampling event. The timed region brackets
exactly --steps steps with inputs resident in HBM.

Prints one JSON line (the metric contract) with:
  value       particle-updates/s (N * steps / seconds)
  roofline    the whole step: algorithmic bytes N (48 + 48 r) per step (r:
              the fraction of steps that resample) over the step time from
              HIP events on the library stream, against 8 TB/s; traffic from
              the HBM counters; roofline_move: the fused move alone
  cpu_baseline  the oracle's OpenMP restatement of the reference path on the
              host cores, on a bounded sample of the same workload
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MOVE_BYTES = 48       # state r/w 16 + 16 (FP32 x4), log-weight r/w 8 + 8
RESAMPLE_BYTES = 48   # log-weights 8, ancestor w/r 4 + 4, state r/w 16 + 16


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=195)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--scheme", default="systematic")
    ap.add_argument("--threshold", type=float, default=0.5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--sections", default="all",
                    help="comma-separated subset of step,extra,dropin,micro,learned,batched,gmm "
                         "(profiling runs); 'all' runs every one")
    return ap.parse_args()


def run_steps(va, f, scheme, thr, t0, k, ox, oy, host_decides=False):
    """k steps of move + resample. With host_decides the resampling decision
    is read on the host (Sampler::do_resample's order) and a step that does
    not resample queues no further work; otherwise the whole event is queued
    and decided on the device."""
    for t in range(t0, t0 + k):
        f.move(t, float(ox[t]), float(oy[t]))
        if host_decides:
            f.resample(scheme, thr)
        else:
            f.resample_async(scheme, thr)
    return t0 + k


def time_config(va, torch, n, scheme, thr, warmup, steps, ox, oy, seed=1,
                host_decides=False):
    f = va.CvFilter(n, "float32", seed=seed)
    f.init(float(ox[0]), float(oy[0]))
    f.resample_async(scheme, thr)
    t = run_steps(va, f, scheme, thr, 1, warmup, ox, oy, host_decides)
    torch.cuda.synchronize()
    r0 = f.resample_count()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    t = run_steps(va, f, scheme, thr, t, steps, ox, oy, host_decides)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    nres = f.resample_count() - r0
    return f, t, ms, nres


def time_iterate(va, torch, n, scheme, thr, warmup, steps, ox, oy, seed=1):
    """The Sampler loop as vsmch_cv_iterate runs it (initialize + warmup
    moves untimed, then `steps` timed moves with their events, histories and
    the "pos" monitor), HIP events on the library stream."""
    f = va.CvFilter(n, "float32", seed=seed)
    h = f.history_buffers(steps)
    f.iterate(0, ox[:warmup + 1], oy[:warmup + 1], scheme, thr, histories=False)
    torch.cuda.synchronize()
    r0 = f.resample_count()
    t0 = warmup + 1
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    f.iterate(t0, ox[t0:t0 + steps], oy[t0:t0 + steps], scheme, thr, histories=h)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    nres = f.resample_count() - r0
    assert int(h["resampled"].sum().item()) == nres
    return f, t0 + steps, ms, nres


def scheme_events(va, torch, n, threshold, ox, oy, steps=20):
    """The iterate loop with each scheme: never resampling, resampling every
    step (systematic), and stratified / multinomial / residual when ESS <
    threshold N. event_us: the step time above the never-resampling loop,
    per resampling event; vs_systematic: that against systematic's."""
    extra = {}
    for name, sch, thr in (("systematic_never", va.SYSTEMATIC, 0.0),
                           ("systematic_always", va.SYSTEMATIC, float("inf")),
                           ("stratified", va.STRATIFIED, threshold),
                           ("multinomial", va.MULTINOMIAL, threshold),
                           ("residual", va.RESIDUAL, threshold)):
        g, _, ms2, nres2 = time_iterate(va, torch, n, sch, thr, 2, steps, ox, oy)
        g.close()
        extra[name] = {"ms_per_step": ms2 / steps, "resampled_steps": nres2,
                       "updates_per_s": n * steps / (ms2 / 1e3),
                       "step_hbm_frac": n * (MOVE_BYTES * steps + RESAMPLE_BYTES * nres2)
                       / (ms2 / 1e3) / 1e9 / HBM_PEAK_GBS}
    base = extra["systematic_never"]["ms_per_step"]
    for name, e in extra.items():
        if e["resampled_steps"]:
            e["event_us"] = (e["ms_per_step"] - base) * steps / e["resampled_steps"] * 1e3
    sys_us = extra["systematic_always"]["event_us"]
    for name, e in extra.items():
        if "event_us" in e:
            e["vs_systematic"] = e["event_us"] / sys_us
    return extra


def move_kernel_ms(va, torch, f, t, reps, ox, oy):
    """Average duration of the fused move (+ its one-block weight merge),
    HIP events on the library stream."""
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    f.move(t, float(ox[t]), float(oy[t]))
    torch.cuda.synchronize()
    e0.record()
    for i in range(reps):
        tt = t + 1 + i
        f.move(tt, float(ox[tt]), float(oy[tt]))
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def copy_bandwidth(torch, nbytes=1 << 30, reps=5):
    """Device-to-device copy of a known size (read + write bytes / time):
    the box's attainable HBM bandwidth, and the reference kernel whose
    FETCH_SIZE / WRITE_SIZE counters scale the traffic figures."""
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2 * nbytes / (ms / 1e3) / 1e9


def _events_ms(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def microbench(va, torch, reps=5):
    """2^28 FP64 U01 draws from Philox4x32-10 and Threefry4x64-20 and 2^28
    FP64 normals (8 bytes written per variate), and resampling of N = 2^24
    log-weights ~ N(0, s^2): set_log + the scheme's counts, ancestors and
    an in-place gather of one FP32 component, every call resampling."""
    out = {}
    n = 1 << 28
    buf = torch.empty(n, dtype=torch.float64, device="cuda")
    for name, fn in (
            ("u01_philox", lambda: va.u01_fill(va.PHILOX4X32, 7, 0, n, out=buf)),
            ("u01_threefry", lambda: va.u01_fill(va.THREEFRY4X64, 7, 0, n, out=buf)),
            ("normal_philox", lambda: va.normal_fill(va.PHILOX4X32, 7, 0, n, out=buf)),
            ("normal_threefry", lambda: va.normal_fill(va.THREEFRY4X64, 7, 0, n, out=buf))):
        ms = _events_ms(torch, fn, reps)
        out[name] = {"n": n, "ms": ms, "variates_per_s": n / (ms / 1e3),
                     "gbs": 8 * n / (ms / 1e3) / 1e9,
                     "hbm_frac": 8 * n / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}
    del buf
    m = 1 << 24
    g = torch.Generator(device="cuda")
    for s in (0.5, 2.0, 8.0):
        g.manual_seed(int(10 * s))
        lw = torch.randn(m, generator=g, dtype=torch.float64, device="cuda") * s
        p = va.Particle(m, 1, "float32", seed=3)
        for sname in ("multinomial", "stratified", "systematic", "residual"):
            sch = va.SCHEMES[sname]

            def one():
                p.set_log(lw)
                p.resample_async(sch, float("inf"))
            ms = _events_ms(torch, one, reps)
            out[f"resample_s{s:g}_{sname}"] = {
                "n": m, "ms": ms, "particles_per_s": m / (ms / 1e3),
                "hbm_frac": m * (8 + 8 + 8 + 4 + 4 + 4 + 4) / (ms / 1e3) / 1e9
                / HBM_PEAK_GBS}
        p.close()
    return out


LG_MODEL = dict(dt=0.1, sd_pos=0.02, sd_vel=0.02, sd_pos0=0.02, sd_vel0=0.02, obs_model=2,
                obs_scale=0.01, obs_df=10.0)
ST_MODEL = dict(dt=0.1, sd_pos=0.02, sd_vel=0.02, sd_pos0=0.02, sd_vel0=0.02, obs_model=1,
                obs_scale=0.1, obs_df=10.0)


def _timed(torch, fn):
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    r = fn()
    e1.record()
    torch.cuda.synchronize()
    return r, e0.elapsed_time(e1) / 1e3


def _train_net(va, torch, model, track, n=1 << 14, steps=600):
    ox, oy = track
    f = va.CvFilter(n, "float32", seed=5, **model)
    (feat, target), s_data = _timed(torch, lambda: f.transitions(ox, oy, va.SYSTEMATIC, 0.5,
                                                                 train_seed=77))
    f.close()
    net = va.ProposalNet()
    net.fit_normalisation(feat)
    net.init_random(1)
    loss, s_train = _timed(torch, lambda: net.train(feat, target, batch=4096, steps=steps,
                                                    lr=2e-3))
    return net, loss, {"samples": int(feat.shape[0]), "steps": steps, "batch": 4096,
                       "data_gpu_s": s_data, "train_gpu_s": s_train,
                       "loss_first": float(loss[0]), "loss_last20": float(loss[-20:].mean()),
                       "gain_nats": net.gain(loss)}


def learned_vs_plain(va, torch, tracks=64, T=50, sizes=(1 << 10, 1 << 12, 1 << 14, 1 << 16,
                                                           1 << 18), t_train=64):
    """log-evidence RMSE against the FP64 Kalman filter vs GPU-seconds per
    filter run, learned proposal vs bootstrap, on the linear-Gaussian CV
    model (obs sd 0.01), T = 50, systematic resampling at ESS < N/2; plus
    the Student-t model with 5% outliers at 10x scale against a 2^22-particle
    bootstrap reference. Each (method, N) runs the tracks as one batch of
    filters (one graph-replayed step for all of them): GPU-seconds per filter
    = the batch's GPU time / tracks. The network trains on the bootstrap
    filter's transitions over a t_train-step track (2^14 particles, (t_train
    - 1) 2^14 >= 10^6 transitions); when it gains less than 0.1 nats per
    transition over the bootstrap proposal the learned run falls back to the
    bootstrap proposal (ProposalNet.select). The bootstrap curve extends to
    16 times the largest learned N, so every learned RMSE has a bootstrap
    point at least as accurate."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_lib as ol
    from vsmc_amd.models import cv_simulate, cv_simulate_many

    def sim(kind, F, seed, steps=T):
        if kind == "lg":
            return cv_simulate_many(F, steps, seed, sd_pos0=0.02, sd_vel0=0.02, obs="gauss",
                                    obs_scale=0.01)
        return cv_simulate_many(F, steps, seed, sd_pos0=0.02, sd_vel0=0.02, obs="student",
                                obs_scale=0.1, obs_df=10.0, outlier_frac=0.05,
                                outlier_mult=10.0)

    def run_batch(model, n, net, ox, oy):
        b = va.CvBatch([n] * tracks, [T] * tracks, list(range(1, tracks + 1)), va.SYSTEMATIC,
                       0.5, **model)
        b.set_obs(list(zip(ox, oy)))
        b.run(net=net, graph=True)
        z = b.histories()["logz"][:, T - 1].copy()
        ms = b.info()["run_ms"]
        b.close()
        return z, ms / 1e3 / tracks

    out = {}
    for kind, model in (("linear_gaussian", LG_MODEL), ("student_t_outliers", ST_MODEL)):
        k = "lg" if kind == "linear_gaussian" else "st"
        tx, ty = sim(k, 1, 11, t_train)
        net, loss, train = _train_net(va, torch, model, (tx[0], ty[0]))
        use = net.select(loss)
        train["proposal"] = "learned" if use is not None else \
            "bootstrap (fallback: gain below 0.1 nats per transition)"
        ox, oy = sim(k, tracks, 1000)
        if k == "lg":
            ref = np.array([ol.kalman_cv(ox[s], oy[s], 0.1, 0.02, 0.02, 0.02, 0.02, 0.01)[0]
                            for s in range(tracks)])
            ref_desc = "FP64 Kalman filter"
        else:
            ref, _ = run_batch(model, 1 << 22, None, ox, oy)
            ref_desc = "bootstrap filter at N = 2^22"
        rows = []
        for n in sizes + (sizes[-1] * 4, sizes[-1] * 16):
            for name, nn in (("bootstrap", None), ("learned", use)):
                if name == "learned" and n > sizes[-1]:
                    continue   # the bootstrap curve extends past the learned sizes
                z, sec = run_batch(model, n, nn, ox, oy)
                errs = z - ref
                rows.append({"method": name, "n": n,
                             "rmse": float(np.sqrt(np.mean(np.square(errs)))),
                             "bias": float(np.mean(errs)),
                             "gpu_s_per_run": float(sec)})
        # GPU-seconds the bootstrap filter needs for each learned RMSE,
        # interpolated log-log along its (gpu_s, rmse) curve
        bs = sorted([r for r in rows if r["method"] == "bootstrap"], key=lambda r: r["n"])
        lx = np.log([r["gpu_s_per_run"] for r in bs])
        ly = np.log([max(r["rmse"], 1e-12) for r in bs])
        ratios = []
        for r in rows:
            if r["method"] != "learned":
                continue
            target = np.log(max(r["rmse"], 1e-12))
            if target >= ly.max() or target <= ly.min():
                ratios.append({"n": r["n"], "rmse": r["rmse"], "gpu_s_ratio": None})
                continue
            order = np.argsort(ly)
            need = float(np.exp(np.interp(target, ly[order], lx[order])))
            ratios.append({"n": r["n"], "rmse": r["rmse"],
                           "bootstrap_gpu_s_same_rmse": need,
                           "gpu_s_ratio": need / r["gpu_s_per_run"]})
        out[kind] = {"reference": ref_desc, "tracks": tracks, "T": T, "training": train,
                     "rows": rows, "equal_rmse": ratios}
        net.close()
    return out


def batched_filters(va, torch, F=1024, seed=2024, plain_sample=24):
    """1,024 independent filters of the CPU configuration's model (Student-t
    observations, stratified resampling at ESS < N/2), N_f = floor(2^U(10,18)),
    T_f ~ U{50..500}, per-filter seeds: one run of the batch (persistent move
    + graph-replayed steps, autotuned move grid) against the plain form (one
    launch sequence per filter, CvFilter.run) on a sample of the filters."""
    import numpy as np
    from vsmc_amd.models import cv_simulate_many
    rs = np.random.default_rng(seed)
    ns = np.floor(2.0 ** rs.uniform(10, 18, F)).astype(np.int64)
    Ts = rs.integers(50, 501, F)
    seeds = 100 + np.arange(F)
    ox, oy = cv_simulate_many(F, int(Ts.max()), seed)
    b = va.CvBatch(ns, Ts, seeds, va.STRATIFIED, 0.5)
    b.set_obs(list(zip(ox, oy)))
    b.run(graph=True)   # first run: allocation, code load
    b.run(graph=True)
    info = b.info()
    h = b.histories()
    b.close()
    updates = float((ns * Ts).sum())
    events = float(sum(int(h["resampled"][f, :Ts[f]].sum()) * ns[f] for f in range(F)))
    secs = info["run_ms"] / 1e3
    idx = rs.choice(F, plain_sample, replace=False)
    p_updates, p_secs = 0.0, 0.0
    for f in idx:
        g = va.CvFilter(int(ns[f]), "float32", seed=int(seeds[f]))
        _, s = _timed(torch, lambda: g.run(ox[f, :Ts[f]], oy[f, :Ts[f]], va.STRATIFIED, 0.5))
        g.close()
        p_updates += float(ns[f] * Ts[f])
        p_secs += s
    return {"filters": F, "particles": int(ns.sum()), "particle_steps": updates,
            "t_max": int(Ts.max()), "batched_s": secs,
            "batched_updates_per_s": updates / secs,
            "batched_hbm_frac": (updates * MOVE_BYTES + events * RESAMPLE_BYTES) / secs / 1e9
            / HBM_PEAK_GBS,
            "move_blocks_autotuned": info["move_blocks"],
            "plain_sample_filters": plain_sample,
            "plain_updates_per_s": p_updates / p_secs,
            "speedup_vs_plain": (updates / secs) / (p_updates / p_secs)}


def tempered_gmm(va, torch, n=1 << 20, moves=5, cpu_budget=4.0):
    """Tempered SMC for the 4-component mixture on 1,000 seeded points:
    N = 2^20 FP32 12-D particles, adaptive tempering to conditional ESS 0.9,
    5 RW-MH moves per step (each move one likelihood evaluation of 1,000
    points x 4 components). The CPU baseline is the oracle's move (FP64) on
    the host cores for a sample of the final particles."""
    import ctypes
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    from vsmc_amd.models import gmm_simulate
    x = gmm_simulate(1000, 2016)
    s = va.GmmSampler(n, x, ncomp=4, seed=101, cess=0.9, moves=moves)
    s.run(max_steps=3)   # code load and allocation
    r, secs = _timed(torch, lambda: s.run(max_steps=1000))
    st = s.state().cpu().numpy().astype(np.float64)
    s.close()
    steps = r["steps"]
    gpu_moves = float(n) * (steps - 1) * moves
    # CPU: the same move on a growing sample of the final particles
    L = ol._gmm_sig()
    L.vo_gmm_moves.restype = ctypes.c_int64
    P = np.ctypeslib.ndpointer
    L.vo_gmm_moves.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                               P(np.float64), P(np.float64), P(np.float64), ctypes.c_double,
                               P(np.float64), P(np.float64), ctypes.c_int64, ctypes.c_int]
    h = ol.gmm_hyper(x)
    threads = min(os.cpu_count() or 1, 16)
    rs = np.random.default_rng(0)
    m = 4096
    while True:
        p = np.ascontiguousarray(st[:, :m].T)
        z = rs.normal(size=(m, 4))
        lu = np.log(rs.random(m))
        t0 = time.perf_counter()
        L.vo_gmm_moves(0, 4, m, p, z, lu, 1.0, h, x, len(x), threads)
        cs = time.perf_counter() - t0
        if cs >= cpu_budget / 4 or m >= n:
            break
        m = min(4 * m, n)
    return {"n": n, "dim": 12, "data_points": len(x), "moves_per_step": moves,
            "steps": steps, "gpu_s": secs, "ms_per_step": secs * 1e3 / max(steps - 1, 1),
            "gpu_moves_per_s": gpu_moves / secs,
            "gpu_component_evals_per_s": gpu_moves * len(x) * 4 / secs,
            "log_evidence_smc": float(r["logz"][-1]),
            "log_evidence_path_sampling": r["path_sampling"],
            "accept_last": [float(a) for a in r["accept"][-1]],
            "cpu_baseline": {"moves_per_s": m / cs, "cores": threads, "kind": "port",
                             "sample": f"oracle mu move (FP64) on {m} final particles, "
                                       f"{cs:.2f}s"}}


def measured_traffic(name="cv_step_traffic.json", key="bytes_per_step"):
    """HBM bytes per step (or per launch) from a committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy kernel), if any."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        return json.load(fh).get(key)


def dropin_path(n=1 << 24, T=24):
    """The drop-in C++ layer at N particles: the reference's PF example
    written with device functors (tests/cpp/pf_hip.cpp: InitializeHIP /
    MoveHIP over a __device__ eval_sp, Weight::add_log from the functor's
    device array, Particle::resample through the library's device event, the
    "pos" monitor), FP64 state, stratified when ESS < N/2; T - 4 timed steps
    after initialize + 3."""
    import subprocess
    import tempfile
    from vsmc_amd.models import cv_simulate
    exe = os.path.join(ROOT, "tests", "cpp", "pf_hip")
    if not os.path.exists(exe):
        return {"error": "tests/cpp/pf_hip not built"}
    _, ox, oy = cv_simulate(T, 20161)
    with tempfile.NamedTemporaryFile("w", suffix=".data", delete=False) as fh:
        for a, b in zip(ox, oy):
            fh.write(f"{float(a)!r} {float(b)!r}\n")
        path = fh.name
    try:
        r = subprocess.run([exe, str(n), str(T), "1", "0.5", "1", path, "time"],
                           capture_output=True, text=True, timeout=600)
    finally:
        os.unlink(path)
    if r.returncode != 0:
        return {"error": r.stderr[-500:]}
    _, steps, ms, nres = r.stdout.strip().split("\n")[-1].split("\t")
    steps, ms, nres = int(steps), float(ms), int(nres)
    return {"n": n, "steps": steps, "ms_per_step": ms / steps,
            "updates_per_s": n * steps / (ms / 1e3), "resampled_steps": nres,
            "state": "fp64 x4 (StateHIP)", "scheme": "stratified", "threshold": 0.5,
            "path": "C++ Sampler: InitializeHIP / MoveHIP device functors, "
                    "Weight::add_log, Particle::resample on the device, monitor"}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(budget_s):
    """The oracle's OpenMP restatement of the reference's Sampler loop on the
    host: the eval_sp loop in parallel, the weight normalisation, resampling,
    rep -> index and copy sequential, as in the reference's OMP backend. The
    threads are this process's CPU share (16 on a one-GPU box, whatever the
    host's core count), with the per-phase split of the time."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    from vsmc_amd.models import cv_simulate
    ncpu = os.cpu_count() or 1
    try:
        share = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        share = ncpu
    threads = min(share, 16)
    n, T = 1 << 18, 10
    _, ox, oy = cv_simulate(200, 20161)
    # grow the sample until it takes a meaningful share of the budget
    r = ol.pf_run(n, T, 1, ol.SYSTEMATIC, 0.5, ox, oy, threads=threads)
    while r["seconds"] < budget_s / 4 and n < (1 << 24):
        n *= 2
        r = ol.pf_run(n, T, 1, ol.SYSTEMATIC, 0.5, ox, oy, threads=threads)
    secs = r["seconds"]
    return {"value": n * T / secs, "unit": "particle-updates/s", "cores": threads,
            "kind": "port", "host_cpus": ncpu, "affinity_cpus": share,
            "cpu_model": _cpu_model(),
            "phase_seconds": r["phase_seconds"],
            "resampled_steps": int(r["resampled"].sum()),
            "sample": f"oracle Sampler restatement, N={n}, T={T}, FP64 state, "
                      f"systematic, ESS<N/2; OpenMP eval_sp loop on {threads} threads, "
                      f"sequential weights/resampling/copy as in the reference; "
                      f"{secs:.2f}s"}


_T0 = time.time()


def progress(section, ms_per_step=None):
    """One line on stderr as each section ends, so a long run shows where it is."""
    extra = "" if ms_per_step is None else f" ({ms_per_step:.4f} ms/step)"
    print(f"[bench] {section} done at {time.time() - _T0:.1f} s{extra}", file=sys.stderr,
          flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU (single-GPU build); --gpus {a.gpus} is not "
              f"supported", file=sys.stderr)
        sys.exit(2)
    import torch
    import vsmc_amd as va
    from vsmc_amd.models import cv_simulate

    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    va.check(va.lib().vsmch_set_stream(stream.cuda_stream))
    scheme = va.SCHEMES[a.scheme]
    T = a.warmup + a.steps + 64
    _, ox, oy = cv_simulate(T, 20161)
    sections = ({"step", "extra", "dropin", "micro", "learned", "batched", "gmm"}
                if a.sections == "all"
                else set(a.sections.split(",")))
    if a.no_extra:
        sections &= {"step"}
    if "step" not in sections:
        # one workload alone (e.g. under a profiler): its figures only
        out = {}
        for name, fn in (("extra", lambda va_, t_: scheme_events(va_, t_, a.n, a.threshold,
                                                                  ox, oy)),
                         ("dropin", lambda va_, t_: dropin_path()),
                         ("micro", microbench), ("learned", learned_vs_plain),
                         ("batched", batched_filters), ("gmm", tempered_gmm)):
            if name in sections and fn is not None:
                out[name] = fn(va, torch)
        print(json.dumps(out), flush=True)
        return

    f, t, ms, nres = time_iterate(va, torch, a.n, scheme, a.threshold, a.warmup, a.steps,
                                  ox, oy)
    f.close()
    # the step-level API (move, then an event that gathers in place)
    f, t, ms_api, nres_api = time_config(va, torch, a.n, scheme, a.threshold, a.warmup,
                                         a.steps, ox, oy)
    reps = min(20, T - t - 2)
    mk_ms = move_kernel_ms(va, torch, f, t, reps, ox, oy)
    f.close()
    torch.cuda.synchronize()
    copy_gbs = copy_bandwidth(torch)

    secs = ms / 1e3
    value = a.n * a.steps / secs
    bytes_alg = a.n * (MOVE_BYTES * a.steps + RESAMPLE_BYTES * nres)
    step_gbs = bytes_alg / secs / 1e9
    achieved = MOVE_BYTES * a.n / (mk_ms / 1e3) / 1e9
    out = {
        "metric": "particle-updates/s at N=2^24 (%HBM roofline)",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state / f64 weights",
        "data": "synthetic (constant-velocity track simulated from seed 20161)",
        "config": {"workload": "cv_bootstrap_pf", "n": a.n, "state": "fp32 SoA x4",
                   "weights": "fp64", "scheme": a.scheme, "threshold": a.threshold,
                   "model": "CV dt=0.1 sd 0.02, Student-t obs df=10 scale 0.1",
                   "parallelism": "single GPU"},
        "step_hbm_frac": step_gbs / HBM_PEAK_GBS,
        "resampled_steps": nres,
        # the whole step: algorithmic bytes N (48 + 48 r_t) summed over the
        # timed steps / the timed region
        "roofline": {"bound": "hbm", "achieved": step_gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
                     "traffic": measured_traffic(),
                     "kernel": "PF step: move + finalize + resampling event (all kernels)",
                     "step_ms": ms / a.steps},
        "roofline_move": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                          "traffic": measured_traffic("cv_move_traffic.json",
                                                      "bytes_per_launch"),
                          "kernel": "k_cv_step<float> + weight merge",
                          "kernel_ms": mk_ms},
        "step_api": {"ms_per_step": ms_api / a.steps, "resampled_steps": nres_api,
                     "updates_per_s": a.n * a.steps / (ms_api / 1e3),
                     "step_hbm_frac": a.n * (MOVE_BYTES * a.steps + RESAMPLE_BYTES * nres_api)
                     / (ms_api / 1e3) / 1e9 / HBM_PEAK_GBS},
        "copy_gbs": copy_gbs,
    }
    progress("step", ms / a.steps)
    for sec, key, fn in (
            ("extra", "extra", lambda: scheme_events(va, torch, a.n, a.threshold, ox, oy)),
            ("dropin", "dropin_functor_path", dropin_path),
            ("micro", "microbench", lambda: microbench(va, torch)),
            ("learned", "learned_proposal", lambda: learned_vs_plain(va, torch)),
            ("batched", "batched_filters", lambda: batched_filters(va, torch)),
            ("gmm", "tempered_gmm", lambda: tempered_gmm(va, torch))):
        if sec in sections:
            out[key] = fn()
            progress(sec)
    out["cpu_baseline"] = None if a.no_cpu else cpu_baseline(a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
