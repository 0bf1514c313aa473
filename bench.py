"""K-build benchmark on one MI355X.

python bench.py [--config C] [--steps K] [--warmup W] [--gpus 1]

A step is one truncated-Coulomb exchange (K) build of the workload: set_P,
execution of its screened call stream (recorded once and kept resident in
HBM as a plan), get_K (H2D of P and D2H of K included).  The timed steps run
with no instrumentation; one further step, serialised and instrumented,
gives the survivor count and the per-kernel times of the roofline.

value = surviving (primitive quartet, lattice cell, class) slots per second,
the reference's per-class Fm / VRR work unit (Fm[0] > 1e-12); ms_per_step is
K-build seconds per SCF step.  Beside it: the per-step cost of the call
stream through the libgint_* C-ABI (what a CP2K caller pays every step,
estimated from a strided sample of set quartets), max|dK| against the FP64
oracle on a strided sample, and the oracle's throughput on that sample on
this host's cores (the CPU baseline).
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector (and FP64 matrix), dense
PEAK_HBM_GBS = 8000.0
METRIC = "K-build s/SCF step + quartets/s at max|dK|<=1e-8 Ha; p50/p99 request latency"


def log(*a):
    print("#", *a, file=sys.stderr, flush=True)


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload`, from the rocprofv3
    FETCH_SIZE and WRITE_SIZE passes summarised under profiles/ (a line
    'hbm_traffic: {json}' per kernel), or (None, None) when none covers them"""
    import glob

    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(f) as fh:
            for line in fh:
                if line.startswith("hbm_traffic: "):
                    d = json.loads(line[len("hbm_traffic: "):])
                    if d.get("kernel") == kernel and d.get("workload") == workload:
                        return d["bytes_per_launch"], os.path.basename(f)
    return None, None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def serving(a):
    """the serving workload: independent periodic systems (20-200 atoms,
    O 3s2p1d / H 2s1p, 27 cells) sending SCF-like K-build requests at
    Poisson arrival times to one device-resident request loop under a
    device budget; p50 / p99 latency and surviving slots per second"""
    from libgint_amd import driver as D, load

    lib = load()
    t0 = time.time()
    systems = D.serving_systems(n_systems=a.clients, seed=3)
    clients = []
    surv = []
    for i, sy in enumerate(systems):
        c = D.ServingClient(lib, sy, steps=10, seed=i)
        # survivors of one K-build of this client (device counters, untimed)
        lib.libgint_amd_enable_stats(c.h, 1)
        lib.libgint_amd_plan_execute(c.h)
        Kw = np.zeros(sy.P_size)
        lib.libgint_get_K(c.h, D._ptr(Kw))
        surv.append(D.stats(lib, "libgint_", c.h)[1])
        lib.libgint_amd_enable_stats(c.h, 0)
        clients.append(c)
        log(f"client {i}: {sy.describe()}; set quartets {c.st.n_sets}, primitive quartets {c.st.n_prm}, "
            f"surviving slots per K-build {surv[-1]}")
    log(f"clients ready in {time.time() - t0:.1f}s")
    done, st = D.serve(lib, clients, n_requests=a.requests, rate=a.rate, slots=a.slots, max_mem_mib=a.budget,
                       seed=11, progress=log)
    ok = [d for d in done if d[3] == 0]
    slots = sum(surv[d[0]] for d in ok)
    wall = st[4]
    lat = np.array([d[4] for d in ok])
    log(f"served {len(ok)} of {len(done)} requests in {wall:.2f}s: p50 {st[1] * 1e3:.1f} ms, p99 {st[2] * 1e3:.1f} ms, "
        f"mean {st[3] * 1e3:.1f} ms; peak resident {st[5] / 2**20:.0f} MiB of {a.budget}; evictions {st[6]:.0f}")
    out = {
        "metric": METRIC, "value": round(slots / wall, 1), "unit": "quartets/s", "n_gpus": 1, "steps": len(ok),
        "warmup": 0, "ms_per_step": round(1e3 * float(np.mean(lat)), 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "p50_ms": round(st[1] * 1e3, 3), "p99_ms": round(st[2] * 1e3, 3),
        "config": {"workload": f"serving: {a.requests} K-build requests of {a.clients} periodic water-like systems "
                               f"(20-200 atoms, 27 cells), Poisson arrivals at {a.rate} req/s, one worker, "
                               f"{a.budget} MiB device budget",
                   "config_index": 3, "requests": len(done), "served": len(ok), "wall_s": round(wall, 3),
                   "peak_resident_MiB": round(st[5] / 2**20, 1), "evictions": int(st[6]), "admissions": int(st[7]),
                   "path": "FP64, one resident plan per client, D and K resident on the device"},
    }
    print(json.dumps(out), flush=True)
    for c in clients:
        lib.libgint_amd_destroy_handle(c.h)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--stride", type=int, default=1, help="every n-th set quartet of the workload (1: all)")
    ap.add_argument("--sample", type=int, default=1,
                    help="record a reproducible 1-in-N hash sample of the set quartets (1: all); the step time "
                         "is reported for the sample and, scaled by N, as an estimate for the whole workload")
    ap.add_argument("--abi-stride", type=int, default=64,
                    help="per-call ABI ingestion is timed on every n-th set quartet and scaled up")
    ap.add_argument("--check-stride", type=int, default=0,
                    help="max|dK| and the CPU baseline use every n-th set quartet (0: sized to ~10 s of CPU)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle (max|dK| and CPU baseline)")
    ap.add_argument("--no-abi", action="store_true", help="skip the per-call ABI ingestion estimate")
    ap.add_argument("--c0-file", default=None, help="CP2K t_c_g.dat to use in place of the generated table")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 3 device copies of MIB MiB (calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--max-mem", type=int, default=65536, help="device budget, MiB")
    ap.add_argument("--roof-stride", type=int, default=16,
                    help="the serialised, instrumented roofline run takes every n-th set quartet")
    ap.add_argument("--mixed", type=float, default=1e-7,
                    help="mixed-precision threshold tau for the extra mixed steps (0: none)")
    ap.add_argument("--clients", type=int, default=16, help="serving workload: systems")
    ap.add_argument("--requests", type=int, default=400, help="serving workload: requests")
    ap.add_argument("--rate", type=float, default=5.0, help="serving workload: mean arrivals per second")
    ap.add_argument("--slots", type=int, default=1,
                    help="serving workload: worker threads asked of the server (it runs one)")
    ap.add_argument("--budget", type=int, default=7000, help="serving workload: device budget, MiB")
    ap.add_argument("--roof-only", action="store_true",
                    help="only the roofline's serialised sample run (for kernel-trace and counter profiles)")
    ap.add_argument("--no-progress", dest="progress", action="store_false",
                    help="no per-quartet-type progress lines (and no host wait per type) in the steps")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)

    if a.copy_probe:
        import torch

        n = a.copy_probe * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 8 * n, "copy_probe_bytes_written": 8 * n, "launches": 3}))
        return

    from libgint_amd import driver as D, load, read_c0

    if a.config == 3:
        serving(a)
        return
    lib = load()
    c0 = read_c0(a.c0_file, 21) if a.c0_file else None
    t0 = time.time()
    s = D.System(a.config)
    st = s.record(sample=a.sample) if a.sample > 1 else s.record()
    t_rec = time.time() - t0
    n_sets = len(range(0, st.n_sets, a.stride))
    log(f"config {a.config}: {s.describe()}; sets {st.n_sets} rows {st.n_rows} primitive quartets {st.n_prm} "
        f"contracted quartets {st.n_qrtt} (screened in {t_rec:.1f}s); "
        f"{'hash sample 1 in %d; ' % a.sample if a.sample > 1 else ''}stride {a.stride}: {n_sets} set quartets")
    P = s.P()
    dp = D._ptr

    # the call stream through the reference ABI (add_qrt / add_qrtt per
    # contracted quartet), as a CP2K caller sends it every step: timed on a
    # strided sample, host only (recording, nothing reaches the device)
    abi = None
    if not a.no_abi:
        h0, t0tab = D.setup_handle(lib, "libgint_", s, P=P, max_mem_mib=a.max_mem, c0=c0)
        lib.libgint_amd_plan_begin(h0)
        sstride = a.stride * a.abi_stride
        t = time.perf_counter()
        st.replay(t0tab, h0, 0, st.n_sets, sstride, False)
        t_abi = time.perf_counter() - t
        nsmp = len(range(0, st.n_sets, sstride))
        lib.libgint_amd_destroy_handle(h0)
        abi = {"s_per_step_est": round(t_abi * n_sets / max(nsmp, 1), 2), "threads": 1,
               "sample": f"every {sstride}th set quartet ({nsmp}), per-call add_qrt / add_qrtt, {t_abi:.2f} s, "
                         f"scaled to {n_sets}"}
        log(f"ABI ingestion: {abi['sample']} -> {abi['s_per_step_est']} s per step on one thread")

    def main_steps():
        h, table = D.setup_handle(lib, "libgint_", s, P=P, max_mem_mib=a.max_mem, c0=c0)
        K = np.zeros(s.P_size)
        t0 = time.time()
        lib.libgint_amd_plan_begin(h)
        nsl = 8
        for i in range(nsl):
            st.replay(table, h, st.n_sets * i // nsl, st.n_sets * (i + 1) // nsl, a.stride, True)
            log(f"ingested {i + 1}/{nsl} ({time.time() - t0:.1f}s)")
        lib.libgint_amd_plan_end(h)
        t_ing = time.time() - t0
        info = (C.c_uint64 * 6)()
        lib.libgint_amd_plan_info(h, info)
        log(f"set-level ingestion {t_ing:.1f}s, resident plan {info[0] / 2**30:.2f} GiB, {info[1]} quartet types, "
            f"{info[2]} chunks")
        # debug bit 2: the engine waits for each quartet type's work and reports
        # it (progress of long K-builds; one host wait per type, ~40 per step)
        lib.libgint_amd_set_debug(h, 4 if a.progress else 0)

        def step(PP=P, out=K):
            lib.libgint_set_P(h, dp(PP), s.P_size)
            lib.libgint_amd_plan_execute(h)
            lib.libgint_get_K(h, dp(out))

        # warm-up steps count the surviving slots (device counters, untimed)
        lib.libgint_amd_enable_stats(h, 1)
        lib.libgint_amd_reset_counters(h)
        for w in range(a.warmup):
            t = time.perf_counter()
            step()
            log(f"warmup {w} {time.perf_counter() - t:.3f}s")
        stats = D.stats(lib, "libgint_", h)
        slots = stats[1] // max(a.warmup, 1)
        lib.libgint_amd_enable_stats(h, 0)
        lib.libgint_amd_synchronize(h)
        times = []
        for k in range(a.steps):
            lib.libgint_amd_synchronize(h)
            t = time.perf_counter()
            step()
            lib.libgint_amd_synchronize(h)
            times.append(time.perf_counter() - t)
            log(f"step {k} {times[-1]:.3f}s")
        tot = sum(times)

        # mixed precision: set quartets whose calibrated bound x max|P| < tau
        # take the single-precision matrix-core contraction; one FP64 calibration
        # execution (untimed), then the same number of timed steps, K compared
        # with the FP64 steps' K of the same density
        mixed = None
        if a.mixed > 0 and a.steps:
            lib.libgint_amd_set_mixed_precision(h, a.mixed)
            Kc = np.zeros(s.P_size)
            t = time.perf_counter()
            step(out=Kc)
            log(f"mixed precision: calibration step {time.perf_counter() - t:.3f}s")
            mtimes = []
            Km = np.zeros(s.P_size)
            for k in range(a.steps):
                lib.libgint_amd_synchronize(h)
                t = time.perf_counter()
                step(out=Km)
                lib.libgint_amd_synchronize(h)
                mtimes.append(time.perf_counter() - t)
                log(f"mixed step {k} {mtimes[-1]:.3f}s")
            mi = (C.c_double * 4)()
            lib.libgint_amd_mixed_info(h, mi)
            dk = float(np.max(np.abs(Km - K)))
            mixed = {"tau": a.mixed, "s_per_step": round(sum(mtimes) / len(mtimes), 4),
                     "speedup_vs_fp64": round((tot / a.steps) / (sum(mtimes) / len(mtimes)), 4),
                     "max_abs_dK_vs_fp64": dk, "bar_Ha": 1e-8, "routed_set_quartets": int(mi[2]),
                     "set_quartets": int(mi[3]),
                     "path": "single-precision matrix-core contraction of the routed set quartets, the rest FP64"}
            log(f"mixed precision tau {a.mixed}: {mixed['s_per_step']}s per step (x{mixed['speedup_vs_fp64']}), "
                f"routed {int(mi[2])} of {int(mi[3])} set quartets, max|dK| vs FP64 {dk:.3e}")
            lib.libgint_amd_set_mixed_precision(h, 0.0)
        tu = (C.c_double * 5)()
        lib.libgint_amd_tune_info(h, tu)
        tms = (C.c_double * 256)()
        tpk = (C.c_int * 256)()
        lib.libgint_amd_tune_detail(h, tms, tpk)
        top = sorted(((tms[i], i) for i in range(256) if tms[i] > 0), reverse=True)[:12]
        log("autotune: classes by sampled time: " + ", ".join(
            f"({i // 64}{(i // 16) % 4}|{(i // 4) % 4}{i % 4}) {ms:.1f}ms/{'tweko'[tpk[i]]}" for ms, i in top))
        tune = {"classes_templated": int(tu[1]), "classes_workgroup": int(tu[2]), "classes_matrix_core": int(tu[3]),
                "seconds": round(tu[4], 2)}
        lib.libgint_amd_destroy_handle(h)
        return slots, tot, info, t_ing, mixed, tune

    if a.roof_only:
        slots, tot, info, t_ing, mixed, tune = 0, 0.0, [0] * 6, 0.0, None, None
        a.steps = 0
    else:
        slots, tot, info, t_ing, mixed, tune = main_steps()

    # the roofline: a stride sample of the same workload as its own plan,
    # executed once more with every kernel serialised on one stream and
    # timed with HIP events around each launch
    rs = a.roof_stride
    hr, tr = D.setup_handle(lib, "libgint_", s, P=P, max_mem_mib=a.max_mem, c0=c0, timing=True, stats=True)
    lib.libgint_amd_plan_begin(hr)
    st.replay(tr, hr, 0, st.n_sets, a.stride * rs, True)
    lib.libgint_amd_plan_end(hr)
    lib.libgint_amd_set_debug(hr, 2)
    lib.libgint_set_P(hr, dp(P), s.P_size)
    lib.libgint_amd_plan_execute(hr)
    lib.libgint_amd_synchronize(hr)   # first execution: buffers, code objects
    lib.libgint_amd_reset_counters(hr)
    # a marker launch (the math self-test kernel): in a kernel trace of this
    # command, the dispatches after it are the timed execution's
    xm = np.ones(1)
    om = np.zeros(4)
    lib.libgint_amd_math_selftest(dp(xm), dp(xm), 1, dp(om))
    Ki = np.zeros(s.P_size)
    t = time.perf_counter()
    lib.libgint_set_P(hr, dp(P), s.P_size)
    lib.libgint_amd_plan_execute(hr)
    lib.libgint_get_K(hr, dp(Ki))
    t_ser = time.perf_counter() - t
    rstats = D.stats(lib, "libgint_", hr)
    tm = (C.c_double * 12)()
    lib.libgint_amd_timing(hr, tm)
    te = (C.c_double * 3)()
    lib.libgint_amd_timing_eco(hr, te)
    nsr = len(range(0, st.n_sets, a.stride * rs))
    log(f"roofline sample (every {a.stride * rs}th set quartet, {nsr}): serialised instrumented step {t_ser:.3f}s; "
        f"kernel ms fm_r {tm[0]:.0f} tq {tm[1]:.0f} gqs {tm[2]:.0f} fixup {tm[3]:.0f} eco {te[0]:.0f}")
    tms = (C.c_double * 256)()
    tpk = (C.c_int * 256)()
    lib.libgint_amd_tune_detail(hr, tms, tpk)
    top = sorted(((tms[i], i) for i in range(256) if tms[i] > 0), reverse=True)[:16]
    log("roofline sample autotune: classes by sampled time: " + ", ".join(
        f"({i // 64}{(i // 16) % 4}|{(i // 4) % 4}{i % 4}) {ms:.1f}ms/{'tweko'[tpk[i]]}" for ms, i in top))
    for cat, nm in ((0, "fm_r by type L"), (1, "tq by type L"), (2, "gqs by type L"), (4, "eco by type L")):
        det = (C.c_double * 256)()
        lib.libgint_amd_timing_detail(hr, cat, det)
        top = sorted(((det[i], i) for i in range(256) if det[i] > 0), reverse=True)[:12]
        log(f"{nm}: " + ", ".join(f"L{i} {ms:.0f}ms" for ms, i in top))
    lib.libgint_amd_destroy_handle(hr)
    value = slots / (tot / a.steps) if a.steps else None
    names = ["fm_r", "tq", "gqs", "fixup", "eco"]
    kms = [tm[i] for i in range(4)] + [te[0]]
    kfl = [float(rstats[6]), tm[9], tm[10], 0.0, te[2]]
    nl = [tm[4 + i] for i in range(4)] + [te[1]]
    dom = int(np.argmax(kms))
    achieved = kfl[dom] / (kms[dom] * 1e-3) / 1e12 if kms[dom] > 0 else 0.0
    traffic, traffic_src = measured_traffic(names[dom], s.describe())
    roof = {"kernel": names[dom], "bound": "fp64-valu", "achieved": round(achieved, 4), "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 5), "traffic": traffic,
            "traffic_source": traffic_src, "launches": nl[dom],
            "avg_launch_ms": round(kms[dom] / max(nl[dom], 1), 4),
            "sample": f"every {a.stride * rs}th set quartet of the workload ({nsr}) as its own resident plan",
            "timing": "HIP events around each launch on its stream, every launch serialised on one stream",
            "kernel_ms_serialised": {n: round(m, 2) for n, m in zip(names, kms)},
            "kernel_tflops": {n: round(f / (m * 1e-3) / 1e12, 4) if m > 0 else 0 for n, f, m in zip(names, kfl, kms)},
            "flops": "fm_r: the reference chain's Fm + VRR FLOPs of the kept (slot, class) pairs; tq: "
                     "contraction 2 n_prm ncp + HRR + SPH + 9 Ns per contracted quartet; eco (matrix cores): the "
                     "contraction 2 n_prm ncp of the classes with many contractions; gqs: the rest of those "
                     "classes and the contraction of the others"}

    # max|dK| and the CPU baseline: a strided sample of the same stream,
    # through the product (stream mode) and the FP64 oracle on all cores
    check = cpu = None
    if not a.no_cpu:
        import oracle

        orc = oracle.load()
        cores = len(os.sched_getaffinity(0))
        omp = int(os.environ.get("OMP_NUM_THREADS", cores))

        def run_oracle(stride):
            ho, to = D.setup_handle(orc, "oracle_", s, P=P, c0=c0)
            t = time.perf_counter()
            st.replay(to, ho, 0, st.n_sets, stride, True)
            Ko = np.zeros(s.P_size)
            orc.oracle_get_K(ho, dp(Ko))
            tc = time.perf_counter() - t
            so = D.stats(orc, "oracle_", ho)
            orc.oracle_destroy_handle(ho)
            return Ko, so, tc

        cs = a.check_stride
        if not cs:
            cs = max(1, st.n_sets // 400)
            _, so, tc = run_oracle(cs)
            if tc < 3.0:
                cs = max(1, int(cs * max(tc, 0.05) / 10.0))
        runs = [run_oracle(cs) for _ in range(3)]
        Ko, so, _ = runs[0]
        tcs = sorted(r[2] for r in runs)
        tc = tcs[1]
        hs, ts = D.setup_handle(lib, "libgint_", s, P=P, max_mem_mib=a.max_mem, c0=c0, stats=True)
        st.replay(ts, hs, 0, st.n_sets, cs, True)
        Kp = np.zeros(s.P_size)
        lib.libgint_get_K(hs, dp(Kp))
        sp = D.stats(lib, "libgint_", hs)
        lib.libgint_amd_destroy_handle(hs)
        err = float(np.max(np.abs(Kp - Ko)))
        nsmp = len(range(0, st.n_sets, cs))
        check = {"max_abs_dK": err, "max_abs_K": float(np.max(np.abs(Ko))), "bar_Ha": 1e-8,
                 "survivors_equal": bool(sp[1] == so[1] and sp[4] == so[4]),
                 "sample": f"every {cs}th set quartet ({nsmp}), FP64 GPU path vs the FP64 oracle"}
        cpu = {"value": round(so[1] / tc, 1), "unit": "quartets/s", "cores": omp, "kind": "port",
               "sample": f"every {cs}th set quartet of the workload ({nsmp} set quartets, {so[1]} surviving slots); "
                         f"median of 3 runs {tc:.2f} s (runs {', '.join('%.2f' % x for x in tcs)}); "
                         f"{omp} OpenMP threads of {cores} cores, {cpu_model()}"}
        log(f"max|dK| {err:.3e} on {check['sample']}; cpu baseline {cpu['sample']}")

    out = {
        "metric": METRIC, "value": round(value, 1) if value else None, "unit": "quartets/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * tot / a.steps, 3) if a.steps else None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": s.describe(), "config_index": a.config, "atoms": s.natom, "nbf": s.nbf,
                   "cells": s.Ng, "set_quartets": n_sets, "stride": a.stride,
                   "primitive_quartets": st.n_prm if a.stride == 1 else None,
                   "contracted_quartets": st.n_qrtt if a.stride == 1 else None,
                   "surviving_slots_per_step": int(slots),
                   "kbuild_s_per_step": round(tot / a.steps, 4) if a.steps else None,
                   "hash_sample": a.sample,
                   "kbuild_s_per_step_whole_workload_est": (round(a.sample * tot / a.steps, 2)
                                                            if a.steps and a.sample > 1 else None),
                   "abi_ingestion": abi, "set_level_ingestion_s": round(t_ing, 2),
                   "resident_plan_GiB": round(info[0] / 2**30, 3),
                   "path": "FP64, resident plan, fused Hermite cell sum per quartet type"},
        "max_dK": check, "mixed_precision": mixed, "autotune": tune, "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
