"""K-build benchmark on one MI355X.

python bench.py [--config C] [--steps K] [--warmup W] [--gpus 1]

A step is one truncated-Coulomb exchange (K) build: set_P, execution of the
workload's screened call stream (recorded once through the libgint_* C-ABI
and kept resident in HBM), get_K (H2D of P and D2H of K included).  The
metric counts surviving (primitive quartet, lattice cell, class) slots
(Fm[0] > 1e-12, the reference's per-class recurrence work unit) per second;
K-build seconds per step is ms_per_step.  The CPU baseline is the FP64
oracle on a strided sample of the same stream, on this host.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector and FP64 matrix, dense
PEAK_HBM_GBS = 8000.0


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--cpu-sample-sets", type=int, default=0, help="sets in the CPU sample (0: auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--c0-file", default=None, help="CP2K t_c_g.dat to use in place of the generated table")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 3 device copies of MIB MiB (calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--max-mem", type=int, default=65536, help="device budget, MiB")
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

    lib = load()
    c0 = read_c0(a.c0_file, 21) if a.c0_file else None
    t0 = time.time()
    s = D.System(a.config)
    st = s.record()
    t_rec = time.time() - t0
    print(f"# config {a.config}: atoms {s.natom} nbf {s.nbf} cells {s.Ng} sets {st.n_sets} rows {st.n_rows} "
          f"primitive quartets {st.n_prm} contracted quartets {st.n_qrtt} (screened in {t_rec:.1f}s)",
          file=sys.stderr, flush=True)
    P = s.P()
    h, table = D.setup_handle(lib, "libgint_", s, P=P, max_mem_mib=a.max_mem, stats=True, timing=True, c0=c0)
    K = np.zeros(s.P_size)
    dp = D._ptr
    # the screened call stream goes through the ABI once and stays resident
    t0 = time.time()
    lib.libgint_amd_plan_begin(h)
    nsl = 16
    for i in range(nsl):
        st.replay(table, h, st.n_sets * i // nsl, st.n_sets * (i + 1) // nsl)
        print(f"# ingested {i + 1}/{nsl} ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    lib.libgint_amd_plan_end(h)
    t_ing = time.time() - t0
    info = (C.c_uint64 * 6)()
    lib.libgint_amd_plan_info(h, info)
    print(f"# ABI ingestion {t_ing:.1f}s, resident plan {info[0] / 2**30:.2f} GiB, {info[1]} quartet types, "
          f"{info[2]} chunks", file=sys.stderr, flush=True)

    def step():
        lib.libgint_set_P(h, dp(P), s.P_size)
        lib.libgint_amd_plan_execute(h)
        lib.libgint_get_K(h, dp(K))

    for w in range(a.warmup):
        step()
        print(f"# warmup {w} done", file=sys.stderr, flush=True)
    lib.libgint_amd_synchronize(h)
    lib.libgint_amd_reset_counters(h)
    times = []
    for k in range(a.steps):
        lib.libgint_amd_synchronize(h)
        t = time.perf_counter()
        step()
        lib.libgint_amd_synchronize(h)
        times.append(time.perf_counter() - t)
        print(f"# step {k} {times[-1]:.3f}s", file=sys.stderr, flush=True)
    stats = D.stats(lib, "libgint_", h)
    tm = (C.c_double * 12)()
    lib.libgint_amd_timing(h, tm)
    tot = sum(times)
    for cat, nm in ((0, "fm_r by type L"), (1, "tq by type L"), (2, "gqs by type L")):
        det = (C.c_double * 256)()
        lib.libgint_amd_timing_detail(h, cat, det)
        top = sorted(((det[i] / a.steps, i) for i in range(256) if det[i] > 0), reverse=True)[:12]
        lab = lambda i: f"L{i}"
        print(f"# {nm}: " + ", ".join(f"{lab(i)} {ms:.1f}ms" for ms, i in top), file=sys.stderr, flush=True)
    slots = stats[1] / a.steps          # surviving (slot, class) pairs per K-build
    value = slots / (tot / a.steps)
    names = ["fm_r", "tq", "gqs", "fixup"]
    cat = [0, 1, 2, 3]
    kms = [tm[i] / a.steps for i in cat]
    kfl = [stats[6] / a.steps, tm[9] / a.steps, tm[10] / a.steps, 0.0]
    nl = [tm[4 + i] / a.steps for i in cat]
    dom = int(np.argmax(kms))
    # average duration of one launch of the dominant kernel, and its FLOPs per launch
    per_launch_ms = kms[dom] / max(nl[dom], 1)
    achieved = kfl[dom] / (kms[dom] * 1e-3) / 1e12 if kms[dom] > 0 else 0.0
    traffic, traffic_src = measured_traffic(names[dom], s.describe())
    roof = {"kernel": names[dom], "bound": "mfma", "achieved": round(achieved, 4), "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 5), "traffic": traffic,
            "traffic_source": traffic_src,
            "launches_per_step": nl[dom], "avg_launch_ms": round(per_launch_ms, 4),
            "kernel_ms": {n: round(m, 3) for n, m in zip(names, kms)},
            "kernel_tflops": {n: round(f / (m * 1e-3) / 1e12, 4) if m > 0 else 0 for n, f, m in zip(names, kfl, kms)}}

    cpu = None
    if not a.no_cpu:
        import oracle

        orc = oracle.load()
        # a strided sample of the same stream, sized to 10-30 s of one core:
        # a short probe first, then the sample the probe's rate allows
        def run_sample(stride):
            ho, to = D.setup_handle(orc, "oracle_", s, P=P, c0=c0)
            t = time.perf_counter()
            st.replay(to, ho, 0, st.n_sets, stride)
            tc = time.perf_counter() - t
            so = D.stats(orc, "oracle_", ho)
            orc.oracle_destroy_handle(ho)
            return so, tc

        stride = max(1, st.n_sets // (a.cpu_sample_sets or 200))
        so, tc = run_sample(stride)
        if not a.cpu_sample_sets and tc < 5.0:
            sets_per_s = (st.n_sets // stride) / max(tc, 1e-3)
            # (the probe's set quartets run about twice as fast as the average)
            stride = max(1, int(st.n_sets / max(1.0, 30.0 * sets_per_s)))
            so, tc = run_sample(stride)
        cpu = {"value": round(so[1] / tc, 1), "unit": "quartets/s", "cores": 1, "kind": "port",
               "sample": f"every {stride}th set quartet of the stream ({st.n_sets // stride} set quartets, "
                         f"{so[1]} surviving slots, {tc:.1f} s on one core)"}
        print(f"# cpu baseline: {cpu['sample']}", file=sys.stderr, flush=True)

    out = {
        "metric": "K-build s/SCF step + quartets/s at max|dK|<=1e-8 Ha; p50/p99 request latency",
        "value": round(value, 1), "unit": "quartets/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(1e3 * tot / a.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": s.describe(), "atoms": s.natom, "nbf": s.nbf, "cells": s.Ng,
                   "set_quartets": st.n_sets, "primitive_quartets": st.n_prm, "contracted_quartets": st.n_qrtt,
                   "surviving_slots": int(slots), "abi_ingestion_s": round(t_ing, 2),
                   "path": "FP64, resident plan, fused Hermite cell sum per quartet type"},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
