#!/usr/bin/env python3
"""Benchmark of the pairwise alignment hot path on one MI355X.

One step = one pass of the path over one batch of synthetic, seeded input that
is already resident in HBM:

  headline  Smith-Waterman, score + span + CIGAR: 50,000 pairs, query 256 bp
            i.i.d. ACGT, target 1,024 bp i.i.d. ACGT with a 10%-mutated copy
            of the query (sub:ins:del = 2:1:1) written in at a uniform offset;
            match 2 / mismatch -3 / gap open 3 / gap extend 1.  value = GCUPS
            (cells = sum of text_len * pattern_len per step / step time).
  secondary WFA edit distance, score + CIGAR: 10,000 pairs, pattern 150 bp,
            text = pattern with 2% per-base edits (2:1:1), smax 64 (pairs/s).
  wfa_mixed WFA over 200,000 pairs, pattern lengths log-uniform 100 bp..10 kbp,
            per-pair divergence uniform 0..15% (2:1:1), smax = min(0.2 len,
            2048), 70% of pairs with CIGAR and 30% score-only (pairs/s), with
            the oracle timed on a random sample of the same pairs.

Also printed on the same JSON line: the dominant kernel's roofline (timed live
with HIP events on the launch stream), per-step latency p50/p99, the number of
results that differ from the CPU oracle on a checked sample, and the oracle
(the CPU restatement of the reference path) timed on this host's cores on a
bounded sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BASELINE_METRIC = "WFA pairs/s + SW GCUPS on 1 MI355X, routed vs plain, 0 mismatches; p99 ms"
SW_SCHEME = (2, -3, 3, 1)
SW_OPS_PER_CELL = 12            # E: 2 sub + 1 max; F: 2 sub + 1 max; H: add + 3 max; best: 1 max
PEAK_INT16_TOPS = 78.6          # gfx950 packed-int16 VALU: 256 CU x 2.4 GHz x 64 lanes x 2
PEAK_HBM_GBS = 8000.0
PEAK_INT32_TOPS = 39.3          # gfx950 int32 VALU: 256 CU x 2.4 GHz x 64 lanes
# WFA, per diagonal update: max(ins, mis), +1, max with del, v = mx - k, two
# bound checks, and the first 16-base compare (xor + test)
WFA_OPS_PER_UPDATE = 8


def wfa_valu_roofline(ps, scores, kernel_ms):
    """Diagonal updates of a WFA batch, sum over pairs of (w + 1)^2 with w the
    score (resolved), smax (unresolved after a full pass) or nothing (length
    difference above smax), against the int32 VALU peak."""
    a = ps.a_len.astype(np.int64)
    b = ps.b_len.astype(np.int64)
    sm = ps.smax.astype(np.int64)
    sc = np.asarray(scores, dtype=np.int64)
    w = np.where(sc >= 0, sc, np.where(np.abs(a - b) <= sm, sm, -1))
    upd = int(((w + 1) ** 2).sum())
    if not kernel_ms:
        return {"bound": "valu", "diagonal_updates": upd}
    tops = WFA_OPS_PER_UPDATE * upd / (kernel_ms * 1e-3) / 1e12
    return {"bound": "valu", "diagonal_updates": upd, "ops_per_update": WFA_OPS_PER_UPDATE,
            "achieved_TOPs": round(tops, 3), "peak_TOPs": PEAK_INT32_TOPS,
            "frac": round(tops / PEAK_INT32_TOPS, 4)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pctl(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else None


def time_steps(run, steps, warmup, lib):
    import torch
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    kern = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record()
        run()
        b.record()
        kern.append(lib.ga_last_kernel_ms())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = [a.elapsed_time(b) for a, b in ev]
    return wall, step_ms, kern


def check_sw(ps, dev, sample, oracle):
    out, cig = dev.results()
    bad = 0
    for i in sample:
        r = oracle.sw_trace(ps.a_seq(i), ps.b_seq(i), SW_SCHEME)
        if tuple(int(x) for x in out[i]) != r[:5] or cig[i].decode() != r[5]:
            bad += 1
    return bad


def check_wfa(ps, dev, sample, oracle):
    sc, cig = dev.results()
    bad = 0
    for i in sample:
        osc, ocg = oracle.wfa(ps.a_seq(i), ps.b_seq(i), int(ps.smax[i]))
        if int(sc[i]) != osc or cig[i].decode() != ocg:
            bad += 1
    return bad


def cpu_baseline_sw(ps, seconds, threads, oracle, tab=None, gaps=None,
                    label="256 x 1024, score + span + CIGAR"):
    """The oracle's SW trace path on `threads` host threads over a sample of the
    same workload, sized to run about `seconds` (score table and gap costs of
    the DNA scheme unless given)."""
    L = oracle.lib()
    if tab is None:
        tab = oracle.mm_table(SW_SCHEME[0], SW_SCHEME[1])
    go, ge = gaps if gaps else (SW_SCHEME[2], SW_SCHEME[3])

    def run(n):
        sub = ps.subset(np.arange(n))
        out = np.zeros(5 * n, dtype=np.int32)
        cig = np.zeros(int((sub.a_len.astype(np.int64) + sub.b_len).sum()) + 1, dtype=np.uint8)
        coff = np.zeros(n, dtype=np.int64)
        coff[1:] = np.cumsum(sub.a_len.astype(np.int64) + sub.b_len)[:-1]
        clen = np.zeros(n, dtype=np.int32)
        t0 = time.perf_counter()
        rc = L.ga_oracle_sw_batch(sub.a.ctypes.data, sub.a_off.ctypes.data, sub.a_len.ctypes.data,
                                  sub.b.ctypes.data, sub.b_off.ctypes.data, sub.b_len.ctypes.data,
                                  n, tab.ctypes.data, go, ge, 1,
                                  out.ctypes.data, cig.ctypes.data, coff.ctypes.data,
                                  clen.ctypes.data, threads)
        assert rc == 0
        return time.perf_counter() - t0, int((sub.a_len.astype(np.int64) * sub.b_len).sum())

    # calibrate on growing samples until one takes ~1 s, then size the timed run
    probe = max(threads * 4, 32)
    dt, _ = run(probe)
    while dt < 1.0 and probe < ps.n:
        probe = int(min(ps.n, probe * max(2.0, 1.2 / max(dt, 1e-3))))
        dt, _ = run(probe)
    n = int(min(ps.n, max(probe, probe * seconds / max(dt, 1e-3))))
    dt, cells = run(n)
    return {"value": round(cells / dt / 1e9, 4), "unit": "GCUPS", "cores": threads,
            "kind": "port",
            "sample": f"{n} of the {ps.n} pairs ({label}), "
                      f"{dt:.1f} s on {threads} threads of the benchmark host"}


def cpu_baseline_wfa(ps, want, seconds, threads, oracle):
    """The oracle's WFA batch (score, CIGAR where wanted) on `threads` host
    threads over a random sample of the pairs, sized to run about `seconds`."""
    L = oracle.lib()
    rng = np.random.default_rng(3)
    perm = rng.permutation(ps.n)

    def run(n):
        idx = np.sort(perm[:n])
        sub = ps.subset(idx)
        w = np.ascontiguousarray(want[idx], dtype=np.uint8)
        slot = sub.a_len.astype(np.int64) + sub.b_len
        cig = np.zeros(int(slot.sum()) + 1, dtype=np.uint8)
        coff = np.zeros(n, dtype=np.int64)
        coff[1:] = np.cumsum(slot)[:-1]
        sc = np.zeros(n, dtype=np.int32)
        clen = np.zeros(n, dtype=np.int32)
        smax = np.ascontiguousarray(sub.smax, dtype=np.int32)
        t0 = time.perf_counter()
        rc = L.ga_oracle_wfa_batch(sub.a.ctypes.data, sub.a_off.ctypes.data, sub.a_len.ctypes.data,
                                   sub.b.ctypes.data, sub.b_off.ctypes.data, sub.b_len.ctypes.data,
                                   smax.ctypes.data, w.ctypes.data, n, sc.ctypes.data,
                                   cig.ctypes.data, coff.ctypes.data, clen.ctypes.data, threads)
        assert rc == 0
        return time.perf_counter() - t0

    probe = max(threads * 8, 64)
    dt = run(probe)
    while dt < 1.0 and probe < ps.n:
        probe = int(min(ps.n, probe * max(2.0, 1.2 / max(dt, 1e-3))))
        dt = run(probe)
    n = int(min(ps.n, max(probe, probe * seconds / max(dt, 1e-3))))
    dt = run(n)
    return {"value": round(n / dt, 1), "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"{n} pairs drawn at random from the {ps.n}, {dt:.1f} s on "
                      f"{threads} threads of the benchmark host"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-wfa", action="store_true")
    ap.add_argument("--no-wfa-mixed", action="store_true")
    ap.add_argument("--wfa-mixed-pairs", type=int, default=200_000)
    ap.add_argument("--no-routing", action="store_true")
    ap.add_argument("--routing-pairs", type=int, default=200_000)
    ap.add_argument("--router-weights", default=None,
                    help="router parameters (.npy); default: the trained file shipped in the package")
    ap.add_argument("--route-q", type=float, default=0.95)
    ap.add_argument("--no-protein", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-fasta", action="store_true")
    ap.add_argument("--no-gene-qc", action="store_true")
    ap.add_argument("--gene-qc-records", type=int, default=20000)
    ap.add_argument("--gene-qc-cpu-records", type=int, default=300)
    ap.add_argument("--fasta-mib", type=int, default=512)
    ap.add_argument("--serve-rate", type=float, default=100.0, help="requests per second (Poisson)")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a kernel of "
                         "known traffic, to scale hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: there is no multi-GPU path "
              f"(--gpus {args.gpus} is not supported; pairs are independent, so a batch would be "
              f"split by cost into one contiguous range per GPU)", file=sys.stderr)
        sys.exit(2)

    import torch
    from genoaligner_amd import _native as N
    from genoaligner_amd import workloads as W
    from genoaligner_amd.device import SwDeviceBatch, WfaDeviceBatch
    from oracle import oracle as O

    lib = N.lib()
    lib.ga_set_kernel_timing(1)
    torch.cuda.init()
    props = torch.cuda.get_device_properties(0)
    dev_name = f"{torch.cuda.get_device_name(0)} ({getattr(props, 'gcnArchName', '?')})"
    log(f"device: {dev_name}")

    if args.copy_probe_mib > 0:
        nbytes = args.copy_probe_mib << 20
        src = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {nbytes} bytes read + {nbytes} bytes written")
        del src, dst

    # ---------------- SW headline ----------------
    t0 = time.perf_counter()
    sw = W.sw_embedded(n=50_000, qlen=256, tlen=1024, rate=0.10)
    cells = int((sw.a_len.astype(np.int64) * sw.b_len).sum())
    swd = SwDeviceBatch(sw, scoring=SW_SCHEME, trace=True)
    log(f"sw workload ready ({time.perf_counter() - t0:.1f} s), {cells / 1e9:.2f} Gcells/step")
    wall, step_ms, kern = time_steps(swd.run, args.steps, args.warmup, lib)
    ms_per_step = wall * 1000.0 / args.steps
    gcups = cells / (wall / args.steps) / 1e9
    kern_ms = float(np.mean([k for k in kern if k > 0])) if any(k > 0 for k in kern) else None
    log(f"sw: {ms_per_step:.2f} ms/step, {gcups:.1f} GCUPS, forward kernel {kern_ms} ms")
    rng = np.random.default_rng(1)
    sample = rng.choice(sw.n, size=min(300, sw.n), replace=False)
    sw_bad = check_sw(sw, swd, sample, O)
    log(f"sw: {sw_bad} mismatches in {len(sample)} checked pairs")

    # HBM bytes per launch of the SW trace kernel, measured with rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes scaled by a copy of known size
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "sw16_trace_sw_embedded_50k_hbm_traffic.md")
    if os.path.exists(tpath):
        import re
        mt = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(tpath).read())
        traffic = float(mt.group(1)) if mt else None
    roof = None
    if kern_ms:
        achieved = SW_OPS_PER_CELL * cells / (kern_ms * 1e-3) / 1e12
        roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_INT16_TOPS,
                "unit": "TOP/s", "frac": round(achieved / PEAK_INT16_TOPS, 4),
                "traffic": traffic, "kernel": "sw16_kernel<true, false, true> (packed forward + fused walk)",
                "kernel_ms": round(kern_ms, 3)}

    # score-only on the same pairs: score + end cell (the lane-rule end)
    swd_s = SwDeviceBatch(sw, scoring=SW_SCHEME, trace=False)
    swall, sstep, skern = time_steps(swd_s.run, args.steps, args.warmup, lib)
    sk = float(np.mean([k for k in skern if k > 0])) if any(k > 0 for k in skern) else None
    out_s, _ = swd_s.results()
    s_bad = 0
    for i in sample[:100]:
        s_, ei, ej = O.sw_score_lane(sw.a_seq(i), sw.b_seq(i), SW_SCHEME)
        if tuple(int(x) for x in out_s[i]) != (s_, -1, -1, ei, ej):
            s_bad += 1
    sw_score_line = {
        "workload": "same pairs, score + end cell only",
        "gcups": round(cells / (swall / args.steps) / 1e9, 3),
        "ms_per_step": round(swall * 1000.0 / args.steps, 3),
        "kernel_ms": round(sk, 4) if sk else None,
        "roofline_frac": round(SW_OPS_PER_CELL * cells / (sk * 1e-3) / 1e12 / PEAK_INT16_TOPS, 4) if sk else None,
        "mismatches": s_bad, "checked": 100,
    }
    log(f"sw score-only: {sw_score_line}")
    del swd_s

    # ---------------- WFA secondary ----------------
    wfa_line = None
    if not args.no_wfa:
        wf = W.wfa_plain(n=10_000, length=150, rate=0.02, smax=64)
        wfd = WfaDeviceBatch(wf, trace=True)
        wwall, wstep, wkern = time_steps(wfd.run, args.steps, args.warmup, lib)
        wsample = rng.choice(wf.n, size=min(1000, wf.n), replace=False)
        wbad = check_wfa(wf, wfd, wsample, O)
        wk = float(np.mean([k for k in wkern if k > 0])) if any(k > 0 for k in wkern) else None
        in_bytes = int(((wf.a_len.astype(np.int64) + 3) // 4 + (wf.b_len.astype(np.int64) + 3) // 4).sum())
        sc, cg = wfd.results()
        out_bytes = 5 * wf.n + int(sum(len(c) for c in cg))
        # the same batch resubmitted as a captured HIP graph (device-planned)
        gb = WfaDeviceBatch(wf, trace=True, classes=False)
        graph = gb.graph()
        gwall, gstep, _ = time_steps(graph.launch, args.steps, args.warmup, lib)
        gsc, gcg = gb.results()
        gdiff = int(sum(1 for i in range(wf.n) if int(gsc[i]) != int(sc[i]) or gcg[i] != cg[i]))
        graph.close()
        del gb
        wfa_line = {
            "workload": "wfa_plain 10000 x 150 bp, 2% edits (2:1:1), smax 64, score + CIGAR",
            "pairs_per_s": round(wf.n / (wwall / args.steps), 1),
            "ms_per_step": round(wwall * 1000.0 / args.steps, 3),
            "graph_pairs_per_s": round(wf.n / (gwall / args.steps), 1),
            "graph_ms_per_step": round(gwall * 1000.0 / args.steps, 3),
            "graph_vs_eager_differences": gdiff,
            "p50_ms": pctl(wstep, 50), "p99_ms": pctl(wstep, 99),
            "kernel_ms": round(wk, 4) if wk else None,
            "hbm_roofline": {"bound": "hbm", "algorithmic_bytes": in_bytes + out_bytes,
                             "achieved_GBs": round((in_bytes + out_bytes) / (wk * 1e-3) / 1e9, 2) if wk else None,
                             "peak_GBs": PEAK_HBM_GBS},
            "valu_roofline": wfa_valu_roofline(wf, sc, wk),
            "mismatches": wbad, "checked": int(len(wsample)),
            "tcups_ref_style": round(float((wf.a_len.astype(np.int64) * wf.b_len).sum())
                                     / (wwall / args.steps) / 1e12, 4),
        }
        log(f"wfa: {wfa_line['pairs_per_s']:.0f} pairs/s, {wbad} mismatches")
        if not args.no_cpu_baseline:
            wfa_line["cpu_baseline"] = cpu_baseline_wfa(wf, np.ones(wf.n, dtype=np.uint8), args.cpu_seconds,
                                                        min(os.cpu_count() or 1, 16), O)
            log(f"wfa cpu baseline: {wfa_line['cpu_baseline']}")

    # ---------------- WFA, mixed lengths / distances ----------------
    mixed_line = None
    if not args.no_wfa_mixed:
        t0 = time.perf_counter()
        wm = W.wfa_mixed(n=args.wfa_mixed_pairs)
        want = (np.random.default_rng(4).random(wm.n) < 0.7).astype(np.uint8)
        wmd = WfaDeviceBatch(wm, trace=True, want=want)
        log(f"wfa_mixed workload ready ({time.perf_counter() - t0:.1f} s)")
        # the autotuner picks waves per workgroup for each size class on the
        # batch itself (a sample has too few pairs per workgroup in the large
        # classes to rank the choices as the whole batch does), then the batch
        # runs with that choice
        from genoaligner_amd.device import wfa_autotune, wfa_set_tuning
        tuned_waves, tuned_ms = wfa_autotune(wmd, reps=1)
        log(f"wfa_mixed autotune: waves per class {tuned_waves}, ms {tuned_ms}")
        mwall, mstep, mkern = time_steps(wmd.run, args.steps, args.warmup, lib)
        msample = rng.choice(wm.n, size=min(300, wm.n), replace=False)
        sc, cg = wmd.results()
        mbad = 0
        for i in msample:
            osc, ocg = O.wfa(wm.a_seq(i), wm.b_seq(i), int(wm.smax[i]))
            if int(sc[i]) != osc or cg[i].decode() != (ocg if want[i] else ""):
                mbad += 1
        mk = float(np.mean([k for k in mkern if k > 0])) if any(k > 0 for k in mkern) else None
        mixed_line = {
            "workload": f"wfa_mixed {wm.n} pairs, lengths log-uniform 100..10000 bp, divergence "
                        "0..15% (2:1:1), smax min(0.2 len, 2048), 70% CIGAR / 30% score-only",
            "pairs_per_s": round(wm.n / (mwall / args.steps), 1),
            "ms_per_step": round(mwall * 1000.0 / args.steps, 3),
            "p50_ms": pctl(mstep, 50), "p99_ms": pctl(mstep, 99),
            "kernel_ms": round(mk, 3) if mk else None,
            "valu_roofline": wfa_valu_roofline(wm, sc, mk),
            "resolved": int((sc >= 0).sum()), "mismatches": mbad, "checked": int(len(msample)),
            "autotuned_waves_per_class": tuned_waves,
        }
        log(f"wfa_mixed: {mixed_line}")
        if not args.no_cpu_baseline:
            mixed_line["cpu_baseline"] = cpu_baseline_wfa(wm, want, args.cpu_seconds,
                                                          min(os.cpu_count() or 1, 16), O)
            log(f"wfa_mixed cpu baseline: {mixed_line['cpu_baseline']}")
        del wmd
        wfa_set_tuning(None)

    # ---------------- WFA routing: plain per-class path vs learned router ----------------
    routing_line = None
    if not args.no_routing:
        from genoaligner_amd import router as R
        t0 = time.perf_counter()
        wr = W.wfa_routing(n=args.routing_pairs)
        rwant = (np.random.default_rng(6).random(wr.n) < 0.7).astype(np.uint8)
        params = torch.from_numpy(R.load(args.router_weights or R.WEIGHTS)).cuda()
        plain = WfaDeviceBatch(wr, trace=True, want=rwant)
        routed = WfaDeviceBatch(wr, trace=True, want=rwant)
        log(f"wfa_routing workload ready ({time.perf_counter() - t0:.1f} s)")
        pwall, pstep, _ = time_steps(plain.run, args.steps, args.warmup, lib)
        stats = []
        rwall, rstep, _ = time_steps(lambda: stats.append(R.run_routed(routed, params, q=args.route_q)),
                                     args.steps, args.warmup, lib)
        psc, pcg = plain.results()
        rsc, rcg = routed.results()
        diff = int(sum(1 for i in range(wr.n) if psc[i] != rsc[i] or pcg[i] != rcg[i]))
        rsample = rng.choice(wr.n, size=min(200, wr.n), replace=False)
        obad = 0
        for i in rsample:
            osc, ocg = O.wfa(wr.a_seq(i), wr.b_seq(i), int(wr.smax[i]))
            if int(rsc[i]) != osc or rcg[i].decode() != (ocg if rwant[i] else ""):
                obad += 1
        st = stats[-1]
        routing_line = {
            "workload": f"wfa_routing {wr.n} pairs, lengths log-uniform 100..10000 bp, divergence 80% "
                        "U[0,0.03] / 15% U[0.03,0.15] / 5% unrelated, smax min(0.2 len, 2048), "
                        "70% CIGAR / 30% score-only",
            "plain_pairs_per_s": round(wr.n / (pwall / args.steps), 1),
            "routed_pairs_per_s": round(wr.n / (rwall / args.steps), 1),
            "routed_over_plain": round(pwall / rwall, 3),
            "plain_ms": round(pwall * 1000 / args.steps, 3), "routed_ms": round(rwall * 1000 / args.steps, 3),
            "routed_p50_ms": pctl(rstep, 50), "routed_p99_ms": pctl(rstep, 99),
            "routed_vs_plain_differences": diff, "oracle_mismatches": obad, "oracle_checked": int(len(rsample)),
            "shortcut": st.shortcut, "routed": st.routed, "plain": st.plain, "rerun": st.rerun,
            "miss_rate": round(st.miss_rate, 4), "route_q": args.route_q,
        }
        log(f"wfa_routing: {routing_line}")
        del plain, routed

    # ---------------- protein SW (BLOSUM62, gap open 11 / extend 1), host entry point ----------------
    protein_line = None
    if not args.no_protein:
        import genoaligner_amd as ga
        pp = W.protein_pairs(n=2000)
        m62 = ga.blosum62(11, 1)
        preqs = [ga.SWRequest(text=pp.a_seq(i), text_len=int(pp.a_len[i]), pattern=pp.b_seq(i),
                              pattern_len=int(pp.b_len[i]), with_cigar=True) for i in range(pp.n)]
        pcells = int((pp.a_len.astype(np.int64) * pp.b_len).sum())
        ga.align_sw_matrix_batch(preqs[:64], m62)
        # the C entry point as a C++ caller uses it: request array and arena
        # built before the clock, as the caller's own data would be
        import ctypes as C
        from genoaligner_amd import api as A, _native as N
        keep: list = []
        parr = (N.SWReq * pp.n)()
        for i, r in enumerate(preqs):
            parr[i] = N.SWReq(A._buf(A._bytes(r.text), keep), int(r.text_len), A._buf(A._bytes(r.pattern), keep),
                              int(r.pattern_len), A._scoring_c(r.scoring), 1)
        pres = (N.SWRes * pp.n)()
        pneed = lib.ga_cigar_arena_bytes_sw(parr, pp.n)
        parena = C.create_string_buffer(max(pneed, 1))
        pok, pun, perr = C.c_int32(), C.c_int32(), C.c_char_p()
        # one untimed full-size call (it sizes the library's reusable device
        # workspace), then the median of timed calls
        pts = []
        for rep in range(1 + max(3, min(args.steps, 5))):
            t0 = time.perf_counter()
            prc = lib.ga_align_sw_matrix_batch(parr, pp.n, C.byref(m62), pres, parena, pneed, C.byref(pok),
                                               C.byref(pun), C.byref(perr))
            if rep > 0:
                pts.append(time.perf_counter() - t0)
        pdt = float(np.median(pts))
        t0 = time.perf_counter()
        pb = ga.align_sw_matrix_batch(preqs, m62)
        pdt_py = time.perf_counter() - t0
        tab = O.matrix_table(m62.alphabet.decode(), [[m62.score[a][b] for b in range(20)] for a in range(20)], -1)
        pbad = 0 if prc == N.GA_OK else pp.n
        raw = parena.raw
        for i in rng.choice(pp.n, size=20, replace=False):
            r = pres[i]
            cg = raw[r.cigar_off: r.cigar_off + r.cigar_len].decode("latin-1") if r.cigar_len > 0 else ""
            e = O.sw_trace(pp.a_seq(i), pp.b_seq(i), tab, 11, 1)
            if (r.score, r.start_i, r.start_j, r.end_i, r.end_j, cg) != e:
                pbad += 1
            q = pb.results[i]
            if (q.score, q.start_i, q.start_j, q.end_i, q.end_j, q.cigar) != e:
                pbad += 1
        protein_line = {
            "workload": "protein_pairs 2000, target 100..2000 aa x query 100..1000 aa, BLOSUM62 background "
                        "frequencies, BLOSUM62 gap open 11 / extend 1, score + span + CIGAR, C entry point "
                        "ga_align_sw_matrix_batch (includes host packing, PCIe copies and CIGAR validation; "
                        "median of timed calls after one untimed full-size call)",
            "gcups": round(pcells / pdt / 1e9, 3), "ms": round(pdt * 1e3, 2), "cells": pcells,
            "python_api_ms": round(pdt_py * 1e3, 2),
            "mismatches": pbad, "checked": 20,
        }
        log(f"protein: {protein_line}")
        if not args.no_cpu_baseline:
            protein_line["cpu_baseline"] = cpu_baseline_sw(
                pp, args.cpu_seconds, min(os.cpu_count() or 1, 16), O, tab=tab, gaps=(11, 1),
                label="protein, BLOSUM62 11/1, score + span + CIGAR")
            log(f"protein cpu baseline: {protein_line['cpu_baseline']}")

    # ---------------- FASTA ingest on the device ----------------
    fasta_line = None
    if not args.no_fasta:
        from genoaligner_amd.fasta import fasta_to_device
        fa = W.fasta_genome(args.fasta_mib << 20, seed=11)
        fa_dev = torch.from_numpy(fa).cuda()
        fasta_to_device(fa_dev)                        # warm-up (kernels, allocator)
        torch.cuda.synchronize()
        fa_ms = []
        for _ in range(5):
            t0 = time.perf_counter()
            d = fasta_to_device(fa_dev)                # returns after the stream has drained
            fa_ms.append((time.perf_counter() - t0) * 1e3)
        # checked against the oracle on the first record's bytes and the totals
        recs, err = O.fasta_parse(fa[: min(len(fa), 8 << 20)].tobytes() + b"\nA")
        first = recs[0][2] if recs else b""
        fa_ok = (err is None and d.seq[: len(first)].cpu().numpy().tobytes() == first
                 and int(d.seq_len.sum().item()) == d.total)
        med = float(np.median(fa_ms))
        moved = len(fa) + d.total + (d.total + 15) // 16 * 4 + 13 * d.n   # file in, bytes + codes + per-record out
        fa_traffic = None     # HBM bytes per call from the FETCH_SIZE / WRITE_SIZE passes, per input byte
        fpath = os.path.join(ROOT, "profiles", "fasta_ingest_genome_256mib_hbm_traffic.md")
        if os.path.exists(fpath):
            import re
            mt = re.search(r"traffic_bytes_per_input_byte:\s*([0-9.eE+]+)", open(fpath).read())
            fa_traffic = round(float(mt.group(1)) * len(fa)) if mt else None
        fasta_line = {
            "workload": f"fasta_genome {len(fa) / 2**20:.0f} MiB, {d.n} records of 10 kbp..2 Mbp, 60-byte "
                        "lines, N runs, soft-masked runs, CRLF records; file resident in HBM",
            "value": round(len(fa) / (med * 1e-3) / 1e9, 2), "unit": "GB/s of FASTA input",
            "ms_per_call": round(med, 3), "records": d.n, "seq_bytes": d.total, "checked": bool(fa_ok),
            "roofline": {"bound": "hbm", "achieved": round(moved / (med * 1e-3) / 1e9, 1), "peak": 8000.0,
                         "unit": "GB/s", "frac": round(moved / (med * 1e-3) / 8e12, 4), "traffic": fa_traffic,
                         "algorithmic_bytes": moved},
        }
        del fa_dev, d
        if not args.no_cpu_baseline:
            sample = W.fasta_genome(64 << 20, seed=12)
            path = os.path.join(tempfile.gettempdir(), "ga_bench_sample.fa")
            sample.tofile(path)
            use_ref = O.ref_available()
            t0, k = time.perf_counter(), 0
            while time.perf_counter() - t0 < min(args.cpu_seconds, 5.0) or k == 0:
                if use_ref:
                    O.ref_fasta_file(path)
                else:
                    O.fasta_parse(sample.tobytes())
                k += 1
            el = time.perf_counter() - t0
            os.remove(path)
            fasta_line["cpu_baseline"] = {
                "value": round(k * len(sample) / el / 1e9, 3), "unit": "GB/s of FASTA input", "cores": 1,
                "kind": "reference" if use_ref else "port",
                "sample": f"{k} x io::read_fasta_file of a 64 MiB file of the same generator"
                          if use_ref else f"{k} x the oracle's reader over a 64 MiB file of the same generator",
            }
        log(f"fasta: {fasta_line}")

    # ---------------- gene QC task: one bait, many records ----------------
    qc_line = None
    if not args.no_gene_qc:
        import io
        from genoaligner_amd import gene_qc as G
        seqs, refs = W.gene_family(args.gene_qc_records, codons=380, seed=21)
        qdir = tempfile.mkdtemp()
        sp, rp = os.path.join(qdir, "seqs.fa"), os.path.join(qdir, "ref.fa")
        open(sp, "wb").write(seqs)
        open(rp, "wb").write(refs)
        G.run(sp, rp, code=2, emit=os.path.join(qdir, "warm.tsv"), out=io.StringIO(), err=io.StringIO())
        qc_ms = []
        for _ in range(3):
            t0 = time.perf_counter()
            G.run(sp, rp, code=2, emit=os.path.join(qdir, "got.tsv"), out=io.StringIO(), err=io.StringIO())
            qc_ms.append((time.perf_counter() - t0) * 1e3)
        med = float(np.median(qc_ms))
        got_tsv = open(os.path.join(qdir, "got.tsv"), encoding="latin-1").read()
        n_rec = args.gene_qc_records
        qc_cells = sum(int(ln.split("\t")[1]) for ln in got_tsv.splitlines()[1:]) * 1141
        qc_line = {
            "workload": f"gene_family {n_rec} records vs a 1,141 bp bait (CYTB-sized, code 2, 3 % "
                        "substitutions, frameshifts, stops, fragments, flanks), FASTA files to TSV",
            "value": round(n_rec / (med * 1e-3), 1), "unit": "records/s", "ms_per_run": round(med, 2),
            "gcups": round(qc_cells / (med * 1e-3) / 1e9, 1),   # record length x bait length, whole run
        }
        if os.path.exists(O.REF_GENE_QC):
            ssub, _ = W.gene_family(args.gene_qc_cpu_records, codons=380, seed=21)
            ssp = os.path.join(qdir, "sample.fa")
            open(ssp, "wb").write(ssub)
            # checked on the sample: the reference tool on the CPU takes ~13 ms a record
            G.run(ssp, rp, code=2, emit=os.path.join(qdir, "got_s.tsv"), out=io.StringIO(), err=io.StringIO())
            t0 = time.perf_counter()
            rc, _, _, ref_tsv = O.ref_gene_qc(ssp, rp, ("--code", "2"), emit=os.path.join(qdir, "ref_s.tsv"))
            el = time.perf_counter() - t0
            qc_line["matches_reference_tool"] = bool(
                rc == 0 and ref_tsv == open(os.path.join(qdir, "got_s.tsv"), encoding="latin-1").read())
            qc_line["checked_records"] = args.gene_qc_cpu_records
            if not args.no_cpu_baseline:
                qc_line["cpu_baseline"] = {
                    "value": round(args.gene_qc_cpu_records / el, 1), "unit": "records/s", "cores": 1,
                    "kind": "reference",
                    "sample": f"the reference gene_qc (tools/gene_qc.cpp on its CPU shim) over the first "
                              f"{args.gene_qc_cpu_records} records of the same family",
                }
        log(f"gene_qc: {qc_line}")

    # ---------------- serving: Poisson arrivals, continuous batching ----------------
    serve_line = None
    if not args.no_serve:
        import genoaligner_amd as ga
        from genoaligner_amd.serve import Server, poisson_load
        srng = np.random.default_rng(5)
        sreqs = [W.serving_request(srng, ga) for _ in range(args.serve_requests)]
        # warm-up: a dozen requests through a server of the same shape first
        # (slot threads' streams, kernels and workspaces), not measured
        warm = Server(budget_bytes=64 << 30, slots=4)
        for t in poisson_load(warm, lambda k: sreqs[k], 2 * args.serve_rate, min(12, len(sreqs)), seed=1):
            t.wait()
        warm.close()
        server = Server(budget_bytes=64 << 30, slots=4)
        t0 = time.perf_counter()
        tickets = poisson_load(server, lambda k: sreqs[k], args.serve_rate, len(sreqs), seed=9)
        for t in tickets:
            t.wait()
        wall = time.perf_counter() - t0
        server.close()
        lat = [t.latency_ms for t in tickets]
        errors = sum(1 for t in tickets if t.error)
        by = {"wfa": [t for t in tickets if t.kind == "wfa"], "sw": [t for t in tickets if t.kind == "sw"]}
        serve_line = {
            "workload": f"{len(sreqs)} requests, Poisson {args.serve_rate}/s, 1..512 pairs each; 60% WFA DNA "
                        "100 bp..5 kbp at 0..10% divergence, 40% protein SW (query 100..1000 aa, target "
                        "100..2000 aa, BLOSUM62 11/1); 70% of requests with CIGAR; 64 GB budget, 4 slots; "
                        "after 12 warm-up requests",
            "p50_ms": pctl(lat, 50), "p99_ms": pctl(lat, 99),
            "wfa_p50_ms": pctl([t.latency_ms for t in by["wfa"]], 50),
            "sw_p50_ms": pctl([t.latency_ms for t in by["sw"]], 50),
            "pairs": int(sum(t.pairs() for t in tickets)), "wall_s": round(wall, 3),
            "pairs_per_s": round(sum(t.pairs() for t in tickets) / wall, 1),
            "batches": len(server.batches), "peak_reserved_gb": round(server.peak_reserved / 2**30, 3),
            "errors": errors,
        }
        log(f"serving: {serve_line}")

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline_sw(sw, args.cpu_seconds, threads, O)
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": BASELINE_METRIC,
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16",
        "data": "synthetic (seeded)",
        "config": {"workload": "sw_embedded: 50000 pairs, query 256 bp, target 1024 bp with a "
                               "10%-mutated query copy (2:1:1), scoring 2/-3/3/1, score + span + CIGAR",
                   "pairs": sw.n, "cells_per_step": cells, "device": dev_name},
        "p50_ms": pctl(step_ms, 50),
        "p99_ms": pctl(step_ms, 99),
        "mismatches": sw_bad,
        "checked": int(len(sample)),
        "roofline": roof,
        "cpu_baseline": cpu,
        "sw_score_only": sw_score_line,
        "wfa": wfa_line,
        "wfa_mixed": mixed_line,
        "wfa_routing": routing_line,
        "sw_protein": protein_line,
        "serving": serve_line,
        "fasta_ingest": fasta_line,
        "gene_qc": qc_line,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
