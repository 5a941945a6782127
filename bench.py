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


def time_steps(run, steps, warmup, lib, kernel_reps=5):
    """W untimed steps, then exactly `steps` timed steps submitted back to back
    (no host wait between them, as a caller streaming batches would submit
    them) between two device synchronisations; per-step spans from events on
    the launch stream.  Then `kernel_reps` more steps, each waited for, give
    the dominant kernels' time from the library's own events on their stream."""
    import torch
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record()
        run()
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = [a.elapsed_time(b) for a, b in ev]
    kern = []
    for _ in range(kernel_reps):
        run()
        kern.append(lib.ga_last_kernel_ms())
    torch.cuda.synchronize()
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


def host_info():
    """The benchmark host's CPU: model, logical CPUs, physical cores, the CPUs
    this process may run on, and the threads the CPU baselines use (this
    process's CPU share, at most 16: the share one GPU's host slice carries)."""
    model, phys = "?", set()
    try:
        cur = {}
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if "core id" in cur:
                    phys.add((cur.get("physical id", "0"), cur["core id"]))
                cur = {}
                continue
            k, v = (x.strip() for x in line.split(":", 1))
            cur[k] = v
            if k == "model name":
                model = v
        if "core id" in cur:
            phys.add((cur.get("physical id", "0"), cur["core id"]))
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "physical_cores": len(phys) or None,
            "allowed_cpus": allowed, "threads": max(1, min(allowed, 16))}


def cpu_timed(prepare, n_total, work_of, unit, label, threads, seed, target_s=1.2, single_s=0.6):
    """The oracle timed as the reference CPU path is timed: a sample sized to
    about `target_s` per repetition on `threads` host threads, one warm-up,
    then the median (and min / max) of 5 repetitions; the same on one thread
    with a sample sized to about `single_s`.  prepare(idx) -> go(threads) ->
    seconds for the pairs idx (a random sample; a repetition passes over it
    several times when the whole workload runs in less than the target);
    work_of(idx) -> units of work (pairs, or cells / 1e9)."""
    perm = np.random.default_rng(seed).permutation(n_total)
    cache = {}

    def timed(k, th):
        m = min(k, n_total)
        if m not in cache:
            cache.clear()
            idx = np.sort(perm[:m])
            cache[m] = (idx, prepare(idx))
        idx, go = cache[m]
        loops = -(-k // m)
        return sum(go(th) for _ in range(loops)), loops * work_of(idx), loops * m

    def reps(th, target):
        k = max(th * 4, 16)
        dt, _, _ = timed(k, th)
        while dt < 0.25:
            k = int(k * max(2.0, 0.3 / max(dt, 1e-4)))
            dt, _, _ = timed(k, th)
        k = max(1, int(k * target / dt))
        timed(k, th)                                   # warm-up
        runs = sorted(timed(k, th)[:2] for _ in range(5))
        ts = [t for t, _ in runs]
        w = runs[0][1]
        return w / ts[2], w / ts[-1], w / ts[0], k, ts[2]

    v, lo, hi, n, med = reps(threads, target_s)
    v1, lo1, hi1, n1, med1 = reps(1, single_s)
    return {"value": round(v, 4), "unit": unit, "cores": threads, "kind": "port",
            "sample": f"{n} pairs of {label} (a random sample, passed over several times if the workload is "
                      f"smaller); median of 5 repetitions ({med:.2f} s each) after one warm-up, on {threads} "
                      "host threads",
            "min": round(lo, 4), "max": round(hi, 4),
            "single_core": {"value": round(v1, 4), "min": round(lo1, 4), "max": round(hi1, 4),
                            "sample": f"{n1} pairs, median of 5 ({med1:.2f} s each) after one warm-up"},
            "host": host_info()}


def cpu_baseline_sw(ps, threads, oracle, tab=None, gaps=None, label="256 x 1024, score + span + CIGAR"):
    """The oracle's SW trace path (score, span, CIGAR) over a random sample of
    the same pairs (score table and gap costs of the DNA scheme unless given)."""
    L = oracle.lib()
    if tab is None:
        tab = oracle.mm_table(SW_SCHEME[0], SW_SCHEME[1])
    go_, ge_ = gaps if gaps else (SW_SCHEME[2], SW_SCHEME[3])

    def prepare(idx):
        sub = ps.subset(idx)
        n = len(idx)
        out = np.zeros(5 * n, dtype=np.int32)
        cig = np.zeros(int((sub.a_len.astype(np.int64) + sub.b_len).sum()) + 1, dtype=np.uint8)
        coff = np.zeros(n, dtype=np.int64)
        coff[1:] = np.cumsum(sub.a_len.astype(np.int64) + sub.b_len)[:-1]
        clen = np.zeros(n, dtype=np.int32)

        def go(th):
            t0 = time.perf_counter()
            rc = L.ga_oracle_sw_batch(sub.a.ctypes.data, sub.a_off.ctypes.data, sub.a_len.ctypes.data,
                                      sub.b.ctypes.data, sub.b_off.ctypes.data, sub.b_len.ctypes.data,
                                      n, tab.ctypes.data, go_, ge_, 1,
                                      out.ctypes.data, cig.ctypes.data, coff.ctypes.data,
                                      clen.ctypes.data, th)
            assert rc == 0
            return time.perf_counter() - t0
        return go

    cells = lambda idx: float((ps.a_len[idx].astype(np.int64) * ps.b_len[idx]).sum()) / 1e9
    return cpu_timed(prepare, ps.n, cells, "GCUPS", label, threads, seed=2)


def cpu_baseline_wfa(ps, want, threads, oracle, label):
    """The oracle's WFA batch (score; CIGAR where wanted) over a random sample
    of the pairs."""
    L = oracle.lib()

    def prepare(idx):
        sub = ps.subset(idx)
        n = len(idx)
        w = np.ascontiguousarray(want[idx], dtype=np.uint8)
        slot = sub.a_len.astype(np.int64) + sub.b_len
        cig = np.zeros(int(slot.sum()) + 1, dtype=np.uint8)
        coff = np.zeros(n, dtype=np.int64)
        coff[1:] = np.cumsum(slot)[:-1]
        sc = np.zeros(n, dtype=np.int32)
        clen = np.zeros(n, dtype=np.int32)
        smax = np.ascontiguousarray(sub.smax, dtype=np.int32)

        def go(th):
            t0 = time.perf_counter()
            rc = L.ga_oracle_wfa_batch(sub.a.ctypes.data, sub.a_off.ctypes.data, sub.a_len.ctypes.data,
                                       sub.b.ctypes.data, sub.b_off.ctypes.data, sub.b_len.ctypes.data,
                                       smax.ctypes.data, w.ctypes.data, n, sc.ctypes.data,
                                       cig.ctypes.data, coff.ctypes.data, clen.ctypes.data, th)
            assert rc == 0
            return time.perf_counter() - t0
        return go

    return cpu_timed(prepare, ps.n, lambda idx: float(len(idx)), "pairs/s", label, threads, seed=3)


SECTIONS = ("sw", "sw_host", "wfa", "mixed", "routing", "protein", "fasta", "gene_qc", "serving", "msa_pp",
            "guide_tree")


def mean_pos(xs):
    xs = [x for x in xs if x and x > 0]
    return float(np.mean(xs)) if xs else None


def _pack_seqs(seqs):
    lens = np.array([len(x) for x in seqs], dtype=np.int32)
    off = np.zeros(len(seqs), dtype=np.int64)
    if len(seqs) > 1:
        off[1:] = np.cumsum(lens.astype(np.int64))[:-1]
    return np.frombuffer(b"".join(seqs), dtype=np.uint8).copy(), off, lens


def wfa_requests_pairset(items):
    """A PairSet (a = pattern, b = text, per-pair smax) of WFA requests."""
    from genoaligner_amd.workloads import PairSet
    return PairSet(*_pack_seqs([bytes(r.pattern) for r in items]), *_pack_seqs([bytes(r.text) for r in items]),
                   np.array([r.smax for r in items], dtype=np.int32))


def sw_requests_pairset(items):
    """A PairSet (a = text, b = pattern) of SW requests."""
    from genoaligner_amd.workloads import PairSet
    return PairSet(*_pack_seqs([bytes(r.text) for r in items]), *_pack_seqs([bytes(r.pattern) for r in items]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--only", default=None,
                    help="comma-separated sections to run (" + ", ".join(SECTIONS) + "); default: all")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-wfa", action="store_true")
    ap.add_argument("--no-wfa-mixed", action="store_true")
    ap.add_argument("--wfa-mixed-pairs", type=int, default=200_000)
    ap.add_argument("--no-routing", action="store_true")
    ap.add_argument("--routing-pairs", type=int, default=200_000)
    ap.add_argument("--router-weights", default=None,
                    help="router parameters (.npy); default: the trained file shipped in the package")
    ap.add_argument("--route-q", type=float, default=0.99)
    ap.add_argument("--no-protein", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-fasta", action="store_true")
    ap.add_argument("--no-gene-qc", action="store_true")
    ap.add_argument("--gene-qc-records", type=int, default=20000)
    ap.add_argument("--gene-qc-cpu-records", type=int, default=300)
    ap.add_argument("--fasta-mib", type=int, default=512)
    ap.add_argument("--serve-rate", type=float, default=100.0, help="requests per second (Poisson), lowest load")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--msa-pairs", type=int, default=4096)
    ap.add_argument("--tree-seqs", type=int, default=1000)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a kernel of "
                         "known traffic, to scale hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: there is no multi-GPU path "
              f"(--gpus {args.gpus} is not supported; pairs are independent, so a batch would be "
              f"split by cost into one contiguous range per GPU)", file=sys.stderr)
        sys.exit(2)
    only = set(args.only.split(",")) if args.only else set(SECTIONS)
    bad = only - set(SECTIONS)
    if bad:
        print(f"unknown sections {sorted(bad)}; known: {', '.join(SECTIONS)}", file=sys.stderr)
        sys.exit(2)
    for flag, sec in ((args.no_wfa, "wfa"), (args.no_wfa_mixed, "mixed"), (args.no_routing, "routing"),
                      (args.no_protein, "protein"), (args.no_serve, "serving"), (args.no_fasta, "fasta"),
                      (args.no_gene_qc, "gene_qc")):
        if flag:
            only.discard(sec)
    cpu_on = not args.no_cpu_baseline

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
    threads = host_info()["threads"]
    rng = np.random.default_rng(1)

    if args.copy_probe_mib > 0:
        nbytes = args.copy_probe_mib << 20
        src = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {nbytes} bytes read + {nbytes} bytes written")
        del src, dst

    # ---------------- SW headline ----------------
    sw = None
    gcups = ms_per_step = step_ms = roof = sw_score_line = cpu = sw_host_line = None
    sw_bad, sample, cells = None, [], None
    if only & {"sw", "sw_host"}:
        t0 = time.perf_counter()
        sw = W.sw_embedded(n=50_000, qlen=256, tlen=1024, rate=0.10)
        cells = int((sw.a_len.astype(np.int64) * sw.b_len).sum())
        log(f"sw workload ready ({time.perf_counter() - t0:.1f} s), {cells / 1e9:.2f} Gcells/step")
        sample = rng.choice(sw.n, size=min(300, sw.n), replace=False)
    if "sw" in only:
        swd = SwDeviceBatch(sw, scoring=SW_SCHEME, trace=True)
        wall, step_ms, kern = time_steps(swd.run, args.steps, args.warmup, lib)
        ms_per_step = wall * 1000.0 / args.steps
        gcups = cells / (wall / args.steps) / 1e9
        kern_ms = mean_pos(kern)
        log(f"sw: {ms_per_step:.2f} ms/step, {gcups:.1f} GCUPS, forward kernel {kern_ms} ms")
        sw_bad = check_sw(sw, swd, sample, O)
        log(f"sw: {sw_bad} mismatches in {len(sample)} checked pairs")
        del swd

        # HBM bytes per launch of the SW trace kernel, measured with rocprofv3
        # FETCH_SIZE / WRITE_SIZE passes scaled by a copy of known size
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "sw16_trace_sw_embedded_50k_hbm_traffic.md")
        if os.path.exists(tpath):
            import re
            mt = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(tpath).read())
            traffic = float(mt.group(1)) if mt else None
        if kern_ms:
            achieved = SW_OPS_PER_CELL * cells / (kern_ms * 1e-3) / 1e12
            roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_INT16_TOPS,
                    "unit": "TOP/s", "frac": round(achieved / PEAK_INT16_TOPS, 4),
                    "traffic": traffic, "kernel": "sw16_kernel<true, false, 2> (packed forward + fused walk, H mod 8 trace record)",
                    "kernel_ms": round(kern_ms, 3)}

        # score-only on the same pairs: score + end cell (the lane-rule end)
        swd_s = SwDeviceBatch(sw, scoring=SW_SCHEME, trace=False)
        swall, sstep, skern = time_steps(swd_s.run, args.steps, args.warmup, lib)
        sk = mean_pos(skern)
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
        if cpu_on:
            cpu = cpu_baseline_sw(sw, threads, O)
            log(f"cpu baseline: {cpu}")

    # ---------------- SW headline pairs through the drop-in host entry point ----------------
    if "sw_host" in only:
        import ctypes as C
        # the C entry point as a C++ caller of align_sw_batch uses it: host
        # sequences, request array and arena built before the clock
        hreqs = (N.SWReq * sw.n)()
        ab, bb = sw.a.ctypes.data, sw.b.ctypes.data
        sc_c = N.SWScoringC(*SW_SCHEME)
        for i in range(sw.n):
            hreqs[i] = N.SWReq(ab + int(sw.a_off[i]), int(sw.a_len[i]), bb + int(sw.b_off[i]), int(sw.b_len[i]),
                               sc_c, 1)
        hres = (N.SWRes * sw.n)()
        hneed = lib.ga_cigar_arena_bytes_sw(hreqs, sw.n)
        harena = C.create_string_buffer(max(hneed, 1))
        hok, hun, herr = C.c_int32(), C.c_int32(), C.c_char_p()
        hts, hk = [], []
        for rep in range(1 + 5):
            t0 = time.perf_counter()
            hrc = lib.ga_align_sw_batch(hreqs, sw.n, hres, harena, hneed, C.byref(hok), C.byref(hun),
                                        C.byref(herr))
            dt = time.perf_counter() - t0
            if rep > 0:
                hts.append(dt)
                hk.append(lib.ga_last_kernel_ms())
        assert hrc == N.GA_OK, herr.value
        raw = harena.raw
        hbad = 0
        for i in sample[:100]:
            r = hres[i]
            cg = raw[r.cigar_off: r.cigar_off + r.cigar_len].decode() if r.cigar_len > 0 else ""
            if (r.score, r.start_i, r.start_j, r.end_i, r.end_j, cg) != O.sw_trace(sw.a_seq(i), sw.b_seq(i),
                                                                                       SW_SCHEME):
                hbad += 1
            if not (r.rescore_ok and r.wellformed_ok):
                hbad += 1
        hmed = float(np.median(hts))
        sw_host_line = {
            "workload": "the headline pairs through ga_align_sw_batch (the drop-in entry point of "
                        "align_sw_batch): host sequences in, host results, CIGARs and validators out; "
                        "packing, PCIe copies both ways and the CIGAR checks included; median of 5 calls "
                        "after one untimed call",
            "gcups": round(cells / hmed / 1e9, 3), "ms": round(hmed * 1e3, 2),
            "p99_ms": round(max(hts) * 1e3, 2), "kernel_ms": round(mean_pos(hk), 3) if mean_pos(hk) else None,
            "mismatches": hbad, "checked": 100,
        }
        log(f"sw host entry: {sw_host_line}")
        del hreqs, hres, harena

    # ---------------- WFA, 10k x 150 bp ----------------
    wfa_line = None
    if "wfa" in only:
        wf = W.wfa_plain(n=10_000, length=150, rate=0.02, smax=64)
        wfd = WfaDeviceBatch(wf, trace=True)
        wwall, wstep, wkern = time_steps(wfd.run, args.steps, args.warmup, lib)
        wsample = rng.choice(wf.n, size=min(1000, wf.n), replace=False)
        wbad = check_wfa(wf, wfd, wsample, O)
        wk = mean_pos(wkern)
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
            "graph_p99_ms": pctl(gstep, 99),
            "kernel_ms": round(wk, 4) if wk else None,
            "roofline": {"bound": "hbm", "achieved": round((in_bytes + out_bytes) / (wk * 1e-3) / 1e9, 2) if wk else None,
                         "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round((in_bytes + out_bytes) / (wk * 1e-3) / 1e9 / PEAK_HBM_GBS, 5) if wk else None,
                         "algorithmic_bytes": in_bytes + out_bytes,
                         "kernel": "wfa_kernel<true, int16_t, true, 1> (smax <= 64 class)"},
            "valu_roofline": wfa_valu_roofline(wf, sc, wk),
            "mismatches": wbad, "checked": int(len(wsample)),
            "tcups_ref_style": round(float((wf.a_len.astype(np.int64) * wf.b_len).sum())
                                     / (wwall / args.steps) / 1e12, 4),
        }
        log(f"wfa: {wfa_line['pairs_per_s']:.0f} pairs/s, {wbad} mismatches")
        # the same pairs through ga_align_batch (the drop-in entry point of
        # align_batch), as a C or C++ caller uses it: requests and arena
        # built before the clock
        import ctypes as C
        wreqs = (N.AlignReq * wf.n)()
        pa, ta = wf.a.ctypes.data, wf.b.ctypes.data
        for i in range(wf.n):
            wreqs[i] = N.AlignReq(pa + int(wf.a_off[i]), int(wf.a_len[i]), ta + int(wf.b_off[i]),
                                  int(wf.b_len[i]), int(wf.smax[i]), 1)
        wres = (N.AlignRes * wf.n)()
        wneed = lib.ga_cigar_arena_bytes_wfa(wreqs, wf.n)
        warena = C.create_string_buffer(max(wneed, 1))
        wok, wun, werr = C.c_int32(), C.c_int32(), C.c_char_p()
        wts = []
        for rep in range(1 + 5):
            t0 = time.perf_counter()
            wrc = lib.ga_align_batch(wreqs, wf.n, wres, warena, wneed, C.byref(wok), C.byref(wun), C.byref(werr))
            if rep > 0:
                wts.append(time.perf_counter() - t0)
        assert wrc == N.GA_OK, werr.value
        wraw = warena.raw
        whbad = 0
        for i in wsample[:100]:
            r = wres[i]
            cg = wraw[r.cigar_off: r.cigar_off + r.cigar_len].decode() if r.cigar_len > 0 else ""
            if (r.score, cg) != O.wfa(wf.a_seq(i), wf.b_seq(i), int(wf.smax[i])) or not (r.rescore_ok and r.wellformed_ok):
                whbad += 1
        wmed = float(np.median(wts))
        wfa_line["host_entry"] = {
            "workload": "the same pairs through ga_align_batch (the drop-in entry point of align_batch): host "
                        "sequences in, host scores, CIGARs and validator flags out; packing, PCIe copies both "
                        "ways and the CIGAR checks included; median of 5 calls after one untimed call",
            "pairs_per_s": round(wf.n / wmed, 1), "ms": round(wmed * 1e3, 3),
            "p99_ms": round(max(wts) * 1e3, 3), "mismatches": whbad, "checked": 100,
        }
        log(f"wfa host entry: {wfa_line['host_entry']}")
        del wreqs, wres, warena
        if cpu_on:
            wfa_line["cpu_baseline"] = cpu_baseline_wfa(wf, np.ones(wf.n, dtype=np.uint8), threads, O,
                                                        "wfa_plain 150 bp, smax 64, score + CIGAR")
            log(f"wfa cpu baseline: {wfa_line['cpu_baseline']}")

    # ---------------- WFA, mixed lengths / distances ----------------
    mixed_line = None
    if "mixed" in only:
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
        mk = mean_pos(mkern)
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
        if cpu_on:
            mixed_line["cpu_baseline"] = cpu_baseline_wfa(wm, want, threads, O,
                                                          "wfa_mixed (70% CIGAR), drawn at random")
            log(f"wfa_mixed cpu baseline: {mixed_line['cpu_baseline']}")
        del wmd
        wfa_set_tuning(None)

    # ---------------- WFA routing: plain per-class path vs learned router ----------------
    routing_line = None
    if "routing" in only:
        from genoaligner_amd import router as R
        t0 = time.perf_counter()
        wr = W.wfa_routing(n=args.routing_pairs)
        rwant = (np.random.default_rng(6).random(wr.n) < 0.7).astype(np.uint8)
        params = torch.from_numpy(R.load(args.router_weights or R.WEIGHTS)).cuda()
        plain = WfaDeviceBatch(wr, trace=True, want=rwant)
        routed = WfaDeviceBatch(wr, trace=True, want=rwant)
        log(f"wfa_routing workload ready ({time.perf_counter() - t0:.1f} s)")
        pwall, pstep, pkern = time_steps(plain.run, args.steps, args.warmup, lib)
        rwall, rstep, rkern = time_steps(lambda: R.run_routed(routed, params, q=args.route_q, stats=False),
                                         args.steps, args.warmup, lib)
        stats = [R.run_routed(routed, params, q=args.route_q)]
        psc, pcg = plain.results()
        rsc, rcg = routed.results()
        diff = int(sum(1 for i in range(wr.n) if psc[i] != rsc[i] or pcg[i] != rcg[i]))
        # the ceiling of any router that only sets bounds: the plain path again
        # with every resolved pair's bound at its exact distance (known here
        # from the plain results, for free), which is the tightest bound that
        # keeps every result
        xs = wr.smax.copy()
        xs[psc >= 0] = psc[psc >= 0]
        exact = WfaDeviceBatch(W.PairSet(wr.a, wr.a_off, wr.a_len, wr.b, wr.b_off, wr.b_len, xs, dict(wr.meta)),
                               trace=True, want=rwant)
        xwall, _, _ = time_steps(exact.run, args.steps, args.warmup, lib, kernel_reps=0)
        xsc, xcg = exact.results()
        xdiff = int(sum(1 for i in range(wr.n) if psc[i] != xsc[i] or pcg[i] != xcg[i]))
        del exact
        rsample = rng.choice(wr.n, size=min(200, wr.n), replace=False)
        obad = 0
        for i in rsample:
            osc, ocg = O.wfa(wr.a_seq(i), wr.b_seq(i), int(wr.smax[i]))
            if int(rsc[i]) != osc or rcg[i].decode() != (ocg if rwant[i] else ""):
                obad += 1
        st = stats[-1]
        pk, rk = mean_pos(pkern), mean_pos(rkern)
        routing_line = {
            "workload": f"wfa_routing {wr.n} pairs, lengths log-uniform 100..10000 bp, divergence 80% "
                        "U[0,0.03] / 15% U[0.03,0.15] / 5% unrelated, smax min(0.2 len, 2048), "
                        "70% CIGAR / 30% score-only",
            "plain_pairs_per_s": round(wr.n / (pwall / args.steps), 1),
            "routed_pairs_per_s": round(wr.n / (rwall / args.steps), 1),
            "routed_over_plain": round(pwall / rwall, 3),
            "plain_ms": round(pwall * 1000 / args.steps, 3), "routed_ms": round(rwall * 1000 / args.steps, 3),
            "plain_p99_ms": pctl(pstep, 99),
            "routed_p50_ms": pctl(rstep, 50), "routed_p99_ms": pctl(rstep, 99),
            "plain_kernel_ms": round(pk, 3) if pk else None,
            "routed_kernel_ms": round(rk, 3) if rk else None,
            # the pairs' algorithmic diagonal updates over the routed path's
            # WFA launches (both passes), against the int32 VALU peak
            "roofline": wfa_valu_roofline(wr, rsc, rk),
            "plain_roofline": wfa_valu_roofline(wr, psc, pk),
            "routed_vs_plain_differences": diff, "oracle_mismatches": obad, "oracle_checked": int(len(rsample)),
            "exact_bound_ms": round(xwall * 1000 / args.steps, 3),
            "exact_bound_over_plain": round(pwall / xwall, 3),
            "exact_bound_differences": xdiff,
            "shortcut": st.shortcut, "routed": st.routed, "plain": st.plain, "rerun": st.rerun,
            "miss_rate": round(st.miss_rate, 4), "route_q": args.route_q,
        }
        log(f"wfa_routing: {routing_line}")
        if cpu_on:
            routing_line["cpu_baseline"] = cpu_baseline_wfa(wr, rwant, threads, O,
                                                            "wfa_routing (70% CIGAR), drawn at random")
            log(f"wfa_routing cpu baseline: {routing_line['cpu_baseline']}")
        del plain, routed

    # ---------------- protein SW (BLOSUM62, gap open 11 / extend 1), host entry point ----------------
    protein_line = None
    m62 = None
    if only & {"protein", "serving"}:
        import genoaligner_amd as ga
        m62 = ga.blosum62(11, 1)
        tab62 = O.matrix_table(m62.alphabet.decode(), [[m62.score[a][b] for b in range(20)] for a in range(20)], -1)
    if "protein" in only:
        import ctypes as C
        import genoaligner_amd as ga
        from genoaligner_amd import api as A
        pp = W.protein_pairs(n=2000)
        preqs = [ga.SWRequest(text=pp.a_seq(i), text_len=int(pp.a_len[i]), pattern=pp.b_seq(i),
                              pattern_len=int(pp.b_len[i]), with_cigar=True) for i in range(pp.n)]
        pcells = int((pp.a_len.astype(np.int64) * pp.b_len).sum())
        ga.align_sw_matrix_batch(preqs[:64], m62)
        # the C entry point as a C++ caller uses it: request array and arena
        # built before the clock, as the caller's own data would be
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
        pts, pks = [], []
        for rep in range(1 + max(3, min(args.steps, 5))):
            t0 = time.perf_counter()
            prc = lib.ga_align_sw_matrix_batch(parr, pp.n, C.byref(m62), pres, parena, pneed, C.byref(pok),
                                               C.byref(pun), C.byref(perr))
            if rep > 0:
                pts.append(time.perf_counter() - t0)
                pks.append(lib.ga_last_kernel_ms())
        pdt = float(np.median(pts))
        pkm = mean_pos(pks)
        py_ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            pb = ga.align_sw_matrix_batch(preqs, m62)
            py_ts.append(time.perf_counter() - t0)
        pdt_py = float(np.median(py_ts))
        pbad = 0 if prc == N.GA_OK else pp.n
        raw = parena.raw
        for i in rng.choice(pp.n, size=20, replace=False):
            r = pres[i]
            cg = raw[r.cigar_off: r.cigar_off + r.cigar_len].decode("latin-1") if r.cigar_len > 0 else ""
            e = O.sw_trace(pp.a_seq(i), pp.b_seq(i), tab62, 11, 1)
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
            "kernel_ms": round(pkm, 3) if pkm else None,
            "kernel_gcups": round(pcells / (pkm * 1e-3) / 1e9, 1) if pkm else None,
            # packed kernel: 12 ops per cell against the packed int16 VALU
            # peak; kernel_ms covers the forward pass and the fused walk
            "roofline": {"bound": "valu", "unit": "TOP/s", "peak": PEAK_INT16_TOPS,
                         "achieved": round(SW_OPS_PER_CELL * pcells / (pkm * 1e-3) / 1e12, 3) if pkm else None,
                         "frac": round(SW_OPS_PER_CELL * pcells / (pkm * 1e-3) / 1e12 / PEAK_INT16_TOPS, 4)
                         if pkm else None,
                         "kernel": "sw16m_kernel<true, false, 8> (packed int16, two pairs per workgroup, "
                                   "8-wave strip pipeline, traceback walk fused into wave 0)"},
            "python_api_ms": round(pdt_py * 1e3, 2),
            "mismatches": pbad, "checked": 20,
        }
        log(f"protein: {protein_line}")
        if cpu_on:
            protein_line["cpu_baseline"] = cpu_baseline_sw(
                pp, threads, O, tab=tab62, gaps=(11, 1), label="protein, BLOSUM62 11/1, score + span + CIGAR")
            log(f"protein cpu baseline: {protein_line['cpu_baseline']}")

    # ---------------- FASTA ingest on the device ----------------
    fasta_line = None
    if "fasta" in only:
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
        if cpu_on:
            # the oracle's reader (a restatement of the reference's io::read_fasta
            # in Python), one host thread, median of 5 after one warm-up
            smp = W.fasta_genome(16 << 20, seed=12).tobytes()
            O.fasta_parse(smp)
            ts = sorted(_timed(lambda: O.fasta_parse(smp)) for _ in range(5))
            fasta_line["cpu_baseline"] = {
                "value": round(len(smp) / ts[2] / 1e9, 4), "unit": "GB/s of FASTA input", "cores": 1,
                "kind": "port", "min": round(len(smp) / ts[-1] / 1e9, 4), "max": round(len(smp) / ts[0] / 1e9, 4),
                "sample": f"the oracle's reader over a 16 MiB file of the same generator, median of 5 "
                          f"({ts[2]:.2f} s each) after one warm-up",
                "host": host_info(),
            }
        log(f"fasta: {fasta_line}")

    # ---------------- gene QC task: one bait, many records ----------------
    qc_line = None
    if "gene_qc" in only:
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
        # checked on a sample against the oracle's pipeline (pinned to the
        # reference tool on the CPU), which is also the CPU baseline
        nq = args.gene_qc_cpu_records
        ssub, _ = W.gene_family(nq, codons=380, seed=21)
        ssp = os.path.join(qdir, "sample.fa")
        open(ssp, "wb").write(ssub)
        G.run(ssp, rp, code=2, emit=os.path.join(qdir, "got_s.tsv"), out=io.StringIO(), err=io.StringIO())
        want_tsv, _ = O.gene_qc_tsv(ssub, refs, code=2)
        qc_line["matches_oracle"] = want_tsv == open(os.path.join(qdir, "got_s.tsv"), encoding="latin-1").read()
        qc_line["checked_records"] = nq
        if cpu_on:
            ts = sorted(_timed(lambda: O.gene_qc_tsv(ssub, refs, code=2)) for _ in range(5))
            qc_line["cpu_baseline"] = {
                "value": round(nq / ts[2], 1), "unit": "records/s", "cores": 1, "kind": "port",
                "min": round(nq / ts[-1], 1), "max": round(nq / ts[0], 1),
                "sample": f"the oracle's gene QC pipeline over the first {nq} records of the same family, "
                          f"median of 5 ({ts[2]:.2f} s each) after the checked run",
                "host": host_info(),
            }
        log(f"gene_qc: {qc_line}")

    # ---------------- MSA: profile-profile Gotoh over a level of pairs ----------------
    msa_line = None
    if "msa_pp" in only:
        msa_line = msa_section(args, W, O, threads, cpu_on, lib)

    # ---------------- MSA guide tree: k-mer distances + neighbour joining ----------------
    tree_line = None
    if "guide_tree" in only:
        tree_line = guide_tree_section(args, W, O, cpu_on, lib)

    # ---------------- serving: Poisson arrivals, continuous batching ----------------
    serve_line = None
    if "serving" in only:
        serve_line = serving_section(args, W, O, tab62, threads, cpu_on)

    line = {
        "metric": BASELINE_METRIC,
        "value": round(gcups, 3) if gcups else None,
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3) if ms_per_step else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16",
        "data": "synthetic (seeded)",
        "config": {"workload": "sw_embedded: 50000 pairs, query 256 bp, target 1024 bp with a "
                               "10%-mutated query copy (2:1:1), scoring 2/-3/3/1, score + span + CIGAR",
                   "pairs": sw.n if sw is not None else None, "cells_per_step": cells, "device": dev_name},
        "p50_ms": pctl(step_ms, 50) if step_ms else None,
        "p99_ms": pctl(step_ms, 99) if step_ms else None,
        "mismatches": sw_bad,
        "checked": int(len(sample)),
        "roofline": roof,
        "cpu_baseline": cpu,
        "sw_score_only": sw_score_line,
        "sw_host_entry": sw_host_line,
        "wfa": wfa_line,
        "wfa_mixed": mixed_line,
        "wfa_routing": routing_line,
        "sw_protein": protein_line,
        "serving": serve_line,
        "fasta_ingest": fasta_line,
        "gene_qc": qc_line,
        "msa_pp": msa_line,
        "guide_tree": tree_line,
    }
    print(json.dumps(line), flush=True)


def _timed(fn):
    t0 = time.perf_counter()
    fn()
    return time.perf_counter() - t0


def msa_section(args, W, O, threads, cpu_on, lib):
    """A level of a progressive MSA: profile pairs of related DNA families
    (600..1,400 columns, 1..4 rows per profile) through the device entry
    point, default genomsa parameters; GCUPS (cells = sum of M x N) and the
    fp32 VALU roofline at 9 ops per cell plus 3 per symbol-fraction term."""
    import torch
    from genoaligner_amd import msa as M
    t0 = time.perf_counter()
    pairs_rows = W.msa_profile_pairs(n=args.msa_pairs)
    P = M.Params()
    profs, idx = [], []
    for a, b in pairs_rows:
        profs += [M.profile_from_rows(a), M.profile_from_rows(b)]
        idx.append((len(profs) - 2, len(profs) - 1))
    dev = M.MsaDeviceBatch(profs, idx, P)
    log(f"msa_pp workload ready ({time.perf_counter() - t0:.1f} s)")
    cells = sum(profs[a].ncols() * profs[b].ncols() for a, b in idx)
    nz = [np.count_nonzero(p.cols, axis=1) for p in profs]
    terms = sum(float(nz[a].sum()) * float(nz[b].sum()) for a, b in idx)
    ops = 9.0 * cells + 3.0 * terms
    wall, step_ms, kern = time_steps(dev.run, args.steps, args.warmup, lib)
    km = mean_pos(kern)
    got = dev.results()
    rng = np.random.default_rng(7)
    sample = rng.choice(len(idx), size=min(40, len(idx)), replace=False)
    OP = O.MsaParams()
    bad = 0
    for q in sample:
        a, b = idx[q]
        w = O.msa_align((profs[a].cols, profs[a].occ), (profs[b].cols, profs[b].occ), OP)
        g = got[q]
        if (np.float32(g.score).tobytes(), (g.ai, g.aj, g.bi, g.bj), g.cigar) != (np.float32(w[0]).tobytes(), w[1], w[2]):
            bad += 1
    line = {
        "workload": f"msa level {len(idx)} profile pairs, DNA families, 600..1400 columns, 1..4 rows per profile, "
                    "default genomsa parameters (semiglobal, position-specific gap penalties), score + span + "
                    "column CIGAR; profiles resident in HBM",
        "gcups": round(cells / (wall / args.steps) / 1e9, 2), "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "cells": cells, "kernel_ms": round(km, 3) if km else None,
        "roofline": {"bound": "valu", "unit": "TOP/s", "peak": PEAK_INT32_TOPS,
                     "ops_per_cell": round(ops / cells, 2),
                     "achieved": round(ops / (km * 1e-3) / 1e12, 3) if km else None,
                     "frac": round(ops / (km * 1e-3) / 1e12 / PEAK_INT32_TOPS, 4) if km else None,
                     "kernel": "msa_pp_kernel<4> (one wave per pair, fp32)"},
        "p50_ms": pctl(step_ms, 50), "p99_ms": pctl(step_ms, 99),
        "mismatches": bad, "checked": int(len(sample)),
    }
    if cpu_on:
        import ctypes as C
        L = O._msa_lib()
        cols = np.concatenate([p.cols for p in profs]).astype(np.float32)
        occ = np.concatenate([p.occ for p in profs]).astype(np.float32)
        lens = np.array([p.ncols() for p in profs], dtype=np.int32)
        coff = np.zeros(len(profs), dtype=np.int64)
        coff[1:] = np.cumsum(lens.astype(np.int64))[:-1]
        ha = np.array([a for a, _ in idx], dtype=np.int32)
        hb = np.array([b for _, b in idx], dtype=np.int32)
        pc = OP.c()

        def prepare(sel):
            a, b = np.ascontiguousarray(ha[sel]), np.ascontiguousarray(hb[sel])
            n = len(sel)
            score = np.zeros(n, np.float32)
            span = np.zeros(4 * n, np.int32)
            slot = lens[a].astype(np.int64) + lens[b]
            so = np.zeros(n, np.int64)
            so[1:] = np.cumsum(slot)[:-1]
            cig = np.zeros(int(slot.sum()) + 1, np.uint8)
            cl = np.zeros(n, np.int32)

            def go(th):
                t0 = time.perf_counter()
                rc = L.ga_oracle_msa_align_batch(cols.ctypes.data, occ.ctypes.data, coff.ctypes.data, lens.ctypes.data,
                                                 a.ctypes.data, b.ctypes.data, n, C.byref(pc), score.ctypes.data,
                                                 span.ctypes.data, cig.ctypes.data, so.ctypes.data, cl.ctypes.data, th)
                assert rc == 0
                return time.perf_counter() - t0
            return go

        cw = lambda sel: float((lens[ha[sel]].astype(np.int64) * lens[hb[sel]]).sum()) / 1e9
        line["cpu_baseline"] = cpu_timed(prepare, len(idx), cw, "GCUPS", "the msa level's profile pairs", threads,
                                         seed=8)
    log(f"msa_pp: {line}")
    del dev
    torch.cuda.synchronize()
    return line


def guide_tree_section(args, W, O, cpu_on, lib):
    """The guide-tree step: 5-mer distances of n related DNA sequences and
    neighbour joining over them (the host entry points: sequences in, tree
    out), with each kernel's time from the library's events; NJ against the
    bytes its rounds read (row sums and the candidate scan read the live
    block of the double matrix, 1.5 m^2 x 8 bytes per round)."""
    from genoaligner_amd import msa as M
    n = args.tree_seqs
    seqs = W.sequence_family_set(n=n)
    M.nj_children(M.kmer_distances(seqs[:50], 5), 50)            # warm-up
    ts_k, ts_n, kk, kn = [], [], [], []
    for _ in range(3):
        t0 = time.perf_counter()
        D = M.kmer_distances(seqs, 5)
        ts_k.append(time.perf_counter() - t0)
        kk.append(lib.ga_last_kernel_ms())
        t0 = time.perf_counter()
        ch = M.nj_children(D, n)
        ts_n.append(time.perf_counter() - t0)
        kn.append(lib.ga_last_kernel_ms())
    want_D = O.kmer_distances(seqs, 5)
    t0 = time.perf_counter()
    want_ch = O.nj_tree(want_D, n)
    cpu_nj = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.kmer_distances(seqs, 5)
    cpu_k = time.perf_counter() - t0
    nj_bytes = sum(1.5 * m * m * 8 for m in range(3, n + 1))
    knm = mean_pos(kn)
    line = {
        "workload": f"guide tree of {n} DNA sequences (40 families, ~1,200 bp, up to 12% divergence): 5-mer "
                    "distances + neighbour joining, host entry points",
        "ms": round(float(np.median(ts_k)) * 1e3 + float(np.median(ts_n)) * 1e3, 2),
        "kmer_ms": round(float(np.median(ts_k)) * 1e3, 3), "kmer_kernel_ms": round(mean_pos(kk), 3) if mean_pos(kk) else None,
        "nj_ms": round(float(np.median(ts_n)) * 1e3, 2), "nj_kernel_ms": round(knm, 2) if knm else None,
        "nj_rounds": n - 2,
        "roofline": {"bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS, "algorithmic_bytes": int(nj_bytes),
                     "achieved": round(nj_bytes / (knm * 1e-3) / 1e9, 1) if knm else None,
                     "frac": round(nj_bytes / (knm * 1e-3) / 1e9 / PEAK_HBM_GBS, 4) if knm else None,
                     "kernel": "nj rounds (rowsum, argmin, merge), latency-bound: 3 launches per round"},
        "distances_equal_oracle": bool(np.array_equal(D, want_D)),
        "tree_equal_oracle": bool(np.array_equal(ch, want_ch)),
    }
    if cpu_on:
        line["cpu_baseline"] = {
            "value": round((cpu_k + cpu_nj) * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
            "kmer_ms": round(cpu_k * 1e3, 1), "nj_ms": round(cpu_nj * 1e3, 1),
            "sample": f"the oracle's kmer_distances and nj_tree on the same {n} sequences, one host thread, one run",
            "host": host_info(),
        }
    log(f"guide_tree: {line}")
    return line


def serving_section(args, W, O, tab62, threads, cpu_on):
    """Poisson arrivals at a rising rate until p99 passes twice its lowest-load
    value (the capacity knee), then every request at once (saturation), with
    per-class work rates against the VALU peaks; the oracle on the same mix."""
    import genoaligner_amd as ga
    from genoaligner_amd.serve import LibraryBackend, Server, poisson_load
    srng = np.random.default_rng(5)
    sreqs = [W.serving_request(srng, ga) for _ in range(args.serve_requests)]
    # warm-up: a dozen requests through a server of the same shape first
    # (slot threads' streams, kernels and workspaces), not measured
    warm = Server(budget_bytes=64 << 30, slots=4)
    for t in poisson_load(warm, lambda k: sreqs[k], 2 * args.serve_rate, min(12, len(sreqs)), seed=1):
        t.wait()
    warm.close()

    def run(rate, route=False):
        server = Server(budget_bytes=64 << 30, slots=4, backend=LibraryBackend(route=route))
        t0 = time.perf_counter()
        if rate is None:
            tickets = [server.submit(*sreqs[k]) for k in range(len(sreqs))]
        else:
            tickets = poisson_load(server, lambda k: sreqs[k], rate, len(sreqs), seed=9)
        for t in tickets:
            t.wait()
        wall = time.perf_counter() - t0
        server.close()
        lat = [t.latency_ms for t in tickets]
        pairs = int(sum(t.pairs() for t in tickets))
        return tickets, {"rate_per_s": rate, "p50_ms": pctl(lat, 50), "p99_ms": pctl(lat, 99),
                         "pairs_per_s": round(pairs / wall, 1), "wall_s": round(wall, 3),
                         "batches": len(server.batches), "errors": sum(1 for t in tickets if t.error),
                         "peak_reserved_gb": round(server.peak_reserved / 2**30, 3)}

    base_t, base = run(args.serve_rate)
    sweep = [base]
    knee = base
    rate = args.serve_rate
    while rate < 64 * args.serve_rate:
        rate *= 1.5
        _, r = run(rate)
        sweep.append(r)
        if r["p99_ms"] > 2 * base["p99_ms"]:
            break
        knee = r
    flood_t, flood = run(None)
    # the same mix with the shipped router loaded into the library (the slots'
    # WFA batches take the routed path; results unchanged)
    _, rbase = run(args.serve_rate, route=True)
    _, rflood = run(None, route=True)
    ga.set_router(None)

    # per-class work of the mix: WFA diagonal updates ((w + 1)^2, w the score,
    # or smax when unresolved within the length bound), SW cells
    def work(tickets):
        wfa_upd = wfa_pairs = sw_cells = sw_pairs = 0
        for t in tickets:
            if t.kind == "wfa":
                for it, r in zip(t.items, t.results or []):
                    m, n, s = int(it.pattern_len), int(it.text_len), int(it.smax)
                    w = r.score if r.score >= 0 else (s if abs(m - n) <= s else -1)
                    wfa_upd += (w + 1) ** 2
                    wfa_pairs += 1
            else:
                for it in t.items:
                    sw_cells += int(it.text_len) * int(it.pattern_len)
                    sw_pairs += 1
        return wfa_upd, wfa_pairs, sw_cells, sw_pairs

    wfa_upd, wfa_pairs, sw_cells, sw_pairs = work(flood_t)
    fw = flood["wall_s"]
    per_class = {
        "wfa": {"pairs": wfa_pairs, "pairs_per_s": round(wfa_pairs / fw, 1), "diagonal_updates": wfa_upd,
                "roofline": {"bound": "valu", "unit": "TOP/s", "peak": PEAK_INT32_TOPS,
                             "achieved": round(WFA_OPS_PER_UPDATE * wfa_upd / fw / 1e12, 4),
                             "frac": round(WFA_OPS_PER_UPDATE * wfa_upd / fw / 1e12 / PEAK_INT32_TOPS, 5)}},
        "sw": {"pairs": sw_pairs, "pairs_per_s": round(sw_pairs / fw, 1), "cells": sw_cells,
               "gcups": round(sw_cells / fw / 1e9, 2),
               "roofline": {"bound": "valu", "unit": "TOP/s", "peak": PEAK_INT16_TOPS,
                            "achieved": round(SW_OPS_PER_CELL * sw_cells / fw / 1e12, 4),
                            "frac": round(SW_OPS_PER_CELL * sw_cells / fw / 1e12 / PEAK_INT16_TOPS, 5)}},
        "note": "work of each class over the saturation run's wall time (the classes share the GPU)",
    }
    by = {"wfa": [t for t in base_t if t.kind == "wfa"], "sw": [t for t in base_t if t.kind == "sw"]}
    line = {
        "workload": f"{len(sreqs)} requests, 1..512 pairs each; 60% WFA DNA 100 bp..5 kbp at 0..10% "
                    "divergence, 40% protein SW (query 100..1000 aa, target 100..2000 aa, BLOSUM62 11/1); "
                    "70% of requests with CIGAR; 64 GB budget, 4 slots; after 12 warm-up requests",
        "p50_ms": base["p50_ms"], "p99_ms": base["p99_ms"], "rate_per_s": args.serve_rate,
        "wfa_p50_ms": pctl([t.latency_ms for t in by["wfa"]], 50),
        "sw_p50_ms": pctl([t.latency_ms for t in by["sw"]], 50),
        "pairs_per_s_offered": base["pairs_per_s"],
        "capacity": {"knee_rate_per_s": knee["rate_per_s"], "knee_p99_ms": knee["p99_ms"],
                     "knee_pairs_per_s": knee["pairs_per_s"],
                     "saturation_pairs_per_s": flood["pairs_per_s"], "saturation_p99_ms": flood["p99_ms"],
                     "rule": "highest Poisson rate (x1.5 steps from the base) whose p99 stays within 2x the "
                             "base p99; saturation = every request submitted at once"},
        "routed": {"p50_ms": rbase["p50_ms"], "p99_ms": rbase["p99_ms"], "rate_per_s": args.serve_rate,
                   "saturation_pairs_per_s": rflood["pairs_per_s"], "saturation_p99_ms": rflood["p99_ms"],
                   "errors": rbase["errors"] + rflood["errors"],
                   "note": "the shipped router network loaded into the library (ga_router_set)"},
        "sweep": sweep + [flood],
        "per_class": per_class,
        "errors": sum(r["errors"] for r in sweep) + flood["errors"],
    }
    if cpu_on:
        # the oracle over the same request mix: each class timed on its own
        # pairs, combined by the mix's pair counts
        wfa_items = [it for k, items, cig in sreqs if k == "wfa" for it in items]
        wfa_want = np.array([1 if cig else 0 for k, items, cig in sreqs if k == "wfa" for _ in items], np.uint8)
        sw_items = [it for k, items, cig in sreqs if k == "sw" for it in items]
        wps = wfa_requests_pairset(wfa_items)
        sps = sw_requests_pairset(sw_items)
        cw = cpu_baseline_wfa(wps, wfa_want, threads, O, "the serving mix's WFA pairs")
        cs = cpu_baseline_sw(sps, threads, O, tab=tab62, gaps=(11, 1), label="the serving mix's protein pairs")
        sw_cells_mix = float((sps.a_len.astype(np.int64) * sps.b_len).sum())
        t_cpu = len(wfa_items) / cw["value"] + sw_cells_mix / (cs["value"] * 1e9)
        t1 = len(wfa_items) / cw["single_core"]["value"] + sw_cells_mix / (cs["single_core"]["value"] * 1e9)
        line["cpu_baseline"] = {
            "value": round((len(wfa_items) + len(sw_items)) / t_cpu, 1), "unit": "pairs/s",
            "cores": threads, "kind": "port",
            "sample": f"the oracle on the mix's {len(wfa_items)} WFA pairs ({cw['value']:.0f} pairs/s) and "
                      f"{len(sw_items)} protein pairs ({cs['value']:.3f} GCUPS), each a median of 5 after one "
                      "warm-up, combined by the mix's counts",
            "single_core": {"value": round((len(wfa_items) + len(sw_items)) / t1, 1)},
            "host": cw["host"],
        }
    log(f"serving: {line}")
    return line


if __name__ == "__main__":
    main()
