#!/usr/bin/env python3
"""Benchmark: ISSL MIT+CFD scoring throughput on one MI355X.

Workload (default): 100,000,000 distinct 20-mer sites at 41 % GC (2 % from
2,000 repeat seeds), 10 seeded slice masks of 10 positions, 10,000 guides
(half from the sites, a quarter repeat neighbourhoods, a quarter random),
MIT+CFD (method mitAndCfd), score threshold disabled, maxDist 4.

One step = scoring the 10,000 guides once, with the index and the guides
already resident in HBM (issl_score_device).  `value` is guides/s of the
BATCHED form (persistent tile kernel, FP32 products with compensated FP32
lane sums combined in FP64); the
PLAIN form (one guide per workgroup, FP64) and the CPU scorer (the oracle's
OpenMP restatement of ISSLScoringModule::run on the host cores, same
workload) are reported beside it.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
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

HBM_PEAK_GBS = 8000.0
PCIE_PEAK_GBS = 64.0   # host link of the MI355X: PCIe 5.0 x16, one direction
TIMING_REPEAT = 10   # dominant-kernel dispatches per HIP event pair
METRIC = "guides/sec at 100M-site index (MIT+CFD); p99 latency; HBM GB/s fraction"


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def host_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def postings_per_guide(ix, guides: np.ndarray) -> np.ndarray:
    """P(g): total length of the slice lists each guide visits."""
    n_lists_before = 0
    total = np.zeros(guides.shape[0], dtype=np.int64)
    for s, m in enumerate(ix.masks.tolist()):
        pos = [j for j in range(ix.seq_len) if (m >> j) & 1]
        key = np.zeros(guides.shape[0], dtype=np.int64)
        for j, p in enumerate(pos):
            key |= ((guides >> np.uint64(2 * p)) & np.uint64(3)).astype(np.int64) << (2 * j)
        total += ix.list_sizes[n_lists_before + key].astype(np.int64)
        n_lists_before += 1 << (2 * len(pos))
    return total


def copy_calibration(hiprt) -> int:
    """A copy of known size: 1 GiB read and 1 GiB written per dispatch."""
    nbytes = 1 << 30
    src = hiprt.DeviceArray(nbytes, np.uint8)
    dst = hiprt.DeviceArray(nbytes, np.uint8)
    for r in range(3):
        t0 = time.perf_counter()
        hiprt.copy_device(dst, src, nbytes)
        hiprt.synchronize()
        log(f"copy {r}: 1 GiB device-to-device in {(time.perf_counter() - t0) * 1e3:.3f} ms")
    return 0


def measured_traffic(workload: str, layout: str):
    """HBM bytes per k_tiles launch from the committed counter summary.

    The summary holds FETCH_SIZE + WRITE_SIZE per dispatch of k_tiles on the
    named workload, for each site-word layout ("inline" or "residual"), each
    scaled by the reading of a 1 GiB copy.  Returns None when no summary
    covers this workload and layout.
    """
    path = os.path.join(ROOT, "profiles", f"k_tiles_{workload}_hbm_traffic.md")
    key = f"traffic_bytes_per_launch_{layout}:"
    try:
        for line in open(path):
            if line.startswith(key):
                return int(line.split(":", 1)[1].strip().replace(",", ""))
    except OSError:
        pass
    return None


def stream_leg(issl, hiprt, hx, guides, ix, args) -> dict:
    """Streamed index over the step's host arrays: every call copies all S
    slices (S * N * 8 bytes) through the slice buffers while scoring."""
    n = hx.n_sites
    reps = (args.stream_guides + guides.shape[0] - 1) // guides.shape[0]
    g = np.ascontiguousarray(np.tile(guides, reps)[:args.stream_guides])
    Qs = g.shape[0]
    sx = issl.Index.from_arrays_streamed(hx, slot_bytes=args.stream_slots * 16 * n)
    slots = sx.stream_stats()[0]
    d_g = hiprt.DeviceArray.from_host(g)
    d_o = {"mit": hiprt.DeviceArray(Qs, np.float64), "cfd": hiprt.DeviceArray(Qs, np.float64),
           "occ_count": hiprt.DeviceArray(Qs, np.uint64),
           "pass": hiprt.DeviceArray(Qs, np.uint8)}
    out = {"slots": int(slots), "slices": int(hx.n_slices), "guides_per_call": int(Qs),
           "bytes_per_call": int(hx.n_slices * n * 8), "pcie_peak_GBps": PCIE_PEAK_GBS}
    for name, kw in (("batched", dict(form=issl.FORM_BATCHED, precision=issl.PREC_FP32C)),
                     ("plain", dict(form=issl.FORM_PLAIN))):
        for _ in range(2):
            sx.score_device(d_g, Qs, d_o, method=issl.METHOD_MIT_AND_CFD, threshold=0.0, **kw)
        hiprt.synchronize()
        _, up0, _ = sx.stream_stats()
        calls = []
        t0 = time.perf_counter()
        for _ in range(args.stream_calls):
            tc = time.perf_counter()
            sx.score_device(d_g, Qs, d_o, method=issl.METHOD_MIT_AND_CFD, threshold=0.0, **kw)
            hiprt.synchronize()
            calls.append((time.perf_counter() - tc) * 1e3)
        dt = (time.perf_counter() - t0) / args.stream_calls
        _, up1, _ = sx.stream_stats()
        gbps = (up1 - up0) / args.stream_calls / dt / 1e9
        out[name] = {"ms_per_call": round(dt * 1e3, 2), "guides_per_s": round(Qs / dt, 1),
                     "h2d_GBps": round(gbps, 2), "pcie_frac": round(gbps / PCIE_PEAK_GBS, 3),
                     "call_ms": [round(c, 1) for c in calls]}
        log(f"streamed {name}: {slots} of {hx.n_slices} slices in HBM, {Qs:,} guides per call, "
            f"{dt * 1e3:.1f} ms/call, {Qs / dt:,.0f} guides/s, H2D {gbps:.1f} GB/s")
    # Results against the resident index on the step's guides.
    q = np.ascontiguousarray(guides)
    a = sx.score(q, method=issl.METHOD_MIT_AND_CFD, form=issl.FORM_PLAIN)
    b = ix.score(q, method=issl.METHOD_MIT_AND_CFD, form=issl.FORM_PLAIN)
    out["plain_bit_identical_to_resident"] = bool(
        (a.cfd.view(np.uint64) == b.cfd.view(np.uint64)).all() and
        (a.mit.view(np.uint64) == b.mit.view(np.uint64)).all() and
        (a.occ_count == b.occ_count).all())
    c = sx.score(q, method=issl.METHOD_MIT_AND_CFD, form=issl.FORM_BATCHED,
                 precision=issl.PREC_FP32C)
    out["batched_max_rel_vs_resident_plain"] = float(
        max(np.max(np.abs(c.mit - b.mit) / b.mit), np.max(np.abs(c.cfd - b.cfd) / b.cfd)))
    sx.close()
    return out


def extract_leg(issl, args, masks) -> dict:
    """Off-target extraction over a synthetic genome (41 % GC, uniform
    bases, 250 Mbase chromosomes): scan, sort and count, the sorted file
    formatted (to /dev/null) and an index built from the sites, all timed
    on the device path; the oracle's single-thread scan on a sample beside
    it."""
    rng = np.random.default_rng(0xE7)
    lut = np.frombuffer(b"A" * 29 + b"C" * 20 + b"G" * 21 + b"T" * 30, dtype=np.uint8)
    n = args.extract_bases
    genome = lut[rng.integers(0, 100, size=n, dtype=np.uint8)].tobytes()
    chrom = 250_000_000
    with issl.Extractor() as warm:  # loads the kernels and the sort's code
        warm.add_sequence(genome[:1_000_000])
        warm.finish()
        warm.write(os.devnull)
    out = {"bases": n, "chromosome_bases": chrom}
    with issl.Extractor() as ex:
        mv = memoryview(genome)
        t0 = time.perf_counter()
        for i in range(0, n, chrom):
            ex.add_sequence(mv[i:i + chrom])
        ex.stats()
        t1 = time.perf_counter()
        sites, distinct = ex.finish()
        t2 = time.perf_counter()
        ex.write(os.devnull)
        t3 = time.perf_counter()
        with ex.build_index(masks) as dix:
            t4 = time.perf_counter()
            info = dix.info()
    out.update({"sites": sites, "distinct": distinct,
                "scan_s": round(t1 - t0, 3), "sort_count_s": round(t2 - t1, 3),
                "write_s": round(t3 - t2, 3), "index_build_s": round(t4 - t3, 3),
                "scan_Gbases_per_s": round(n / (t1 - t0) / 1e9, 2),
                "genome_to_sorted_sites_Gbases_per_s": round(n / (t2 - t0) / 1e9, 2),
                "index_slices": int(info.n_slices)})
    log(f"extraction: {n / 1e6:.0f} Mbases -> {sites:,} sites ({distinct:,} distinct): scan "
        f"{t1 - t0:.3f} s, sort+count {t2 - t1:.3f} s, format {t3 - t2:.3f} s, index "
        f"{t4 - t3:.3f} s")
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        sample = genome[:args.extract_cpu_bases]
        t0 = time.perf_counter()
        oracle_lib.lib().or_extract(sample, len(sample), None, 0)
        dt = time.perf_counter() - t0
        out["cpu_scan"] = {"bases_per_s": round(len(sample) / dt, 1), "cores": 1, "kind": "port",
                           "sample": f"the genome's first {len(sample):,} bases, the oracle's "
                                     f"pattern scan only (no sort)"}
        log(f"extraction CPU scan: {len(sample) / dt / 1e6:.1f} Mbases/s on one core")
    return out


def serve_poisson(issl, srv, index_ids, pools, rate, n_requests, size_of, rng):
    """Open-loop Poisson arrivals at `rate` req/s through the request loop.

    Each request picks an index id, a size from size_of(rng) and a random run
    of that index's guide pool; latency is submit -> results in host buffers,
    from the server's own per-request clock."""
    srv.stats(reset=True)
    gaps = rng.exponential(1.0 / rate, size=n_requests)
    reqs = []
    t0 = time.perf_counter()
    t_next = t0
    for i in range(n_requests):
        t_next += gaps[i]
        k = int(rng.integers(0, len(index_ids)))
        pool = pools[k]
        n = size_of(rng)
        lo = int(rng.integers(0, max(1, pool.shape[0] - n)))
        g = pool[lo:lo + n] if lo + n <= pool.shape[0] else np.resize(pool[lo:], n)
        while True:
            now = time.perf_counter()
            if now >= t_next:
                break
            time.sleep(min(t_next - now, 0.001))
        reqs.append(srv.submit(index_ids[k], g, method=issl.METHOD_MIT_AND_CFD, threshold=0.0,
                               form=issl.FORM_BATCHED, precision=issl.PREC_FP32C))
    for r in reqs:
        r.wait(timeout=120.0)
    wall = time.perf_counter() - t0
    st = srv.stats(reset=True)
    return {"rate_rps": rate, "requests": int(st.requests_done),
            "failed": int(st.requests_failed), "guides": int(st.guides_done),
            "achieved_rps": round(st.requests_done / wall, 1),
            "guides_per_s": round(st.guides_done / wall, 1),
            "p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3),
            "max_ms": round(st.max_ms, 3), "mean_ms": round(st.mean_ms, 3),
            "launches": int(st.launches), "mean_batch_guides": round(st.mean_batch_guides, 1)}


def serving_main(args) -> int:
    """Serving workload: 100 M- and 300 M-site indexes resident together,
    Poisson arrivals at several rates, request sizes log-uniform in
    1..50,000 guides, a random index per request."""
    import cracklingplusplus_gpu_amd as issl
    from cracklingplusplus_gpu_amd import workloads
    threads = host_threads()
    srv = issl.Server(max_batch=65536, hbm_reserve=8 << 30)
    ids, pools = [], []
    for name in ("genome100m_10x10", "genome300m_10x10"):
        log(f"generating {name}")
        sigs, occ, masks, guides, kinds = workloads.make(name, threads=threads)
        hx = issl.build_lists(sigs, occ, masks, threads=threads)
        del sigs, occ
        log(f"admitting {name}: {hx.n_sites:,} sites")
        ids.append(srv.add_index(hx))
        pools.append(guides)
        del hx
    st = srv.stats()
    log(f"resident {st.resident_bytes / 1e9:.1f} GB, free {st.hbm_free / 1e9:.1f} GB")
    rng = np.random.default_rng(0x5E4E)
    hi = np.log(50000.0)

    def size_of(r):
        return max(1, int(np.exp(r.uniform(0.0, hi))))

    # warm both indexes (workspaces, tuning defaults) outside the measurement
    serve_poisson(issl, srv, ids, pools, 50.0, 10, size_of, rng)
    out = []
    for rate in args.serve_rates:
        n_req = max(20, int(rate * args.serve_seconds))
        log(f"rate {rate} req/s: {n_req} requests")
        res = serve_poisson(issl, srv, ids, pools, rate, n_req, size_of, rng)
        log(f"  p50 {res['p50_ms']} ms  p99 {res['p99_ms']} ms  {res['guides_per_s']:,.0f} guides/s")
        out.append(res)
    # Baseline: the PLAIN kernel, one request at a time (closed loop), same size mix.
    plain_lat, plain_guides = [], 0
    t0 = time.perf_counter()
    for i in range(args.plain_requests):
        k = int(rng.integers(0, len(ids)))
        n = size_of(rng)
        lo = int(rng.integers(0, max(1, pools[k].shape[0] - n)))
        g = pools[k][lo:lo + n] if lo + n <= pools[k].shape[0] else np.resize(pools[k][lo:], n)
        t1 = time.perf_counter()
        srv.submit(ids[k], g, form=issl.FORM_PLAIN, precision=issl.PREC_FP64).wait(timeout=300.0)
        plain_lat.append((time.perf_counter() - t1) * 1e3)
        plain_guides += n
    wall = time.perf_counter() - t0
    plain_lat = np.array(plain_lat)
    plain = {"requests": len(plain_lat), "guides_per_s": round(plain_guides / wall, 1),
             "p50_ms": round(float(np.percentile(plain_lat, 50)), 3),
             "p99_ms": round(float(np.percentile(plain_lat, 99)), 3),
             "mode": "PLAIN form, one request at a time"}
    log(f"plain one-at-a-time: p50 {plain['p50_ms']} ms  p99 {plain['p99_ms']} ms  "
        f"{plain['guides_per_s']:,.0f} guides/s")
    line = {"metric": "serving latency, two resident indexes (100 M + 300 M sites)",
            "unit": "ms", "data": "synthetic", "n_gpus": 1,
            "config": {"workload": "serve_genome100m+300m", "sizes": "log-uniform 1..50000",
                       "arrivals": "Poisson", "method": "mitAndCfd", "threshold": 0,
                       "form": "batched", "precision": "fp32c", "max_batch": 65536,
                       "resident_GB": round(st.resident_bytes / 1e9, 2)},
            "rates": out, "plain_one_at_a_time": plain}
    print(json.dumps(line), flush=True)
    srv.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="genome100m_10x10")
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--cpu-sample", type=int, default=10000,
                    help="guides timed on the host CPU scorer")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--latency-requests", type=int, default=200)
    ap.add_argument("--tune-cache", default=os.path.join(tempfile.gettempdir(), "issl_tune.txt"))
    ap.add_argument("--serve-rate", type=float, default=500.0,
                    help="Poisson arrival rate of the 1,000-guide serving leg (req/s)")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serving", action="store_true",
                    help="only the two-index serving workload (100 M + 300 M sites)")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[10, 50, 100, 250, 500])
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--plain-requests", type=int, default=100,
                    help="closed-loop PLAIN requests in the serving workload's baseline")
    ap.add_argument("--graph", action="store_true",
                    help="replay the step's kernels from a captured hipGraph")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="minimum CPU scoring time (the batch is repeated)")
    ap.add_argument("--stream-slots", type=int, default=4,
                    help="slice buffers of the streamed-index leg (0: skip the leg)")
    ap.add_argument("--stream-guides", type=int, default=100_000,
                    help="guides per call of the streamed-index leg")
    ap.add_argument("--stream-calls", type=int, default=3)
    ap.add_argument("--extract-bases", type=int, default=500_000_000,
                    help="synthetic genome of the extraction leg (0: skip the leg)")
    ap.add_argument("--extract-cpu-bases", type=int, default=50_000_000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device three times (scales the "
                         "FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the scorer "
              "runs on one device and this benchmark reports that device's throughput.",
              file=sys.stderr)
        return 2

    import cracklingplusplus_gpu_amd as issl
    from cracklingplusplus_gpu_amd import hiprt, workloads

    if issl.device_count() <= 0:
        print("no HIP device", file=sys.stderr)
        return 1
    if args.copy_calibration:
        return copy_calibration(hiprt)
    if args.serving:
        return serving_main(args)
    threads = host_threads()
    t_setup = time.time()
    log(f"generating workload {args.workload} with {threads} threads")
    sigs, occ, masks, guides, kinds = workloads.make(args.workload, threads=threads)
    log(f"sites {sigs.shape[0]:,}  guides {guides.shape[0]:,}; building slice lists")
    t_b = time.perf_counter()
    hx = issl.build_lists(sigs, occ, masks, threads=threads)
    host_build_s = time.perf_counter() - t_b
    log("uploading index")
    ix = issl.Index.from_arrays(hx)
    info = ix.info()
    log(f"index resident: {info.device_bytes / 1e9:.2f} GB of HBM")
    Q = guides.shape[0]
    S = hx.n_slices
    P = postings_per_guide(hx, guides)
    p_total = int(P.sum())
    log(f"postings per step {p_total:,} (mean {P.mean():.0f}, max {P.max():,})")

    d_guides = hiprt.DeviceArray.from_host(guides)
    d_out = {"mit": hiprt.DeviceArray(Q, np.float64), "cfd": hiprt.DeviceArray(Q, np.float64),
             "occ_count": hiprt.DeviceArray(Q, np.uint64),
             "pass": hiprt.DeviceArray(Q, np.uint8)}
    log("autotuning the batched kernel")
    os.makedirs(os.path.dirname(args.tune_cache), exist_ok=True)
    tn = ix.autotune(guides, args.tune_cache)
    log(f"tuning: chunk {tn.chunk} depth {tn.depth} grid/CU {tn.grid_per_cu} "
        f"stream depth {tn.direct_depth} "
        f"heavy {tn.heavy_kpost} ({tn.us_per_guide * 1e3:.2f} ns/guide)")
    setup_s = time.time() - t_setup

    kw = dict(method=issl.METHOD_MIT_AND_CFD, threshold=0.0, form=issl.FORM_BATCHED,
              precision=issl.PREC_FP32C)
    # With --graph the step's launch sequence is captured in a hipGraph by the
    # first warm-up call and replayed afterwards (measured slower than eager
    # launches for this three-kernel sequence, so off by default).
    kw_step = dict(kw, flags=issl.FLAG_GRAPH if args.graph else 0)

    # ---- timed region: batched, inputs resident in HBM
    for _ in range(max(1, args.warmup)):
        ix.score_device(d_guides, Q, d_out, **kw_step)
    hiprt.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ix.score_device(d_guides, Q, d_out, **kw_step)
    hiprt.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = Q * args.steps / (t1 - t0)
    log(f"batched: {ms_per_step:.3f} ms/step, {value:,.0f} guides/s")

    # ---- dominant kernel time (k_tiles): HIP events on the index's stream around
    # TIMING_REPEAT back-to-back dispatches of k_tiles per step (it is idempotent)
    ix.timing_enable(True, repeat=TIMING_REPEAT)
    for _ in range(args.steps):
        ix.score_device(d_guides, Q, d_out, **kw)
    kms, nl = ix.timing_read()
    ix.timing_enable(False)
    k_avg_ms = kms / max(nl, 1)
    res_built, res_used, _ = ix.residual()
    # Algorithmic bytes per k_tiles dispatch: the reference's method reads an
    # 8 B posting entry and gathers an 8 B signature per posting, plus two 8 B
    # list bounds per (guide, slice).  That figure does not depend on how this
    # build lays the sites out; the bytes its own layout streams (one 4 B
    # list-residual or 8 B signature-inline word per posting, plus the bounds)
    # and the measured HBM traffic are reported beside it.
    word = 4 if res_used else 8
    bytes_layout = word * p_total + 16 * S * Q
    bytes_inline = 8 * p_total + 16 * S * Q
    bytes_ref_layout = 16 * p_total + 16 * S * Q
    achieved = bytes_ref_layout / (k_avg_ms * 1e-3) / 1e9
    # The timed dispatch: the tile kernel, the one-wave-per-guide kernel, or
    # both together when the tuning sends only long guides to the tile kernel.
    kernel = ("k_tiles" if tn.heavy_kpost == 0 else
              "k_stream" if tn.heavy_kpost == issl._capi.HEAVY_NEVER else "k_stream+k_tiles")
    # (the committed counter summaries cover the tile kernel's dispatches)
    traffic = (measured_traffic(args.workload, "residual" if res_used else "inline")
               if kernel == "k_tiles" else None)
    log(f"{kernel} avg {k_avg_ms * 1e3:.1f} us/launch -> {achieved:,.0f} GB/s")

    res_batched = {k: v.to_host() for k, v in d_out.items()}

    # ---- the same step with the other site-word layout at the same launch
    # shape (the autotuner picks the layout with the shape); results must be
    # the same bits
    other_leg = None
    if res_built:
        ix.set_residual(not res_used)
        for _ in range(max(1, args.warmup)):
            ix.score_device(d_guides, Q, d_out, **kw)
        hiprt.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ix.score_device(d_guides, Q, d_out, **kw)
        hiprt.synchronize()
        ms_i = (time.perf_counter() - t0) * 1e3 / args.steps
        ix.timing_enable(True, repeat=TIMING_REPEAT)
        for _ in range(args.steps):
            ix.score_device(d_guides, Q, d_out, **kw)
        kms_i, nl_i = ix.timing_read()
        ix.timing_enable(False)
        k_i = kms_i / max(nl_i, 1)
        same_i = all(np.array_equal(d_out[k].to_host().view(np.uint8), res_batched[k].view(np.uint8))
                     for k in d_out)
        word_o = 8 if res_used else 4
        other_leg = {"site_words": "signature-inline, 8 B" if res_used else "list-residual, 4 B",
                     "value": round(Q / (ms_i * 1e-3), 1), "ms_per_step": round(ms_i, 4),
                     "kernel_us": round(k_i * 1e3, 2),
                     "layout_GBps": round((word_o * p_total + 16 * S * Q) / (k_i * 1e-3) / 1e9, 1),
                     "same_bits": bool(same_i)}
        log(f"other layout: {ms_i:.3f} ms/step, k_tiles {k_i * 1e3:.1f} us, same bits {same_i}")
        ix.set_residual(res_used)

    # ---- plain form, same workload
    kwp = dict(kw, form=issl.FORM_PLAIN, precision=issl.PREC_FP64)
    ix.score_device(d_guides, Q, d_out, **kwp)
    hiprt.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        ix.score_device(d_guides, Q, d_out, **kwp)
    hiprt.synchronize()
    t1 = time.perf_counter()
    plain_ms = (t1 - t0) * 1e3 / args.plain_steps
    ix.timing_enable(True, repeat=TIMING_REPEAT)
    for _ in range(args.plain_steps):
        ix.score_device(d_guides, Q, d_out, **kwp)
    pkms, pnl = ix.timing_read()
    ix.timing_enable(False)
    plain_k_ms = pkms / max(pnl, 1)
    res_plain = {k: v.to_host() for k, v in d_out.items()}
    log(f"plain: {plain_ms:.3f} ms/step, kernel {plain_k_ms * 1e3:.1f} us")

    # ---- precision of the batched FP32C form against the FP64 in-order path
    rel = {}
    for f in ("mit", "cfd"):
        a, b = res_batched[f], res_plain[f]
        rel[f] = np.abs(a - b) / np.abs(b)
    same_counts = bool((res_batched["occ_count"] == res_plain["occ_count"]).all())

    # ---- request latency: 1,000-guide requests, host buffers, one at a time
    lat = []
    rng = np.random.default_rng(7)
    for i in range(args.latency_requests):
        lo = int(rng.integers(0, max(1, Q - 1000)))
        req = guides[lo:lo + 1000]
        t0 = time.perf_counter()
        ix.score(req, method=issl.METHOD_MIT_AND_CFD, threshold=0.0,
                 form=issl.FORM_BATCHED, precision=issl.PREC_FP32C)
        lat.append((time.perf_counter() - t0) * 1e3)
    lat = np.array(lat)

    # ---- the whole step through host buffers (issl_score): guides uploaded,
    # results copied back, both ways over PCIe; reported beside `value`
    hres = ix.results(Q)
    for _ in range(max(1, args.warmup)):
        ix.score(guides, out=hres, **kw)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ix.score(guides, out=hres, **kw)
    host_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    host_same = bool(np.array_equal(hres.cfd.view(np.uint64), res_batched["cfd"].view(np.uint64)))
    log(f"host buffers: {host_ms:.3f} ms/call, {Q / host_ms * 1e3:,.0f} guides/s")

    # ---- index construction on the device (sites and occurrences uploaded,
    # lists sorted in HBM) against the host builder above
    t_b = time.perf_counter()
    with issl.Index.build(sigs, occ, masks) as dix:
        device_build_s = time.perf_counter() - t_b
        dsizes, _ = dix.export(entries=False)
    del occ
    build = {"host_lists_s": round(host_build_s, 3), "host_threads": threads,
             "device_build_s": round(device_build_s, 3),
             "list_sizes_identical": bool(np.array_equal(dsizes, hx.list_sizes))}
    log(f"index build: host lists {host_build_s:.2f} s ({threads} threads), device "
        f"(upload + sort + postings) {device_build_s:.2f} s, sizes identical "
        f"{build['list_sizes_identical']}")

    # ---- streamed index: postings in host memory, slices streamed through
    # args.stream_slots HBM slice buffers on every call
    streamed = None
    if args.stream_slots > 0:
        streamed = stream_leg(issl, hiprt, hx, guides, ix, args)

    # ---- off-target extraction (the step before the index) on a synthetic genome
    extraction = extract_leg(issl, args, masks) if args.extract_bases > 0 else None

    # ---- serving: Poisson arrivals of 1,000-guide requests through the request loop
    log(f"serving: {args.serve_requests} requests of 1,000 guides at {args.serve_rate:g} req/s")
    srv = issl.Server(max_batch=65536)
    sid = srv.add_index(hx)
    srv_rng = np.random.default_rng(0x5E4E)
    serve_poisson(issl, srv, [sid], [guides], args.serve_rate, 20, lambda r: 1000, srv_rng)
    serving = serve_poisson(issl, srv, [sid], [guides], args.serve_rate, args.serve_requests,
                            lambda r: 1000, srv_rng)
    srv.close()
    log(f"serving: p50 {serving['p50_ms']} ms, p99 {serving['p99_ms']} ms, "
        f"mean batch {serving['mean_batch_guides']:.0f} guides")

    # ---- CPU baseline: the oracle's OpenMP scorer on the host cores
    cpu = None
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        n_cpu = min(args.cpu_sample, Q)
        sample = guides[:n_cpu]
        log(f"CPU scorer: {n_cpu} guides on {threads} threads, repeated for "
            f">= {args.cpu_seconds:g} s")
        t0 = time.perf_counter()
        o = oracle_lib.score(hx, sample, issl.METHOD_MIT_AND_CFD, 0.0, threads=threads)
        passes = 1
        while time.perf_counter() - t0 < args.cpu_seconds:
            oracle_lib.score(hx, sample, issl.METHOD_MIT_AND_CFD, 0.0, threads=threads)
            passes += 1
        t1 = time.perf_counter()
        cpu_rate = n_cpu * passes / (t1 - t0)
        cpu_match = bool((o["cfd"].view(np.uint64) ==
                          res_plain["cfd"][:n_cpu].view(np.uint64)).all())
        cpu = {"value": round(cpu_rate, 1), "unit": "guides/s", "cores": threads,
               "kind": "port",
               "sample": f"the step's first {n_cpu} guides scored {passes} times, index "
                         f"preloaded, scoring loop only ({t1 - t0:.1f} s), {cpu_model()}",
               "plain_gpu_bit_exact": cpu_match}
        log(f"CPU: {cpu_rate:,.0f} guides/s (GPU plain bit-exact: {cpu_match})")

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "guides/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64+f32c",
        "data": "synthetic",
        "config": {"workload": args.workload, "n_sites": int(hx.n_sites), "n_slices": S,
                   "slice_width": 10 if args.workload.startswith("genome") else None,
                   "guides_per_step": Q, "method": "mitAndCfd", "threshold": 0,
                   "max_dist": 4, "form": "batched", "precision": "fp32c",
                   "launch": "hipGraph replay" if args.graph else "eager",
                   "postings_per_step": p_total,
                   "site_words": "list-residual, 4 B" if res_used else "signature-inline, 8 B",
                   "tuning": {"chunk": tn.chunk, "depth": tn.depth,
                              "grid_per_cu": tn.grid_per_cu,
                              "stream_depth": tn.direct_depth,
                              "heavy_kpost": tn.heavy_kpost}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": kernel,
                     "kernel_us": round(k_avg_ms * 1e3, 2),
                     "bytes_per_launch": bytes_ref_layout,
                     "bytes_basis": "16 B per posting (8 B entry + 8 B signature gather) "
                                    "+ 16 B list bounds per (guide, slice)",
                     "layout_bytes_per_launch": bytes_layout,
                     "layout_GBps": round(bytes_layout / (k_avg_ms * 1e-3) / 1e9, 1),
                     "moved_GBps": (round(traffic / (k_avg_ms * 1e-3) / 1e9, 1)
                                    if traffic else None),
                     "moved_frac": (round(traffic / (k_avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                    if traffic else None)},
        "cpu_baseline": cpu,
        "other_layout": other_leg,
        "plain": {"value": round(Q / (plain_ms * 1e-3), 1), "unit": "guides/s",
                  "ms_per_step": round(plain_ms, 4), "kernel": "k_plain",
                  "kernel_us": round(plain_k_ms * 1e3, 2),
                  "achieved_GBps": round(bytes_ref_layout / (plain_k_ms * 1e-3) / 1e9, 1),
                  "layout_GBps": round(bytes_inline / (plain_k_ms * 1e-3) / 1e9, 1),
                  "batched_speedup": round(plain_ms / ms_per_step, 3)},
        "precision": {"fp32c_vs_fp64_max_rel": {f: float(rel[f].max()) for f in rel},
                      "fp32c_vs_fp64_mean_rel": {f: float(rel[f].mean()) for f in rel},
                      "counts_identical": same_counts},
        "host_buffers": {"value": round(Q / host_ms * 1e3, 1), "unit": "guides/s",
                         "ms_per_call": round(host_ms, 4), "calls": args.steps,
                         "same_as_device_buffers": host_same,
                         "mode": "issl_score: guides from host memory, the seven result "
                                 "columns to host memory (PCIe both ways)"},
        "latency_ms": {"request_guides": 1000, "requests": int(lat.size),
                       "p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "mode": "one request at a time, host buffers in/out"},
        "serving": dict(serving, request_guides=1000,
                        mode="request loop, Poisson arrivals, host buffers in/out"),
        "index_build": build,
        "streamed": streamed,
        "extraction": extraction,
        "setup_s": round(setup_s, 1),
    }
    if cpu:
        line["gpu_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)
    ix.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
