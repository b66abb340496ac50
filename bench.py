#!/usr/bin/env python3
"""Benchmark: ISSL MIT+CFD scoring throughput on one MI355X.

Workload (default): 100,000,000 distinct 20-mer sites at 41 % GC (2 % from
2,000 repeat seeds), 10 seeded slice masks of 10 positions, batches of 10,000
guides (half from the sites, a quarter repeat neighbourhoods, a quarter
random), MIT+CFD (method mitAndCfd), score threshold disabled, maxDist 4.

One step = scoring one 10,000-guide batch, with the index and the guides
already resident in HBM (issl_score_device).  The timed steps rotate over
--batches disjoint batches of that mix (batch 0 is the workload's own 10,000
guides), so that the lists the steps stream do not fit the 256 MB Infinity
Cache together; the autotuner picks the launch shape on a further, held-out
batch.  `value` is guides/s of the BATCHED form in FP64 (persistent tile
kernel, FP64 terms and sums: the reference's arithmetic).  Beside it: the
same steps in compensated FP32 (FP32C), the PLAIN form (one guide per
workgroup, FP64, bit-exact), the other site-word layout, the CPU scorer (the
oracle's OpenMP restatement of ISSLScoringModule::run on the host cores),
100,000 guides in one launch on the same index, and the serving loop with the
100 M- and 300 M-site indexes resident together.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
    python bench.py --headline-only        (the timed steps and kernel timing only)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

# The host cores this process may use, read before any OpenMP runtime loads:
# with OMP_PROC_BIND set, the runtime binds the main thread to one core when it
# starts, and the affinity mask then shows that core alone.
try:
    AFFINITY_CPUS = len(os.sched_getaffinity(0))
except AttributeError:
    AFFINITY_CPUS = os.cpu_count() or 1
# The CPU scorer's OpenMP threads stay on neighbouring cores; the runtime reads
# this when it loads, so it is set before any library that links it.
os.environ.setdefault("OMP_PROC_BIND", "close")

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PCIE_PEAK_GBS = 64.0   # host link of the MI355X: PCIe 5.0 x16, one direction
INFINITY_CACHE_BYTES = 256 << 20
METRIC = "guides/sec at 100M-site index (MIT+CFD); p99 latency; HBM GB/s fraction"


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_info() -> dict:
    """Host cores this process may use (its affinity mask at start-up), with
    the cgroup CPU quota and OMP_NUM_THREADS recorded beside it."""
    aff = AFFINITY_CPUS
    quota = None
    try:
        quota = open("/sys/fs/cgroup/cpu.max").read().strip()
    except OSError:
        pass
    return {"threads": aff, "affinity_cpus": aff, "cgroup_cpu_max": quota,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"),
            "system_cpus": os.cpu_count(), "model": cpu_model(),
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND")}


def host_threads() -> int:
    return cpu_info()["threads"]


def list_index(ix, guides: np.ndarray) -> np.ndarray:
    """Global list number (slice-major over all slices' 4^w lists) of every
    (guide, slice): shape (n, S)."""
    out = np.zeros((guides.shape[0], ix.n_slices), dtype=np.int64)
    base = 0
    for s, m in enumerate(ix.masks.tolist()):
        pos = [j for j in range(ix.seq_len) if (m >> j) & 1]
        key = np.zeros(guides.shape[0], dtype=np.int64)
        for j, p in enumerate(pos):
            key |= ((guides >> np.uint64(2 * p)) & np.uint64(3)).astype(np.int64) << (2 * j)
        out[:, s] = base + key
        base += 1 << (2 * len(pos))
    return out


def postings_per_guide(ix, guides: np.ndarray) -> np.ndarray:
    """P(g): total length of the slice lists each guide visits."""
    return ix.list_sizes[list_index(ix, guides)].astype(np.int64).sum(axis=1)


def footprint_bytes(ix, batches, word: int) -> int:
    """Bytes of site words the batches stream together: every list any of
    them visits, once."""
    ids = np.unique(np.concatenate([list_index(ix, b).ravel() for b in batches]))
    return int(ix.list_sizes[ids].astype(np.int64).sum()) * word


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




def measured_traffic(kernel_tag: str, workload: str, layout: str):
    """HBM bytes per dispatch of a kernel from its committed counter summary.

    profiles/<kernel_tag>_<workload>_hbm_traffic.md holds FETCH_SIZE +
    WRITE_SIZE per dispatch on the named workload, for each site-word layout
    ("inline" or "residual"), each scaled by the reading of a 1 GiB copy.
    Returns None when no summary covers them."""
    path = os.path.join(ROOT, "profiles", f"{kernel_tag}_{workload}_hbm_traffic.md")
    key = f"traffic_bytes_per_launch_{layout}:"
    try:
        for line in open(path):
            if line.startswith(key):
                return int(line.split(":", 1)[1].strip().replace(",", ""))
    except OSError:
        pass
    return None


class Rotation:
    """Disjoint guide batches resident in HBM, each with its own device result
    columns; step k scores batch k mod R."""

    def __init__(self, hiprt, batches):
        self.batches = [np.ascontiguousarray(b) for b in batches]
        self.Q = self.batches[0].shape[0]
        self.d_g = [hiprt.DeviceArray.from_host(b) for b in self.batches]
        self.d_out = [{"mit": hiprt.DeviceArray(self.Q, np.float64),
                       "cfd": hiprt.DeviceArray(self.Q, np.float64),
                       "occ_count": hiprt.DeviceArray(self.Q, np.uint64),
                       "site_count": hiprt.DeviceArray(self.Q, np.uint64),
                       "pass": hiprt.DeviceArray(self.Q, np.uint8)} for _ in self.batches]

    def __len__(self):
        return len(self.batches)

    def run(self, ix, steps: int, start: int, streams=None, **kw) -> None:
        """Steps start .. start+steps-1; step k on streams[k mod len] (None:
        the index's own stream)."""
        for k in range(steps):
            b = (start + k) % len(self)
            s = streams[k % len(streams)] if streams else None
            ix.score_device(self.d_g[b], self.Q, self.d_out[b], stream=s, **kw)

    def results(self):
        return [{k: v.to_host() for k, v in o.items()} for o in self.d_out]


def timed(hiprt, rot, ix, steps, warmup, streams=None, **kw):
    """Warm-up steps, then `steps` timed steps between two device
    synchronizations (hipDeviceSynchronize: every stream); ms per step."""
    rot.run(ix, max(1, warmup), 0, streams, **kw)
    hiprt.synchronize()
    t0 = time.perf_counter()
    rot.run(ix, steps, 0, streams, **kw)
    hiprt.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def pick_streams(hiprt, ix, g_tune, bound, **kw):
    """Streams the steps run on: the index's own stream alone; the two bound
    streams (each with its own workspace); or the index's stream and both
    bound ones; with the tile kernels' workgroups per CU on bound streams
    (ovl_grid) 2 to 5.  (Three streams of the process plus the null stream
    fill the four hardware queues HIP gives a process by default; more
    streams would share queues.)  Each choice is timed on the held-out batch
    and the fastest kept (its ovl_grid set).  Returns (streams, ovl_grid,
    {choice: us per step})."""
    rot = Rotation(hiprt, [g_tune])
    us1 = timed(hiprt, rot, ix, 30, 6, None, **kw) * 1e3
    sweep = {"1": round(us1, 2)}
    best = (us1, [None], 0)
    for label, ss in (("2 bound", bound[:2]), ("index + 2 bound", [None] + bound[:2])):
        for grid in (2, 3, 4, 5):
            ix.set_overlap_grid(grid)
            us = timed(hiprt, rot, ix, 30, 6, ss, **kw) * 1e3
            sweep[f"{label}, ovl_grid {grid}"] = round(us, 2)
            if us < best[0]:
                best = (us, ss, grid)
    ix.set_overlap_grid(best[2])
    return best[1], best[2], sweep


def kernel_us(hiprt, rot, ix, steps, **kw):
    """Average duration of the dominant kernel's dispatches over `steps`
    rotated steps: a HIP event pair attached to each dispatch on the stream
    it runs on (one dispatch per pair, no repeats of a batch)."""
    ix.timing_enable(True, repeat=1)
    rot.run(ix, steps, 0, **kw)
    hiprt.synchronize()
    kms, nl = ix.timing_read()
    ix.timing_enable(False)
    return kms * 1e3 / max(nl, 1), nl


def max_rel(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b))) if a.size else 0.0


def kernel_name(issl, tn):
    return ("k_tiles" if tn.heavy_kpost == 0 else
            "k_stream" if tn.heavy_kpost == issl._capi.HEAVY_NEVER else "k_stream+k_tiles")


def tail_leg(issl, hiprt, ix, hx, batches, kinds_all, args):
    """A heavy tail in one launch: the posting volume of batch 0's guides, and
    the tile kernel's time for that mixed batch against its two parts
    (repeat-neighbourhood guides, unique guides) scored alone, and against a
    tail-free control: guides of the other batches without the repeat
    neighbourhoods, taken in order until their postings equal the mixed
    batch's.  BATCHED FP64."""
    Q = batches[0].shape[0]
    batch, kinds = batches[0], kinds_all[:Q]
    P = postings_per_guide(hx, batch)
    med = float(np.median(P))
    heavy = kinds == 3
    out = {"guides": int(batch.shape[0]), "heavy_guides": int(heavy.sum()),
           "postings_median": med,
           "postings_p50_p99_max": [int(np.percentile(P, 50)), int(np.percentile(P, 99)),
                                    int(P.max())],
           "heavy_postings_p50_max_over_median": [round(float(np.median(P[heavy])) / med, 1),
                                                  round(float(P[heavy].max()) / med, 1)],
           "heavy_share_of_postings": round(float(P[heavy].sum()) / float(P.sum()), 3)}
    kw = dict(method=issl.METHOD_MIT_AND_CFD, threshold=0.0, form=issl.FORM_BATCHED,
              precision=issl.PREC_FP64)
    pool = np.concatenate([b[kinds_all[i * Q:(i + 1) * Q] != 3] for i, b in enumerate(batches) if i])
    Pc = np.cumsum(postings_per_guide(hx, pool))
    n_ctl = min(int(np.searchsorted(Pc, P.sum())) + 1, pool.shape[0])
    control = pool[:n_ctl]
    out.update({"control_guides": n_ctl, "control_postings": int(Pc[n_ctl - 1]),
                "mixed_postings": int(P.sum())})
    times, steps = {}, {}
    for name, g in (("mixed", batch), ("heavy_only", batch[heavy]), ("unique_only", batch[~heavy]),
                    ("control", control)):
        rot = Rotation(hiprt, [g])
        steps[name] = timed(hiprt, rot, ix, args.steps, 2, **kw)
        times[name], _ = kernel_us(hiprt, rot, ix, args.steps, **kw)
    out.update({f"{k}_kernel_us": round(v, 2) for k, v in times.items()})
    out.update({f"{k}_step_ms": round(v, 4) for k, v in steps.items()})
    out["mixed_over_control_step"] = round(steps["mixed"] / steps["control"], 3)
    # Hits (distinct sites scored) carry the stage-A/B work, which postings do not count.
    hits = {name: int(ix.score(g, **kw).site_count.sum()) for name, g in (("mixed", batch), ("control", control))}
    out.update({f"{k}_hits": v for k, v in hits.items()})
    out["mixed_over_parts"] = round(times["mixed"] / (times["heavy_only"] + times["unique_only"]), 3)
    out["mixed_over_control"] = round(times["mixed"] / times["control"], 3)
    log(f"tail: heavy guides at {out['heavy_postings_p50_max_over_median']}x the median postings, "
        f"{out['heavy_share_of_postings']:.0%} of the postings; mixed {times['mixed']:.1f} us, "
        f"heavy alone {times['heavy_only']:.1f} us, unique alone {times['unique_only']:.1f} us; "
        f"control ({n_ctl:,} tail-free guides, {out['control_postings']:,} postings) "
        f"{times['control']:.1f} us: mixed/control {out['mixed_over_control']} (steps "
        f"{steps['mixed']:.4f} / {steps['control']:.4f} ms = {out['mixed_over_control_step']}); "
        f"hits mixed {hits['mixed']:,}, control {hits['control']:,}")
    return out


def large_batch_leg(issl, hiprt, ix, hx, args, threads):
    """100,000 guides in one launch on the same index (30 % repeat
    neighbourhoods, 70 % unique sites), BATCHED FP64, tuned on a held-out
    batch of the same mix; PLAIN FP64 on the same launch beside it."""
    from cracklingplusplus_gpu_amd import workloads
    name = "genome100m_10x10_q100k"
    log(f"{name}: generating guides")
    sigs2, _, masks2, guides2, _ = workloads.make(name, threads=threads, guide_batches=2)
    if not (np.array_equal(sigs2, hx.sigs) and np.array_equal(masks2, hx.masks)):
        raise RuntimeError(f"{name} does not share the index of the headline workload")
    del sigs2
    Q2 = guides2.shape[0] // 2
    g_time, g_tune = guides2[:Q2], guides2[Q2:]
    tn = ix.autotune(g_tune, args.tune_cache, precision=issl.PREC_FP64)
    rot = Rotation(hiprt, [g_time])
    kw = dict(method=issl.METHOD_MIT_AND_CFD, threshold=0.0, form=issl.FORM_BATCHED,
              precision=issl.PREC_FP64)
    ms = timed(hiprt, rot, ix, args.large_steps, 1, **kw)
    k_us, _ = kernel_us(hiprt, rot, ix, args.large_steps, **kw)
    res_b = rot.results()[0]
    P = postings_per_guide(hx, g_time)
    S = hx.n_slices
    bytes_ref = 16 * int(P.sum()) + 16 * S * Q2
    _, res_used, _ = ix.residual()
    word = 4 if res_used else 8
    kwp = dict(kw, form=issl.FORM_PLAIN)
    ms_p = timed(hiprt, rot, ix, 2, 1, **kwp)
    res_p = rot.results()[0]
    out = {"workload": name, "guides_per_launch": Q2, "value": round(Q2 / (ms * 1e-3), 1),
           "unit": "guides/s", "ms_per_step": round(ms, 4), "precision": "fp64",
           "kernel": kernel_name(issl, tn), "kernel_us": round(k_us, 2),
           "site_words": "list-residual, 4 B" if res_used else "signature-inline, 8 B",
           "postings_per_launch": int(P.sum()),
           "postings_p50_p99_max": [int(np.percentile(P, 50)), int(np.percentile(P, 99)),
                                    int(P.max())],
           "roofline_achieved_GBps": round(bytes_ref / (k_us * 1e-6) / 1e9, 1),
           "roofline_frac": round(bytes_ref / (k_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
           "layout_GBps": round((word * int(P.sum()) + 16 * S * Q2) / (k_us * 1e-6) / 1e9, 1),
           "tuning": {"chunk": tn.chunk, "depth": tn.depth, "grid_per_cu": tn.grid_per_cu,
                      "heavy_kpost": tn.heavy_kpost},
           "plain": {"ms_per_step": round(ms_p, 4), "value": round(Q2 / (ms_p * 1e-3), 1)},
           "fp64_vs_plain_max_rel": max(max_rel(res_b["mit"], res_p["mit"]),
                                        max_rel(res_b["cfd"], res_p["cfd"])),
           "counts_identical": bool((res_b["occ_count"] == res_p["occ_count"]).all())}
    log(f"{name}: {ms:.3f} ms/step, {out['value']:,.0f} guides/s, kernel {k_us:.1f} us, "
        f"plain {ms_p:.3f} ms")
    return out


def serving_c4_leg(issl, hx1, args, threads):
    """The 100 M- and 300 M-site indexes resident together in one server:
    Poisson arrivals of log-uniform 1..50,000-guide requests against a random
    index, and 1,000-guide requests at a rate at which arriving requests queue
    behind busy slots and are merged into shared launches."""
    from cracklingplusplus_gpu_amd import workloads
    srv = issl.Server(max_batch=65536, hbm_reserve=8 << 30)
    ids, pools = [srv.add_index(hx1)], []
    g1 = workloads.make("genome100m_10x10", threads=threads, guide_batches=4)[3]
    pools.append(g1)
    log(f"serving: generating genome300m_10x10 with {threads} threads")
    sigs, occ, masks, g3, _ = workloads.make("genome300m_10x10", threads=threads)
    log("serving: building the 300 M-site slice lists")
    hx3 = issl.build_lists(sigs, occ, masks, threads=threads)
    del sigs, occ
    log("serving: admitting the 300 M-site index")
    ids.append(srv.add_index(hx3))
    pools.append(g3)
    del hx3
    st = srv.stats()
    log(f"serving: resident {st.resident_bytes / 1e9:.1f} GB, free {st.hbm_free / 1e9:.1f} GB")
    rng = np.random.default_rng(0x5E4E)
    hi = np.log(50000.0)

    def log_uniform(r):
        return max(1, int(np.exp(r.uniform(0.0, hi))))

    serve_poisson(issl, srv, ids, pools, 50.0, 10, log_uniform, rng)   # warm both indexes
    legs = []
    for rate in args.c4_rates:
        n_req = max(20, int(rate * args.serve_seconds))
        res = serve_poisson(issl, srv, ids, pools, rate, n_req, log_uniform, rng)
        res["request_guides"] = "log-uniform 1..50000"
        log(f"serving C4 {rate} req/s: p50 {res['p50_ms']} ms p99 {res['p99_ms']} ms, "
            f"{res['launches']} launches for {res['requests']} requests")
        legs.append(res)
    for rate in args.merge_rates:
        res = serve_poisson(issl, srv, ids, pools, rate, args.merge_requests, lambda r: 1000, rng)
        res["request_guides"] = 1000
        log(f"serving C4 1,000-guide requests at {rate} req/s (achieved {res['achieved_rps']}): "
            f"p50 {res['p50_ms']} ms p99 {res['p99_ms']} ms, {res['launches']} launches for "
            f"{res['requests']} requests (mean batch {res['mean_batch_guides']})")
        legs.append(res)
    srv.close()
    return {"workload": "genome100m_10x10 + genome300m_10x10 resident in one server",
            "resident_GB": round(st.resident_bytes / 1e9, 2), "form": "batched",
            "precision": "fp32c", "max_batch": 65536, "slots": 3, "legs": legs}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="genome100m_10x10")
    ap.add_argument("--batches", type=int, default=12,
                    help="disjoint guide batches the timed steps rotate over (12 C1 batches "
                         "stream about 590 MB of list-residual words, 1.2 GB of "
                         "signature-inline words: twice the 256 MB Infinity Cache or more)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the timed steps and their kernel timing (for profiling)")
    ap.add_argument("--form", choices=["batched", "plain", "ordered"], default="batched",
                    help="form of the timed steps with --headline-only (profiling)")
    ap.add_argument("--precision", choices=["fp64", "fp32c"], default="fp64",
                    help="precision of the timed steps with --headline-only (profiling)")
    ap.add_argument("--threshold", type=float, default=0.0,
                    help="score threshold of --headline-only runs (75 arms the early exit)")
    ap.add_argument("--plain-steps", type=int, default=6)
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
                    help="replay the step's kernels from a captured hipGraph (one stream)")
    ap.add_argument("--streams", type=int, default=0, choices=[0, 1, 2, 3],
                    help="streams the rotated steps go round (0: the count, 1-3, that is "
                         "fastest on the held-out batch)")
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
    ap.add_argument("--large-steps", type=int, default=5,
                    help="steps of the 100,000-guide leg (0: skip it)")
    ap.add_argument("--c4", type=int, default=1,
                    help="1: serving leg with the 100 M- and 300 M-site indexes resident")
    ap.add_argument("--c4-rates", type=float, nargs="+", default=[50, 500])
    ap.add_argument("--merge-rates", type=float, nargs="+", default=[20000])
    ap.add_argument("--merge-requests", type=int, default=4000)
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
    cpu = cpu_info()
    threads = cpu["threads"]
    log(f"host: {cpu}")
    t_setup = time.time()
    R = max(1, args.batches)
    log(f"generating workload {args.workload}: {R} timed guide batches and one held out, "
        f"{threads} threads")
    sigs, occ, masks, guides_all, kinds = workloads.make(args.workload, threads=threads,
                                                         guide_batches=R + 1)
    Q = guides_all.shape[0] // (R + 1)
    batches = [guides_all[b * Q:(b + 1) * Q] for b in range(R)]
    g_tune = guides_all[R * Q:]
    log(f"sites {sigs.shape[0]:,}, {R} x {Q:,} guides; building slice lists")
    t_b = time.perf_counter()
    hx = issl.build_lists(sigs, occ, masks, threads=threads)
    host_build_s = time.perf_counter() - t_b
    log("uploading index")
    ix = issl.Index.from_arrays(hx)
    info = ix.info()
    log(f"index resident: {info.device_bytes / 1e9:.2f} GB of HBM")
    S = hx.n_slices
    P = [postings_per_guide(hx, b) for b in batches]
    p_step = [int(p.sum()) for p in P]
    P0 = P[0]
    log(f"postings per step {np.mean(p_step):,.0f} (batch 0: mean {P0.mean():.0f}, "
        f"p99 {np.percentile(P0, 99):.0f}, max {P0.max():,})")

    log("autotuning the batched kernel (FP64) on the held-out batch")
    os.makedirs(os.path.dirname(args.tune_cache), exist_ok=True)
    tn = ix.autotune(g_tune, args.tune_cache, precision=issl.PREC_FP64)
    log(f"tuning: chunk {tn.chunk} depth {tn.depth} grid/CU {tn.grid_per_cu} "
        f"stream depth {tn.direct_depth} heavy {tn.heavy_kpost} "
        f"({tn.us_per_guide * 1e3:.2f} ns/guide)")
    res_built, res_used, _ = ix.residual()
    word = 4 if res_used else 8
    foot = footprint_bytes(hx, batches, word)
    log(f"site words streamed by the {R} batches together: {foot / 1e6:.0f} MB "
        f"(Infinity Cache {INFINITY_CACHE_BYTES >> 20} MB)")
    rot = Rotation(hiprt, batches)
    setup_s = time.time() - t_setup

    kw = dict(method=issl.METHOD_MIT_AND_CFD, threshold=0.0, form=issl.FORM_BATCHED,
              precision=issl.PREC_FP64)
    if args.headline_only:
        kw.update(form={"plain": issl.FORM_PLAIN, "ordered": issl.FORM_ORDERED,
                        "batched": issl.FORM_BATCHED}[args.form],
                  precision=issl.PREC_FP32C if args.precision == "fp32c" else issl.PREC_FP64,
                  threshold=args.threshold)
    elif (args.form, args.precision, args.threshold) != ("batched", "fp64", 0.0):
        raise SystemExit("--form / --precision / --threshold apply to --headline-only runs only")
    # With --graph each batch's launch sequence is captured in a hipGraph by its
    # first warm-up call and replayed afterwards (slower than eager launches for
    # this three-kernel sequence, so off by default).
    kw_step = dict(kw, flags=issl.FLAG_GRAPH if args.graph else 0)

    # ---- batches in flight: consecutive steps go round 2 or 3 streams bound
    # to the index (issl_index_bind_stream), each with its own scoring
    # workspace, so that one step's planner, finalizer and last waves overlap
    # the next steps' tile kernels; launches on bound streams run the tile
    # kernel at ovl_grid workgroups per CU so that several fit at once.  One
    # stream is the index's own, at the tuned grid.  The choice is made on
    # the held-out batch unless --streams gives the count (then ovl_grid 3).
    bound = [hiprt.Stream() for _ in range(2)]
    for s_ in bound:
        ix.bind_stream(s_)
    if args.streams or args.graph:
        ns_ = 1 if args.graph else args.streams
        streams, ovl, sweep = ([None], bound[:2], [None] + bound)[ns_ - 1], 0, None
    else:
        streams, ovl, sweep = pick_streams(hiprt, ix, g_tune, bound, **kw_step)
    n_streams = len(streams)
    log(f"streams: {n_streams}, ovl_grid {ovl or 'default'}" +
        (f" (held-out batch, us per step: {sweep})" if sweep else ""))

    # ---- timed region: BATCHED FP64, inputs resident in HBM, rotating batches
    warm = max(args.warmup, R if args.graph else 0)
    ms_per_step = timed(hiprt, rot, ix, args.steps, warm, streams, **kw_step)
    value = Q / (ms_per_step * 1e-3)
    log(f"batched fp64: {ms_per_step:.4f} ms/step, {value:,.0f} guides/s ({n_streams} streams)")
    # the same steps one at a time on the index's stream
    ms_serial = (timed(hiprt, rot, ix, args.steps, warm, None, **kw_step) if n_streams > 1
                 else ms_per_step)
    log(f"batched fp64, one stream: {ms_serial:.4f} ms/step, {Q / (ms_serial * 1e-3):,.0f} guides/s")

    # ---- dominant kernel time, HIP events on each of K rotated steps' dispatches
    k_us, n_disp = kernel_us(hiprt, rot, ix, args.steps, **kw)
    kernel = ("k_plain" if kw["form"] == issl.FORM_PLAIN else
              "k_tiles (in order)" if kw["form"] == issl.FORM_ORDERED else kernel_name(issl, tn))
    steps_bytes = [16 * p_step[k % R] + 16 * S * Q for k in range(args.steps)]
    bytes_ref = int(np.mean(steps_bytes))
    bytes_layout = int(np.mean([word * p_step[k % R] + 16 * S * Q for k in range(args.steps)]))
    achieved = bytes_ref / (k_us * 1e-6) / 1e9
    traffic = (measured_traffic("k_tiles_fp64", args.workload, "residual" if res_used else "inline")
               if kernel == "k_tiles" else None)
    log(f"{kernel} avg {k_us:.2f} us over {n_disp} dispatches -> {achieved:,.0f} GB/s "
        f"(frac {achieved / HBM_PEAK_GBS:.3f})")
    res_fp64 = rot.results()

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
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "n_sites": int(hx.n_sites), "n_slices": S,
                   "slice_width": 10 if args.workload.startswith("genome") else None,
                   "guides_per_step": Q, "rotation_batches": R,
                   "rotation_footprint_bytes": foot,
                   "tuning_sample": "a further batch of the same mix, not timed",
                   "method": "mitAndCfd", "threshold": 0, "max_dist": 4, "form": "batched",
                   "precision": "fp64 (FP64 per-site terms and sums; integer signature work)",
                   "launch": "hipGraph replay" if args.graph else "eager",
                   "postings_per_step": int(np.mean(p_step)),
                   "site_words": "list-residual, 4 B" if res_used else "signature-inline, 8 B",
                   "streams": n_streams,
                   "ovl_grid": ovl,
                   "streams_note": "consecutive steps round streams bound to the index "
                                   "(issl_index_bind_stream), each with its own workspace, tile "
                                   "kernels at ovl_grid workgroups per CU; count and ovl_grid "
                                   "picked on the held-out batch",
                   "stream_pick_us_per_step": sweep,
                   "one_stream": {"value": round(Q / (ms_serial * 1e-3), 1),
                                  "ms_per_step": round(ms_serial, 4)},
                   "tuning": {"chunk": tn.chunk, "depth": tn.depth,
                              "grid_per_cu": tn.grid_per_cu,
                              "stream_depth": tn.direct_depth,
                              "heavy_kpost": tn.heavy_kpost,
                              "ord_depth": tn.ord_depth, "ord_grid": tn.ord_grid,
                              "ord_split": tn.ord_split}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": kernel, "kernel_us": round(k_us, 2),
                     "kernel_dispatches_timed": n_disp,
                     "bytes_per_launch": bytes_ref,
                     "bytes_basis": "16 B per posting (8 B entry + 8 B signature gather) "
                                    "+ 16 B list bounds per (guide, slice)",
                     "layout_bytes_per_launch": bytes_layout,
                     "layout_GBps": round(bytes_layout / (k_us * 1e-6) / 1e9, 1),
                     "moved_GBps": (round(traffic / (k_us * 1e-6) / 1e9, 1) if traffic else None),
                     "moved_frac": (round(traffic / (k_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
                                    if traffic else None)},
    }
    if args.headline_only:
        line["config"].update(form=args.form, precision=args.precision, threshold=args.threshold)
        print(json.dumps(line), flush=True)
        ix.close()
        return 0

    # ---- compensated FP32 (the north star's precision form), same steps and shape
    kwc = dict(kw, precision=issl.PREC_FP32C)
    ms_c = timed(hiprt, rot, ix, args.steps, args.warmup, streams, **kwc)
    kc_us, _ = kernel_us(hiprt, rot, ix, args.steps, **kwc)
    res_fp32c = rot.results()
    log(f"batched fp32c: {ms_c:.4f} ms/step, kernel {kc_us:.2f} us")

    # ---- the other site-word layout at the same shape (FP64); same bits expected
    other_leg = None
    if res_built:
        ix.set_residual(not res_used)
        ms_o = timed(hiprt, rot, ix, args.steps, args.warmup, streams, **kw)
        ko_us, _ = kernel_us(hiprt, rot, ix, args.steps, **kw)
        same_o = all(np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8))
                     for a, b in zip(rot.results(), res_fp64) for k in a)
        word_o = 8 if res_used else 4
        lb_o = np.mean([word_o * p + 16 * S * Q for p in p_step])
        other_leg = {"site_words": "signature-inline, 8 B" if res_used else "list-residual, 4 B",
                     "value": round(Q / (ms_o * 1e-3), 1), "ms_per_step": round(ms_o, 4),
                     "kernel_us": round(ko_us, 2),
                     "layout_GBps": round(lb_o / (ko_us * 1e-6) / 1e9, 1),
                     "same_bits": bool(same_o)}
        log(f"other layout: {ms_o:.4f} ms/step, kernel {ko_us:.2f} us, same bits {same_o}")
        ix.set_residual(res_used)

    # ---- PLAIN form (bit-exact FP64), same rotation
    kwp = dict(kw, form=issl.FORM_PLAIN)
    plain_ms = timed(hiprt, rot, ix, args.plain_steps, 1, **kwp)
    plain_k_us, _ = kernel_us(hiprt, rot, ix, args.plain_steps, **kwp)
    # PLAIN scores batches 0 .. plain_steps-1 only (its warm-up step is batch 0):
    # every comparison with it covers those batches.
    covered = min(R, max(1, args.plain_steps))
    res_plain = rot.results()[:covered]
    log(f"plain: {plain_ms:.4f} ms/step, kernel {plain_k_us:.1f} us")

    # ---- ORDERED form (bit-exact FP64 in the tile kernel, one wave per guide)
    kwo = dict(kw, form=issl.FORM_ORDERED)
    if args.streams:
        ord_streams, ord_ovl, ord_sweep = streams, ovl, None
    else:
        ord_streams, ord_ovl, ord_sweep = pick_streams(hiprt, ix, g_tune, bound, **kwo)
    ord_ns = len(ord_streams)
    ord_ms = timed(hiprt, rot, ix, args.steps, args.warmup, ord_streams, **kwo)
    ord_ms1 = (timed(hiprt, rot, ix, args.steps, args.warmup, None, **kwo) if ord_ns > 1
               else ord_ms)
    ord_k_us, _ = kernel_us(hiprt, rot, ix, args.steps, **kwo)
    ord_same = all(np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8))
                   for a, b in zip(rot.results()[:covered], res_plain) for k in a)
    log(f"ordered: {ord_ms:.4f} ms/step ({ord_ns} streams; one stream {ord_ms1:.4f}), "
        f"kernel {ord_k_us:.1f} us, bits of plain {ord_same}")

    # ---- threshold 75 (the module's usual setting): ORDERED and PLAIN take the
    # reference's early exit, hit by hit
    kw75 = dict(kw, threshold=75.0)
    ord75_ms = timed(hiprt, rot, ix, args.steps, args.warmup, **dict(kw75, form=issl.FORM_ORDERED))
    ord75 = rot.results()[:covered]
    plain75_ms = timed(hiprt, rot, ix, args.plain_steps, 1, **dict(kw75, form=issl.FORM_PLAIN))
    plain75 = rot.results()[:covered]
    ord75_same = all(np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8))
                     for a, b in zip(ord75, plain75) for k in a)
    bat75_ms = timed(hiprt, rot, ix, args.steps, args.warmup, **kw75)
    bat75 = rot.results()[:covered]
    # threshold decisions (pass/fail calls) of the BATCHED forms against the
    # reference's (PLAIN, bit-exact), on the batches PLAIN scored
    rot.run(ix, covered, 0, **dict(kw75, precision=issl.PREC_FP32C))
    hiprt.synchronize()
    f75 = rot.results()[:covered]
    calls = np.concatenate([r["pass"] for r in plain75])
    agree75 = {name: float(np.mean(np.concatenate([r["pass"] for r in rs]) == calls))
               for name, rs in (("batched_fp64", bat75), ("batched_fp32c", f75))}
    agree75["guides"] = int(calls.shape[0])
    log(f"threshold 75: ordered {ord75_ms:.4f} ms/step (bits of plain {ord75_same}), "
        f"plain {plain75_ms:.4f}, batched {bat75_ms:.4f}; calls agreeing with plain {agree75}")

    # ---- precision against the in-order FP64 path (PLAIN), every batch
    cat = lambda rs, f: np.concatenate([r[f] for r in rs])  # noqa: E731
    prec = {}
    for name, rs in (("fp64", res_fp64[:covered]), ("fp32c", res_fp32c[:covered])):
        rel = {f: np.abs(cat(rs, f) - cat(res_plain, f)) / np.abs(cat(res_plain, f))
               for f in ("mit", "cfd")}
        prec[name] = {"max_rel": {f: float(v.max()) for f, v in rel.items()},
                      "mean_rel": {f: float(v.mean()) for f, v in rel.items()},
                      "counts_identical": bool(
                          (cat(rs, "occ_count") == cat(res_plain, "occ_count")).all() and
                          (cat(rs, "site_count") == cat(res_plain, "site_count")).all())}

    # ---- heavy-tail workloads: the mixed batch against its parts
    tail = (tail_leg(issl, hiprt, ix, hx, batches, kinds, args)
            if args.workload.endswith("_heavy") else None)

    # ---- request latency: 1,000-guide requests, host buffers, one at a time (FP64)
    g0 = batches[0]
    lat = []
    rng = np.random.default_rng(7)
    for i in range(args.latency_requests):
        lo = int(rng.integers(0, max(1, Q - 1000)))
        t0 = time.perf_counter()
        ix.score(g0[lo:lo + 1000], **kw)
        lat.append((time.perf_counter() - t0) * 1e3)
    lat = np.array(lat)

    # ---- the step through host buffers (issl_score): PCIe both ways, never `value`
    hres = ix.results(Q)
    for _ in range(max(1, args.warmup)):
        ix.score(g0, out=hres, **kw)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ix.score(g0, out=hres, **kw)
    host_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    host_same = bool(np.array_equal(hres.cfd.view(np.uint64), res_fp64[0]["cfd"].view(np.uint64)))
    log(f"host buffers: {host_ms:.3f} ms/call, {Q / host_ms * 1e3:,.0f} guides/s")

    # ---- index construction on the device against the host builder above
    t_b = time.perf_counter()
    with issl.Index.build(sigs, occ, masks) as dix:
        device_build_s = time.perf_counter() - t_b
        dsizes, _ = dix.export(entries=False)
    del occ, sigs
    build = {"host_lists_s": round(host_build_s, 3), "host_threads": threads,
             "device_build_s": round(device_build_s, 3),
             "list_sizes_identical": bool(np.array_equal(dsizes, hx.list_sizes))}
    log(f"index build: host lists {host_build_s:.2f} s, device {device_build_s:.2f} s, "
        f"sizes identical {build['list_sizes_identical']}")

    streamed = stream_leg(issl, hiprt, hx, g0, ix, args) if args.stream_slots > 0 else None
    extraction = extract_leg(issl, args, masks) if args.extract_bases > 0 else None

    # ---- serving: Poisson arrivals of 1,000-guide requests through the request loop
    log(f"serving: {args.serve_requests} requests of 1,000 guides at {args.serve_rate:g} req/s")
    srv = issl.Server(max_batch=65536)
    sid = srv.add_index(hx)
    srv_rng = np.random.default_rng(0x5E4E)
    serve_poisson(issl, srv, [sid], [g0], args.serve_rate, 20, lambda r: 1000, srv_rng)
    serving = serve_poisson(issl, srv, [sid], [g0], args.serve_rate, args.serve_requests,
                            lambda r: 1000, srv_rng)
    srv.close()
    log(f"serving: p50 {serving['p50_ms']} ms, p99 {serving['p99_ms']} ms, "
        f"mean batch {serving['mean_batch_guides']:.0f} guides")

    # ---- 100,000 guides in one launch (FP64) on the same index
    # (both legs below use the default workload's index)
    c1_index = args.workload == "genome100m_10x10"
    large = (large_batch_leg(issl, hiprt, ix, hx, args, threads)
             if args.large_steps > 0 and c1_index else None)

    # ---- CPU baseline: the oracle's OpenMP scorer on the host cores
    cpu_base = None
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        n_cpu = min(args.cpu_sample, Q)
        sample = g0[:n_cpu]
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
                          res_plain[0]["cfd"][:n_cpu].view(np.uint64)).all() and
                         (o["mit"].view(np.uint64) ==
                          res_plain[0]["mit"][:n_cpu].view(np.uint64)).all())
        cpu_base = {"value": round(cpu_rate, 1), "unit": "guides/s", "cores": threads,
                    "kind": "port",
                    "sample": f"batch 0's first {n_cpu} guides scored {passes} times, index "
                              f"preloaded, scoring loop only ({t1 - t0:.1f} s); {threads} OpenMP "
                              f"threads (affinity {cpu['affinity_cpus']} CPUs, cgroup cpu.max "
                              f"{cpu['cgroup_cpu_max']}, OMP_PROC_BIND={cpu['omp_proc_bind']}), "
                              f"{cpu['model']}",
                    "host": cpu, "plain_gpu_bit_exact": cpu_match}
        log(f"CPU: {cpu_rate:,.0f} guides/s (GPU plain bit-exact: {cpu_match})")

    c4 = serving_c4_leg(issl, hx, args, threads) if args.c4 and c1_index else None

    line.update({
        "cpu_baseline": cpu_base,
        "fp32c": {"value": round(Q / (ms_c * 1e-3), 1), "unit": "guides/s",
                  "ms_per_step": round(ms_c, 4), "kernel_us": round(kc_us, 2),
                  "mode": "BATCHED, FP32 per-site products, compensated FP32 lane sums"},
        "other_layout": other_leg,
        "ordered": {"value": round(Q / (ord_ms * 1e-3), 1), "unit": "guides/s",
                    "ms_per_step": round(ord_ms, 4), "kernel": "k_tiles (in order)",
                    "kernel_us": round(ord_k_us, 2),
                    "achieved_GBps": round(bytes_ref / (ord_k_us * 1e-6) / 1e9, 1),
                    "same_bits_as_plain": bool(ord_same), "batches_compared": covered,
                    "plain_speedup": round(plain_ms / ord_ms, 3),
                    "streams": ord_ns, "ovl_grid": ord_ovl,
                    "stream_pick_us_per_step": ord_sweep,
                    "one_stream": {"value": round(Q / (ord_ms1 * 1e-3), 1),
                                   "ms_per_step": round(ord_ms1, 4)},
                    "mode": "ORDERED: FP64 hits added in the reference's order, one wave "
                            "per guide in the persistent tile kernel"},
        "threshold_75": {"ordered": {"value": round(Q / (ord75_ms * 1e-3), 1),
                                     "ms_per_step": round(ord75_ms, 4),
                                     "same_bits_as_plain": bool(ord75_same)},
                         "plain": {"value": round(Q / (plain75_ms * 1e-3), 1),
                                   "ms_per_step": round(plain75_ms, 4)},
                         "batched_fp64": {"value": round(Q / (bat75_ms * 1e-3), 1),
                                          "ms_per_step": round(bat75_ms, 4)},
                         "call_agreement_vs_plain": agree75,
                         "mode": "MIT+CFD at threshold 75: the exit is armed (cut 33.3)"},
        "plain": {"value": round(Q / (plain_ms * 1e-3), 1), "unit": "guides/s",
                  "ms_per_step": round(plain_ms, 4), "kernel_us": round(plain_k_us, 2),
                  "achieved_GBps": round(bytes_ref / (plain_k_us * 1e-6) / 1e9, 1),
                  "batched_fp64_speedup": round(plain_ms / ms_per_step, 3),
                  "batched_fp32c_speedup": round(plain_ms / ms_c, 3)},
        "precision": prec,
        "host_buffers": {"value": round(Q / host_ms * 1e3, 1), "unit": "guides/s",
                         "ms_per_call": round(host_ms, 4), "calls": args.steps,
                         "same_as_device_buffers": host_same,
                         "mode": "issl_score: guides from host memory, the seven result "
                                 "columns to host memory (PCIe both ways)"},
        "latency_ms": {"request_guides": 1000, "requests": int(lat.size),
                       "p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "mode": "one request at a time, host buffers in/out, batched fp64"},
        "serving": dict(serving, request_guides=1000, precision="fp32c",
                        mode="request loop, Poisson arrivals, host buffers in/out"),
        "serving_c4": c4,
        "tail": tail,
        "large_batch": large,
        "index_build": build,
        "streamed": streamed,
        "extraction": extraction,
        "setup_s": round(setup_s, 1),
    })
    if cpu_base:
        line["gpu_vs_cpu"] = round(value / cpu_base["value"], 1)
    print(json.dumps(line), flush=True)
    ix.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
