"""Benchmark of the vector/segment primitive hot path on one MI355X.

Headline workload (the HBM-scale vector-primitive configuration):
inclusive_scan (add, wrapping) over 2^28 u32 uniform in [0, 1024) from seed 2.
One step is one scan over the device-resident input; `value` is its
throughput in GB/s of algorithmic traffic (8 bytes per element: read 4,
write 4), and `roofline` relates it to the 8 TB/s HBM3E peak (HIP events on
the executor's stream around each launch). The same JSON line carries, under
`extra`, the other primitives of the configurations measured the same way:

* sort_by_key: 2^27 u64 keys (uniform seed 6, Zipf s=1.1 over 2^20 values,
  top 8 bits constant) + u32 payload; 2^28 u32 keys + u32 payload; the plain
  2^20 u32 keys + iota values case (Gkeys/s);
* reduce f32 2^28, copy_where 2^28 u32 at 10% / 50% (stencil form and fused
  predicate form), gather / scatter through a seeded random permutation of
  2^28, merge of 2 x 2^27 sorted u32, set ops at 25% overlap (2^27 each);
* seg inclusive_scan (u32) / reduce (f32) / sort (u32 key, f32 value) over
  2^26 elements in 2^20 log-normal segments.

cpu_baseline: the C oracle (a restatement of the reference's sequential CPU
path) timed on this host, one thread, on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--only a,b] [--no-cpu]

`--only copy` adds the calibration copy (torch's copy kernel over 2^28 u32),
which the default run leaves out.
"""
from __future__ import annotations

import argparse
import json
import math
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "radix sort Gkeys/s + scan GB/s (frac of HBM peak) @2^28; served p50/p99 latency"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def timed(ex, fn, steps: int, warmup: int):
    """Runs fn() warmup+steps times; per-step HIP-event milliseconds on ex's stream."""
    from massively_amd.hiptime import Event

    for _ in range(warmup):
        fn()
    ex.sync()
    evs = [(Event(), Event()) for _ in range(steps)]
    for a, b in evs:
        a.record(ex.stream)
        fn()
        b.record(ex.stream)
    ex.sync()
    return [a.elapsed_ms(b) for a, b in evs]


def summarize(name, ms, bytes_=None, items=None, **kw):
    avg = statistics.fmean(ms)
    d = dict(op=name, ms=round(avg, 5), ms_median=round(statistics.median(ms), 5), **kw)
    if bytes_ is not None:
        gbs = bytes_ / (avg * 1e-3) / 1e9
        d.update(gbs=round(gbs, 1), frac_hbm=round(gbs / HBM_PEAK_GBS, 4), bytes=int(bytes_))
    if items is not None:
        d["gitems_per_s"] = round(items / (avg * 1e-3) / 1e9, 3)
    log(f"[bench] {name}: {avg:.4f} ms" + (f"  {d['gbs']:.1f} GB/s ({d['frac_hbm']:.3f})" if bytes_ else "")
        + (f"  {d['gitems_per_s']:.3f} G/s" if items else ""))
    return d


class Ctx:
    def __init__(self, ex, m, steps, warmup):
        import numpy as np

        self.ex, self.m, self.np = ex, m, np
        self.steps, self.warmup = steps, warmup
        from massively_amd import random as mr

        self.mr = mr


def w_scan(c, log2n=28):
    n = 1 << log2n
    x = c.ex.alloc(c.np.uint32, n)
    c.mr.uniform_u32(c.ex, 0, 1023, 2, x)
    ms = timed(c.ex, lambda: c.m.vector.inclusive_scan(c.ex, x, c.m.op.Add), c.steps, c.warmup)
    return summarize("inclusive_scan_u32_add_2^%d" % log2n, ms, 8.0 * n, n, n=n)


def w_reduce(c):
    """f32 sum of 2^28 N(0,1) values (Box-Muller in f64 over the seed-3 u64
    stream, rounded to f32), checked against the FP64 sum and its bound
    (ceil(log2 n) + 2) * 2^-24 * sum|x| outside the timed region."""
    np = c.np
    n = 1 << 28
    hx = c.mr.normal_f32_host(c.ex, 3, n)
    xf = c.ex.to_device(hx)
    res = []
    ms = timed(c.ex, lambda: res.append(c.m.vector.reduce(c.ex, xf, 0.0, c.m.op.Add)), c.steps, c.warmup)
    s64 = a64 = 0.0
    for lo in range(0, n, 1 << 24):
        part = hx[lo:lo + (1 << 24)].astype(np.float64)
        s64 += float(part.sum())
        a64 += float(np.abs(part).sum())
    bound = (math.ceil(math.log2(n)) + 2) * 2.0**-24 * a64
    repeatable = len({np.float32(r).tobytes() for r in res}) == 1
    return summarize("reduce_f32_add_normal_2^28", ms, 4.0 * n, n,
                     note="includes the scalar device-to-host copy",
                     abs_err_vs_fp64=abs(float(res[-1]) - s64), fp64_bound=bound,
                     within_bound=abs(float(res[-1]) - s64) <= bound, bitwise_repeatable=repeatable)


def w_copy_where(c):
    n = 1 << 28
    np, m = c.np, c.m
    x = c.ex.alloc(np.uint32, n)
    c.mr.uniform_u32(c.ex, 0, 2**32 - 1, 4, x)
    u = c.ex.alloc(np.uint32, n)
    c.mr.uniform_u32(c.ex, 0, 99, 5, u)
    out = {}
    for p in (10, 50):
        # Stencil flags = (u < p), built on the device by a table lookup.
        table = c.ex.to_device((np.arange(100) < p).astype(np.uint32))
        st = m.vector.gather(c.ex, table, u)
        ms = timed(c.ex, lambda: m.vector.copy_where(c.ex, x, st), c.steps, c.warmup)
        out[f"stencil_{p}pct"] = summarize(f"copy_where_stencil_{p}pct_2^28", ms, (8.0 + 4.0 * p / 100) * n, n)
        thr = int(p / 100 * 2**32)
        ms = timed(c.ex, lambda: m.vector.copy_where_if(c.ex, x, x, m.op.Lt(thr)), c.steps, c.warmup)
        out[f"predicate_{p}pct"] = summarize(f"copy_where_pred_{p}pct_2^28", ms, (4.0 + 4.0 * p / 100) * n, n)
    return out


def random_permutation(c, n, seed):
    keys = c.ex.alloc(c.np.uint32, n)
    c.mr.uniform_u32(c.ex, 0, 2**32 - 1, seed, keys)
    iota = c.ex.alloc(c.np.uint32, n)
    c.mr.iota_u32(c.ex, iota)
    return c.m.vector.radix_sort_by_key(c.ex, keys, iota)


def w_gather_scatter(c):
    n = 1 << 28
    np, m = c.np, c.m
    idx = random_permutation(c, n, 9)
    v = c.ex.alloc(np.uint32, n)
    c.mr.iota_u32(c.ex, v)
    ms = timed(c.ex, lambda: m.vector.gather(c.ex, v, idx), c.steps, c.warmup)
    g = summarize("gather_u32_perm_2^28", ms, 12.0 * n, n)
    out = c.ex.alloc(np.uint32, n)
    ms = timed(c.ex, lambda: m.vector.scatter(c.ex, v, idx, out), c.steps, c.warmup)
    s = summarize("scatter_u32_perm_2^28", ms, 12.0 * n, n)
    return {"gather": g, "scatter": s}


def _sort_case(c, name, keys, vals, n, key_bytes, val_bytes, passes):
    m = c.m
    ms = timed(c.ex, lambda: m.vector.radix_sort_by_key(c.ex, keys, vals), c.steps, c.warmup)
    model = n * (key_bytes + passes * 2 * (key_bytes + val_bytes))
    return summarize(name, ms, model, n, model_bytes_per_key=model / n)


def w_sort(c):
    np, m = c.np, c.m
    out = {}
    n = 1 << 27
    keys = c.ex.alloc(np.uint64, n)
    c.mr.uniform_u64(c.ex, 0, 2**64 - 1, 6, keys)
    vals = c.ex.alloc(np.uint32, n)
    c.mr.iota_u32(c.ex, vals)
    out["u64_uniform_2^27"] = _sort_case(c, "sort_by_key_u64_uniform_2^27", keys, vals, n, 8, 4, 8)
    # The reduce-then-scan schedule on the same input (the default is
    # onesweep), for comparison.
    c.ex.set_option(m.Executor.OPT_RADIX_ONESWEEP, 0)
    out["u64_uniform_2^27_reduce_then_scan"] = _sort_case(
        c, "sort_by_key_u64_uniform_2^27_reduce_then_scan", keys, vals, n, 8, 4, 8)
    c.ex.set_option(m.Executor.OPT_RADIX_ONESWEEP, -1)
    ref = c.ex.to_host(m.vector.radix_sort_by_key(c.ex, keys, vals))
    out["rocprim_u64_uniform_2^27"] = rocprim_sort_case(c, "sort_pairs_u64_u32_2^27", keys, vals, n, 8,
                                                        check=ref)
    del ref
    hk = c.ex.to_host(keys)
    rng = np.random.default_rng(6)
    # Zipf s=1.1 over 2^20 distinct values: inverse-CDF sampling of ranks.
    ranks = np.arange(1, (1 << 20) + 1, dtype=np.float64)
    cdf = np.cumsum(ranks ** -1.1)
    cdf /= cdf[-1]
    r = np.searchsorted(cdf, rng.random(n)).astype(np.uint64)
    zipf = (r * np.uint64(0x9E3779B97F4A7C15)) ^ np.uint64(0xD1B54A32D192ED03)
    zk = c.ex.to_device(zipf)
    out["u64_zipf_2^27"] = _sort_case(c, "sort_by_key_u64_zipf1.1_2^27", zk, vals, n, 8, 4, 8)
    c.ex.set_option(m.Executor.OPT_RADIX_ONESWEEP, 0)
    out["u64_zipf_2^27_reduce_then_scan"] = _sort_case(
        c, "sort_by_key_u64_zipf1.1_2^27_reduce_then_scan", zk, vals, n, 8, 4, 8)
    c.ex.set_option(m.Executor.OPT_RADIX_ONESWEEP, -1)
    del zk, zipf, r
    top = (np.uint64(0xA5) << np.uint64(56)) | (hk & np.uint64((1 << 56) - 1))
    tk = c.ex.to_device(top)
    out["u64_top8const_2^27"] = _sort_case(c, "sort_by_key_u64_top8const_2^27", tk, vals, n, 8, 4, 7)
    del tk, top, hk
    n28 = 1 << 28
    k28 = c.ex.alloc(np.uint32, n28)
    c.mr.uniform_u32(c.ex, 0, 2**32 - 1, 1, k28)
    v28 = c.ex.alloc(np.uint32, n28)
    c.mr.iota_u32(c.ex, v28)
    out["u32_2^28"] = _sort_case(c, "sort_by_key_u32_2^28", k28, v28, n28, 4, 4, 4)
    # The reduce-then-scan schedule (upsweep + count scan + downsweep per
    # digit), for comparison with the default onesweep schedule.
    c.ex.set_option(m.Executor.OPT_RADIX_ONESWEEP, 0)
    out["u32_2^28_reduce_then_scan"] = _sort_case(c, "sort_by_key_u32_2^28_reduce_then_scan", k28, v28,
                                                  n28, 4, 4, 4)
    c.ex.set_option(m.Executor.OPT_RADIX_ONESWEEP, -1)
    out["rocprim_u32_2^28"] = rocprim_sort_case(c, "sort_pairs_u32_u32_2^28", k28, v28, n28, 4)
    out["reference_schedule_proxy_u32_2^28"] = refproxy_case(c, k28, v28, n28, reps=2)
    del k28, v28
    n20 = 1 << 20
    k20 = c.ex.alloc(np.uint32, n20)
    c.mr.uniform_u32(c.ex, 0, 2**32 - 1, 1, k20)
    v20 = c.ex.alloc(np.uint32, n20)
    c.mr.iota_u32(c.ex, v20)
    ms = timed(c.ex, lambda: m.vector.sort_by_key(c.ex, k20, v20, m.op.Less), max(c.steps, 20), c.warmup)
    out["u32_iota_2^20"] = summarize("sort_by_key_u32_iota_2^20", ms, 36.0 * n20, n20)
    ref = m.vector.sort_by_key(c.ex, k20, v20, m.op.Less)
    out["reference_schedule_proxy_u32_iota_2^20"] = refproxy_case(c, k20, v20, n20, reps=20,
                                                                  check=c.ex.to_host(ref))
    return out


def rocprim_lib():
    """The vendor yardstick library (bench/vendor/), or None if not built."""
    import ctypes

    path = ROOT / "bench" / "vendor" / "librocprim_yardstick.so"
    if not path.exists():
        return None
    lib = ctypes.CDLL(str(path))
    P, U32, I = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
    for name in ("rocprim_sort_pairs_u32_u32_ms", "rocprim_sort_pairs_u64_u32_ms"):
        getattr(lib, name).restype = ctypes.c_double
        getattr(lib, name).argtypes = [P, P, U32, P, P, I]
    lib.rocprim_segmented_sort_pairs_u32_ms.restype = ctypes.c_double
    lib.rocprim_segmented_sort_pairs_u32_ms.argtypes = [P, P, U32, P, U32, P, P, I]
    lib.rocprim_inclusive_scan_u32_ms.restype = ctypes.c_double
    lib.rocprim_inclusive_scan_u32_ms.argtypes = [P, U32, P, I]
    return lib


def rocprim_sort_case(c, name, keys, vals, n, key_bytes, reps=5, check=None):
    """rocPRIM radix_sort_pairs on the same device-resident keys and values
    (a labelled vendor yardstick; never part of the library)."""
    lib = rocprim_lib()
    if lib is None:
        return None
    ok, ov = c.ex.alloc(keys.dtype, n), c.ex.alloc(c.np.uint32, n)
    c.ex.sync()
    f = lib.rocprim_sort_pairs_u32_u32_ms if key_bytes == 4 else lib.rocprim_sort_pairs_u64_u32_ms
    ms = f(keys.device_ptr, vals.device_ptr, n, ok.device_ptr, ov.device_ptr, reps)
    if ms < 0:
        return None
    d = summarize(f"rocprim_{name}", [ms], None, n, label="vendor yardstick (rocPRIM radix_sort_pairs)")
    if check is not None:
        d["same_values_as_msv"] = bool(c.ex.to_host(ov).tobytes() == check.tobytes())
    return d


def refproxy_case(c, keys, vals, n, reps, check=None):
    """The reference's radix schedule (4-bit LSD, 4 launches per pass, key
    gathers through the permutation, final gathers), timed on the same GPU
    as a stand-in for its CubeCL-HIP build; see bench/proxy/."""
    import ctypes

    path = ROOT / "bench" / "proxy" / "libmsv_refproxy.so"
    if not path.exists():
        return None
    lib = ctypes.CDLL(str(path))
    f = lib.refproxy_sort_by_key_u32_ms
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_int]
    ok, ov = c.ex.alloc(c.np.uint32, n), c.ex.alloc(c.np.uint32, n)
    c.ex.sync()
    ms = f(keys.device_ptr, vals.device_ptr, n, ok.device_ptr, ov.device_ptr, reps)
    if ms < 0:
        return None
    d = summarize(f"reference_schedule_proxy_sort_by_key_u32_{n}", [ms], None, n)
    if check is not None:
        d["same_values_as_msv"] = bool(c.ex.to_host(ov).tobytes() == check.tobytes())
    return d


def w_merge_sets(c):
    np, m = c.np, c.m
    n = 1 << 27
    a = c.ex.alloc(np.uint32, n)
    c.mr.uniform_u32(c.ex, 0, (1 << 26) - 1, 11, a)
    b = c.ex.alloc(np.uint32, n)
    c.mr.uniform_u32(c.ex, 0, (1 << 26) - 1, 12, b)
    sa, sb = m.vector.sort(c.ex, a, m.op.Less), m.vector.sort(c.ex, b, m.op.Less)
    out = {}
    ms = timed(c.ex, lambda: m.vector.merge(c.ex, sa, sb, m.op.Less), c.steps, c.warmup)
    out["merge"] = summarize("merge_u32_2x2^27", ms, 8.0 * 2 * n, 2 * n)
    # 25% overlap: sorted pcg images of [0, n) and [0.75n, 1.75n) (one
    # stream of 1.75n values of seed 0; pcg_hash32 is a bijection).
    stream = c.ex.alloc(np.uint32, 7 * n // 4)
    c.mr.uniform_u32(c.ex, 0, 2**32 - 1, 0, stream)
    A = m.vector.sort(c.ex, stream.slice(0, n), m.op.Less)
    B = m.vector.sort(c.ex, stream.slice(3 * n // 4, 7 * n // 4), m.op.Less)
    del stream
    for kind, fn, outlen in (("intersection", m.vector.set_intersection, n // 4),
                             ("difference", m.vector.set_difference, 3 * n // 4),
                             ("union", m.vector.set_union, 7 * n // 4)):
        ms = timed(c.ex, lambda: fn(c.ex, A, B, m.op.Less), c.steps, c.warmup)
        out[kind] = summarize(f"set_{kind}_25pct_2x2^27", ms, 4.0 * (2 * n) + 4.0 * outlen, 2 * n)
    return out


def lognormal_offsets(np, S=1 << 20, total=1 << 26, seed=7):
    rng = np.random.default_rng(seed)
    L = np.minimum(np.round(rng.lognormal(math.log(32), 1.1774, S)), 1 << 16).astype(np.int64)
    L[rng.random(S) < 0.02] = 0
    L = np.floor(L * (total / L.sum())).astype(np.int64)
    L[-1] += total - int(L.sum())
    return np.concatenate([[0], np.cumsum(L)]).astype(np.uint32)


def w_seg(c):
    np, m = c.np, c.m
    from massively_amd import seg

    off = lognormal_offsets(np)
    n, S = int(off[-1]), off.size - 1
    sg = seg.Segmentation.from_offsets(c.ex, c.ex.to_device(off))
    x = c.ex.alloc(np.uint32, n)
    c.mr.uniform_u32(c.ex, 0, 2**32 - 1, 7, x)
    f = c.ex.alloc(np.uint32, n)
    c.mr.uniform_u32(c.ex, 0x3F000000, 0x3F800000, 8, f)
    ff = c.ex.wrap(f.device_ptr, np.float32, n)
    out = {}
    it = seg.SegmentIterator(x, sg)
    ms = timed(c.ex, lambda: seg.ForEachSegment(seg.InclusiveScan(m.op.Add)).run(c.ex, it), c.steps, c.warmup)
    out["inclusive_scan_u32"] = summarize("seg_inclusive_scan_u32_2^26_in_2^20", ms, 8.0 * n + 4.0 * (S + 1), n)
    itf = seg.SegmentIterator(ff, sg)
    ms = timed(c.ex, lambda: seg.ForEachSegment(seg.Reduce(m.op.Add, 0.0)).run(c.ex, itf), c.steps, c.warmup)
    out["reduce_f32"] = summarize("seg_reduce_f32_2^26_in_2^20", ms, 4.0 * n + 4.0 * (S + 1) + 4.0 * S, n)
    its = seg.SegmentIterator((x, ff), sg)
    # seg::Sort: one persistent launch over all size classes (tile shape chosen
    # by the autotuner on the warm-up call) vs one launch per size class.
    run_sort = lambda: seg.ForEachSegment(seg.Sort(m.op.Less)).run(c.ex, its)
    for launch, tag in ((c.ex.SEG_LAUNCH_PERSISTENT, ""), (c.ex.SEG_LAUNCH_PER_CLASS, "_per_class")):
        c.ex.set_option(c.ex.OPT_SEG_LAUNCH, launch)
        ms = timed(c.ex, run_sort, c.steps, 1)
        d = summarize("seg_sort_u32key_f32val_2^26_in_2^20" + tag, ms, 16.0 * n + 4.0 * (S + 1), n)
        d["tuned_config"] = m.segsort_configs()[max(0, m.segsort_tuned_config(4, 4, launch, n, S))]
        # Every candidate tile shape, forced (the autotuner's table).
        cands = {}
        for k, cfg in enumerate(m.segsort_configs()):
            c.ex.set_option(c.ex.OPT_SEGSORT_CONFIG, k)
            cms = timed(c.ex, run_sort, 3, 1)
            cands["x".join(map(str, cfg))] = round(statistics.fmean(cms), 4)
        c.ex.set_option(c.ex.OPT_SEGSORT_CONFIG, -1)
        d["candidates_ms"] = cands
        log(f"[bench]   tuned {d['tuned_config']}, candidates (block x ipt x digit bits): {cands}")
        out["sort_u32_f32" + tag] = d
    c.ex.set_option(c.ex.OPT_SEG_LAUNCH, c.ex.SEG_LAUNCH_PERSISTENT)
    # The short segments packed into tile-sized windows instead of sorted
    # inside one wave each (the default), for comparison.
    c.ex.set_option(c.ex.OPT_SEGSORT_WAVE, 0)
    ms = timed(c.ex, run_sort, c.steps, 1)
    out["sort_u32_f32_windows"] = summarize("seg_sort_u32key_f32val_2^26_in_2^20_windows", ms,
                                            16.0 * n + 4.0 * (S + 1), n)
    c.ex.set_option(c.ex.OPT_SEGSORT_WAVE, 1)
    lib = rocprim_lib()
    if lib is not None:
        # Vendor yardstick: rocPRIM segmented_radix_sort_pairs on the same
        # segments (u32 keys, the f32 values moved as 4-byte payloads).
        doff = c.ex.to_device(off)
        ok, ov = c.ex.alloc(np.uint32, n), c.ex.alloc(np.uint32, n)
        c.ex.sync()
        ms = lib.rocprim_segmented_sort_pairs_u32_ms(x.device_ptr, f.device_ptr, n, doff.device_ptr, S,
                                                     ok.device_ptr, ov.device_ptr, 3)
        if ms > 0:
            out["rocprim_seg_sort_u32"] = summarize("rocprim_segmented_sort_pairs_u32_2^26_in_2^20", [ms],
                                                    16.0 * n + 4.0 * (S + 1), n,
                                                    label="vendor yardstick (rocPRIM segmented_radix_sort_pairs)")
    return out


def w_bykey(c):
    """The by-key family over 2^28 u32 values whose u32 keys form runs of
    random length (sorted uniform keys in [0, 2^24), mean run 16)."""
    np, m = c.np, c.m
    n = 1 << 28
    k = c.ex.alloc(np.uint32, n)
    c.mr.uniform_u32(c.ex, 0, (1 << 24) - 1, 21, k)
    ks = m.vector.sort(c.ex, k, m.op.Less)
    x = c.ex.alloc(np.uint32, n)
    c.mr.uniform_u32(c.ex, 0, 2**32 - 1, 22, x)
    out = {}
    ms = timed(c.ex, lambda: m.vector.inclusive_scan_by_key(c.ex, ks, x, m.op.Add), c.steps, c.warmup)
    out["inclusive_scan_by_key"] = summarize("inclusive_scan_by_key_u32_2^28", ms, 12.0 * n, n)
    ms = timed(c.ex, lambda: m.vector.reduce_by_key(c.ex, ks, x, 0, m.op.Add), c.steps, c.warmup)
    out["reduce_by_key"] = summarize("reduce_by_key_u32_2^28", ms, 8.0 * n + 8.0 * n / 16, n)
    ms = timed(c.ex, lambda: m.vector.unique_by_key(c.ex, ks, x), c.steps, c.warmup)
    out["unique_by_key"] = summarize("unique_by_key_u32_2^28", ms, 8.0 * n + 4.0 * n / 16, n)
    return out


def w_columns(c):
    """radix_sort_by_key over zipped columns, 2^26 rows: (u32, u32) keys (one
    packed 64-bit sort) and (u32, u32, u32) keys (two sorts), each with three
    u32 value columns gathered by the row permutation."""
    np, m = c.np, c.m
    n = 1 << 26
    cols = []
    for s in range(3):
        k = c.ex.alloc(np.uint32, n)
        c.mr.uniform_u32(c.ex, 0, 2**32 - 1, 40 + s, k)
        cols.append(k)
    vals = []
    for s in range(3):
        v = c.ex.alloc(np.uint32, n)
        c.mr.iota_u32(c.ex, v)
        vals.append(v)
    out = {}
    for nk in (2, 3):
        keys = tuple(cols[:nk])
        ms = timed(c.ex, lambda: m.vector.radix_sort_by_key(c.ex, keys, tuple(vals)), c.steps, c.warmup)
        # model: one 64-bit (key + permutation) sort per 64-bit key group
        # (12 B read + 12 B written per pass, 8 passes), the packing reads, and
        # three value gathers (4 B index + 4 B value read + 4 B written).
        groups = 1 if nk == 2 else 2
        b = n * (groups * (8 * 24 + 4.0 * nk) + 3 * 12.0)
        out[f"keys{nk}_vals3"] = summarize(f"radix_sort_by_key_{nk}xu32_keys_3xu32_values_2^26", ms, b, n)
    # Zipped rows moved by one launch that reads the indices / stencil once,
    # against one call per column: gather of 4 u32 columns through a random
    # permutation of 2^26, and copy_where of 4 u32 columns by a 50% stencil.
    zc = tuple(cols) + (vals[0],)
    perm = c.ex.alloc(np.uint32, n)
    c.mr.uniform_u32(c.ex, 0, n - 1, 47, perm)
    st = c.ex.alloc(np.uint32, n)
    c.mr.uniform_u32(c.ex, 0, 1, 48, st)
    gb = n * (4.0 + 4 * (4.0 + 4.0))  # indices once, per column a 4 B read and a 4 B write
    ms = timed(c.ex, lambda: m.vector.gather(c.ex, zc, perm), c.steps, c.warmup)
    out["gather_4cols_one_pass"] = summarize("gather_4xu32_columns_2^26_one_pass", ms, gb, n)
    ms = timed(c.ex, lambda: [m.vector.gather(c.ex, col, perm) for col in zc], c.steps, c.warmup)
    out["gather_4cols_per_column"] = summarize("gather_4xu32_columns_2^26_per_column", ms, gb, n)
    # The same with coherent indices (every 4th row, ascending): the index
    # reads are then a quarter of each column's traffic.
    from massively_amd import lazy
    from massively_amd.op import UnaryOp

    sidx = m.vector.map(c.ex, lazy.stride(0, 4).take(n // 4), UnaryOp("x"))
    sb = (n // 4) * (4.0 + 4 * (4.0 + 4.0))
    ms = timed(c.ex, lambda: m.vector.gather(c.ex, zc, sidx), c.steps, c.warmup)
    out["gather_4cols_strided_one_pass"] = summarize("gather_4xu32_columns_stride4_2^24_one_pass", ms, sb, n // 4)
    ms = timed(c.ex, lambda: [m.vector.gather(c.ex, col, sidx) for col in zc], c.steps, c.warmup)
    out["gather_4cols_strided_per_column"] = summarize("gather_4xu32_columns_stride4_2^24_per_column", ms, sb,
                                                       n // 4)
    cb = n * (4.0 + 4 * (4.0 + 0.5 * 4.0))  # stencil once, per column a read and the kept half written
    ms = timed(c.ex, lambda: m.vector.copy_where(c.ex, zc, st), c.steps, c.warmup)
    out["copy_where_4cols_one_pass"] = summarize("copy_where_4xu32_columns_2^26_one_pass", ms, cb, n)
    ms = timed(c.ex, lambda: [m.vector.copy_where(c.ex, col, st) for col in zc], c.steps, c.warmup)
    out["copy_where_4cols_per_column"] = summarize("copy_where_4xu32_columns_2^26_per_column", ms, cb, n)
    return out


def w_compact(c):
    """seg::{Filter, Unique, Take} over the seg workload's segments (2^26 u32
    values in [0, 4), 2^20 log-normal segments), partition of 2^28 u32 at 50%,
    and scatter_reduce add of 2^28 u32 (and 2^26 f32) values into 2^24 uniform
    random destinations."""
    np, m = c.np, c.m
    from massively_amd import seg

    off = lognormal_offsets(np)
    n, S = int(off[-1]), off.size - 1
    sg = seg.Segmentation.from_offsets(c.ex, c.ex.to_device(off))
    x = c.ex.alloc(np.uint32, n)
    c.mr.uniform_u32(c.ex, 0, 3, 31, x)
    it = seg.SegmentIterator(x, sg)
    out = {}
    for name, alg in (("filter_lt2", seg.Filter(m.op.Lt(2))), ("unique", seg.Unique()),
                      ("take16", seg.Take(16))):
        kept = [0]

        def run():
            r = seg.ForEachSegment(alg).run(c.ex, it)
            kept[0] = r.offsets.value_count()

        ms = timed(c.ex, run, c.steps, c.warmup)
        # read the values and offsets, write the kept values and new offsets
        out["seg_" + name] = summarize(f"seg_{name}_u32_2^26_in_2^20", ms,
                                       4.0 * n + 4.0 * kept[0] + 8.0 * (S + 1), n, kept=kept[0])
    N = 1 << 28
    y = c.ex.alloc(np.uint32, N)
    c.mr.uniform_u32(c.ex, 0, 2**32 - 1, 32, y)
    ms = timed(c.ex, lambda: m.vector.partition(c.ex, y, m.op.Lt(1 << 31)), c.steps, c.warmup)
    # count pass (read) + partition pass (read + write)
    out["partition"] = summarize("partition_u32_50pct_2^28", ms, 12.0 * N, N)
    D = 1 << 24
    idx = c.ex.alloc(np.uint32, N)
    c.mr.uniform_u32(c.ex, 0, D - 1, 33, idx)
    dst = c.ex.alloc(np.uint32, D)
    ms = timed(c.ex, lambda: m.vector.scatter_reduce(c.ex, y, idx, 0, m.op.Add, dst), c.steps, c.warmup)
    out["scatter_reduce_u32"] = summarize("scatter_reduce_add_u32_2^28_into_2^24", ms,
                                          8.0 * N + 8.0 * D, N)
    Nf = 1 << 26
    fv = c.ex.alloc(np.uint32, Nf)
    c.mr.uniform_u32(c.ex, 0x3F000000, 0x3F7FFFFF, 34, fv)
    fvf = c.ex.wrap(fv.device_ptr, np.float32, Nf)
    fi = c.ex.wrap(idx.device_ptr, np.uint32, Nf)
    fdst = c.ex.alloc(np.float32, D)
    ms = timed(c.ex, lambda: m.vector.scatter_reduce(c.ex, fvf, fi, 0.0, m.op.Add, fdst), c.steps, c.warmup)
    out["scatter_reduce_f32"] = summarize("scatter_reduce_add_f32_2^26_into_2^24_sorted", ms,
                                          8.0 * Nf + 8.0 * D, Nf)
    return out


def w_lazy(c):
    """Lazy inputs fused into the consuming kernel, 2^28 elements: reduce of
    map(x, x*x) over f32 (against writing the map out and reducing it), the
    max of counting(0).take(2^32 - 1) (no memory traffic), and the scan and
    copy_where of map(y, x*3+1) over u32."""
    np, m = c.np, c.m
    from massively_amd import lazy
    from massively_amd.op import UnaryOp

    N = 1 << 28
    raw = c.ex.alloc(np.uint32, N)
    c.mr.uniform_u32(c.ex, 0x3F000000, 0x3F7FFFFF, 41, raw)
    x = c.ex.wrap(raw.device_ptr, np.float32, N)
    sq = lazy.map(x, UnaryOp("x * x"))
    out = {}
    ms = timed(c.ex, lambda: m.vector.reduce(c.ex, sq, 0.0, m.op.Add), c.steps, c.warmup)
    out["reduce_map_f32"] = summarize("reduce_map_sq_f32_2^28_fused", ms, 4.0 * N, N)
    ms = timed(c.ex, lambda: m.vector.reduce(c.ex, sq.materialize(c.ex), 0.0, m.op.Add),
               c.steps, c.warmup)
    out["reduce_map_f32_unfused"] = summarize("reduce_map_sq_f32_2^28_written_out", ms, 4.0 * N, N)
    cnt = lazy.counting(0).take(0xFFFFFFFF)
    ms = timed(c.ex, lambda: m.vector.reduce(c.ex, cnt, 0, m.op.Max), c.steps, c.warmup)
    out["reduce_counting_max"] = summarize("reduce_max_counting_2^32-1", ms, None, 0xFFFFFFFF)
    y = c.ex.alloc(np.uint32, N)
    c.mr.uniform_u32(c.ex, 0, 2**32 - 1, 42, y)
    t = lazy.map(y, UnaryOp("x * 3u + 1u"))
    ms = timed(c.ex, lambda: m.vector.inclusive_scan(c.ex, t, m.op.Add), c.steps, c.warmup)
    out["scan_map_u32"] = summarize("inclusive_scan_map_u32_2^28_fused", ms, 8.0 * N, N)
    last = [None]

    def sel():
        last[0] = m.vector.copy_where_if(c.ex, t, t, m.op.Lt(1 << 31))

    ms = timed(c.ex, sel, c.steps, c.warmup)
    kept = last[0].len()  # the device-resident length, read after the timed region
    out["copy_where_map_u32"] = summarize("copy_where_lt_map_u32_2^28_fused", ms,
                                          4.0 * N + 4.0 * kept, N, kept=kept)
    return out


def served_trace(np, jobs=500, seed=10, log2_max=26):
    """The served mixed stream: kinds uniform over sort_by_key / inclusive_scan /
    reduce / copy_where, sizes log-uniform in 2^10..2^log2_max, key dtypes
    uniform over u32 / u64 / f32, one seeded input per job."""
    from massively_amd import op, serve

    rng = np.random.default_rng(seed)
    reqs, bytes_ = [], []
    for i in range(jobs):
        kind = ("sort_by_key", "inclusive_scan", "reduce", "copy_where")[rng.integers(0, 4)]
        dtype = (np.uint32, np.uint64, np.float32)[rng.integers(0, 3)]
        n = int(round(2 ** rng.uniform(10, log2_max)))
        es = np.dtype(dtype).itemsize
        if kind == "sort_by_key":
            r = serve.Request(kind, dtype, n, seed=1000 + i, op=op.Less)
            passes = 2 * es
            b = n * (es + passes * 2 * (es + 4))
        elif kind == "inclusive_scan":
            r = serve.Request(kind, dtype, n, seed=1000 + i, op=op.Add)
            b = 2 * n * es
        elif kind == "reduce":
            r = serve.Request(kind, dtype, n, seed=1000 + i, op=op.Add)
            b = n * es
        else:
            thr = 0.75 if dtype == np.float32 else (1 << (8 * es - 1))
            r = serve.Request(kind, dtype, n, seed=1000 + i, pred=op.Lt(thr))
            b = 1.5 * n * es
        reqs.append(r)
        bytes_.append(b)
    return reqs, bytes_


def run_served(np, mode, reqs, arrivals_us=None):
    """Submits the requests (at the given arrival times, or all at once),
    returns (makespan_us, latencies_us, service_us): latency is submit to
    done, service is admit to done."""
    from massively_amd import serve

    srv = serve.Server(streams=4, max_batch=256, mode=mode, hbm_budget=64 << 30)
    try:
        # Warm-up: one pass over the requests, all at once (allocator and
        # kernel caches; the batched grids' autotuning of the batch shapes).
        for t in [srv.submit(r) for r in reqs]:
            srv.wait(t)
            srv.release(t)
        t0 = srv.now_us()
        tickets = []
        for i, r in enumerate(reqs):
            if arrivals_us is not None:
                due = t0 + arrivals_us[i]
                while srv.now_us() < due:
                    pass
            tickets.append(srv.submit(r))
        lat, svc, last = [], [], t0
        for t in tickets:
            info = srv.wait(t)
            lat.append(info["done_us"] - info["submit_us"])
            svc.append(info["done_us"] - info["admit_us"])
            last = max(last, info["done_us"])
            srv.release(t)
        return last - t0, lat, svc
    finally:
        srv.close()


SIZE_BANDS = [(10, 14), (14, 18), (18, 22), (22, 27)]


def band_percentiles(np, reqs, lat):
    """p50 / p99 latency (us) per size band [2^lo, 2^hi) of the request length."""
    out = {}
    n = np.array([r.n for r in reqs])
    lat = np.asarray(lat)
    for lo, hi in SIZE_BANDS:
        m = (n >= (1 << lo)) & (n < (1 << hi))
        if m.any():
            out[f"2^{lo}..2^{hi}"] = dict(jobs=int(m.sum()),
                                         p50_us=round(float(np.percentile(lat[m], 50)), 1),
                                         p99_us=round(float(np.percentile(lat[m], 99)), 1))
    return out


def w_served(c):
    """Served mixed stream: capacity (all requests at once) and Poisson
    arrivals at 80% of the serial capacity, for continuous batching against
    two baselines: one request at a time through the device-wide kernels
    ("serial") and one request per launch sequence on the same four streams
    ("per_job"). The Amdahl bound on batching's gain over serial comes from the
    serial service times: batching can remove at most the time of the
    requests the batched kernels take (all requests of at most 2^22
    elements), never that of the larger ones."""
    np = c.np
    reqs, bytes_ = served_trace(np)
    total_bytes = float(sum(bytes_))
    out = {"jobs": len(reqs), "algorithmic_bytes": total_bytes}
    cap = {}
    modes = ("batched", "per_job", "serial")
    for mode in modes:
        span, lat, svc = run_served(np, mode, reqs)
        cap[mode] = len(reqs) / (span * 1e-6)
        out[f"capacity_{mode}"] = dict(makespan_ms=round(span / 1e3, 3), jobs_per_s=round(cap[mode], 1),
                                       gbs=round(total_bytes / (span * 1e-6) / 1e9, 1))
        log(f"[bench] served capacity {mode}: {span / 1e3:.2f} ms for {len(reqs)} jobs = "
            f"{cap[mode]:.0f} jobs/s, {total_bytes / (span * 1e-6) / 1e9:.1f} GB/s")
        if mode == "serial":
            svc = np.asarray(svc)
            big = np.array([r.n > (1 << 22) for r in reqs])
            out["serial_service_ms"] = {"all": round(float(svc.sum()) / 1e3, 3),
                                        "over_2^22": round(float(svc[big].sum()) / 1e3, 3)}
            out["amdahl_bound_vs_serial"] = round(float(svc.sum() / max(svc[big].sum(), 1e-9)), 3)
            log(f"[bench] served serial service: {svc.sum() / 1e3:.2f} ms in all, "
                f"{svc[big].sum() / 1e3:.2f} ms for requests over 2^22 "
                f"(bound on the batched gain {out['amdahl_bound_vs_serial']:.2f}x)")
    rate = 0.8 * cap["serial"]
    rng = np.random.default_rng(10)
    arrivals = np.cumsum(rng.exponential(1e6 / rate, len(reqs)))
    out["poisson_rate_jobs_per_s"] = round(rate, 1)
    for mode in modes:
        span, lat, _ = run_served(np, mode, reqs, arrivals)
        p50, p99 = np.percentile(lat, 50), np.percentile(lat, 99)
        out[f"poisson_{mode}"] = dict(p50_us=round(float(p50), 1), p99_us=round(float(p99), 1),
                                      mean_us=round(float(np.mean(lat)), 1),
                                      jobs_per_s=round(len(reqs) / (span * 1e-6), 1),
                                      by_size=band_percentiles(np, reqs, lat))
        log(f"[bench] served poisson {mode} @ {rate:.0f} jobs/s: p50 {p50:.0f} us, p99 {p99:.0f} us")
    out["capacity_ratio_batched_over_serial"] = round(cap["batched"] / cap["serial"], 3)
    out["capacity_ratio_batched_over_per_job"] = round(cap["batched"] / cap["per_job"], 3)
    # The launch-bound part of the stream on its own: the same mix with sizes
    # 2^10..2^16 only (a separate, labelled measurement).
    sreqs, sbytes = served_trace(np, jobs=2000, seed=11, log2_max=16)
    scap = {}
    for mode in modes:
        span, lat, _ = run_served(np, mode, sreqs)
        scap[mode] = len(sreqs) / (span * 1e-6)
        out[f"small_jobs_capacity_{mode}"] = dict(jobs=len(sreqs), makespan_ms=round(span / 1e3, 3),
                                                  jobs_per_s=round(scap[mode], 1),
                                                  gbs=round(sum(sbytes) / (span * 1e-6) / 1e9, 2))
        log(f"[bench] served small-jobs capacity {mode}: {scap[mode]:.0f} jobs/s")
    out["small_jobs_capacity_ratio"] = round(scap["batched"] / scap["serial"], 3)
    out["small_jobs_capacity_ratio_over_per_job"] = round(scap["batched"] / scap["per_job"], 3)
    return out


def m_np():
    import numpy

    return numpy


def host_cpu():
    """The host's CPU model and logical CPU counts (all, and usable by this process)."""
    import os

    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return dict(cpu_model=model, nproc=os.cpu_count(), nproc_usable=usable)


def cpu_baseline(n_sample: int, reps: int):
    import oracle

    x = oracle.uniform_u32(0, 1023, 2, n_sample)
    oracle.inclusive_scan(x[:1024], "add")
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        oracle.inclusive_scan(x, "add")
        t.append(time.perf_counter() - t0)
    sec = statistics.median(t)
    return dict(value=round(8.0 * n_sample / sec / 1e9, 3), unit="GB/s", cores=1, kind="port",
                sample=f"inclusive_scan add over {n_sample} u32 (seed 2), median of {reps} "
                       f"single-thread runs of the C oracle (restatement of the reference CPU path; "
                       f"{sum(t):.1f} s of CPU work in all)",
                runs=reps, min_s=round(min(t), 4), median_s=round(sec, 4), **host_cpu())


def cpu_sort_baseline(n_sample: int, reps: int):
    import numpy as np

    import oracle

    keys = oracle.uniform_u64(0, 2**64 - 1, 6, n_sample)
    vals = np.arange(n_sample, dtype=np.uint32)
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        oracle.radix_sort_by_key(keys, vals)
        t.append(time.perf_counter() - t0)
    sec = statistics.median(t)
    return dict(value=round(n_sample / sec / 1e9, 5), unit="Gkeys/s", cores=1, kind="port",
                sample=f"stable sort_by_key of {n_sample} u64 keys (seed 6) + u32 payload, median "
                       f"of {reps} single-thread runs of the C oracle",
                runs=reps, **host_cpu())


def w_copy(c):
    """Calibration leg: a plain device copy of 2^28 u32 (1 GiB read + 1 GiB
    written) by torch's copy kernel. Its FETCH_SIZE / WRITE_SIZE counters
    scale the counters of the other kernels to bytes, and its rate is the
    practical streaming ceiling the scan is compared against."""
    import torch

    n = 1 << 28
    x = torch.arange(n, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(c.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(c.steps)]
    for a, b in evs:
        a.record()
        y.copy_(x)
        b.record()
    torch.cuda.synchronize()
    ms = [a.elapsed_time(b) for a, b in evs]
    return summarize("device_copy_u32_2^28", ms, 8.0 * n, n)


def load_traffic(name: str):
    """HBM bytes per launch of a kernel from a committed counter summary
    (a `traffic_bytes_per_launch: <int>` line in profiles/<name>.md), or None."""
    f = ROOT / "profiles" / f"{name}.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":", 1)[1].strip())
    return None


WORKLOADS = {
    "copy": w_copy,
    "reduce": w_reduce,
    "copy_where": w_copy_where,
    "gather_scatter": w_gather_scatter,
    "sort": w_sort,
    "merge_sets": w_merge_sets,
    "seg": w_seg,
    "bykey": w_bykey,
    "compact": w_compact,
    "columns": w_columns,
    "lazy": w_lazy,
    "served": w_served,
}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--only", default="", help="comma list of: scan," + ",".join(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {a.gpus} is not supported (the multi-GPU "
              "sharding is a design, not a measured path)", file=sys.stderr)
        return 2

    import massively_amd as m

    ex = m.Executor(0)
    c = Ctx(ex, m, a.steps, a.warmup)
    only = set(filter(None, a.only.split(",")))
    n = 1 << 28
    log(f"[bench] scan n=2^28 steps={a.steps} warmup={a.warmup}")
    scan = w_scan(c)
    extra = {}
    lib = rocprim_lib()
    if lib is not None and (not only or "vendor" in only):
        x = ex.alloc(m_np().uint32, n)
        c.mr.uniform_u32(ex, 0, 1023, 2, x)
        y = ex.alloc(m_np().uint32, n)
        ex.sync()
        ms = lib.rocprim_inclusive_scan_u32_ms(x.device_ptr, n, y.device_ptr, 10)
        if ms > 0:
            extra["rocprim_inclusive_scan_u32_2^28"] = summarize(
                "rocprim_inclusive_scan_u32_add_2^28", [ms], 8.0 * n, n,
                label="vendor yardstick (rocPRIM inclusive_scan)")
        del x, y
    for name, fn in WORKLOADS.items():
        if (only and name not in only) or (not only and name == "copy"):
            continue
        log(f"[bench] workload {name}")
        extra[name] = fn(c)
        ex.trim()

    cpu = None
    if not a.no_cpu:
        log("[bench] cpu baseline (C oracle, 1 thread)")
        cpu = cpu_baseline(n, 15)
        log(f"[bench] cpu scan: {cpu['value']:.2f} GB/s")
        if not only or "sort" in only:
            extra["cpu_sort_baseline"] = cpu_sort_baseline(1 << 23, 5)
            log(f"[bench] cpu sort: {extra['cpu_sort_baseline']['value']:.4f} Gkeys/s")

    sort_g = extra.get("sort", {}).get("u32_2^28", {}).get("gitems_per_s")
    proxy_g = (extra.get("sort", {}).get("reference_schedule_proxy_u32_2^28") or {}).get("gitems_per_s")
    served = extra.get("served", {})
    line = {
        "metric": METRIC,
        "value": scan["gbs"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": scan["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded PCG streams of the reference generator, generated on device)",
        "config": {"workload": "inclusive_scan add over 2^28 u32 uniform in [0,1024) (seed 2), "
                               "wrapping; device-resident input", "n": n},
        "roofline": {"bound": "hbm", "achieved": scan["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": scan["frac_hbm"], "traffic": load_traffic("pmc_inclusive_scan_u32_2^28")},
        "cpu_baseline": cpu,
        "sort_gkeys_per_s_u32_2^28": sort_g,
        "sort_vs_reference_schedule_proxy": round(sort_g / proxy_g, 2) if sort_g and proxy_g else None,
        "served_p50_p99_us": [served.get("poisson_batched", {}).get("p50_us"),
                              served.get("poisson_batched", {}).get("p99_us")] if served else None,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    ex.sync()
    ex.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
