"""GPU: library calls captured into a HIP graph and replayed.

Each replay runs after eager work that rewrites the inputs (the library's
generator), and every replay's outputs are checked against the oracle, so a
replay that reused stale lookback state or scratch would fail.
"""
import numpy as np
import pytest

import massively_amd as m
import oracle
from massively_amd import random as mr
from massively_amd import seg, serve

pytestmark = pytest.mark.gpu


def test_graph_replays_scan_seg_sort_and_jobs(exec_):
    ex = exec_
    n = 300000
    x = ex.alloc(np.uint32, n)
    lengths = np.random.default_rng(5).integers(0, 3000, 200)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.uint32)
    ns = int(off[-1])
    sg = seg.Segmentation.from_offsets(ex, ex.to_device(off))
    k = ex.alloc(np.uint32, ns)
    v = ex.alloc(np.uint32, ns)
    mr.iota_u32(ex, v)
    small = ex.alloc(np.uint32, 5000)
    small_out = ex.alloc(np.uint32, 5000)
    skeys = ex.alloc(np.uint32, 3000)
    skeys_out = ex.alloc(np.uint32, 3000)
    jobs = [serve.Job("exclusive_scan", small, small_out, op=m.op.Add, init=7),
            serve.Job("sort_by_key", skeys, None, op=m.op.Less, out2=skeys_out)]
    ex.sync()
    with ex.capture() as cap:
        y = m.vector.inclusive_scan(ex, x, m.op.Add)
        srt = seg.ForEachSegment(seg.Sort(m.op.Less)).run(ex, seg.SegmentIterator((k, v), sg))
        serve.run_jobs(ex, jobs)
    g = cap.graph
    with pytest.raises(m.Error):
        with ex.capture():
            ex.to_host(x)
    for rep in range(4):
        mr.uniform_u32(ex, 0, 1000, 40 + rep, x)
        mr.uniform_u32(ex, 0, 50, 60 + rep, k)
        mr.uniform_u32(ex, 0, 2**32 - 1, 80 + rep, small)
        mr.uniform_u32(ex, 0, 99, 90 + rep, skeys)
        ex.sync()
        g.launch()
        ex.sync()
        hx = oracle.uniform_u32(0, 1000, 40 + rep, n)
        assert ex.to_host(y).tobytes() == oracle.inclusive_scan(hx, "add").tobytes(), rep
        hk = oracle.uniform_u32(0, 50, 60 + rep, ns)
        perm = oracle.seg_sort_perm(hk, off, "less")
        assert ex.to_host(srt.values[0]).tobytes() == hk[perm].tobytes(), rep
        assert ex.to_host(srt.values[1]).tobytes() == perm.astype(np.uint32).tobytes(), rep
        hs = oracle.uniform_u32(0, 2**32 - 1, 80 + rep, 5000)
        assert ex.to_host(small_out).tobytes() == oracle.exclusive_scan(hs, 7, "add").tobytes(), rep
        hk2 = oracle.uniform_u32(0, 99, 90 + rep, 3000)
        assert ex.to_host(skeys_out).tobytes() == np.sort(hk2, kind="stable").tobytes(), rep
    g.close()
