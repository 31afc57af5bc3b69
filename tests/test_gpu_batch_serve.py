"""GPU parity of the batched-jobs path (msv_run_jobs) and the serving loop
(msv_server_*), through the C ABI.

* Integer scans / reduce, copy_where and sorts are bit-exact vs the oracle,
  for jobs on the batched (one-workgroup) kernels and on the device-wide ones.
* f32 scan and reduce jobs are deterministic; a reduce lies within
  (17 + 2) * eps * sum|x| of the FP64 sum (the batched tree has 4 + 6 + 2 + 5
  levels for up to 2^17 elements, the device-wide one fewer), each inclusive
  scan output within (c + 16) * eps * prefix_sum|x| of the FP64 prefix
  (c = number of 2048-element chunks carried before it).
* A job's result is the same alone and in any mix (run_jobs and the server).
"""
import numpy as np
import pytest

import massively_amd as m
import oracle
from massively_amd import serve

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(77)
EPS32 = 2.0**-24
SIZES = [0, 1, 100, 2047, 2048, 2049, 5000, 1 << 16, (1 << 17) - 1, 1 << 17, (1 << 17) + 1, 300000]
SORT_SIZES = [0, 1, 2, 255, 4095, 4096, 4097, 20000, 1 << 15, (1 << 15) + 1, 100000]


def _data(dtype, n):
    dt = np.dtype(dtype)
    if dt.kind == "f":
        return RNG.standard_normal(n).astype(dt)
    if dt.itemsize == 4:
        return RNG.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32).view(dt)
    return RNG.integers(0, 2**63, n, dtype=np.uint64).view(dt)


def _check_scan(got, x, op, excl, init):
    if x.dtype.kind == "f":
        exact = np.cumsum(x.astype(np.float64))
        mag = np.cumsum(np.abs(x.astype(np.float64)))
        if excl:
            exact = np.concatenate([[0.0], exact[:-1]]) + float(init)
            mag = np.concatenate([[0.0], mag[:-1]]) + abs(float(init))
        chunks = np.arange(x.size) // 2048
        assert np.all(np.abs(got.astype(np.float64) - exact) <= (chunks + 16) * EPS32 * mag + 1e-30)
    else:
        want = oracle.exclusive_scan(x, init, op) if excl else oracle.inclusive_scan(x, op)
        assert got.tobytes() == want.tobytes()


def _jobs(ex):
    """A mix of jobs over every path, with their host inputs."""
    jobs, host = [], []
    for n in SIZES:
        for dtype, op, name in ((np.uint32, m.op.Add, "add"), (np.int64, m.op.Max, "max"),
                                (np.float32, m.op.Add, "add"), (np.uint64, m.op.Xor, "xor")):
            x = _data(dtype, n)
            d = ex.to_device(x) if n else ex.alloc(dtype, 0)
            kind = ("inclusive_scan", "exclusive_scan")[len(jobs) % 2]
            init = 0 if kind == "inclusive_scan" else (3 if np.dtype(dtype).kind != "f" else 0.5)
            jobs.append(serve.Job(kind, d, ex.alloc(dtype, max(n, 1)), op=op, init=init))
            host.append((kind, x, name, init))
            jobs.append(serve.Job("reduce", d, ex.alloc(dtype, 1), op=op, init=init))
            host.append(("reduce", x, name, init))
        x = _data(np.uint32, n)
        d = ex.to_device(x) if n else ex.alloc(np.uint32, 0)
        jobs.append(serve.Job("copy_where", d, ex.alloc(np.uint32, max(n, 1)), pred=m.op.Lt(1 << 30),
                              out2=ex.alloc(np.uint32, 1)))
        host.append(("copy_where", x, None, None))
    for n in SORT_SIZES:
        for kd, vd, cmp in ((np.uint32, np.uint32, None), (np.float32, np.uint32, m.op.Less),
                            (np.uint64, np.uint32, m.op.Greater), (np.int32, None, m.op.Less)):
            k = _data(kd, n)
            if np.dtype(kd).kind == "f":
                k = np.round(k * 4).astype(kd)
            else:
                k = (k % 1000).astype(kd)
            v = np.arange(n, dtype=np.uint32) if vd is not None else None
            dk = ex.to_device(k) if n else ex.alloc(kd, 0)
            dv = (ex.to_device(v) if n else ex.alloc(np.uint32, 0)) if vd is not None else None
            jobs.append(serve.Job("sort_by_key", dk, ex.alloc(np.uint32, max(n, 1)) if vd is not None else None,
                                  op=cmp, inp2=dv, out2=ex.alloc(kd, max(n, 1))))
            host.append(("sort_by_key", k, cmp, v))
    return jobs, host


def _results(ex, jobs, host):
    out = []
    for j, (kind, x, a, b) in zip(jobs, host):
        n = x.size
        if kind == "reduce":
            out.append(ex.to_host(j.out)[:1])
        elif kind == "copy_where":
            cnt = int(ex.to_host(j.out2)[0])
            out.append(ex.to_host(j.out)[:cnt])
        elif kind == "sort_by_key":
            ks = ex.to_host(j.out2)[:n]
            vs = ex.to_host(j.out)[:n] if j.out is not None else None
            out.append((ks, vs))
        else:
            out.append(ex.to_host(j.out)[:n])
    return out


def _check(host, res):
    for (kind, x, a, b), r in zip(host, res):
        n = x.size
        if kind in ("inclusive_scan", "exclusive_scan"):
            _check_scan(r, x, a, kind == "exclusive_scan", b)
        elif kind == "reduce":
            if x.dtype.kind == "f":
                s, mag = oracle.sum_f32_f64(x)
                levels = max(17, int(np.ceil(np.log2(max(n, 2)))))
                assert abs(float(r[0]) - (s + float(b))) <= (levels + 2) * EPS32 * (mag + abs(float(b))) + 1e-30
            else:
                assert r[0] == oracle.reduce(x, b, a), (n, x.dtype)
        elif kind == "copy_where":
            want = x[x < (1 << 30)]
            assert r.tobytes() == want.tobytes(), n
        else:
            mode = "total" if a is None else ("less" if a is m.op.Less else "greater")
            perm = oracle.sort_perm(x, mode)
            assert r[0].tobytes() == x[perm].tobytes(), (n, x.dtype, mode)
            if b is not None:
                assert r[1].tobytes() == b[perm].tobytes(), (n, x.dtype, mode)


def test_run_jobs_mixed_batch(exec_):
    jobs, host = _jobs(exec_)
    serve.run_jobs(exec_, jobs)
    res = _results(exec_, jobs, host)
    _check(host, res)
    # Deterministic, and each job alone gives the same bytes as in the mix.
    serve.run_jobs(exec_, jobs)
    again = _results(exec_, jobs, host)
    for i in range(0, len(jobs), 7):
        serve.run_jobs(exec_, [jobs[i]])
        alone = _results(exec_, [jobs[i]], [host[i]])[0]
        for r in (again[i], alone):
            if isinstance(r, tuple):
                assert r[0].tobytes() == res[i][0].tobytes()
                assert (r[1] is None) or r[1].tobytes() == res[i][1].tobytes()
            else:
                assert r.tobytes() == res[i].tobytes(), (i, host[i][0])


def test_run_jobs_rejects_bad_jobs(exec_):
    x = exec_.to_device(np.arange(10, dtype=np.float32))
    with pytest.raises(m.Error):
        serve.run_jobs(exec_, [serve.Job("reduce", x, exec_.alloc(np.float32, 1), op=m.op.Xor)])
    with pytest.raises(m.Error):
        serve.run_jobs(exec_, [serve.Job("copy_where", x, exec_.alloc(np.float32, 10), pred=m.op.Lt(1.0))])


def _host_input(req):
    """The server's seeded input for a request (the library generator)."""
    dt = np.dtype(req.dtype)
    if dt == np.float32:
        return oracle.uniform_u32(0x3F000000, 0x3F7FFFFF, req.seed, req.n).view(np.float32)
    if dt.itemsize == 4:
        return oracle.uniform_u32(0, 2**32 - 1, req.seed, req.n).view(dt)
    return oracle.uniform_u64(0, 2**64 - 1, req.seed, req.n).view(dt)


def _requests():
    reqs = []
    seed = 100
    for n in (1000, 5000, 1 << 15, (1 << 15) + 7, 1 << 17, 200000, 1 << 20):
        for dtype in (np.uint32, np.uint64, np.float32):
            seed += 1
            reqs.append(serve.Request("sort_by_key", dtype, n, seed=seed, op=m.op.Less))
            reqs.append(serve.Request("inclusive_scan", dtype, n, seed=seed, op=m.op.Add))
            reqs.append(serve.Request("reduce", dtype, n, seed=seed, op=m.op.Add))
            thr = 0.75 if dtype == np.float32 else (1 << 31 if dtype == np.uint32 else 1 << 63)
            reqs.append(serve.Request("copy_where", dtype, n, seed=seed, pred=m.op.Lt(thr)))
    return reqs


def _fetch(srv, t):
    info = srv.wait(t)
    r = srv.result(t, info["out_len"])
    srv.release(t)
    return r


def _same(a, b):
    if isinstance(a, tuple):
        return a[0].tobytes() == b[0].tobytes() and a[1].tobytes() == b[1].tobytes()
    if isinstance(a, np.ndarray):
        return a.tobytes() == b.tobytes()
    return np.array([a]).tobytes() == np.array([b]).tobytes()


@pytest.mark.parametrize("mode", ["batched", "serial", "per_job"])
def test_server_alone_equals_mix(mode):
    reqs = _requests()
    with serve.Server(streams=4, max_batch=64, mode=mode, hbm_budget=8 << 30) as srv:
        alone = [_fetch(srv, srv.submit(r)) for r in reqs]
        tickets = [srv.submit(r) for r in reqs]
        mixed = [_fetch(srv, t) for t in tickets]
    for r, a, b in zip(reqs, alone, mixed):
        assert _same(a, b), (r.kind, r.dtype, r.n)
        x = _host_input(r)
        if r.kind == "sort_by_key":
            perm = oracle.sort_perm(x, "less")
            assert a[0].tobytes() == x[perm].tobytes()
            assert a[1].tobytes() == perm.astype(np.uint32).tobytes()
        elif r.kind == "copy_where":
            assert a.tobytes() == x[x < r.pred.a].tobytes()
        elif r.kind == "inclusive_scan":
            _check_scan(a, x, "add", False, 0)
        else:
            if x.dtype.kind == "f":
                s, mag = oracle.sum_f32_f64(x)
                assert abs(float(a) - s) <= 19 * EPS32 * mag
            else:
                assert a == oracle.reduce(x, 0, "add")


def test_server_modes_agree_on_exact_jobs():
    reqs = [r for r in _requests() if r.kind in ("sort_by_key", "copy_where") or r.dtype != np.float32]
    out = {}
    for mode in ("batched", "serial", "per_job"):
        with serve.Server(streams=3, max_batch=32, mode=mode, hbm_budget=8 << 30) as srv:
            ts = [srv.submit(r) for r in reqs]
            out[mode] = [_fetch(srv, t) for t in ts]
    for r, a, b, c in zip(reqs, out["batched"], out["serial"], out["per_job"]):
        assert _same(a, b) and _same(b, c), (r.kind, r.dtype, r.n)


def test_server_budget_admission():
    """Requests larger than the budget are refused; the rest complete one
    budget-full at a time as outputs are released."""
    n = 1 << 20
    with serve.Server(streams=2, max_batch=8, hbm_budget=24 << 20) as srv:
        with pytest.raises(m.Error):
            srv.submit(serve.Request("sort_by_key", np.uint64, 1 << 22, seed=1))
        ts = [srv.submit(serve.Request("inclusive_scan", np.uint32, n, seed=i, op=m.op.Add)) for i in range(12)]
        for i, t in enumerate(ts):
            got = _fetch(srv, t)
            x = _host_input(serve.Request("inclusive_scan", np.uint32, n, seed=i))
            assert got.tobytes() == oracle.inclusive_scan(x, "add").tobytes()


def test_server_large_requests_alone_equals_mix():
    """Requests of 2^21..2^26 elements (the upper half of the served stream's
    size range): the same bytes alone and in a mix with small requests, and
    each checked against the oracle."""
    big = [serve.Request("sort_by_key", np.uint32, 1 << 21, seed=501, op=m.op.Less),
           serve.Request("sort_by_key", np.uint64, 1 << 22, seed=502, op=m.op.Less),
           serve.Request("sort_by_key", np.float32, 1 << 23, seed=503, op=m.op.Less),
           serve.Request("inclusive_scan", np.uint32, 1 << 26, seed=504, op=m.op.Add),
           serve.Request("inclusive_scan", np.float32, 1 << 24, seed=505, op=m.op.Add),
           serve.Request("reduce", np.float32, 1 << 26, seed=506, op=m.op.Add),
           serve.Request("reduce", np.uint64, 1 << 25, seed=507, op=m.op.Add),
           serve.Request("copy_where", np.uint64, 1 << 26, seed=508, pred=m.op.Lt(1 << 63)),
           serve.Request("copy_where", np.float32, 1 << 21, seed=509, pred=m.op.Lt(0.75))]
    small = [r for r in _requests() if r.n <= (1 << 17)][:40]
    with serve.Server(streams=4, max_batch=64, mode="batched", hbm_budget=16 << 30) as srv:
        alone = [_fetch(srv, srv.submit(r)) for r in big]
        mix = big[:]
        for i, r in enumerate(small):
            mix.insert((i * 7) % (len(mix) + 1), r)
        tickets = [srv.submit(r) for r in mix]
        mixed = {id(r): _fetch(srv, t) for r, t in zip(mix, tickets)}
    for r, a in zip(big, alone):
        assert _same(a, mixed[id(r)]), (r.kind, r.dtype, r.n)
        x = _host_input(r)
        if r.kind == "sort_by_key":
            perm = oracle.sort_perm(x, "less")
            assert a[0].tobytes() == x[perm].tobytes()
            assert a[1].tobytes() == perm.astype(np.uint32).tobytes()
        elif r.kind == "copy_where":
            assert a.tobytes() == x[x < r.pred.a].tobytes()
        elif r.kind == "inclusive_scan":
            if x.dtype.kind == "f":
                exact = np.cumsum(x.astype(np.float64))
                mag = np.cumsum(np.abs(x.astype(np.float64)))
                # The device-wide float scan's bound (tests/test_gpu_scan.py).
                lv = np.ceil(np.log2(x.size))
                assert np.all(np.abs(a.astype(np.float64) - exact) <= (lv + 16) * EPS32 * mag)
            else:
                assert a.tobytes() == oracle.inclusive_scan(x, "add").tobytes()
        else:
            if x.dtype.kind == "f":
                s, mag = oracle.sum_f32_f64(x)
                assert abs(float(a) - s) <= oracle.f32_sum_bound(x.size, mag)
            else:
                assert a == oracle.reduce(x, 0, "add")


def test_run_jobs_grid_autotune(exec_, tmp_path):
    """The batched kernels' grid is tuned on the first batch of a (family, job
    count, mean length) class: the choice is recorded, survives a save / clear
    / load of the tuning table, and the results are the bytes of a run with
    autotuning off (each job runs inside one workgroup whatever the grid)."""
    m.tuning_clear()
    n_jobs, n = 300, 3000
    xs = [_data(np.uint32, n) for _ in range(n_jobs)]
    ks = [(_data(np.uint32, n) % 5000).astype(np.uint32) for _ in range(n_jobs)]
    vs = np.arange(n, dtype=np.uint32)
    dx = [exec_.to_device(x) for x in xs]
    dk = [exec_.to_device(k) for k in ks]
    dv = exec_.to_device(vs)

    def batch():
        jobs = [serve.Job("inclusive_scan", d, exec_.alloc(np.uint32, n), op=m.op.Add) for d in dx]
        jobs += [serve.Job("sort_by_key", d, exec_.alloc(np.uint32, n), op=m.op.Less, inp2=dv,
                           out2=exec_.alloc(np.uint32, n)) for d in dk]
        return jobs

    def results(jobs):
        return [exec_.to_host(j.out).tobytes() + (exec_.to_host(j.out2).tobytes() if j.out2 is not None
                                                   else b"") for j in jobs]

    exec_.set_option(m.Executor.OPT_AUTOTUNE, 0)
    try:
        plain = batch()
        serve.run_jobs(exec_, plain)
        want = results(plain)
        assert m.batch_tuned_grid(0, n_jobs, n) == -1
    finally:
        exec_.set_option(m.Executor.OPT_AUTOTUNE, 1)
    tuned = batch()
    serve.run_jobs(exec_, tuned)
    assert results(tuned) == want
    c0, c2 = m.batch_tuned_grid(0, n_jobs, n), m.batch_tuned_grid(2, n_jobs, n)
    assert 0 <= c0 < 4 and 0 <= c2 < 4
    path = tmp_path / "tune.txt"
    m.tuning_save(path)
    m.tuning_clear()
    assert m.batch_tuned_grid(0, n_jobs, n) == -1
    m.tuning_load(path)
    assert (m.batch_tuned_grid(0, n_jobs, n), m.batch_tuned_grid(2, n_jobs, n)) == (c0, c2)
    again = batch()
    serve.run_jobs(exec_, again)
    assert results(again) == want
    for i in (0, 7, n_jobs - 1):
        assert results([again[i]])[0] == oracle.inclusive_scan(xs[i], "add").tobytes()
        perm = oracle.sort_perm(ks[i], "less")
        assert results([again[n_jobs + i]])[0] == vs[perm].tobytes() + ks[i][perm].tobytes()
    # In-place jobs are not rerunnable: their class stays untuned, results exact.
    m.tuning_clear()
    inplace = [serve.Job("inclusive_scan", d, d, op=m.op.Add) for d in dx[:64]]
    serve.run_jobs(exec_, inplace)
    assert m.batch_tuned_grid(0, 64, n) == -1
    for i in (0, 63):
        assert exec_.to_host(dx[i]).tobytes() == oracle.inclusive_scan(xs[i], "add").tobytes()
