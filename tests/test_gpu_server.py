"""Continuous-batching server: every request's result equals the plain path
bit for bit, alone and in the mix, with FP32 and FP64 grids mixed; oversize
requests are rejected; frameworks stay resident and are shared; a launch
that cannot be enqueued fails its requests (and only those); a Poisson
stream of requests from the batching distribution matches the oracle."""
import threading
import time

import numpy as np
import pytest

from emptyroom_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

CELLS = [(10.0, 11.0, 12.0, 85.0, 95.0, 100.0), (9.0, 14.0, 8.5, 70.0, 110.0, 65.0),
         (16.0, 16.0, 16.0, 90.0, 90.0, 90.0)]


def requests(precisions=(api.ER_FP32,)):
    out = []
    for i in range(9):
        f = synth.small_framework(100 + i, n_atoms=20 + 15 * i, cell=CELLS[i % 3])
        kind = i % 3  # 0 grid, 1 Henry, 2 both
        out.append(dict(f=f, grid_n=[9 + i, 10, 8] if kind != 1 else [0, 0, 0],
                        precision=precisions[i % len(precisions)],
                        n_insertions=0 if kind == 0 else 3000 + 500 * i,
                        temperature=250.0 + 10 * i, seed=1000 + i))
    return out


def plain(ctx, r):
    f = r["f"]
    a = f.particles()
    fw = api.Framework(ctx, a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2)
    g = api.compute_grid(ctx, fw, r["grid_n"], r.get("precision", api.ER_FP32)) \
        if r["grid_n"][0] else None
    w = api.widom(ctx, fw, r["n_insertions"], r["temperature"], r["seed"]) \
        if r["n_insertions"] else None
    fw.close()
    return g, w


def submit(srv, r):
    f = r["f"]
    a = f.particles()
    return srv.submit(a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2, grid_n=r["grid_n"],
                      precision=r.get("precision", api.ER_FP32),
                      n_insertions=r["n_insertions"], temperature=r["temperature"],
                      seed=r["seed"])


def check_equal(srv, tk, r, g, w):
    st = srv.poll(tk, wait_ms=-1)
    assert st.state == api.RQ_DONE, st.status
    assert st.t_done >= st.t_admit >= st.t_submit
    if g is not None:
        assert np.array_equal(srv.fetch_grid(tk, g.shape, r.get("precision", api.ER_FP32)), g)
    if w is not None:
        assert st.widom.sum_b == w[0] and st.widom.sum_eb == w[1] and st.widom.n == w[2]


@pytest.mark.parametrize("items_per_launch,streams", [(16, 3), (0, 1)])
def test_server_matches_plain(ctx, items_per_launch, streams):
    print(f"server items_per_launch={items_per_launch} streams={streams}", flush=True)
    rs = requests()
    ref = [plain(ctx, r) for r in rs]
    sctx = api.Context(0)
    srv = api.Server(sctx, streams=streams, items_per_launch=items_per_launch)
    tickets = []
    for r in rs:
        code, tk = submit(srv, r)
        assert code == 0
        tickets.append(tk)
    for r, tk, (g, w) in zip(rs, tickets, ref):
        check_equal(srv, tk, r, g, w)
        srv.release(tk)
    s = srv.stats()
    assert s["submitted"] == len(rs) and s["completed"] == len(rs) and s["rejected"] == 0
    assert s["launches"] >= 1
    srv.close()
    sctx.close()


@pytest.mark.parametrize("items_per_launch", [8, 0])
def test_server_fp64_and_fp32_grids_mixed(ctx, items_per_launch):
    """FP64-grid requests (the drop-in writegrid's default precision) interleaved
    with FP32 grids and Henry runs: each launch computes one pair-math variant,
    picked by the oldest request's pending work, and every result equals the
    plain path bit for bit."""
    print(f"server fp64/fp32 mix items_per_launch={items_per_launch}", flush=True)
    rs = requests(precisions=(api.ER_FP64, api.ER_FP32))
    ref = [plain(ctx, r) for r in rs]
    sctx = api.Context(0)
    srv = api.Server(sctx, streams=3, items_per_launch=items_per_launch)
    tickets = [submit(srv, r) for r in rs]
    assert all(code == 0 for code, _ in tickets)
    for r, (_, tk), (g, w) in zip(rs, tickets, ref):
        check_equal(srv, tk, r, g, w)
        srv.release(tk)
    assert srv.stats()["completed"] == len(rs)
    srv.close()
    sctx.close()


def test_server_rejects_oversize_request():
    sctx = api.Context(0, hbm_cap=64 << 20)
    srv = api.Server(sctx, streams=1)
    f = synth.small_framework(7, n_atoms=10)
    a = f.particles()
    code, tk = srv.submit(a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2, grid_n=[400, 400, 400])
    assert code == 6
    assert srv.poll(tk).state == api.RQ_REJECTED
    code, tk2 = srv.submit(a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2, grid_n=[8, 8, 8])
    assert code == 0 and srv.poll(tk2, wait_ms=-1).state == api.RQ_DONE
    assert srv.stats()["rejected"] == 1
    srv.close()
    sctx.close()


@pytest.mark.parametrize("cache_bytes", [-1, 0])
def test_server_framework_residency(ctx, cache_bytes):
    """Requests naming the same framework share one resident upload; with
    cache_bytes = 0 a framework is evicted as soon as no request holds it.
    Results are the plain path's either way."""
    print(f"server residency cache_bytes={cache_bytes}", flush=True)
    f1 = synth.small_framework(301, n_atoms=60, cell=CELLS[0])
    f2 = synth.small_framework(302, n_atoms=45, cell=CELLS[1])
    rs = []
    for i in range(6):
        f = f1 if i % 3 != 2 else f2
        rs.append(dict(f=f, grid_n=[10, 9, 8] if i % 2 == 0 else [0, 0, 0],
                       n_insertions=2000 + 100 * i, temperature=260.0 + 5 * i, seed=50 + i))
    ref = [plain(ctx, r) for r in rs]
    sctx = api.Context(0)
    srv = api.Server(sctx, streams=2, items_per_launch=64, cache_bytes=cache_bytes)
    for r, (g, w) in zip(rs, ref):  # one at a time: every request after the first of
        code, tk = submit(srv, r)   # its framework finds it resident (or evicted)
        assert code == 0
        check_equal(srv, tk, r, g, w)
        srv.release(tk)
    c = srv.cache_stats()
    if cache_bytes < 0:
        assert c["misses"] == 2 and c["hits"] == 4 and c["evictions"] == 0
        assert c["resident_frameworks"] == 2 and c["resident_bytes"] > 0
    else:
        assert c["hits"] == 0 and c["misses"] == 6 and c["evictions"] == 6
        assert c["resident_frameworks"] == 0 and c["resident_bytes"] == 0
    srv.close()
    sctx.close()


def test_server_failed_launch_fails_only_its_requests(ctx):
    """A launch refused on the host (the fail_launch hook) fails the requests
    holding its items with ER_E_HIP; every other request completes and equals
    the plain path, and the server keeps serving."""
    rs = requests()
    ref = [plain(ctx, r) for r in rs]
    sctx = api.Context(0)
    srv = api.Server(sctx, streams=1, items_per_launch=16, fail_launch=2)
    tickets = []
    for r in rs:
        code, tk = submit(srv, r)
        assert code == 0
        tickets.append(tk)
    failed = 0
    for r, tk, (g, w) in zip(rs, tickets, ref):
        st = srv.poll(tk, wait_ms=-1)
        if st.state == api.RQ_FAILED:
            assert st.status == 5
            failed += 1
        else:
            check_equal(srv, tk, r, g, w)
        srv.release(tk)
    assert 1 <= failed < len(rs)
    code, tk = submit(srv, rs[0])  # still serving
    assert code == 0
    check_equal(srv, tk, rs[0], *ref[0])
    srv.close()
    sctx.close()


def test_server_poisson_stream_against_oracle(ctx):
    """72 requests from the batching distribution (atoms 50-5000, cells 8-40 A,
    angles 60-120 deg, spacings 0.1/0.2/0.5 A, CH4/Xe/Kr), Poisson arrivals at
    200 requests/s, T ~ U[250, 350] K, grid and Henry requests mixed.  Every
    Henry result (20000 insertions) is checked against the oracle at the same
    points (K_H to 1e-4, both exactly 0 allowed); every FP32 grid at 300
    sampled points against the oracle's FP64 energies (1e-3 K where E < 1e4 K,
    1e-6 relative above); every 6th request also alone, bit for bit."""
    n_req, rate = 72, 200.0
    arrivals, kinds, temps = synth.config4_stream(rate, n_req)
    rs = []
    for i in range(n_req):
        f = synth.config3_framework(i % 48, config=4)
        grid = kinds[i] == 0
        rs.append(dict(f=f, grid_n=f.grid_n() if grid else [0, 0, 0],
                       n_insertions=0 if grid else 20000, temperature=float(temps[i]),
                       seed=f.seed + i))
    sctx = api.Context(0)
    srv = api.Server(sctx, streams=3)
    tickets = [None] * n_req

    def feed():
        t0 = time.perf_counter()
        for i, r in enumerate(rs):
            dt = t0 + arrivals[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            tickets[i] = submit(srv, r)

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    rng = np.random.default_rng(11)
    n_grid = n_kh = 0
    for i, (r, (code, tk)) in enumerate(zip(rs, tickets)):
        assert code == 0
        st = srv.poll(tk, wait_ms=-1)
        assert st.state == api.RQ_DONE, (i, st.status)
        f = r["f"]
        a = f.particles()
        print(f"request {i} {f.name}: {len(a)} atoms, grid {r['grid_n']}, "
              f"insertions {r['n_insertions']}", flush=True)
        if r["grid_n"][0]:
            n = r["grid_n"]
            g = srv.fetch_grid(tk, n).astype(np.float64)
            idx = np.stack([rng.integers(0, n[d], 300) for d in range(3)], 1)
            ref = oracle.lj_points(f.t, a, f.rep, f.rc2, 1e-24, idx / (np.array(n) - 1.0))
            got = g[idx[:, 0], idx[:, 1], idx[:, 2]]
            lo = ref < 1e4
            if lo.any():
                assert np.max(np.abs(got[lo] - ref[lo])) <= 1e-3
            hi = ~lo & (ref < 3.0e38)
            if hi.any():
                assert np.max(np.abs(got[hi] - ref[hi]) / ref[hi]) < 1e-6
            n_grid += 1
        else:
            rb, _ = oracle.widom(f.t, a, f.rep, f.rc2, r["temperature"], r["seed"], 0,
                                 r["n_insertions"])
            sb = st.widom.sum_b
            if rb > 0:
                assert abs(sb - rb) / rb <= 1e-4
            else:
                assert sb == 0.0
            n_kh += 1
        if i % 6 == 0:
            g, w = plain(ctx, r)
            check_equal(srv, tk, r, g, w)
        srv.release(tk)
    assert n_grid > 20 and n_kh > 20
    s = srv.stats()
    assert s["completed"] == n_req and s["rejected"] == 0
    c = srv.cache_stats()
    assert c["hits"] + c["misses"] == n_req and c["misses"] <= 48
    srv.close()
    sctx.close()
