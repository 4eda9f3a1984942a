"""Continuous-batching server: every request's result equals the plain path
bit for bit, alone and in the mix; oversize requests are rejected."""
import numpy as np
import pytest

from emptyroom_amd import api, synth

pytestmark = pytest.mark.gpu

CELLS = [(10.0, 11.0, 12.0, 85.0, 95.0, 100.0), (9.0, 14.0, 8.5, 70.0, 110.0, 65.0),
         (16.0, 16.0, 16.0, 90.0, 90.0, 90.0)]


def requests():
    out = []
    for i in range(9):
        f = synth.small_framework(100 + i, n_atoms=20 + 15 * i, cell=CELLS[i % 3])
        kind = i % 3  # 0 grid, 1 Henry, 2 both
        out.append(dict(f=f, grid_n=[9 + i, 10, 8] if kind != 1 else [0, 0, 0],
                        n_insertions=0 if kind == 0 else 3000 + 500 * i,
                        temperature=250.0 + 10 * i, seed=1000 + i))
    return out


def plain(ctx, r):
    f = r["f"]
    a = f.particles()
    fw = api.Framework(ctx, a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2)
    g = api.compute_grid(ctx, fw, r["grid_n"], api.ER_FP32) if r["grid_n"][0] else None
    w = api.widom(ctx, fw, r["n_insertions"], r["temperature"], r["seed"]) \
        if r["n_insertions"] else None
    fw.close()
    return g, w


@pytest.mark.parametrize("items_per_launch,streams", [(16, 3), (0, 1)])
def test_server_matches_plain(ctx, items_per_launch, streams):
    print(f"server items_per_launch={items_per_launch} streams={streams}", flush=True)
    rs = requests()
    ref = [plain(ctx, r) for r in rs]
    sctx = api.Context(0)
    srv = api.Server(sctx, streams=streams, items_per_launch=items_per_launch)
    tickets = []
    for r in rs:
        a = r["f"].particles()
        code, tk = srv.submit(a[:, :3], a[:, 3], a[:, 4], r["f"].t, r["f"].rep, r["f"].rc2,
                              grid_n=r["grid_n"], n_insertions=r["n_insertions"],
                              temperature=r["temperature"], seed=r["seed"])
        assert code == 0
        tickets.append(tk)
    for r, tk, (g, w) in zip(rs, tickets, ref):
        st = srv.poll(tk, wait_ms=-1)
        assert st.state == api.RQ_DONE, st.status
        assert st.t_done >= st.t_admit >= st.t_submit
        if g is not None:
            assert np.array_equal(srv.fetch_grid(tk, g.shape), g)
        if w is not None:
            assert st.widom.sum_b == w[0] and st.widom.sum_eb == w[1] and st.widom.n == w[2]
        srv.release(tk)
    s = srv.stats()
    assert s["submitted"] == len(rs) and s["completed"] == len(rs) and s["rejected"] == 0
    assert s["launches"] >= 1
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
