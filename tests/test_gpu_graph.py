"""The batched finite-source run captured in a HIP graph: every replay, with
eager runs and host synchronizations between replays, gives bitwise the eager
result for the same curves, also after the curves change between replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _subset(n_curves, seed):
    from lcbinint_amd.workloads import finite_source_batch
    curves, offsets, times, nbin = finite_source_batch()
    rng = np.random.default_rng(seed)
    pick = np.sort(rng.choice(len(curves), n_curves, replace=False))
    cs = [curves[i] for i in pick]
    ts = [times[offsets[i]:offsets[i + 1]] for i in pick]
    off = np.concatenate([[0], np.cumsum([len(t) for t in ts])]).astype(np.int64)
    return cs, off, np.concatenate(ts), nbin[pick]


def test_graph_replays_match_eager(gpu):
    import torch
    from lcbinint_amd import capi
    cs, off, t, nbin = _subset(96, 31)
    cs2 = [dict(c, u0=c["u0"] * 0.7, rho=c["rho"] * 1.3) for c in cs]
    recs = [capi.Curve(**c) for c in cs]
    recs2 = [capi.Curve(**c) for c in cs2]
    n = len(t)
    stream = torch.cuda.current_stream().cuda_stream
    d_off = torch.from_numpy(off).cuda()
    d_t = torch.from_numpy(t).cuda()

    def eager(rs):
        b = capi.Batch(rs)
        b.set_source_bins(nbin)
        m = torch.empty(n, dtype=torch.float64, device="cuda")
        k = torch.empty(n, dtype=torch.int32, device="cuda")
        capi.finite_source(b, d_off.data_ptr(), d_t.data_ptr(), n, m.data_ptr(), k.data_ptr(), stream)
        torch.cuda.synchronize()
        out = m.cpu().numpy(), k.cpu().numpy()
        b.close()
        return out

    ref1 = eager(recs)
    ref2 = eager(recs2)
    assert not np.array_equal(ref1[0], ref2[0])
    g = capi.Batch(recs)
    g.set_source_bins(nbin)
    gm = torch.empty(n, dtype=torch.float64, device="cuda")
    gk = torch.empty(n, dtype=torch.int32, device="cuda")
    print(f"capturing a {n}-epoch finite-source run", flush=True)
    capi.capture_finite_source(g, d_off.data_ptr(), d_t.data_ptr(), n, gm.data_ptr(), gk.data_ptr(), stream)
    for rep, (rs, ref) in enumerate([(recs, ref1), (recs, ref1), (recs2, ref2), (recs, ref1)]):
        g.update_curves(rs, stream)
        gm.fill_(-1.0)
        capi.launch_graph(g, stream)
        torch.cuda.synchronize()
        print(f"replay {rep}", flush=True)
        assert np.array_equal(gk.cpu().numpy(), ref[1]), rep
        assert np.array_equal(gm.cpu().numpy().view(np.int64), ref[0].view(np.int64)), rep
        # eager work of the same batch between replays
        em = torch.empty(n, dtype=torch.float64, device="cuda")
        ek = torch.empty(n, dtype=torch.int32, device="cuda")
        capi.finite_source(g, d_off.data_ptr(), d_t.data_ptr(), n, em.data_ptr(), ek.data_ptr(), stream)
        torch.cuda.synchronize()
        assert np.array_equal(em.cpu().numpy().view(np.int64), ref[0].view(np.int64)), rep
    g.close()


def test_autotuner_keeps_results(gpu):
    # the schedule changes which waves run which epochs, never the values
    import torch
    from lcbinint_amd import capi
    cs, off, t, nbin = _subset(128, 32)
    n = len(t)
    stream = torch.cuda.current_stream().cuda_stream
    d_off = torch.from_numpy(off).cuda()
    d_t = torch.from_numpy(t).cuda()
    b = capi.Batch([capi.Curve(**c) for c in cs])
    b.set_source_bins(nbin)
    m0 = torch.empty(n, dtype=torch.float64, device="cuda")
    k0 = torch.empty(n, dtype=torch.int32, device="cuda")
    capi.finite_source(b, d_off.data_ptr(), d_t.data_ptr(), n, m0.data_ptr(), k0.data_ptr(), stream)
    torch.cuda.synchronize()
    m1 = torch.empty_like(m0)
    k1 = torch.empty_like(k0)
    print("autotuning", flush=True)
    best, ms = capi.autotune(b, d_off.data_ptr(), d_t.data_ptr(), n, m1.data_ptr(), k1.data_ptr(), stream,
                             max_runs=10)
    torch.cuda.synchronize()
    print("schedule", best.as_dict(), "run ms", ms, flush=True)
    assert ms > 0.0
    # the tuned schedule stays on the batch; its results equal the default's
    capi.finite_source(b, d_off.data_ptr(), d_t.data_ptr(), n, m1.data_ptr(), k1.data_ptr(), stream)
    torch.cuda.synchronize()
    assert torch.equal(k0, k1)
    assert np.array_equal(m0.cpu().numpy().view(np.int64), m1.cpu().numpy().view(np.int64))
    # an extreme schedule: one wave per CU everywhere, the farm taking every scan
    capi.set_schedule(b, capi.Schedule(seed_waves_per_cu=1, cart_waves_per_cu=1, polar_waves_per_cu=1,
                                       heavy_budget=1, lead_items=1))
    capi.finite_source(b, d_off.data_ptr(), d_t.data_ptr(), n, m1.data_ptr(), k1.data_ptr(), stream)
    torch.cuda.synchronize()
    assert torch.equal(k0, k1)
    assert np.array_equal(m0.cpu().numpy().view(np.int64), m1.cpu().numpy().view(np.int64))
    b.close()
