"""The request loop (osc_server_*): results are the same alone and in a mix of
concurrent requests (the kernels are deterministic, so bit for bit), equal
to the single-call API, FP32 requests stay within 1e-5 of FP64, requests of
one geometry spread over the workers, the memory budget admits requests one
at a time when it must, and errors are reported per request."""
import numpy as np
import pytest

import openspacecharge_amd as osc
from oracle.oracle import gaussian_bunch

pytestmark = pytest.mark.gpu

GAMMA = 5.0e6 / 0.51099891e6


def _mix():
    """A small mixed stream: free space, cathode images, pipe; several grids."""
    import torch
    dev = torch.device("cuda:0")
    reqs = []
    specs = [("free", (32, 32, 32), 200_000, (1e-3, 1e-3, 1e-4), (0, 0, 0)),
             ("free", (48, 48, 48), 300_000, (0.5e-3, 0.5e-3, 2e-3), (0, 0, 0)),
             ("cathode", (32, 32, 64), 250_000, (1e-3, 1e-3, 0.5e-3), (0, 0, 2.5e-3)),
             ("pipe", (32, 32, 48), 150_000, (1e-3, 0.6e-3, 3e-3), (0, 0, 0)),
             ("free", (64, 64, 64), 400_000, (1e-3, 1e-3, 1e-3), (0, 0, 0)),
             ("free", (32, 32, 32), 100_000, (2e-3, 1e-3, 0.2e-3), (0, 0, 0))]
    for k, (solver, nhi, n, sig, ctr) in enumerate(specs):
        x, y, z = gaussian_bunch(n, sig, 200 + k, center=ctr)
        if solver == "cathode":
            x, y, z = (a[z > 0].copy() for a in (x, y, z))
        t = [torch.from_numpy(a).to(dev) for a in (x, y, z)]
        kw = dict(nhi=nhi, gamma=GAMMA, total_charge=1e-9, solver=solver)
        if solver == "pipe":
            a, b, sz = 10e-3, 6e-3, 3e-3
            kw.update(mesh_min=(-a / 2, -b / 2, -6 * sz), mesh_max=(a / 2, b / 2, 6 * sz), apipe=a,
                      bpipe=b)
        reqs.append((t, kw))
    return reqs


def _run(server, reqs, precision, concurrent):
    import torch
    outs = []
    tickets = []
    for t, kw in reqs:
        n = t[0].shape[0]
        E = torch.empty((n, 3), dtype=torch.float64, device=t[0].device)
        B = torch.empty((n, 3), dtype=torch.float64, device=t[0].device)
        tk = server.submit(*t, E, B, precision=precision, **kw)
        if not concurrent:
            server.wait(tk)
        outs.append((E, B))
        tickets.append(tk)
    if concurrent:
        for tk in tickets:
            server.wait(tk)
    return [(E.cpu().numpy(), B.cpu().numpy()) for E, B in outs]


def test_server_same_alone_and_in_mix():
    reqs = _mix()
    alone = osc.Server(nworkers=1)
    a = _run(alone, reqs, osc.OSC_PREC_FP64, concurrent=False)
    alone.close()
    mix = osc.Server(nworkers=4)
    b = _run(mix, reqs, osc.OSC_PREC_FP64, concurrent=True)
    b2 = _run(mix, reqs, osc.OSC_PREC_FP64, concurrent=True)  # second pass: cached spectra
    mix.close()
    for (Ea, Ba), (Eb, Bb), (Ec, Bc) in zip(a, b, b2):
        assert np.array_equal(Ea, Eb) and np.array_equal(Ba, Bb)
        assert np.array_equal(Ea, Ec) and np.array_equal(Ba, Bc)


def test_server_equals_single_call_api():
    import torch
    reqs = _mix()
    srv = osc.Server(nworkers=2)
    res = _run(srv, reqs, osc.OSC_PREC_FP64, concurrent=True)
    srv.close()
    ctx = osc.Context(0)
    for (t, kw), (E, B) in zip(reqs, res):
        if kw["solver"] == "pipe":
            continue
        m = osc.Mesh3D(nhi=kw["nhi"], gamma=GAMMA)
        Er, Br = osc.space_charge_step(*[a.cpu().numpy() for a in t], m, total_charge=1e-9,
                                       at_cathode=kw["solver"] == "cathode", calc_bfield=True,
                                       ctx=ctx)
        assert np.array_equal(E, Er) and np.array_equal(B, Br)
    ctx.close()


def test_server_fp32_within_1e5():
    reqs = _mix()
    srv = osc.Server(nworkers=4)
    r64 = _run(srv, reqs, osc.OSC_PREC_FP64, concurrent=True)
    r32 = _run(srv, reqs, osc.OSC_PREC_FP32, concurrent=True)
    srv.close()
    for (E64, B64), (E32, B32) in zip(r64, r32):
        for c in range(3):
            err = np.max(np.abs(E32[:, c] - E64[:, c])) / np.max(np.abs(E64[:, c]))
            assert err < 1e-5, err


def test_server_tiny_budget_serialises():
    reqs = _mix()
    srv = osc.Server(nworkers=4, max_bytes=1 << 20)  # every request exceeds it: one at a time
    res = _run(srv, reqs, osc.OSC_PREC_FP64, concurrent=True)
    srv.close()
    ref_srv = osc.Server(nworkers=1)
    ref = _run(ref_srv, reqs, osc.OSC_PREC_FP64, concurrent=False)
    ref_srv.close()
    for (E, B), (Er, Br) in zip(res, ref):
        assert np.array_equal(E, Er)


def test_server_is_work_conserving():
    """Requests of one geometry are not tied to one worker: with 4 workers they
    run concurrently, later ones on workers that served the geometry before
    (their Green spectra cached), and the results stay bit-identical to one
    worker running them in order."""
    import torch
    (t, kw) = _mix()[1]
    kw = dict(kw, mesh_min=(-2e-3, -2e-3, -8e-3), mesh_max=(2e-3, 2e-3, 8e-3))
    reqs = [(t, kw)] * 12
    srv = osc.Server(nworkers=4)
    res = _run(srv, reqs, osc.OSC_PREC_FP64, concurrent=True)
    st = srv.stats()
    srv.close()
    assert st["done"] == 12 and st["max_running"] >= 2, st
    # E x, y, z spectra: at most two builds of each per worker (a spectrum is
    # cached from its second miss on)
    assert st["green_misses"] <= 3 * 2 * 4 and st["green_hits"] >= 3 * (12 - 2 * 4), st
    ref_srv = osc.Server(nworkers=1)
    ref = _run(ref_srv, reqs[:1], osc.OSC_PREC_FP64, concurrent=False)
    ref_srv.close()
    for E, B in res:
        assert np.array_equal(E, ref[0][0]) and np.array_equal(B, ref[0][1])


def test_server_green_admission_and_clear():
    """A geometry's Green spectra are cached from its second miss on (the
    first build goes back to the buffer pool), hit from the third request;
    osc_server_clear_caches drops them and the keys seen, so the next request
    of that geometry builds again.  Results bit-identical throughout."""
    (t, kw) = _mix()[1]
    kw = dict(kw, mesh_min=(-2e-3, -2e-3, -8e-3), mesh_max=(2e-3, 2e-3, 8e-3))
    srv = osc.Server(nworkers=1)
    counts, res = [], []
    for k in range(5):
        if k == 3:
            srv.clear_caches()
        res += _run(srv, [(t, kw)], osc.OSC_PREC_FP64, concurrent=False)
        st = srv.stats()
        counts.append((st["green_misses"], st["green_hits"]))
    srv.close()
    m = counts[0][0]
    assert m > 0 and counts == [(m, 0), (2 * m, 0), (2 * m, m), (3 * m, m), (4 * m, m)], counts
    for E, B in res[1:]:
        assert np.array_equal(E, res[0][0]) and np.array_equal(B, res[0][1])


def _new_shape_allocs(admit):
    """Device allocations of a 1-worker server after each of eight FP32
    requests on a 48^3 mesh that follows its bunch, the bunch length changing
    every time (a new Green key per request)."""
    import os
    import torch
    dev = torch.device("cuda:0")
    old = os.environ.get("OSC_GREEN_ADMIT")
    os.environ["OSC_GREEN_ADMIT"] = "1" if admit else "0"
    try:
        srv = osc.Server(nworkers=1)  # contexts read the switch at creation
    finally:
        if old is None:
            del os.environ["OSC_GREEN_ADMIT"]
        else:
            os.environ["OSC_GREEN_ADMIT"] = old
    allocs, res = [], []
    for k in range(8):
        x, y, z = gaussian_bunch(100_000, (1e-3, 1e-3, 1e-3 * (0.5 + 0.25 * k)), 300 + k)
        t = [torch.from_numpy(a).to(dev) for a in (x, y, z)]
        kw = dict(nhi=(48, 48, 48), gamma=GAMMA, total_charge=1e-9, solver="free")
        res += _run(srv, [(t, kw)], osc.OSC_PREC_FP32, concurrent=False)
        allocs.append(srv.stats()["allocs"])
    st = srv.stats()
    srv.close()
    assert st["green_hits"] == 0 and st["green_misses"] == 3 * 8, st
    return allocs, res


def test_server_new_bunch_shapes_allocate_nothing():
    """Spectra seen once are built into pooled buffers: once the first
    requests have filled the worker's pool, a request whose spacing is new
    allocates no device memory, where keeping every spectrum allocates for
    each one.  Results equal either way."""
    allocs, res = _new_shape_allocs(True)
    keep, res_keep = _new_shape_allocs(False)
    print("admit", allocs, "keep all", keep)
    assert allocs[-1] == allocs[2], allocs
    assert keep[-1] - keep[2] >= 3 * 5, keep
    for (E, B), (Ek, Bk) in zip(res, res_keep):
        assert np.array_equal(E, Ek) and np.array_equal(B, Bk)


def test_server_reports_request_errors():
    import torch
    (t, kw) = _mix()[3]  # the pipe request
    kw = dict(kw)
    kw.pop("mesh_min")
    kw.pop("mesh_max")
    srv = osc.Server(nworkers=1)
    E = torch.empty((t[0].shape[0], 3), dtype=torch.float64, device=t[0].device)
    tk = srv.submit(*t, E, **kw)  # pipe without a fixed mesh
    with pytest.raises(osc.OSCError):
        srv.wait(tk)
    # the server keeps serving after a failed request
    (t2, kw2) = _mix()[0]
    E2 = torch.empty((t2[0].shape[0], 3), dtype=torch.float64, device=t2[0].device)
    srv.wait(srv.submit(*t2, E2, **kw2))
    srv.close()


def test_server_stats_report_pools():
    """The statistics report the workers' device allocations (a running server
    reuses its pooled buffers: a second pass over the same requests allocates
    far less than the first) and the pool trims (a budget well below what the
    workers hold makes them give buffers back; results stay bit-identical)."""
    reqs = _mix()
    srv = osc.Server(nworkers=2)
    a = _run(srv, reqs, osc.OSC_PREC_FP64, concurrent=True)
    s1 = srv.stats()
    b = _run(srv, reqs, osc.OSC_PREC_FP64, concurrent=True)
    s2 = srv.stats()
    srv.close()
    assert s1["allocs"] > 0 and s1["alloc_us"] >= 0
    assert s2["allocs"] - s1["allocs"] < s1["allocs"], (s1, s2)
    for (Ea, Ba), (Eb, Bb) in zip(a, b):
        assert np.array_equal(Ea, Eb) and np.array_equal(Ba, Bb)
    small = osc.Server(nworkers=2, max_bytes=256 << 20)
    c = _run(small, reqs, osc.OSC_PREC_FP64, concurrent=True)
    st = small.stats()
    small.close()
    assert st["trims"] > 0 and st["trimmed_bytes"] > 0, st
    for (Ea, Ba), (Ec, Bc) in zip(a, c):
        assert np.array_equal(Ea, Ec) and np.array_equal(Ba, Bc)
