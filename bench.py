#!/usr/bin/env python3
"""Headline benchmark: OpenSC free-space IGF space-charge solve on one MI355X.

Workload (the 128^3 configuration of the metric): a 128x128x128 mesh (FFT on
the 256^3 doubled grid), 10,000,000 macroparticles of a seeded 3-D Gaussian
(sigma = 0.5, 0.5, 2 mm, 4-sigma cut), 1 nC, gamma = 9.78476 (5 MeV).
One step = deposit_particles (auto mesh) -> space_charge_3d (E and B) ->
interpolate_field of E and B at all 10M particles.  Particles, meshes and
fields are resident in HBM before the timed region; the Green function is
rebuilt in every step, as the upstream solver does (no cached spectra).

Each precision runs the step K times back to back between two device
synchronisations (the solves/s figures), then K times each waited for (p50,
p99 latency), then K times with per-stage HIP-event timing (stage times and
the roofline kernel's launch duration).

Prints ONE JSON line.  `value` is FP64 drop-in solves/s; the same line
carries the FP32-convolution rate, its field error against the FP64 result,
the per-stage times, the roofline of the dominant kernel and the CPU baseline
(the upstream Fortran library compiled with amdflang, run on this host).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

# the request loop runs one HIP stream per worker and leaves HIP's default of
# 4 hardware queues per process: with 8 workers on the batch stream 16 queues
# measured 3882 against 4283 FP32 requests/s (the served mix 805 against 798)

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "mixed-size solves/s + p99 latency; 128³ solve ms, FP32 vs FP64 @1e-5 rel err"
HBM_PEAK_GBS = 8000.0
NX = NY = NZ = 128
NPART = 10_000_000
SIGMA = (0.5e-3, 0.5e-3, 2e-3)
GAMMA = 5.0e6 / 0.510998910e6  # test_opensc.f90:64,66,102
QTOT = 1e-9


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--particles", type=int, default=NPART)
    p.add_argument("--grid", type=int, default=NX)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--host-steps", type=int, default=3,
                   help="steps of the three drop-in calls on host arrays (PCIe included); 0 skips")
    p.add_argument("--cpu-timeout", type=float, default=240.0)
    p.add_argument("--cpu-child", nargs=2, metavar=("IN", "OUT"), help=argparse.SUPPRESS)
    p.add_argument("--calibrate-copy", action="store_true",
                   help="only run a 1 GiB device copy (profiler counter calibration)")
    p.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    p.add_argument("--no-traffic", action="store_true",
                   help="skip the rocprofv3 counter passes (roofline traffic = null)")
    p.add_argument("--traffic-timeout", type=float, default=120.0)
    p.add_argument("--batch-bunches", type=int, default=64,
                   help="bunches of the batched mixed-grid leg (0 skips it)")
    p.add_argument("--served-requests", type=int, default=64,
                   help="requests of the served mixed stream (0 skips it)")
    p.add_argument("--batch-workers", type=int, default=8,
                   help="request-loop workers of the batch stream's served runs (meshes up to "
                        "128^3: 8 workers 3.12x, 4 workers 2.94-2.96x)")
    p.add_argument("--served-workers", type=int, default=4,
                   help="request-loop workers of the served mix (meshes up to 256^3, 2e7 "
                        "particles: with 8 the pools' device allocations stalled for tens of ms "
                        "each)")
    p.add_argument("--stream-reps", type=int, default=3,
                   help="timed runs of each stream configuration on one warmed server, "
                        "its Green caches cleared before each; the median wall is reported")
    p.add_argument("--poisson-requests", type=int, default=1000,
                   help="requests of each Poisson-arrival run of the served mix (three loads)")
    return p.parse_args()


def bunch(n, seed=20261202):
    from oracle.oracle import gaussian_bunch  # generator only (no oracle compute)
    return gaussian_bunch(n, SIGMA, seed)


# algorithmic bytes per launch of each stage (compulsory traffic model)
def stage_bytes(stage, s, n3, L, P, nout, nspec):
    nx, ny, nz = n3
    n = nx * ny * nz
    c = 2 * s
    KX = L[0] // 2 + 1
    half_rows_in = KX * ny * nz
    if stage == "fft.fwd_x":
        return 8 * n + c * half_rows_in
    if stage == "fft.fwd_y":
        return c * KX * ny * nz + c * KX * L[1] * nz
    if stage == "fft.fwd_z":
        return c * KX * L[1] * nz + c * KX * L[1] * L[2]
    if stage == "fft.inv_zmul":
        # fused z convolution: rho_hat's nz stored planes once, the z-real
        # Green arrays (of the three E components two are stored as halves
        # along z, one whole), each output's nz planes
        return c * KX * L[1] * nz * (1 + nout) + c * KX * L[1] * 2 * L[2] * nspec // 3
    if stage == "fft.inv_y":
        return nout * (c * KX * L[1] * nz + c * KX * ny * nz)
    if stage == "fft.inv_x":
        return nout * c * KX * ny * nz + 8 * n * 5
    if stage == "deposit.cells":
        # B_deposit: the particle coordinates once, the charge grid once; the
        # kernel of this stage is the one that sums the particles' terms (from
        # their records binned by tile, or in cell order)
        return 24 * P + s * n
    if stage == "gather":
        # compulsory part of the field gather (the return to particle order
        # is a separate pass)
        return 24 * P + 48 * P + 48 * n
    if stage == "gather.out":
        # one 64-byte record per particle in, E and B (48 B) and the flag out
        return 64 * P + 48 * P + P
    return None


def fft_tune_measuring():
    """FFT pass kinds whose tile width the tuner is still measuring."""
    import ctypes as C
    from openspacecharge_amd import _capi as K
    st = (C.c_int64 * 2)()
    K.load().osc_fft_tune_stats(st)
    return int(st[1])


def gpu_leg(args, x, y, z):
    import torch
    import openspacecharge_amd as osc

    dev = torch.device("cuda:0")
    xs, ys, zs = (torch.from_numpy(a).to(dev) for a in (x, y, z))
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream().cuda_stream
    res = {}
    fields = {}
    for prec_name, prec in (("fp64", osc.OSC_PREC_FP64), ("fp32", osc.OSC_PREC_FP32)):
        ctx = osc.Context(0, precision=prec)
        ctx.set_stream(stream)
        ctx.set_async(True)
        ctx.set_green_cache(False)  # rebuild the Green function every step, as upstream
        mesh = osc.Mesh3D(nhi=(args.grid,) * 3, gamma=GAMMA, device="cuda")

        def step():
            # deposit -> solve (E, B) -> gather at all particles, one call
            return osc.space_charge_step(xs, ys, zs, mesh, total_charge=QTOT, calc_bfield=True,
                                         ctx=ctx)

        def step3():
            # the same step as three drop-in calls (upstream's deposit_particles
            # returns no dropped count, so none is asked for)
            osc.deposit_particles(xs, ys, zs, mesh, total_charge=QTOT, ctx=ctx, count_dropped=False)
            osc.space_charge_3d(mesh, calc_bfield=True, ctx=ctx)
            return osc.interpolate_field(xs, ys, zs, mesh, E=True, B=True, ctx=ctx)

        # first use of each FFT pass: the tile-width tuner measures its
        # candidates on real steps (osc_fft_tuning); settle it before timing
        for _ in range(24):
            step()
            step3()
            torch.cuda.synchronize()
            if fft_tune_measuring() == 0:
                break
        for _ in range(args.warmup):
            step3()
        torch.cuda.synchronize()
        t30 = time.perf_counter()
        for _ in range(args.steps):
            step3()
        torch.cuda.synchronize()
        t31 = time.perf_counter()
        for _ in range(args.warmup):
            step()
        # throughput (the headline): K steps back to back
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        # latency: each step waited for
        lat = []
        for _ in range(args.steps):
            ts = time.perf_counter()
            step()
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - ts)
        # per-stage HIP-event timing (events recorded on the context's stream)
        ctx.profile_read()
        ctx.set_profiling(True)
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        prof = ctx.profile_read()
        ctx.set_profiling(False)
        # cached-Green variant (repeated mesh geometry), untimed by the headline
        ctx.set_green_cache(True)
        step()
        torch.cuda.synchronize()
        tc0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        tc1 = time.perf_counter()
        host = None
        if prec_name == "fp64" and args.host_steps > 0:
            # the three drop-in calls on host (numpy) arrays, as a Fortran
            # caller's allocatables arrive: x, y, z cross PCIe in the deposit
            # and again in the gather; rho, the fields and E, B come back
            hctx = osc.Context(0, precision=prec)
            hctx.set_green_cache(False)
            hmesh = osc.Mesh3D(nhi=(args.grid,) * 3, gamma=GAMMA)

            def step_host():
                osc.deposit_particles(x, y, z, hmesh, total_charge=QTOT, ctx=hctx)
                osc.space_charge_3d(hmesh, calc_bfield=True, ctx=hctx)
                return osc.interpolate_field(x, y, z, hmesh, E=True, B=True, ctx=hctx)

            step_host()
            th0 = time.perf_counter()
            for _ in range(args.host_steps):
                step_host()
            th1 = time.perf_counter()
            hctx.close()
            host = dict(solves_per_s=args.host_steps / (th1 - th0),
                        ms_per_step=1e3 * (th1 - th0) / args.host_steps, steps=args.host_steps)
            log(f"host arrays: {host['ms_per_step']:.1f} ms/step")
        fields[prec_name] = (mesh.efield.cpu().numpy().copy(order="F"), mesh.rho.cpu().numpy())
        log(f"{prec_name}: {1e3 * (t1 - t0) / args.steps:.3f} ms/step")
        res[prec_name] = dict(
            solves_per_s=args.steps / (t1 - t0), ms_per_step=1e3 * (t1 - t0) / args.steps,
            p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)),
            cached_green_solves_per_s=args.steps / (tc1 - tc0),
            three_call_solves_per_s=args.steps / (t31 - t30), host_arrays=host,
            stages_ms={k: round(v[0] / args.steps, 4) for k, v in prof.items()},
            stage_launches={k: v[1] // args.steps for k, v in prof.items()},
            prof=prof, mesh=dict(min=list(mesh.min), max=list(mesh.max), delta=list(mesh.delta)))
        ctx.close()
    e64, e32 = fields["fp64"][0], fields["fp32"][0]
    errs = []
    for cmp in range(3):
        d = e32[..., cmp] - e64[..., cmp]
        m = np.max(np.abs(e64[..., cmp]))
        errs.append((float(np.max(np.abs(d)) / m), float(np.sqrt(np.mean(d * d)) / m)))
    res["fp32_vs_fp64"] = dict(max_rel=max(e[0] for e in errs), rms_rel=max(e[1] for e in errs))
    return res, fields


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# mixed-size streams: batched free-space bunches and a served request mix
# ---------------------------------------------------------------------------
EDGES = sorted(e for e in {2 ** a * 3 ** b for a in range(9) for b in range(6)} if 32 <= e <= 256)
PIPE_A, PIPE_B = 10e-3, 6e-3


def gpu_bunch(n, sigma, seed, center=(0.0, 0.0, 0.0), zpos=False):
    """Seeded truncated (4 sigma) Gaussian generated on the GPU (torch Philox)."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(int(seed))
    out = []
    have = 0
    while have < n:
        r = torch.randn((int((n - have) * 1.2) + 64, 3), generator=g, device="cuda",
                        dtype=torch.float64)
        r = r[(r * r).sum(1) <= 16.0]
        r = r * torch.tensor(sigma, dtype=torch.float64, device="cuda") + torch.tensor(
            center, dtype=torch.float64, device="cuda")
        if zpos:
            r = r[r[:, 2] > 0]
        out.append(r)
        have += r.shape[0]
    r = torch.cat(out)[:n]
    return r[:, 0].contiguous(), r[:, 1].contiguous(), r[:, 2].contiguous()


def make_stream(kind, count, seed):
    """Request list.  kind "batch": free-space bunches with grids from
    {32^3, 48^3, 64^3, 96x96x128, 128^3}, 1e5-4e6 particles (log-uniform),
    sigma_z/sigma_x log-uniform in [0.05, 20].  kind "served": 70/20/10 %
    free/cathode/pipe, grid edge log-uniform over 32..256 snapped to 2^a 3^b,
    1e5-2e7 particles (log-uniform), varied sigma."""
    rng = np.random.Generator(np.random.Philox(seed))
    reqs = []
    for k in range(count):
        if kind == "batch":
            solver = "free"
            nhi = [(32,) * 3, (48,) * 3, (64,) * 3, (96, 96, 128), (128,) * 3][rng.integers(5)]
            n = int(np.exp(rng.uniform(np.log(1e5), np.log(4e6))))
        else:
            u = rng.random()
            solver = "free" if u < 0.7 else ("cathode" if u < 0.9 else "pipe")
            e = np.exp(rng.uniform(np.log(32), np.log(256)))
            edge = min(EDGES, key=lambda v: abs(np.log(v / e)))
            nhi = (edge,) * 3
            n = int(np.exp(rng.uniform(np.log(1e5), np.log(2e7))))
        ratio = float(np.exp(rng.uniform(np.log(0.05), np.log(20))))
        sx = 1e-3
        sig = (sx, sx, sx * ratio)
        kw = dict(nhi=nhi, gamma=GAMMA, total_charge=QTOT, solver=solver)
        center = (0.0, 0.0, 0.0)
        if solver == "cathode":
            center = (0.0, 0.0, 1.5 * sig[2] + 0.5e-3)
        if solver == "pipe":
            sig = (1e-3, 0.6e-3, min(sig[2], 3e-3))
            kw.update(mesh_min=(-PIPE_A / 2, -PIPE_B / 2, -6 * sig[2]),
                      mesh_max=(PIPE_A / 2, PIPE_B / 2, 6 * sig[2]), apipe=PIPE_A, bpipe=PIPE_B)
        reqs.append(dict(n=n, sigma=sig, center=center, seed=seed * 1000 + k,
                         zpos=solver == "cathode", kw=kw))
    return reqs


def run_stream(reqs, parts, nworkers, precision, green_cache, sequential, warm, reps=1):
    """Times the requests `reps` times on one Server, after an untimed run of
    `warm` = (requests, bunches) of another seed, so that the timed runs find
    the device buffer pools of a running server (its meshes differ, so it
    leaves no Green spectra the timed stream could reuse).  Before each timed
    run the server's Green caches are cleared (their buffers stay pooled: no
    device memory is freed), so every run builds every spectrum, as a stream
    of new bunches does.  sequential: each request waits for the previous one
    (plain per-request solving).  Returns the median wall (s), every wall, and
    the last run's per-request records, outputs and server counters."""
    import openspacecharge_amd as osc
    srv = osc.Server(nworkers=nworkers, green_cache=green_cache)
    _drive(srv, warm[0], warm[1], precision, sequential)
    walls, recs, outs = [], None, None
    for _ in range(max(1, reps)):
        srv.clear_caches()
        outs = None
        st0 = srv.stats()
        wall, recs, outs = _drive(srv, reqs, parts, precision, sequential)
        st1 = srv.stats()
        walls.append(wall)
    srv.close()
    st = {k: st1[k] - st0[k] for k in ("done", "affine_starts", "bypasses", "green_hits",
                                       "green_misses", "trims", "allocs", "alloc_us")}
    st["max_running"] = st1["max_running"]
    return sorted(walls)[len(walls) // 2], walls, recs, outs, st


def long_stream(srv, reqs, parts, nreq, rate, precision, seed=5):
    """nreq requests on the running server srv with Poisson arrivals at
    `rate`/s: request k is request k % len(reqs) with its bunch stretched by
    its own factors in [0.9, 1.1] per axis (z only, and the fixed mesh with
    it, for the pipe), so that every mesh differs and no Green spectrum is
    reused.  Requests are waited for in submission order by a second thread,
    which drops them once done.  Returns (wall s, latencies ms, the server's
    counters over the run)."""
    import threading
    import queue as qmod
    import torch
    st0 = srv.stats()
    rng = np.random.Generator(np.random.Philox(seed))
    pend = qmod.Queue()
    lat = []

    def reaper():
        while True:
            item = pend.get()
            if item is None:
                return
            lat.append(srv.wait(item[0])["total_ms"])

    th = threading.Thread(target=reaper)
    th.start()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t_next = t0
    for k in range(nreq):
        r = reqs[k % len(reqs)]
        x, y, z = parts[k % len(reqs)]
        f = rng.uniform(0.9, 1.1, 3)
        kw = dict(r["kw"])
        if kw["solver"] == "pipe":
            f[0] = f[1] = 1.0
            kw["mesh_min"] = tuple(kw["mesh_min"][:2]) + (kw["mesh_min"][2] * f[2],)
            kw["mesh_max"] = tuple(kw["mesh_max"][:2]) + (kw["mesh_max"][2] * f[2],)
        xs, ys, zs = x * float(f[0]), y * float(f[1]), z * float(f[2])
        E = torch.empty((x.shape[0], 3), dtype=torch.float64, device="cuda")
        torch.cuda.current_stream().synchronize()  # the stretched copies exist before submit
        t_next += rng.exponential(1.0 / rate)
        while time.perf_counter() < t_next:
            time.sleep(max(0.0, min(1e-3, t_next - time.perf_counter())))
        tk = srv.submit(xs, ys, zs, E, precision=precision, **kw)
        pend.put((tk, xs, ys, zs, E))
    pend.put(None)
    th.join()
    wall = time.perf_counter() - t0
    st1 = srv.stats()
    st = {k: st1[k] - st0[k] for k in ("done", "bypasses", "green_hits", "trims", "allocs",
                                       "alloc_us")}
    st["max_running"] = st1["max_running"]
    return wall, lat, st


def _drive(srv, reqs, parts, precision, sequential, rate=None, seed=1):
    import torch
    outs, tickets, recs = [], [], []
    rng = np.random.Generator(np.random.Philox(seed))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t_next = t0
    for r, (x, y, z) in zip(reqs, parts):
        if rate:
            t_next += rng.exponential(1.0 / rate)
            while time.perf_counter() < t_next:
                time.sleep(max(0.0, min(1e-3, t_next - time.perf_counter())))
        E = torch.empty((x.shape[0], 3), dtype=torch.float64, device="cuda")
        tk = srv.submit(x, y, z, E, precision=precision, **r["kw"])
        if sequential:
            recs.append(srv.wait(tk))
        else:
            tickets.append(tk)
        outs.append(E)
    for tk in tickets:
        recs.append(srv.wait(tk))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, recs, outs


def stream_legs(args):
    import torch
    import openspacecharge_amd as osc
    res = {}
    for kind, count in (("batch", args.batch_bunches), ("served", args.served_requests)):
        if count <= 0:
            continue
        reqs = make_stream(kind, count, seed=7 if kind == "batch" else 11)
        parts = [gpu_bunch(r["n"], r["sigma"], r["seed"], r["center"], r["zpos"]) for r in reqs]
        wreqs = make_stream(kind, max(8, count // 2), seed=101 if kind == "batch" else 111)
        wparts = [gpu_bunch(r["n"], r["sigma"], r["seed"], r["center"], r["zpos"]) for r in wreqs]
        warm = (wreqs, wparts)
        torch.cuda.synchronize()
        reps = args.stream_reps
        nw = args.batch_workers if kind == "batch" else args.served_workers
        # plain FP64 per request: one at a time, the Green function rebuilt for every request
        w64, l64, _, o64, _ = run_stream(reqs, parts, 1, osc.OSC_PREC_FP64, False, True, warm,
                                         reps)
        # served: concurrent workers, Green spectra kept resident
        ws64, ls64, _, _, st64 = run_stream(reqs, parts, nw, osc.OSC_PREC_FP64, True, False, warm,
                                            reps)
        ws32, ls32, recs32, o32, st32 = run_stream(reqs, parts, nw, osc.OSC_PREC_FP32, True, False,
                                                   warm, reps)
        err = 0.0
        for a, b in zip(o32, o64):
            a, b = a.cpu().numpy(), b.cpu().numpy()
            err = max(err, max(float(np.max(np.abs(a[:, c] - b[:, c])) / np.max(np.abs(b[:, c])))
                               for c in range(3)))
        out = {"requests": count, "particles_total": int(sum(r["n"] for r in reqs)),
               "fp64_per_request_solves_per_s": round(count / w64, 2),
               "fp64_served_solves_per_s": round(count / ws64, 2),
               "fp32_served_solves_per_s": round(count / ws32, 2),
               "fp32_served_vs_fp64_per_request": round(w64 / ws32, 3),
               "target_3x": "pass" if w64 / ws32 >= 3.0 else "fail",
               "max_rel_err_fp32_vs_fp64": err,
               "repetitions": max(1, reps), "workers": nw,
               "walls_ms": {"fp64_per_request": [round(w * 1e3, 2) for w in l64],
                            "fp64_served": [round(w * 1e3, 2) for w in ls64],
                            "fp32_served": [round(w * 1e3, 2) for w in ls32]},
               "server_fp64": st64, "server_fp32": st32}
        if kind == "served":
            out["mix"] = "70% free / 20% cathode / 10% pipe, grid edge 32..256 (2^a 3^b), 1e5-2e7 particles"
            cap = count / ws32
            poisson = []
            if args.poisson_requests > 0:
                # one running server for the three loads, as a service would be:
                # warmed by the request set and an untimed Poisson run, so that
                # its buffer pools hold the size classes of the stream
                psrv = osc.Server(nworkers=nw, green_cache=True)
                _drive(psrv, reqs, parts, osc.OSC_PREC_FP32, False)
                long_stream(psrv, reqs, parts, args.poisson_requests // 2, 0.5 * cap,
                            osc.OSC_PREC_FP32, seed=4)
                for k, load in enumerate((0.5, 0.7, 0.9)):
                    wl, lat, stl = long_stream(psrv, reqs, parts, args.poisson_requests,
                                               load * cap, osc.OSC_PREC_FP32, seed=5 + k)
                    poisson.append({"load": load, "rate_per_s": round(load * cap, 2),
                                    "requests": len(lat), "wall_s": round(wl, 2),
                                    "solves_per_s": round(len(lat) / wl, 2),
                                    "p50_ms": round(float(np.percentile(lat, 50)), 3),
                                    "p99_ms": round(float(np.percentile(lat, 99)), 3),
                                    "max_running": stl["max_running"],
                                    "bypasses": stl["bypasses"],
                                    "green_hits": stl["green_hits"],
                                    "device_allocs": stl["allocs"],
                                    "device_alloc_ms": round(stl["alloc_us"] / 1e3, 1)})
                    log(f"poisson {load}: {poisson[-1]}")
                psrv.close()
            out["fp32_poisson"] = poisson
        else:
            out["mix"] = "free space, grids {32^3,48^3,64^3,96x96x128,128^3}, 1e5-4e6 particles"
        res[kind] = out
        log(f"{kind}: {out}")
        del parts, wparts, warm, o64, o32
        torch.cuda.empty_cache()
    return res


def cpu_leg(args, x, y, z):
    """Runs cpu_child() in a child process started before this process touches
    the GPU, with the stack limit raised (upstream's FFT transposes keep large
    automatic arrays on the stack).  Returns (cpu_baseline dict, E, deposit)."""
    import resource
    import subprocess
    import tempfile
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    if hard == resource.RLIM_INFINITY or hard > soft:
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))
    with tempfile.TemporaryDirectory() as td:
        inp, out = os.path.join(td, "bunch.npz"), os.path.join(td, "cpu.npz")
        np.savez(inp, x=x, y=y, z=z)
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", inp, out, "--grid",
               str(args.grid)]
        r = subprocess.run(cmd, timeout=args.cpu_timeout, capture_output=True, text=True)
        if r.returncode != 0:
            log(f"cpu leg failed (exit {r.returncode}): {r.stderr[-2000:]}")
            return None, None, None
        info = json.loads(r.stdout.strip().splitlines()[-1])
        with np.load(out, allow_pickle=False) as f:
            e = f["e"]
            d = {k: f[k] for k in ("rho", "min", "max", "delta")}
    return info, e, d


def cpu_child(args):
    """Upstream Fortran library (oracle/_ref) on the host: one full step of the
    same workload; the C restatement if it is not built."""
    inp, out = args.cpu_child
    with np.load(inp, allow_pickle=False) as f:
        x, y, z = f["x"], f["y"], f["z"]
    from oracle.oracle import REF_SO, Reference, Oracle
    ncore = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nlo, nhi = (1, 1, 1), (args.grid,) * 3
    if os.path.exists(REF_SO):
        R = Reference()
        t0 = time.perf_counter()
        d = R.deposit_particles(x, y, z, nlo, nhi, total_charge=QTOT)
        t1 = time.perf_counter()
        e, b = R.space_charge_3d(d["rho"], nlo, nhi, d["min"], d["max"], d["delta"], GAMMA,
                                 calc_bfield=True)
        t2 = time.perf_counter()
        R.interpolate_field(nlo, nhi, d["min"], d["max"], d["delta"], e, b, x, y, z)
        t3 = time.perf_counter()
        kind = "reference"
        what = "upstream Fortran library (amdflang -O2 -fopenmp; OpenMP only in its Green loops)"
    else:
        O = Oracle()
        t0 = time.perf_counter()
        d = O.deposit_particles(x, y, z, nlo, nhi, total_charge=QTOT)
        t1 = time.perf_counter()
        e, b = O.space_charge_3d(d["rho"], nlo, nhi, d["min"], d["max"], d["delta"], GAMMA)
        t2 = time.perf_counter()
        O.interpolate_field(nlo, nhi, d["min"], d["delta"], e, b, x, y, z)
        t3 = time.perf_counter()
        kind = "port"
        ncore = 1
        what = "C restatement of the upstream algorithm, one thread"
    total = t3 - t0
    info = dict(value=1.0 / total, unit="solves/s", cores=ncore, kind=kind,
                sample=f"one full step of the same workload ({args.grid}^3 mesh, {x.size} particles): "
                       f"deposit {t1 - t0:.2f} s, solve E+B {t2 - t1:.2f} s, gather E,B to all "
                       f"particles {t3 - t2:.2f} s; {what}, OMP threads={ncore}")
    np.savez(out, e=e, rho=d["rho"], min=d["min"], max=d["max"], delta=d["delta"])
    print(json.dumps(info))


# stage of the per-stage HIP-event timing -> the kernel it times (the tiled
# deposit's and gather's first, the cell-sorted ones' second: meshes whose
# tiles do not fit take those)
STAGE_KERNEL = {"gather": ("k_tile_gather<false>", "k_interp_cells"), "gather.out": ("k_unpack",),
                "fft.inv_zmul": ("k_zconv<double",), "fft.inv_x": ("k_xc2r<double",),
                "fft.fwd_x": ("k_xr2c<double",),
                "deposit.cells": ("k_tile_accum<false>", "k_seg_reduce<true>")}


def traffic_child(args):
    """A few fused FP64 steps (the workload of the headline), for profiler
    counter passes."""
    import torch
    import openspacecharge_amd as osc
    x, y, z = bunch(args.particles)
    dev = torch.device("cuda:0")
    xs, ys, zs = (torch.from_numpy(a).to(dev) for a in (x, y, z))
    ctx = osc.Context(0)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.set_async(True)
    ctx.set_green_cache(False)
    mesh = osc.Mesh3D(nhi=(args.grid,) * 3, gamma=GAMMA, device="cuda")
    for _ in range(2):
        osc.space_charge_step(xs, ys, zs, mesh, total_charge=QTOT, calc_bfield=True, ctx=ctx)
    torch.cuda.synchronize()
    ctx.close()


def _counter_csv(d):
    import glob
    f = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    return f[0] if f else None


def traffic_leg(args):
    """HBM bytes per launch of each kernel of the fused step: two rocprofv3
    counter passes (FETCH_SIZE, WRITE_SIZE) over a child running the step,
    each scaled by the same counter's reading for a 1 GiB device copy.
    Returns ({kernel name: (read bytes, write bytes)}, scales) or None."""
    import csv
    import shutil
    import signal
    import subprocess
    import tempfile
    rp = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(rp):
        return None
    per = {}
    scale = {}
    with tempfile.TemporaryDirectory() as td:
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            for what, extra in (("step", ["--traffic-child"]), ("copy", ["--calibrate-copy"])):
                d = os.path.join(td, f"{ctr}_{what}")
                cmd = [rp, "--pmc", ctr, "--output-format", "csv", "-d", d, "-o", "p", "--",
                       sys.executable, os.path.abspath(__file__), *extra, "--particles",
                       str(args.particles), "--grid", str(args.grid)]
                p = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                                     start_new_session=True)
                try:
                    p.wait(timeout=args.traffic_timeout)
                except subprocess.TimeoutExpired:
                    os.killpg(p.pid, signal.SIGKILL)
                    p.wait()
                    log(f"traffic pass {ctr}/{what} timed out")
                    return None
                f = _counter_csv(d)
                if p.returncode != 0 or f is None:
                    log(f"traffic pass {ctr}/{what} failed (exit {p.returncode})")
                    return None
                vals = {}
                for r in csv.DictReader(open(f)):
                    vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]) * 1024.0)
                if what == "copy":
                    k = [k for k in vals if "copyBuffer" in k]
                    if not k:
                        return None
                    scale[ctr] = (1 << 30) / (sum(vals[k[0]]) / len(vals[k[0]]))
                else:
                    for k, v in vals.items():
                        per.setdefault(k, {})[ctr] = sum(v) / len(v)
            log(f"traffic pass {ctr} done")
    out = {k: (v.get("FETCH_SIZE", 0.0) * scale["FETCH_SIZE"], v.get("WRITE_SIZE", 0.0) * scale["WRITE_SIZE"])
           for k, v in per.items()}
    return out, scale


def calibrate_copy():
    """A plain device copy of known size, for scaling the FETCH_SIZE and
    WRITE_SIZE counters of the same profiler run to bytes."""
    import torch
    nbytes = 1 << 30
    a = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1)
    for _ in range(4):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "copies": 4}))


def main():
    args = parse()
    if args.cpu_child:
        cpu_child(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); independent solves "
              "shard across GPUs with no exchange, so the whole-job rate is N x the 1-GPU value.",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.traffic_child:
        traffic_child(args)
        return
    x, y, z = bunch(args.particles)
    log(f"bunch ready: {x.size} particles")
    cpu = e_ref = d_ref = None
    if not args.no_cpu:
        # before any GPU work: the child is started from a process with no GPU state
        t = time.perf_counter()
        cpu, e_ref, d_ref = cpu_leg(args, x, y, z)
        log(f"cpu leg done in {time.perf_counter() - t:.1f} s")
    traffic = None
    if not args.no_traffic:
        t = time.perf_counter()
        traffic = traffic_leg(args)
        log(f"traffic passes done in {time.perf_counter() - t:.1f} s")
    res, fields = gpu_leg(args, x, y, z)
    log("gpu legs done")
    streams = stream_legs(args)
    r64 = res["fp64"]
    # roofline of the dominant memory-bound stage
    n3 = (args.grid,) * 3
    L = tuple(2 * v for v in n3)
    st = {k: v for k, v in r64["prof"].items()}
    cand = [(ms, k) for k, (ms, cnt) in st.items() if stage_bytes(k, 8, n3, L, args.particles, 3, 3)]
    ms_tot, dom = max(cand)
    launches = st[dom][1]
    per_launch_s = ms_tot / launches / 1e3
    alg = stage_bytes(dom, 8, n3, L, args.particles, 3, 3)
    achieved = alg / per_launch_s / 1e9
    knames = STAGE_KERNEL.get(dom, ())
    kname = knames[0] if knames else None
    traffic_bytes = None
    if traffic:
        for cand_name in knames:
            hits = [v for k, v in traffic[0].items() if cand_name in k]
            if hits:
                kname = cand_name
                traffic_bytes = round(hits[0][0] + hits[0][1])
                break
    out = {
        "metric": METRIC, "value": round(r64["solves_per_s"], 3), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r64["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Philox Gaussian bunch)",
        "config": {"workload": "free-space IGF solve, 128^3 mesh (256^3 doubled FFT), 1e7 particles, "
                               "deposit + E/B solve + gather to all particles",
                   "mesh": list(n3), "particles": args.particles, "sigma_m": list(SIGMA),
                   "gamma": GAMMA, "green": "rebuilt every step"},
        "p50_ms": round(r64["p50_ms"], 3), "p99_ms": round(r64["p99_ms"], 3),
        "stages_ms_fp64": r64["stages_ms"],
        "fp32": {"solves_per_s": round(res["fp32"]["solves_per_s"], 3),
                 "ms_per_step": round(res["fp32"]["ms_per_step"], 3),
                 "p99_ms": round(res["fp32"]["p99_ms"], 3),
                 "max_rel_err_vs_fp64": res["fp32_vs_fp64"]["max_rel"],
                 "rms_rel_err_vs_fp64": res["fp32_vs_fp64"]["rms_rel"],
                 "stages_ms": res["fp32"]["stages_ms"]},
        "cached_green_solves_per_s": {"fp64": round(r64["cached_green_solves_per_s"], 3),
                                      "fp32": round(res["fp32"]["cached_green_solves_per_s"], 3)},
        "three_call_solves_per_s": {"fp64": round(r64["three_call_solves_per_s"], 3),
                                    "fp32": round(res["fp32"]["three_call_solves_per_s"], 3)},
        "host_arrays_three_call_fp64": None if not r64["host_arrays"] else {
            "solves_per_s": round(r64["host_arrays"]["solves_per_s"], 3),
            "ms_per_step": round(r64["host_arrays"]["ms_per_step"], 2),
            "steps": r64["host_arrays"]["steps"],
            "note": "deposit_particles, space_charge_3d, interpolate_field on numpy arrays: "
                    "PCIe transfers of x, y, z (twice), rho, fields, E and B included"},
        "roofline": {"bound": "hbm", "kernel": f"{kname or dom} (stage {dom})",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_bytes,
                     "algorithmic_bytes_per_launch": alg,
                     "avg_launch_ms": round(per_launch_s * 1e3, 5),
                     "traffic_note": "FETCH_SIZE x1024 x %.3f + WRITE_SIZE x1024 x %.3f bytes per launch "
                                     "(scales from a 1 GiB device copy)" % (
                                         traffic[1]["FETCH_SIZE"] if traffic else float("nan"),
                                         traffic[1]["WRITE_SIZE"] if traffic else float("nan"))},
    }
    out.update(streams)
    if cpu is not None:
        out["cpu_baseline"] = cpu
        # full-size FP64 parity against the upstream library on the same bunch
        e64, rho64 = fields["fp64"]
        e32 = fields["fp32"][0]
        err = lambda a: max(float(np.max(np.abs(a[..., k] - e_ref[..., k])) / np.max(np.abs(e_ref[..., k])))
                            for k in range(3))
        out["vs_upstream_fullsize"] = {"rho_max_rel": float(np.max(np.abs(rho64 - d_ref["rho"])) /
                                                            np.max(np.abs(d_ref["rho"]))),
                                       "E_fp64_max_rel": err(e64), "E_fp32_max_rel": err(e32)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
