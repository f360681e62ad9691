"""Benchmark of the per-step PIC hot path on one MI355X.

Workload (the plain method at scale): a 3D periodic uniform thermal plasma
on 256x128x128 cells (dx = d_e = 1, natural units c = eps0 = m_e = |q_e| =
n0 = 1), electrons and ions (mi/me = 25), 100 particles per cell per
species (838,860,800 particles), uth_e = 0.1c, uth_i = 0.02c, equal
weights n0*dV/ppc, dt = 0.95 of the Courant limit, fields initially zero,
sort every 20 steps.  A step is one pass of the whole hot path:
sort_p (every 20th step), clear/advance_p x2/reduce, unload + synchronize_jf,
advance_b / advance_e / advance_b, load_interpolator.

value = particle-steps/s over the timed steps (whole job, data resident in
HBM).  roofline = the main push kernel (k_advance_p_tiled, the partitioned
push every timed step runs): algorithmic 64 B per particle (one 32-B
particle read, one 32-B write) per launch over its mean launch time from
hipEvents on its stream; traffic = its HBM bytes per launch from the
FETCH_SIZE / WRITE_SIZE counters (profiles/, calibrated on a copy kernel).  cpu_baseline = the reference's
own advance_p (oracle/_ref, compiled from the lanl/vpic sources; else the
oracle's restatement) on the host cores, on a bounded sample of the same
plasma.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("particle-steps/s on 1 MI355X, % HBM roofline; ensemble runs/s; "
          "adjoint/fwd time")
HBM_PEAK_GBS = 8000.0
BYTES_PER_PARTICLE = 64


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def set_tile(spec):
    """Push tile "tx,ty,tz" (one-cell window halo) or "tx,ty,tz,halo"."""
    from vpic_amd import _lib
    v = [int(x) for x in spec.split(",")]
    if len(v) == 4:
        _lib.call("vpic_hip_push_tile_window", *v)
    else:
        _lib.call("vpic_hip_push_tile", *v)


def build_sim(nx, ny, nz, ppc, seed=2026, sort_interval=20):
    from vpic_amd import sim as S
    from vpic_amd import types as T
    dt = 0.95 * T.courant_length(nx, ny, nz, nx, ny, nz)
    g = T.grid_periodic(0, 0, 0, nx, ny, nz, nx, ny, nz, dt)
    s = S.Simulation(g)
    w = 1.0 / ppc  # n0 * dV / ppc with n0 = dV = 1
    S.uniform_plasma(s, ppc, -1.0, 1.0, 0.1, seed, "electron", w=w,
                     sort_interval=sort_interval)
    S.uniform_plasma(s, ppc, 1.0, 25.0, 0.02, seed + 1, "ion", w=w,
                     sort_interval=sort_interval)
    s.load_interpolator_array()
    return s


def host_info():
    """The host's CPU model (lscpu), nproc, and the CPUs this process may
    run on."""
    import subprocess
    fields = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True,
                             timeout=20).stdout
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                fields[k.strip()] = v.strip()
    except Exception:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    nproc = os.cpu_count() or 1
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True,
                                   timeout=20).stdout.strip())
    except Exception:
        pass

    def num(k):
        try:
            return int(fields.get(k, ""))
        except ValueError:
            return None
    cores, sockets = num("Core(s) per socket"), num("Socket(s)")
    return {"lscpu_model": fields.get("Model name", ""),
            "lscpu_cpus": num("CPU(s)"),
            "host_physical_cores": (cores * sockets) if cores and sockets
            else None,
            "nproc": nproc, "cpus_allowed": allowed}


def cpu_threads(info):
    """Threads for the CPU baseline: every CPU this process may use, capped
    at the host's share per GPU (OMP_NUM_THREADS where set; the GPU hosts
    here allot 16 CPUs to each GPU)."""
    n = info["cpus_allowed"]
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(n, cap) if cap > 0 else n)


def cpu_baseline(s, budget_s=12.0):
    """The reference's own advance_p (oracle/_ref: advance_p_pipeline.cc
    compiled from the lanl/vpic sources; else the oracle's restatement) on
    the host cores, on the benchmark's own arrays: the electron species of
    the 256x128x128 plasma (419,430,400 particles) as the GPU left it, with
    the GPU's current interpolators.  N pthreads, each one DISTRIBUTE(np,16)
    block with a private accumulator copy, the caller on the tail, then the
    copies reduced (advance_p_pipeline.cc:253-340,
    reduce_array_pipeline.cc); 3 warm-up calls, then timed calls until
    budget_s, as sample/bench/advance_p:46-52 times it."""
    import numpy as np
    import torch
    from oracle.oracle import Oracle, Ref, host_grid
    from vpic_amd import types as T

    info = host_info()
    threads = cpu_threads(info)
    sp = s.species[0]
    torch.cuda.synchronize()
    log("cpu baseline: copying %d particles to the host" % sp.np)
    p = sp.p.download()[:sp.np].copy()
    fi = s.interp.download().copy()
    g = s.g
    hg = host_grid(g)
    use_ref = Ref.available()
    if use_ref:
        ref = Ref.get()
        stride = (g.nv + 1) & ~1
        acc = np.zeros(((threads + 1) * stride, 12), dtype=np.float32)
        pm = np.zeros(max(2 * len(p) // 25, 16 * 273), dtype=T.MOVER)

        def push():
            ref.advance_p_pipelines(p, sp.q, sp.m, acc, fi, hg, threads, pm)
    else:
        orc = Oracle.get()
        acc = np.zeros(((threads + 1) * g.nv, 12), dtype=np.float32)

        def push():
            orc.advance_p_threaded(p, sp.q, sp.m, acc, fi, hg, threads,
                                   max(2 * len(p) // 25, 16 * 273))
    for k in range(3):
        push()
        log("cpu baseline: warm-up %d" % (k + 1))
    n_step, t0 = 0, time.perf_counter()
    while True:
        push()
        n_step += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n_step >= 50:
            break
    out = {"value": len(p) * n_step / el, "unit": "particle-steps/s",
           "cores": threads, "kind": "reference" if use_ref else "port",
           "sample": ("advance_p of the electron species of the benchmark "
                      "plasma (%d particles, 256x128x128 cells) as the GPU "
                      "left it, with its interpolators; %d timed calls after "
                      "3 warm-ups, %.1f s" % (len(p), n_step, el))}
    out.update(info)
    out["threads_note"] = ("threads = the CPUs this process may use "
                           "(nproc: the host's CPU share for this GPU), "
                           "capped at OMP_NUM_THREADS; the host's physical "
                           "cores beyond that share are not available to "
                           "this process")
    return out


def ensemble_probe(nbox=256, seed=2026, steps=500, tune_steps=None,
                   sort_interval=20):
    """The batched ensemble: nbox counter-streaming beam boxes drawn from
    the seed (sides from {32,64,96,128}, 16-256 ppc, drift 0.05-0.3c,
    spread 0.01c, immobile ions), every box sorted every sort_interval
    steps, tuned (each trial a whole sort interval from a sort, by
    default, so that the drift between sorts counts), then `steps`
    graph-replayed steps timed; runs/s and aggregate particle-steps/s."""
    if tune_steps is None:
        tune_steps = sort_interval if sort_interval > 0 else 3
    import torch
    from vpic_amd import ensemble as E
    cfg = E.beam_config(nbox, seed)
    t0 = time.perf_counter()
    e = E.beam_ensemble(cfg, seed)
    e.sort_interval(sort_interval)
    npart = sum(e.np)
    log("ensemble: %d boxes, %d particles, loaded in %.1f s"
        % (nbox, npart, time.perf_counter() - t0))
    shape, tuned_ms = e.tune(tune_steps)
    tile = e.tile()
    log("ensemble: tuned shape %s, tile %s (%.2f ms per step)"
        % (shape, tile, tuned_ms))
    e.step(2)  # settle
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.step(steps)
    el = time.perf_counter() - t0
    e.close()
    return {"boxes": nbox, "particles": npart, "steps_per_run": steps,
            "runs_per_s": 1.0 / el, "s_per_run": el,
            "particle_steps_per_s": npart * steps / el,
            "ms_per_step": el * 1e3 / steps,
            "tuned_shape": {"push_bs": shape[0], "push_chunk": shape[1],
                            "field_bs": shape[2], "field_chunk": shape[3],
                            "blocks_per_cu": shape[4],
                            "tile": list(tile[:3]),
                            "window_halo": tile[3]},
            "sort_interval": sort_interval,
            "config": "256 2D periodic counter-streaming electron-beam boxes, "
                      "seed %d, hipGraph-replayed step, sort_p of every box "
                      "every %d steps" % (seed, sort_interval)}


def adjoint_probe(n=64, ppc=32, steps=200, ckpt=20):
    """The differentiable Weibel run: n^3 cells (dx = 0.1 d_e), ppc
    bi-Maxwellian electrons (T_perp/T_par = 4, uth_par = 0.05c), immobile
    ions, `steps` steps; the forward alone and the forward with the
    checkpointed adjoint (gradient of the final magnetic energy with respect
    to every initial momentum), FP32, timed on the device."""
    import ctypes
    import numpy as np
    import torch
    from vpic_amd import _lib
    from vpic_amd import device
    from vpic_amd import types as T
    L = n * 0.1
    dt = 0.95 * T.courant_length(L, L, L, n, n, n)
    g = T.grid_periodic(0, 0, 0, L, L, L, n, n, n, dt)
    npart = ppc * n ** 3
    p = device.particles(npart)
    _lib.call("vpic_hip_load_uniform_plasma", p.ptr, ctypes.byref(g), ppc,
              0.05, 0.1, 0.1, 0.0, 0.0, 0.0, 0.001 / ppc, 17, None)
    f = device.fields(g)
    f.zero()
    ub = torch.empty((npart, 3), dtype=torch.float32, device="cuda")
    en = ctypes.c_double()

    def call(grads):
        _lib.call("vpic_hip_adjoint_bfield_energy", p.ptr, npart, -1.0, 1.0,
                  f.ptr, ctypes.byref(g), steps, ckpt, 4, ctypes.byref(en),
                  ub.data_ptr() if grads else None, None, None, None)
    call(False)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    call(False)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    call(True)
    t_adj = time.perf_counter() - t0
    gnorm = float(ub.double().norm())
    ub32 = ub.double()
    # the same run in FP64 (the kernels over double) from the widened state
    from vpic_amd import adjoint as A
    p64 = torch.from_numpy(A.host_particles64(p.download())).cuda()
    f64 = torch.zeros((g.nv, 18), dtype=torch.float64, device="cuda")
    ub64 = torch.empty((npart, 3), dtype=torch.float64, device="cuda")

    def call64(grads):
        _lib.call("vpic_hip_adjoint_bfield_energy", p64.data_ptr(), npart,
                  -1.0, 1.0, f64.data_ptr(), ctypes.byref(g), steps, ckpt, 8,
                  ctypes.byref(en), ub64.data_ptr() if grads else None, None,
                  None, None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    call64(False)
    t_fwd64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    call64(True)
    t_adj64 = time.perf_counter() - t0
    diff = float((ub32 - ub64).norm() / ub64.norm())
    return {"cells": n ** 3, "particles": npart, "steps": steps,
            "checkpoint_interval": ckpt, "forward_s": t_fwd,
            "forward_and_adjoint_s": t_adj,
            "adjoint_over_forward": t_adj / t_fwd,
            "energy": en.value, "grad_norm": gnorm,
            "fp64": {"forward_s": t_fwd64, "forward_and_adjoint_s": t_adj64,
                     "adjoint_over_forward": t_adj64 / t_fwd64},
            "fp32_grad_vs_fp64_rel_norm": diff,
            "config": "3D Weibel 64^3, 32 ppc electrons, T_perp/T_par=4, "
                      "uth_par=0.05c, immobile ions, FP32 (and the same run "
                      "in FP64)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--ppc", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the batched-ensemble measurement")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the differentiable-run measurement")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--tile", type=str, default=None,
                    help="push tile tx,ty,tz (cells) or tx,ty,tz,halo")
    ap.add_argument("--direct-runs", type=int, default=None,
                    help="push deposit threshold (runs per wavefront)")
    ap.add_argument("--sweep", type=str, default=None,
                    help="tune: 'tx,ty,tz:runs[:group[:bin_mode]];...' each "
                         "timed in turn on one plasma (one JSON line per "
                         "candidate)")
    ap.add_argument("--stream-probe", action="store_true",
                    help="only time an in-place read+write of 13.4 GB (one "
                         "species' particle array) with a plain elementwise "
                         "kernel: the stream rate the push is compared to")
    ap.add_argument("--collision-probe", action="store_true",
                    help="only time the Takizuka-Abe operator on the sorted "
                         "plasma (electron-ion and electron-electron), one "
                         "JSON line")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 4 GiB device-to-device copy three times "
                         "(calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the hot path on a "
              "single MI355X (multi-GPU sharding is design-only); use "
              "--gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from vpic_amd import _lib
    from vpic_amd import sim as S

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    _lib.call("vpic_hip_init", 0)
    if args.copy_probe:
        n = 1 << 30  # floats = 4 GiB
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)  # an elementwise read+write kernel
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n,
                          "copy_probe_bytes_written": 4 * n}))
        return
    if args.collision_probe:
        # the benchmark plasma, both species sorted; each operator timed
        # over 5 calls (a new stream counter per call) after one warm-up
        s = build_sim(args.nx, args.ny, args.nz, args.ppc)
        el, io = s.species
        for sp in (el, io):
            s.sort_p(sp)
        torch.cuda.synchronize()
        out = {"collision_probe": "takizuka_abe", "cells": args.nx * args.ny *
               args.nz, "ppc": args.ppc}
        for name, a, b in (("electron_ion", el, io),
                           ("electron_electron", el, el)):
            def call(k):
                _lib.call("vpic_hip_takizuka_abe", a.p.ptr, a.partition.ptr,
                          a.np, a.q, a.m, b.p.ptr, b.partition.ptr, b.np,
                          b.q, b.m, 1e-3, 1, 7, k, s._G(), s.stream)
            call(0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for k in range(1, 6):
                call(k)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / 5 * 1e3
            npairs = a.np if a is not b else a.np // 2
            out[name] = {"ms_per_call": ms,
                         "collisions_per_s": npairs / (ms * 1e-3)}
            log("%s: %.2f ms per call" % (name, ms))
        print(json.dumps(out))
        return
    if args.stream_probe:
        n = 419430400 * 8  # floats: 32 B x 419,430,400 particles
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        for _ in range(2):
            a.mul_(1.0)
        e0, e1 = torch.cuda.Event(True), torch.cuda.Event(True)
        e0.record()
        for _ in range(5):
            a.mul_(1.0)  # reads and writes every byte in place
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 5
        print(json.dumps({"stream_probe_bytes": 8 * n, "ms": ms,
                          "tb_per_s": 8 * n / (ms * 1e-3) / 1e12}))
        return
    if args.sweep:
        # one plasma, several push settings in turn (each timed over the
        # same number of steps, sorts included)
        s = build_sim(args.nx, args.ny, args.nz, args.ppc)
        npart = sum(sp.np for sp in s.species)
        for _ in range(args.warmup):
            s.advance()
        for cand in args.sweep.split(";"):
            f = cand.split(":")
            tile, runs = f[0], int(f[1])
            gs = int(f[2]) if len(f) > 2 else 16
            bmode = int(f[3]) if len(f) > 3 else 0
            set_tile(tile)
            _lib.call("vpic_hip_push_direct_runs", runs)
            _lib.call("vpic_hip_push_group", gs)
            _lib.call("vpic_hip_push_bin_mode", bmode)
            s.advance()  # settle the new setting
            _lib.call("vpic_hip_push_timing", None, None, None)
            _lib.call("vpic_hip_push_timing_enable", 1)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                s.advance()
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            _lib.call("vpic_hip_push_timing_enable", 0)
            launches, push_ms = ctypes.c_int64(), ctypes.c_double()
            _lib.call("vpic_hip_push_timing", ctypes.byref(launches),
                      ctypes.byref(push_ms), None)
            print(json.dumps({"tile": tile, "direct_runs": runs, "group": gs,
                              "bin_mode": bmode,
                              "ms_per_step": el * 1e3 / args.steps,
                              "push_ms": push_ms.value / launches.value,
                              "value": npart * args.steps / el}), flush=True)
        return
    if args.tile:
        set_tile(args.tile)
    if args.direct_runs:
        _lib.call("vpic_hip_push_direct_runs", args.direct_runs)
    log("building %dx%dx%d, %d ppc/species" % (args.nx, args.ny, args.nz,
                                               args.ppc))
    s = build_sim(args.nx, args.ny, args.nz, args.ppc)
    npart = sum(sp.np for sp in s.species)
    torch.cuda.synchronize()
    log("particles: %d" % npart)
    for _ in range(args.warmup):
        s.advance()
    torch.cuda.synchronize()
    nm, ign, bad = [sum(v) for v in zip(*[s.check_movers(sp) for sp in s.species])]
    assert bad == 0 and nm == 0, (nm, ign, bad)
    log("warmup done (%d steps)" % args.warmup)

    timer = S.PhaseTimer()
    _lib.call("vpic_hip_push_timing", None, None, None)  # reset
    _lib.call("vpic_hip_push_timing_enable", 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        s.advance(timer)
        if (k + 1) % 50 == 0:
            log("step %d" % (k + 1))
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    _lib.call("vpic_hip_push_timing_enable", 0)
    launches, push_ms, push_parts = (ctypes.c_int64(), ctypes.c_double(),
                                     ctypes.c_int64())
    _lib.call("vpic_hip_push_timing", ctypes.byref(launches),
              ctypes.byref(push_ms), ctypes.byref(push_parts))
    phases = timer.totals_ms()
    nm, ign, bad = [sum(v) for v in zip(*[s.check_movers(sp) for sp in s.species])]
    assert bad == 0

    value = npart * args.steps / el
    tile = (ctypes.c_int32 * 4)()
    _lib.call("vpic_hip_push_get_tile", tile)
    stale = ctypes.c_int64()
    _lib.call("vpic_hip_sort_keys_stale", ctypes.byref(stale))
    avg_launch_s = push_ms.value / 1e3 / max(launches.value, 1)
    bytes_per_launch = BYTES_PER_PARTICLE * push_parts.value / max(
        launches.value, 1)
    achieved = bytes_per_launch / avg_launch_s / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("3D periodic uniform thermal plasma %dx%dx%d cells, "
                         "electrons+ions mi/me=25, %d ppc/species "
                         "(%d particles), uth_e=0.1c uth_i=0.02c, dt=0.95 "
                         "Courant, sort every 20 steps"
                         % (args.nx, args.ny, args.nz, args.ppc, npart)),
            "particles": npart,
            "cells": args.nx * args.ny * args.nz,
            "sort_interval": 20,
        },
        "phases_ms_per_step": {k: v / args.steps for k, v in phases.items()},
        "push_kernel": {
            "name": "k_advance_p_tiled",
            "tile": list(tile[:3]),
            "window_halo": tile[3],
            "sort_keys_stale": stale.value,
            "launches": launches.value,
            "avg_launch_ms": avg_launch_s * 1e3,
            "particle_steps_per_s": push_parts.value / (push_ms.value / 1e3),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
    }
    pmc = os.path.join(ROOT, "profiles",
                       "k_advance_p_tiled_plasma256x128x128_traffic.json")
    if os.path.exists(pmc) and (args.nx, args.ny, args.nz, args.ppc) == (
            256, 128, 128, 100):
        with open(pmc) as fh:
            out["roofline"]["traffic"] = json.load(fh).get("bytes_per_launch")
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(s, budget_s=args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    del s
    torch.cuda.empty_cache()
    if not args.no_ensemble:
        log("ensemble ...")
        out["ensemble"] = ensemble_probe()
    if not args.no_adjoint:
        log("adjoint ...")
        out["adjoint"] = adjoint_probe()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
