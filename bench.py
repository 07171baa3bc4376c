#!/usr/bin/env python3
"""Benchmark of the cloning-sampler inner loop on one MI355X.

Workload: channel flow 400x100 past a square of side 20 centred at
(100, 50), tau 0.53, inlet u 0.05, 4096 clones, 200 LBM steps per cloning
interval, bias k 0.005, population noise 1e-4 after each copy.  A "step" of
this benchmark is one full cloning interval of all 4096 clones: the
persistent step kernel (200 LBM steps with fused drag and FP64 weights), the
selection and the on-device copy/noise, replayed from one HIP graph.

The headline value is the clone-MLUPS of the FP16-shifted storage path
(FP32 collide, FP64 drag/weights/selection).  The plain FP64 path and the
FP32 / BF16-shifted paths are timed on the same workload in the same run.
Each low-precision path's per-interval drag, log-mean-weight and
scaled-cumulant errors are measured against FP64 over several cloning
steps from identical developed states, with both runs making the FP64
run's selections (shadow mode).

The copy kernel is also timed on its own, with HIP events, on selections
that copy half the clones ("gather_roofline"): at this workload the weights
are close enough that a cloning step copies only a few clones.

--workload runs one of the other channel workloads instead (the same
measurements, with the storages listed for it): channel200x50 (16 clones,
FP64), channel400x100-n512 (512 clones, FP64), batched8 (8 samplers of four
lattice sizes and two biases in one launch, FP32/FP64/FP16) and
channel800x200 (2048 clones, 400 steps, FP16/FP64), ref129, the
reference's own mode (periodic box with body force, stress-integral drag and
snapshot perturbation) at its L = 32 size, 4096 clones, FP64/FP16/FP32, with
the reference's own kernels as the CPU baseline, and forcing129, its forcing
driver at the same size (random-field body force in the first half of every
interval, fields drawn on the device).

Prints one JSON line.  Only N=1 is supported (one GPU).
"""
from __future__ import annotations

import argparse
import glob
import json
import math
import os
import re
import subprocess
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cloning-lbm-mpi_amd"))
import lbc  # noqa: E402

HBM_PEAK_GBS = 8000.0


def _inst(Lx, Ly, side, cx, cy, tau, u_in, n_clones, steps, k, sigma):
    return dict(Lx=Lx, Ly=Ly, side=side, cx=cx, cy=cy, tau=tau, u_in=u_in,
                n_clones=n_clones, steps=steps, k=k, sigma=sigma)


# Channel workloads (square centred at (cx, cy)); the first storage listed
# gives the headline value.  The default is the 4096-clone 400x100 sweep
# over all four storages.
WORKLOADS = {
    "channel400x100": dict(
        desc="channel 400x100, square side 20 at (100,50), tau 0.53, u_in "
             "0.05, 4096 clones, 200 LBM steps per cloning interval, k 0.005, "
             "noise sigma 1e-4",
        instances=[_inst(400, 100, 20, 100, 50, 0.53, 0.05, 4096, 200, 0.005,
                         1e-4)],
        storages=["f16", "f64", "f32", "bf16"]),
    "channel200x50": dict(
        desc="channel 200x50, square side 10 at (50,25), tau 0.515, u_in 0.05, "
             "16 clones, 100 LBM steps per cloning interval, k 0.01, noise "
             "sigma 1e-4, seed 0",
        instances=[_inst(200, 50, 10, 50, 25, 0.515, 0.05, 16, 100, 0.01,
                         1e-4)],
        storages=["f64"]),
    "channel400x100-n512": dict(
        desc="channel 400x100, square side 20 at (100,50), tau 0.53, u_in "
             "0.05, 512 clones, 200 LBM steps per cloning interval, k 0.005, "
             "noise sigma 1e-4",
        instances=[_inst(400, 100, 20, 100, 50, 0.53, 0.05, 512, 200, 0.005,
                         1e-4)],
        storages=["f64"]),
    "batched8": dict(
        desc="8 samplers in one launch: {200x50, 300x75, 400x100, 600x150} x "
             "k {0.002, 0.01}, square side Ly/5 at (Lx/4, Ly/2), tau 0.53, "
             "u_in 0.05, {1024, 512, 384, 128} clones per lattice (4096 in "
             "all), 200 LBM steps per cloning interval, noise sigma 1e-4",
        instances=[_inst(Lx, Ly, Ly // 5, Lx // 4, Ly // 2, 0.53, 0.05, n, 200,
                         k, 1e-4)
                   for k in (0.002, 0.01)
                   for (Lx, Ly, n) in ((200, 50, 1024), (300, 75, 512),
                                       (400, 100, 384), (600, 150, 128))],
        storages=["f32", "f64", "f16"]),
    "channel800x200": dict(
        desc="channel 800x200, square side 40 at (200,100), tau 0.52, u_in "
             "0.04, 2048 clones, 400 LBM steps per cloning interval, k 0.002, "
             "noise sigma 5e-5",
        instances=[_inst(800, 200, 40, 200, 100, 0.52, 0.04, 2048, 400, 0.002,
                         5e-5)],
        storages=["f16", "f64"]),
    # the reference's own production mode: periodic box with body force,
    # stress-integral drag and snapshot perturbation of the copies
    "ref129": dict(
        desc="reference box: square side L 32 in a periodic-x 129x129 box "
             "with walls, body force beta0 for U0 0.1, tau 0.6, 4096 clones, "
             "320 LBM steps per cloning interval (dT 1 at T0 = L/U0), alpha "
             "50, copies perturbed by mixing 10 snapshots at eps 0.01",
        mode="ref",
        ref=dict(L=32, tau=0.6, U0=0.1, n_clones=4096, dT=1.0, alpha=50.0,
                 eps=0.01, nn=10),
        storages=["f64", "f16", "f32"]),
    # the reference's forcing driver at the same size: the first half of
    # every interval under a spatial body force (1 + eps*map)*beta0, map a
    # Gaussian random field drawn per clone and interval on the device
    "forcing129": dict(
        desc="reference forcing mode: square side L 32 in a periodic-x "
             "129x129 box with walls, U0 0.1, tau 0.6, 4096 clones, 320 LBM "
             "steps per cloning interval of which the first 160 (dT0 0.5 at "
             "T0 = L/U0) run under (1 + 0.1*map)*beta0 with a fresh random "
             "field per clone and interval (covariance length 32, hipFFT), "
             "alpha 50",
        mode="ref",
        forcing=dict(L=32, tau=0.6, U0=0.1, n_clones=4096, dT=1.0, dT0=0.5,
                     eps=0.1, alpha=50.0),
        storages=["f64", "f16", "f32"]),
}
DEFAULT_WORKLOAD = "channel400x100"


def mode_of(workload: str) -> int:
    return (lbc.MODE_REF if WORKLOADS[workload].get("mode") == "ref"
            else lbc.MODE_CHANNEL)


def specs_of(workload: str, max_clones: int | None = None
             ) -> list[lbc.InstanceSpec]:
    """The workload's instances; max_clones caps each instance's clones."""
    wl = WORKLOADS[workload]
    if "ref" in wl:
        r = wl["ref"]
        n = r["n_clones"] if max_clones is None else min(max_clones,
                                                         r["n_clones"])
        steps = math.floor(r["dT"] * r["L"] / r["U0"])
        return [lbc.ref_box(r["L"], r["tau"], r["U0"], n, steps, r["alpha"],
                            eps=r["eps"], n_snapshots=r["nn"])]
    if "forcing" in wl:
        r = wl["forcing"]
        n = r["n_clones"] if max_clones is None else min(max_clones,
                                                         r["n_clones"])
        return [lbc.forcing_box(r["L"], r["tau"], r["U0"], n, r["dT"],
                                r["dT0"], r["eps"], r["alpha"])]
    out = []
    for w in wl["instances"]:
        n = w["n_clones"] if max_clones is None else min(max_clones,
                                                         w["n_clones"])
        out.append(lbc.channel(w["Lx"], w["Ly"], w["side"], w["cx"], w["cy"],
                               w["tau"], w["u_in"], n, w["steps"], w["k"],
                               sigma=w["sigma"]))
    return out


def prepare(s: "lbc.Sampler", specs: list, amp: float = 1e-3) -> None:
    """Seeded equilibrium states; in the reference mode the snapshot set of
    the perturbation is the first clones' initial states, as the reference
    driver takes its snapshots from the initial pops files."""
    for i, sp in enumerate(specs):
        s.init_equilibrium(i, amp)
        if s.mode == lbc.MODE_REF and sp.n_snapshots > 0:
            s.set_snapshots(i, s.store_states(i, 0, sp.n_snapshots))


def hbm_sweeps(M: int, P: int, M1: int = 0, E: int = 0) -> tuple[int, int]:
    """HBM sweeps of one M-step interval run P passes per sweep, and how many
    of them collide forced steps (M1 forced steps; output every E steps).
    Pass 0 collides only, pass M streams only, pass n collides step n; a
    sweep groups P passes 1..M-1 unless the group would straddle the end of
    the forced steps or an output step (the kernel's own grouping rule,
    fused_group_at in csrc/lbc_internal.h)."""
    def grouped(n):
        return (n >= 1 and n + P - 1 <= M - 1 and
                (n + P - 1 < M1 or n >= M1) and
                (E <= 0 or (n + E - 1) // E * E > n + P - 1))
    n = sweeps = forced = 0
    while n <= M:
        sweeps += 1
        if n < M1 and n < M:
            forced += 1
        n += P if grouped(n) else 1
    return sweeps, forced


def time_storage(storage: int, K: int, W: int, specs: list, mode: int,
                 seed: int = 0, shape=None, G: int = 3) -> dict:
    """W untimed warm-up intervals then K timed ones (graph replays), then
    G selections and copies of the last interval's clones with the copy
    kernel timed alone.  shape: (threads, sites per thread, non-temporal,
    LBM passes per HBM sweep[, clones per workgroup]) or None (tuned at
    context creation)."""
    kw = {}
    if shape:
        kw = dict(threads=shape[0], vec=shape[1], nt=bool(shape[2]),
                  passes=shape[3] if len(shape) > 3 else 1,
                  cpw=shape[4] if len(shape) > 4 else 1)
    with lbc.Sampler(specs, mode=mode, storage=storage, seed=seed,
                     **kw) as s:
        prepare(s, specs)
        if W:
            s.advance(W)
        s.sync()
        t0 = time.perf_counter()
        s.advance(K)
        s.sync()
        t1 = time.perf_counter()
        kms, nk = s.step_kernel_ms()
        R = s.R_history(0)
        threads = s.threads
        vec = s.vec
        nt = s.nt
        passes = s.passes
        cpw = s.cpw
        # the copy kernel on its own, through the API (the kernels the graph
        # replays), each launch timed with events: G selections on the last
        # interval's weights, then G on weights that copy every other clone
        # (1 and 1e-9 alternating: nb = 2, 0, 2, 0, ...), which loads the
        # kernel with N/2 copies -- at these workloads the weights are so
        # close that a cloning step copies only a few clones
        gather, gather_half = [], []
        for runs, half in ((gather, False), (gather_half, True)):
            for _ in range(G):
                copies = 0
                for i, sp in enumerate(specs):
                    w = (np.where(np.arange(sp.n_clones) % 2 == 0, 1.0, 1e-9)
                         if half else None)
                    copies += len(s.select(i, weights=w).pairs)
                s.gather()
                gms, gbytes = s.gather_kernel_ms()
                runs.append((gms, gbytes, copies))
        # the same copies collided into the next interval's pass 0 (channel
        # noise, as lbc_advance does between intervals): alternating
        # selections without copies (equal weights) and copying every other
        # clone, each followed by gather_fused and an interval, both kernels
        # timed with events; the copy cost is the difference of their sums
        fused = None
        if mode == lbc.MODE_CHANNEL and all(sp.noise_sigma > 0
                                            for sp in specs):
            rec = {False: [], True: []}
            for half in (False, True) * G:
                copies = 0
                for i, sp in enumerate(specs):
                    w = (np.where(np.arange(sp.n_clones) % 2 == 0, 1.0, 1e-9)
                         if half else np.ones(sp.n_clones))
                    copies += len(s.select(i, weights=w).pairs)
                s.gather_fused()
                gms, _ = s.gather_kernel_ms()
                s.run_interval()
                sms, _ = s.step_kernel_ms()
                rec[half].append((gms, sms, copies))
            fused = fused_copy_stats(rec)
    secs = t1 - t0
    lup1 = sum(sp.n_clones * sp.sites * sp.steps for sp in specs)
    lups = lup1 * K
    bpl = lbc.bytes_per_lup(storage)
    # plus, in forced steps, the clone's random-field map: 8 B per site in
    # FP64 storage, 4 B (its FP32 copy) in the shifted storages
    map_b = 8 if storage == lbc.STORE_F64 else 4
    alg_bytes = lup1 * bpl + sum(map_b * sp.n_clones * sp.sites *
                                 sp.forced_steps for sp in specs)
    achieved = alg_bytes / (kms * 1e-3) / 1e9 if kms > 0 else float("nan")
    # the bytes the blocked kernel moves at least: every HBM sweep (P LBM
    # passes) reads and writes every population once, and a forced sweep
    # reads the clone's field map once
    sweep_bytes = 0
    for sp in specs:
        n_sw, n_forced = hbm_sweeps(sp.steps, passes, sp.forced_steps,
                                    sp.macro_every)
        # bpl = 2*9*b: one read and one write of the nine populations
        sweep_bytes += sp.n_clones * sp.sites * (bpl * n_sw + map_b * n_forced)
    sweep_gbs = sweep_bytes / (kms * 1e-3) / 1e9 if kms > 0 else float("nan")
    return dict(
        storage=lbc.STORAGE_NAMES[storage],
        clone_mlups=lups / secs / 1e6,
        ms_per_interval=secs / K * 1e3,
        kernel_ms=kms,
        kernel_launches=nk,
        kernel_clone_mlups=lup1 / (kms * 1e-3) / 1e6,
        bytes_per_lup=bpl,
        alg_bytes_per_launch=alg_bytes,
        achieved_gbs=achieved,
        roofline_frac=achieved / HBM_PEAK_GBS,
        sweep_bytes_per_launch=sweep_bytes,
        sweep_gbs=sweep_gbs,
        sweep_frac=sweep_gbs / HBM_PEAK_GBS,
        threads=threads,
        sites_per_thread=vec,
        nontemporal=nt,
        passes_per_sweep=passes,
        clones_per_workgroup=cpw,
        logR_last=float(math.log(R[-1])) if len(R) else None,
        gather_workload=gather_stats(gather),
        gather=gather_stats(gather_half),
        gather_fused=fused,
    )


def fused_copy_stats(rec: dict) -> dict:
    """Copies collided into pass 0 (lbc_gather_fused): mean copy-kernel and
    step-kernel ms without copies and with them, and what the copies add to
    the two together, also scaled to 2048 copies."""
    def mean(v, k):
        return sum(r[k] for r in v) / len(v)
    none, half = rec[False], rec[True]
    n_copies = mean(half, 2)
    extra = (mean(half, 0) + mean(half, 1)) - (mean(none, 0) + mean(none, 1))
    return dict(
        copies=int(round(n_copies)), repeats=len(half),
        gather_ms_no_copies=mean(none, 0), gather_ms=mean(half, 0),
        step_ms_no_copies=mean(none, 1), step_ms=mean(half, 1),
        copy_cost_ms=extra,
        copy_cost_ms_per_2048=extra * 2048.0 / n_copies if n_copies else None)


def gather_stats(runs: list) -> dict | None:
    """Copy kernel (k_gather): mean device time and population bytes over
    the timed gathers; achieved = bytes read + written / time."""
    if not runs:
        return None
    ms = sum(r[0] for r in runs) / len(runs)
    nbytes = sum(r[1] for r in runs) / len(runs)
    gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else float("nan")
    return dict(kernel="k_gather", launches=len(runs),
                copies=[int(r[2]) for r in runs],
                avg_ms=ms, bytes_per_launch=nbytes, achieved_gbs=gbs,
                peak=HBM_PEAK_GBS, frac=gbs / HBM_PEAK_GBS)


def log_mean_exp(v: np.ndarray) -> float:
    m = float(v.max())
    return m + math.log(float(np.mean(np.exp(v - m))))


def scgf_last(logR: list, dT: float) -> float:
    """Last value of the running estimate of post_processing/read_rvalue.py
    (sum of log R over intervals 1..n-2, over T = (n-1) dT)."""
    n = len(logR)
    s = 0.0
    T = 0.0
    lam = 0.0
    for i in range(1, n):
        if i >= 2:
            s = s + logR[i - 1]
        T = T + dT
        lam = (1.0 / T) * s
    return lam


def accuracy(storage: int, workload: str, n: int = 256, warm: int = 10,
             intervals: int = 8, seed: int = 1) -> dict:
    """Errors of `storage` against FP64 from identical developed states
    (at most n clones per instance after `warm` FP64 intervals): over the
    first interval (the stated bar), and over `intervals` cloning steps in
    shadow mode, where both samplers select with the FP64 weights and the
    same draws, so they copy the same clones (with the same noise) and stay
    paired while their states drift apart.  Per interval and clone: the
    interval drag integral's relative error; per interval and instance: the
    log mean weight error; at the end: the running scaled-cumulant estimate
    from each sampler's own mean weights (the worst instance)."""
    specs = specs_of(workload, max_clones=n)
    mode = mode_of(workload)
    ni = len(specs)
    rng = np.random.default_rng(seed)
    with lbc.Sampler(specs, mode=mode, storage=lbc.STORE_F64,
                     seed=seed) as sw:
        prepare(sw, specs)
        sw.advance(warm)
        states = [sw.store_states(i) for i in range(ni)]
    drag, dlr = [], []
    lr64 = [[] for _ in range(ni)]
    lrlp = [[] for _ in range(ni)]
    with lbc.Sampler(specs, mode=mode, storage=lbc.STORE_F64,
                     seed=seed) as s64, \
            lbc.Sampler(specs, mode=mode, storage=storage,
                        seed=seed) as slp:
        for i, sp in enumerate(specs):
            s64.load_states(i, 0, states[i])
            slp.load_states(i, 0, states[i])
            if mode == lbc.MODE_REF and sp.n_snapshots > 0:
                s64.set_snapshots(i, states[i][:sp.n_snapshots])
                slp.set_snapshots(i, states[i][:sp.n_snapshots])
        for it in range(intervals):
            if intervals > 8 and it % 10 == 0:
                print(f"bench: accuracy {lbc.STORAGE_NAMES[storage]}: shadow "
                      f"interval {it}/{intervals}", file=sys.stderr, flush=True)
            s64.run_interval()
            slp.run_interval()
            d_int, l_int = [], []
            for i, sp in enumerate(specs):
                o64, w64 = s64.logweights(i)
                olp, wlp = slp.logweights(i)
                d_int.append(np.abs(olp - o64) / np.abs(o64))
                lr64[i].append(log_mean_exp(w64))
                lrlp[i].append(log_mean_exp(wlp))
                l_int.append(abs(lrlp[i][-1] - lr64[i][-1]))
                wsel = np.exp(w64 - w64.max())
                u = rng.random(sp.n_clones)
                r = rng.integers(0, 2**31 - 1, sp.n_clones,
                                 dtype=np.int64).astype(np.int32)
                s64.select(i, weights=wsel, u=u, rint=r)
                slp.select(i, weights=wsel, u=u, rint=r)
            drag.append(np.concatenate(d_int))
            dlr.append(max(l_int))
            s64.gather()
            slp.gather()
    first = drag[0]
    drag = np.concatenate(drag)
    psi = []
    for i, sp in enumerate(specs):
        dT = sp.steps * sp.dt
        psi.append((scgf_last(lr64[i], dT), scgf_last(lrlp[i], dT)))
    psi64, psilp = max(psi, key=lambda p: abs(p[1] - p[0]))
    return dict(storage=lbc.STORAGE_NAMES[storage],
                clones=sum(sp.n_clones for sp in specs),
                warmup_intervals=warm,
                # the first interval, from identical states
                drag_rel_err_max=float(first.max()),
                drag_rel_err_median=float(np.median(first)),
                logR_abs_err=float(dlr[0]),
                # the whole shadow run
                shadow=dict(
                    intervals=intervals,
                    drag_rel_err_max=float(drag.max()),
                    drag_rel_err_median=float(np.median(drag)),
                    logR_abs_err_max=float(max(dlr)),
                    scgf_fp64=psi64, scgf=psilp,
                    scgf_abs_err=abs(psilp - psi64),
                    scgf_rel_err=(abs(psilp - psi64) / abs(psi64)
                                  if psi64 else None)))


def host_cpus() -> dict:
    """What the host offers this process: logical CPUs (nproc), the CPUs in
    its affinity mask, the cgroup's CPU quota (cpu.max, as CPUs) and the
    CPU model."""
    out = {"nproc": os.cpu_count() or 1}
    try:
        out["affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        out["affinity"] = out["nproc"]
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        out["cgroup_cpus"] = (None if quota == "max"
                              else round(int(quota) / int(period), 2))
    except (OSError, ValueError):
        out["cgroup_cpus"] = None
    out["model"] = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    out["model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return out


# CPU share of one GPU on the GPU boxes: worker pools there are sized to it,
# whatever nproc shows
GPU_BOX_CPU_SHARE = 16


def cpu_baseline(workload: str, lups_per_core: float = 2.4e8,
                 steps: int = 200) -> dict:
    """The CPU path timed on the box's host cores, one worker thread per
    core with its clones run serially, like one MPI rank per core, on a
    bounded sample of the workload: clones of its largest lattice, `steps`
    steps each, about lups_per_core lattice updates per thread (30 clones of
    400x100).  In the reference and forcing modes the workers run the
    reference's own kernels (oracle/_ref/libref.so, its sources at its -O3,
    in its driver's step loop, uniform force: a forced step adds one
    multiply-add per site) when that build is present: kind "reference".
    Otherwise, and for the channel workloads, which have no reference code,
    they run the oracle's port of the reference's AoS push step: kind
    "port"."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    wl = WORKLOADS[workload]
    ref_mode = "ref" in wl or "forcing" in wl
    if ref_mode:
        r = wl.get("ref") or wl["forcing"]
        g = O.ref_box_geom(r["L"], r["tau"], r["U0"])
        base = np.broadcast_to(O.W, (g.Lx, g.Ly, 9)).copy()
    else:
        w = max(wl["instances"], key=lambda i: i["Lx"] * i["Ly"])
        g = O.channel_geom(w["Lx"], w["Ly"], w["side"], w["cx"], w["cy"],
                           w["tau"], w["u_in"])
        base = O.init_channel(g, 1e-3, 0, 0)
    kind = "reference" if ref_mode and O.ref_available() else "port"
    sample_clones_per_core = max(1, round(lups_per_core /
                                          (g.Lx * g.Ly * steps)))
    host = host_cpus()
    # one worker per CPU this process may run on (affinity mask, cgroup
    # quota), within the GPU's CPU share of the box
    usable = host["affinity"]
    if host["cgroup_cpus"]:
        usable = min(usable, max(1, int(host["cgroup_cpus"])))
    cores = max(1, min(usable, GPU_BOX_CPU_SHARE))
    states = [np.stack([base] * sample_clones_per_core) for _ in range(cores)]
    O.lib()
    if kind == "reference":
        O.ref_set_geometry(g)  # fixed while the workers run

    def work(i):
        if kind == "reference":
            O.ref_run_clones(g, states[i], steps)
        else:
            O.run_clones(g, states[i], steps)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(cores)]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    secs = time.perf_counter() - t0
    lups = cores * sample_clones_per_core * g.Lx * g.Ly * steps
    cpu = host["model"]
    what = ("the reference's own step kernels in its driver's step loop, "
            "FP64 AoS" if kind == "reference" else "FP64 AoS push oracle")
    out = dict(value=lups / secs / 1e6, unit="clone-MLUPS", cores=cores,
               kind=kind, host=host, cpu_share=GPU_BOX_CPU_SHARE,
               sample=(f"{cores} threads x {sample_clones_per_core} clones x "
                       f"{steps} steps of {g.Lx}x{g.Ly} ({what}), "
                       f"{secs:.1f} s on {cpu}"))
    cal = calibrate_port(O)
    if cal:
        out["calibration"] = cal
    return out


def calibrate_port(O, L: int = 50, steps: int = 60) -> dict | None:
    """One thread of the oracle's REF step against the reference's own step
    kernels (oracle/_ref/libref.so: streamCollCompute.cpp, walls, square,
    interior reset and force, compiled from the reference's sources with
    its -O3) on the same (4L+1)^2 lattice: the port's speed relative to the
    reference's.  None when the reference build is absent."""
    if not O.ref_available():
        return None
    g = O.ref_box_geom(L, 0.53, 0.05)
    f0 = np.broadcast_to(O.W, (g.Lx, g.Ly, 9)).copy()
    rates = {}
    for name in ("reference", "port"):
        a, b = f0.copy(), np.zeros_like(f0)
        t0 = time.perf_counter()
        for _ in range(steps):
            if name == "port":
                O.step(g, a, b)
            else:
                O.ref_step(g, a, b)
            a, b = b, a
        rates[name] = g.Lx * g.Ly * steps / (time.perf_counter() - t0) / 1e6
    return dict(lattice=f"{g.Lx}x{g.Ly} REF, {steps} steps, 1 thread",
                reference_kernels_mlups=round(rates["reference"], 2),
                port_mlups=round(rates["port"], 2),
                port_over_reference=round(rates["port"] / rates["reference"], 3))


# the step kernel's device code (lbc_kernels.hip and lbc_forced.hip only
# instantiate and launch it)
KERNEL_SOURCES = ("csrc/lbc_step.h", "csrc/lbc_internal.h")


def kernel_source_sha16() -> str:
    """Hash of the device sources the step kernel is compiled from; a
    traffic summary under profiles/ records the hash of the build it
    measured, and its figures are used only while the hash matches."""
    import hashlib

    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "cloning-lbm-mpi_amd", rel), "rb") as f:
            h.update(rel.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def measured_traffic(run: dict, workload: str) -> tuple[int | None, dict]:
    """HBM bytes per step-kernel launch for this workload, storage and
    number of LBM passes per sweep, from the FETCH_SIZE / WRITE_SIZE
    summaries under profiles/ (copy-calibrated there), and where they came
    from.  A summary counts only if its `source_sha16` line names the
    current kernel sources; otherwise the bytes are None and the reason is
    given."""
    key = (f"workload={workload} storage={run['storage']} "
           f"passes={run['passes_per_sweep']}")
    pat = re.compile(r"^traffic_bytes_per_launch (.*): (\d+)\s*$")
    want = kernel_source_sha16()
    stale = None
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.txt"))):
        sha, hit = None, None
        with open(path) as f:
            for line in f:
                line = line.strip()
                if line.startswith("source_sha16:"):
                    sha = line.split(":", 1)[1].strip()
                m = pat.match(line)
                if m and m.group(1) == key:
                    hit = int(m.group(2))
        rel = os.path.relpath(path, ROOT)
        if hit is not None and sha == want:
            return hit, {"file": rel, "source_sha16": sha}
        if hit is not None:
            stale = {"file": rel, "source_sha16": sha,
                     "dropped": f"profiled build {sha} is not this build {want}"}
    return None, stale or {"dropped": f"no traffic summary for {key}"}


def stream_ceilings() -> dict | None:
    """Measured streaming rates of the profiled box (GB/s, read + write) from
    the stream_probe summary under profiles/: plain copy, read, write, and
    the step kernel's FP16 access pattern without arithmetic; None if
    absent."""
    pat = re.compile(r"^stream_ceiling (\w+): (\d+)\s*$")
    out = {}
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.txt"))):
        with open(path) as f:
            for line in f:
                m = pat.match(line.strip())
                if m:
                    out[m.group(1)] = int(m.group(2))
    if out:
        out["source"] = "profiles/stream_ceilings.txt"
    return out or None


def live_stream_ceilings() -> dict | None:
    """The same rates measured now, on this box, by running stream_probe as
    a child process (a few seconds); None if it is not built or fails.
    Boxes differ: the copy rate has ranged over 5.4-5.7 TB/s."""
    exe = os.path.join(ROOT, "cloning-lbm-mpi_amd", "stream_probe")
    if not os.path.exists(exe):
        return None
    try:
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    except (OSError, subprocess.TimeoutExpired):
        return None
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return None
    out = {k: v for k, v in json.loads(lines[-1]).items() if k != "unit"}
    out["source"] = "measured in this run"
    return out


def roofline(run: dict, workload: str, live: dict | None = None) -> dict:
    """HBM roofline of the step kernel.

    achieved: the bytes the kernel's algorithm must move per launch (one
    read and one write of every population per HBM sweep of P LBM passes,
    plus the field map once per forced sweep) over its average launch time
    from HIP events; frac = achieved / 8 TB/s, which cannot exceed 1.
    traffic: HBM bytes per launch counted by rocprofv3 (copy-calibrated
    FETCH_SIZE + WRITE_SIZE) for this build, with dram_achieved / dram_frac
    from them.  alg_achieved / alg_frac: the per-lattice-update model
    (2*9*b bytes per update); with P passes per sweep the kernel moves
    about 1/P of those bytes, so alg_frac exceeds 1 and is not a roofline
    fraction."""
    traffic, source = measured_traffic(run, workload)
    frac = run["sweep_frac"]
    if not frac <= 1.0:
        raise SystemExit(f"bench.py: roofline fraction {frac:.3f} > 1: the "
                         "byte model does not describe this kernel")
    out = {
        "bound": "hbm",
        "achieved": round(run["sweep_gbs"], 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(frac, 4),
        "traffic": traffic,
        "traffic_source": source,
        "bytes_per_launch": run["sweep_bytes_per_launch"],
        "passes_per_sweep": run["passes_per_sweep"],
        "alg_bytes_per_launch": run["alg_bytes_per_launch"],
        "alg_achieved": round(run["achieved_gbs"], 1),
        "alg_frac": round(run["roofline_frac"], 4),
    }
    if traffic:
        dram = traffic / (run["kernel_ms"] * 1e-3) / 1e9
        out["dram_achieved"] = round(dram, 1)
        out["dram_frac"] = round(dram / HBM_PEAK_GBS, 4)
        out["traffic_over_bytes"] = round(traffic / run["sweep_bytes_per_launch"],
                                          4)
        ceil = live or stream_ceilings()
        if ceil:
            out["stream_ceilings"] = ceil
            if "copy" in ceil:
                out["dram_frac_of_copy"] = round(dram / ceil["copy"], 4)
            if "pattern_f16" in ceil and run["storage"] in ("f16", "bf16"):
                # the 2-byte pull pattern with no arithmetic
                out["dram_frac_of_pattern"] = round(dram / ceil["pattern_f16"],
                                                    4)
    return out


def calibrate_copy() -> int:
    """A plain device copy of 2^28 FP32 (1 GiB read, 1 GiB written), for
    scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5,
                    help="timed cloning intervals")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--acc-clones", type=int, default=256,
                    help="clones per instance in the accuracy comparison")
    ap.add_argument("--acc-intervals", type=int, default=8,
                    help="cloning steps of the accuracy comparison's shadow "
                         "run (the workloads' full runs are 40-50)")
    ap.add_argument("--stream-ceiling", action="store_true",
                    help="run cloning-lbm-mpi_amd/stream_probe (streaming "
                         "ceilings for the roofline) and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (torch) so that "
                         "HBM counters can be scaled against known bytes")
    ap.add_argument("--shape", default="",
                    help="threads,sites_per_thread,nontemporal,passes_per_sweep"
                         "[,clones_per_workgroup] for every storage (skips "
                         "the tuning at context creation)")
    ap.add_argument("--only", default="",
                    help="comma list of storages to time (f64,f32,f16,bf16); "
                         "default: the workload's list")
    ap.add_argument("--no-stream-probe", action="store_true",
                    help="skip measuring this box's streaming ceilings (the "
                         "roofline then quotes the profiled box's); runs "
                         "restricted with --only or --shape, such as the "
                         "profiling runs, skip it too")
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD,
                    choices=sorted(WORKLOADS),
                    help="channel workload to run (default %(default)s)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (N=1); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2
    if args.calibrate_copy:
        return calibrate_copy()
    if args.stream_ceiling:
        exe = os.path.join(ROOT, "cloning-lbm-mpi_amd", "stream_probe")
        return subprocess.run([exe]).returncode
    K, W = args.steps, args.warmup
    wl = args.workload
    specs = specs_of(wl)
    want = ([x for x in args.only.split(",") if x]
            or WORKLOADS[wl]["storages"])
    codes = {"f64": lbc.STORE_F64, "f32": lbc.STORE_F32, "f16": lbc.STORE_F16,
             "bf16": lbc.STORE_BF16}
    live = None
    if not (args.no_stream_probe or args.only or args.shape):
        print("bench: streaming ceilings of this box", file=sys.stderr,
              flush=True)
        live = live_stream_ceilings()
    runs = {}
    for name in want:
        print(f"bench: {wl}: timing {name} ({K} intervals after {W} warm-up)",
              file=sys.stderr, flush=True)
        shape = [int(v) for v in args.shape.split(",")] if args.shape else None
        runs[name] = time_storage(codes[name], K, W, specs, mode_of(wl),
                                  shape=shape)
    acc = {}
    for name in ([] if args.no_accuracy else want):
        if name == "f64":
            continue
        print(f"bench: {wl}: accuracy {name} vs f64", file=sys.stderr,
              flush=True)
        acc[name] = accuracy(codes[name], wl, n=args.acc_clones,
                             intervals=args.acc_intervals)
    cpu = None
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(wl)
    head = runs[want[0]]
    config = {
        "workload": WORKLOADS[wl]["desc"] + "; one step = one cloning "
                    "interval of every clone (step kernel + selection + copy)",
        "name": wl,
        "clones": sum(sp.n_clones for sp in specs),
    }
    if len(specs) == 1:
        config["lattice"] = [specs[0].Lx, specs[0].Ly]
        config["lbm_steps_per_interval"] = specs[0].steps
    else:
        config["instances"] = [dict(lattice=[sp.Lx, sp.Ly], clones=sp.n_clones,
                                    lbm_steps_per_interval=sp.steps, k=sp.bias)
                               for sp in specs]
    out = {
        "metric": "clone-MLUPS + % HBM peak, FP16-shifted/FP32 vs plain FP64, "
                  "at stated drag error",
        "value": round(head["clone_mlups"], 1),
        "unit": "clone-MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(head["ms_per_interval"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head["storage"] == "f64" else "f32",
        "storage": (head["storage"] if head["storage"] == "f64"
                    else head["storage"] + "-shifted"),
        "data": "synthetic",
        "config": config,
        "roofline": roofline(head, wl, live),
        "gather_roofline": head.get("gather"),
        "gather_fused": head.get("gather_fused"),
        "cpu_baseline": cpu,
        "paths": runs,
        "accuracy_vs_fp64": acc,
    }
    for name in ("f16", "f32", "bf16"):
        if "f64" in runs and name in runs:
            out[f"speedup_{name}_vs_f64"] = round(
                runs[name]["clone_mlups"] / runs["f64"]["clone_mlups"], 3)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
