#!/usr/bin/env python3
"""Throughput of the HIP cloud-resolving model on one MI355X.

Workload: 4096 2D CRM instances of 64 columns x 58 levels (dx 1 km, dt 5 s, model top 40 km,
stretched dz 50 m .. 1 km), one 1200 s GCM step per call = 240 substeps; per-instance seeded
tropical soundings (SST ~ U[290,304] K, surface RH ~ U[0.6,0.9], shear ~ U[0,10] m/s per
10 km) with +/-0.1 K noise in the lowest 5 levels.  One "step" is one crm() call
(crm/crm_module.F90:45-1718) over the whole batch with every array already resident in HBM
(crm_hip_run_device); consecutive steps continue from the state the previous one left, as
consecutive GCM steps do.

Printed: one JSON line with the throughput in CRM column-substeps per second, the roofline of
the dominant kernel (its average launch time measured with HIP events on the library's own
stream over the timed steps), and the reference CRM timed on this host's CPU on a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "acme-ecp_amd"))
# The uneven batch runs five classes on concurrent streams; with HIP's default of 4 hardware
# queues per process two of them share a queue and run one after the other, which two
# depending on every stream created before.  At least eight queues (set before the GPU is
# first used) give each class its own.  One stream, as in the other measurements, is
# unaffected.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

import numpy as np  # noqa: E402

from crm_inputs import add_convection, make_case  # noqa: E402

# The operational SP1V1 CRM (cam/cime_config/config_component.xml:100,106): 64 x 1 columns,
# 58 levels, dx 1 km, dt 5 s, radiation on 16 column groups of 4.
WORKLOAD = dict(ncrms=4096, nx=64, nz=58, dx=1000.0, dt=5.0, dt_gl=1200.0, nx_rad=16)
HBM_PEAK_GBS = 8000.0
# Compulsory HBM words per grid cell of one launch of each fused kernel: every field the
# kernel reads counted once, every field it writes counted once (2D, FP64 words).
KERNEL_WORDS = {
    "kurant": 3,     # u, w, tkh
    "pre": 14,       # R tabs qv qcl qci qpl qpi u w t qt; W t qt dudt dwdt
    "sgs": 13,       # R u w t tabs qv qcl qci qpl qpi tk; W tk tkh tke
    "mom": 7,        # R u w tk dudt dwdt; W dudt dwdt
    "press": 14,     # R u w dudt(3) dwdt(3); W dudt(na,nc) dwdt(na,nc) u w
    "mpdata": 8,     # R/W t qt qp, R u w
    "precip": 5,     # per sedimentation sweep: R qp tabs t, W qp t
    "diffuse": 7,    # R/W t qt qp, R tkh
    "micro": 13,     # R t qt qp u; W qt qp qn tabs qv qcl qci qpl qpi
}
# Whole-step compulsory state traffic per column-substep: prognostics u w t qt qp read and
# written once plus Adams-Bashforth history (read 2, write 1 per velocity component) = 16
# words per level.
STATE_WORDS_PER_LEVEL = 16


# HBM bytes per launch measured with rocprofv3 PMC counters (FETCH_SIZE, WRITE_SIZE), each
# scaled by what the same counter reads for a device copy of known size; written next to the
# command that produced it.  Lines: "<kernel class> <fetch bytes> <write bytes>", the FP32
# (mixed-precision) engine's kernels prefixed "mixed.".
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_4096x64x58.txt")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel class from the committed counter summary, or None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        parts = line.split("#")[0].split()
        if len(parts) == 3 and parts[0] == kernel:
            return float(parts[1]) + float(parts[2])
    return None


def _cpu_worker(first, count, tag, barrier, results):
    """One CPU-baseline process: instances first..first+count-1 of the workload through the
    reference CRM build `tag` (oracle/_ref), timed from a barrier that every process passes
    together.  The reference keeps its state in module variables, so each process is one
    independent crm() caller, as each MPI rank of the GCM is."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_lib
    w = WORKLOAD
    case = make_case(count, w["nx"], w["nz"], w["dx"], w["dt"], randomize=True, seed_base=first,
                     nx_rad=w["nx_rad"])
    kind = "reference"
    try:
        path = oracle_lib.ref_so_path(case)
        if tag:
            path = path[:-3] + tag + ".so"
        runner = oracle_lib._Runner(path, "ref_crm", "ref_get")
    except (FileNotFoundError, OSError):
        if tag:
            results.put((first, None, "missing", 0.0))
            barrier.wait()
            return
        runner, kind = oracle_lib.oracle(), "port"
    barrier.wait()
    t0 = time.perf_counter()
    rc = runner.run(case, w["dt_gl"])
    sec = time.perf_counter() - t0
    results.put((first, rc, kind, sec))


def cpu_baseline(per_proc, procs, tag=""):
    """The reference CRM (compiled from the reference sources, oracle/_ref) on `procs` host
    cores at once, each process advancing its own `per_proc` instances of the workload by one
    GCM step; the wall time is the slowest process's.  Falls back to the C restatement where
    the reference build is missing.  Processes are started fresh (spawn): none of them touches
    the GPU."""
    import multiprocessing as mp
    w = WORKLOAD
    ctx = mp.get_context("spawn")
    barrier = ctx.Barrier(procs)
    results = ctx.Queue()
    ps = [ctx.Process(target=_cpu_worker, args=(p * per_proc, per_proc, tag, barrier, results))
          for p in range(procs)]
    for p in ps:
        p.start()
    import queue
    got = []
    deadline = time.time() + 600
    while len(got) < procs and time.time() < deadline:
        try:
            got.append(results.get(timeout=5))
        except queue.Empty:
            if any(p.exitcode not in (None, 0) for p in ps):
                break  # a worker died (e.g. an instruction set this host lacks)
    for p in ps:
        p.join(timeout=10)
        if p.is_alive():
            p.terminate()
    if len(got) < procs:
        log(f"CPU baseline{tag}: {procs - len(got)} of {procs} processes gave no result")
        return None
    if any(r[2] == "missing" for r in got):
        return None
    if any(r[1] != 0 for r in got):
        raise RuntimeError(f"CPU baseline returned status {[r[1] for r in got]}")
    sec = max(r[3] for r in got)
    kind = got[0][2]
    nstop = int(w["dt_gl"] / w["dt"])
    colsteps = procs * per_proc * w["nx"] * nstop
    return {"value": colsteps / sec, "unit": "column-steps/s", "cores": procs, "kind": kind,
            "sample": f"{procs * per_proc} of the {w['ncrms']} instances in {procs} processes "
                      f"of {per_proc} (one per core), full 1200 s GCM step ({nstop} substeps); "
                      f"slowest process {sec:.1f} s, fastest {min(r[3] for r in got):.1f} s",
            "build": "flang -O3 -march=x86-64-v3" if tag else "flang -O2"}


def measure(prec, case, args, torch):
    """Warmup and timed crm() calls of one precision on a fresh device copy of case, then a
    separate pass with every kernel class timed by HIP events for the roofline."""
    from crm_hip import KERNELS, CrmHip, DeviceCase
    w = WORKLOAD
    nstop = int(w["dt_gl"] / w["dt"])
    dcase = DeviceCase(case, "cuda:0")
    torch.cuda.synchronize()
    hip = CrmHip(0, prec)
    # the workgroup sizes and launch variants the autotuner picks for this class, from trial
    # calls of 12 substeps on a scratch copy of the state
    scratch = DeviceCase(case.copy(), "cuda:0")
    sched = hip.autotune(scratch, 12 * w["dt"])
    del scratch
    tuned = {"threads_per_workgroup": {k: v for k, v in sched.items()
                                       if k not in ("setup", "setperturb", "precip_init", "press")},
             "variants": {k: v for k, v in hip.variants().items() if k in ("press", "mpdata")}}
    log(f"[{prec}] autotuned: {tuned}")
    # warmup on the path a caller gets (the default: one captured HIP graph per call, no
    # events); the first call also sizes the device pool and captures the graph
    for n in range(max(1, args.warmup)):
        hip.run_device(dcase, w["dt_gl"])
        log(f"[{prec}] warmup step {n + 1}/{max(1, args.warmup)} done")
    torch.cuda.synchronize()
    g0 = hip.graph_launches()
    t0 = time.perf_counter()
    ncyc, nprec = [], []
    for n in range(args.steps):
        hip.run_device(dcase, w["dt_gl"])
        st = hip.stats()
        ncyc.append(st.ncycle_max)
        nprec.append(st.nprec_max)
        log(f"[{prec}] step {n + 1}/{args.steps} done")
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    graph_substeps = hip.graph_launches() - g0
    # roofline pass: plain launches with a HIP-event pair around every kernel on the library's
    # own stream (graphs are off while events are on); the state keeps advancing
    hip.profile_kernels(KERNELS)
    for n in range(args.profile_steps):
        hip.run_device(dcase, w["dt_gl"])
        log(f"[{prec}] roofline pass step {n + 1}/{args.profile_steps} done")
    times = hip.kernel_times()
    hip.profile_kernels([])
    hip.close()
    dominant = max((k for k in times if k in KERNEL_WORDS), key=lambda k: times[k][0])
    log(f"[{prec}] kernel time over {args.profile_steps} steps [ms/launches]: " +
        ", ".join(f"{k}={v[0]:.1f}/{v[1]}" for k, v in times.items()))
    kms, klaunch = times[dominant]
    value = case.ncrms * w["nx"] * nstop * args.steps / sec
    word = 8 if prec == "fp64" else 4
    bytes_launch = KERNEL_WORDS[dominant] * word * case.ncrms * w["nx"] * w["nz"]
    avg_ms = kms / klaunch
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    per_kernel = {}
    for k, (ms, n) in times.items():
        e = {"ms_per_step": ms / args.profile_steps, "avg_launch_us": 1e3 * ms / n}
        if k in KERNEL_WORDS:
            b = KERNEL_WORDS[k] * word * case.ncrms * w["nx"] * w["nz"]
            e["hbm_frac_algorithmic"] = b / (ms / n * 1e-3) / 1e9 / HBM_PEAK_GBS
        per_kernel[k] = e
    return {"value": value, "ms_per_step": 1e3 * sec / args.steps, "max_ncycle": max(ncyc), "schedule": tuned,
            "max_nprec": max(nprec), "graph_substeps_timed": graph_substeps,
            "dcase": dcase, "kernels": per_kernel,
            "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": measured_traffic(dominant if prec == "fp64" else "mixed." + dominant),
                         "avg_launch_ms": avg_ms, "launches": klaunch,
                         "algorithmic_bytes_per_launch": bytes_launch,
                         "timing": f"HIP events on the library stream, separate pass of "
                                   f"{args.profile_steps} GCM steps after the timed ones"}}


# Uneven batch: one 1200 s GCM step of several CRM size classes advanced together
# (crm_hip_run_many), each class with the workgroup schedule the autotuner chose for it.
# (ncrms, nx, ny, nz, dx, dt); 128-column CRMs run on two waves per instance.
UNEVEN = [(2048, 16, 1, 28, 4000.0, 20.0), (1024, 32, 1, 28, 4000.0, 20.0),
          (1024, 64, 1, 58, 1000.0, 5.0), (256, 128, 1, 58, 1000.0, 5.0),
          (128, 8, 8, 28, 4000.0, 20.0)]
UNEVEN_NAMES = [f"{n}x{nx}x{ny}x{nz}" for n, nx, ny, nz, dx, dt in UNEVEN]


def measure_uneven(prec, args, torch):
    from crm_hip import CrmHip, DeviceCase, run_many, tune_many
    dt_gl = WORKLOAD["dt_gl"]
    cases = [make_case(n, nx, nz, dx, dt, ny=ny, randomize=True) for n, nx, ny, nz, dx, dt in UNEVEN]
    hips = [CrmHip(0, prec) for _ in UNEVEN]
    scheds, variants = [], []
    for h, c, cls in zip(hips, cases, UNEVEN):
        scratch = DeviceCase(c.copy(), "cuda:0")
        # tuning calls of 3 substeps on a scratch copy: workgroup size and launch variant
        sched = h.autotune(scratch, 3 * cls[5])
        # classes with a fixed launch shape (one instance per workgroup for the pressure solve,
        # one thread per instance or level for the set-up kernels) are not reported
        scheds.append({k: v for k, v in sched.items()
                       if k not in ("setup", "setperturb", "precip_init", "press")})
        v = h.variants()
        variants.append({"press": v["press"], "mpdata": v["mpdata"]})
        del scratch
    dcs = [DeviceCase(c, "cuda:0") for c in cases]
    # the pressure variants again, timing the whole batch from this state (restored after)
    t_tune = time.perf_counter()
    chosen, trials = tune_many(hips, dcs, dt_gl)
    tune_s = time.perf_counter() - t_tune
    batch_variants = [{"press": v["press"], "mpdata": v["mpdata"]} for v in chosen]
    log(f"[uneven {prec}] batch tuning: {trials} trials in {tune_s:.1f} s, variants {batch_variants}")
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        run_many(hips, dcs, dt_gl)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for n in range(args.steps):
        run_many(hips, dcs, dt_gl)
        log(f"[uneven {prec}] step {n + 1}/{args.steps} done")
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    colsteps = sum(n * nx * ny * int(dt_gl / dt) for n, nx, ny, nz, dx, dt in UNEVEN)
    graphs = sum(h.graph_launches() for h in hips)
    # each class alone on the device (its own whole-call graph): the critical path of the batch
    alone = []
    for j, (h, dc) in enumerate(zip(hips, dcs)):
        h.run_device(dc, dt_gl)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.alone_steps):
            h.run_device(dc, dt_gl)
        torch.cuda.synchronize()
        alone.append(1e3 * (time.perf_counter() - t1) / args.alone_steps)
        log(f"[uneven {prec}] {UNEVEN_NAMES[j]} alone: {alone[-1]:.1f} ms per GCM step")
    # the classes' whole calls captured into one graph (CRM_GRAPHS_JOINT) instead of one graph
    # per class on its own stream
    for h in hips:
        h.set_graphs(3)
    run_many(hips, dcs, dt_gl)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.alone_steps):
        run_many(hips, dcs, dt_gl)
    torch.cuda.synchronize()
    joint_ms = 1e3 * (time.perf_counter() - t1) / args.alone_steps
    log(f"[uneven {prec}] classes in one joint graph: {joint_ms:.1f} ms per GCM step")
    for h in hips:
        h.close()
    levels = sum(n * nx * ny * nz * int(dt_gl / dt) for n, nx, ny, nz, dx, dt in UNEVEN)
    return {"value": colsteps * args.steps / sec, "unit": "column-steps/s",
            "ms_per_step": 1e3 * sec / args.steps,
            "column_level_steps_per_s": levels * args.steps / sec,
            "classes": [f"{n}x{nx}x{ny}x{nz} dx {dx / 1000:g} km dt {dt:g} s" for n, nx, ny, nz, dx, dt in UNEVEN],
            "schedule_threads_per_workgroup": scheds, "class_tuned_variants": variants,
            "schedule_variants": batch_variants, "batch_tuning_trials": trials,
            "graph_substeps": graphs,
            "class_alone_ms_per_step": alone,
            "sum_alone_ms_per_step": sum(alone),
            "critical_path_class": UNEVEN_NAMES[max(range(len(alone)), key=lambda j: alone[j])],
            "joint_graph_ms_per_step": joint_ms,
            "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
            "note": "one 1200 s GCM step of every class per step (crm_hip_run_many): each "
                    "class replays its own whole-call HIP graph on its own stream, "
                    "concurrently; class_tuned_variants are the launch variants each class's "
                    "own tuning picked, schedule_variants those kept after timing the whole "
                    "batch with each pressure variant (tune_many); "
                    "class_alone_ms_per_step is each class run by itself; "
                    "joint_graph_ms_per_step is the classes' calls captured into one graph "
                    "(CRM_GRAPHS_JOINT)"}


def _clone(dc):
    """A device-side copy of a DeviceCase's arrays (same host case)."""
    from crm_hip import DeviceCase
    c = DeviceCase.__new__(DeviceCase)
    c.case, c.torch = dc.case, dc.torch
    c.arrays = {g: {k: t.clone() for k, t in d.items()} for g, d in dc.arrays.items()}
    return c


def measure_learned(args, torch):
    """The learned microphysics at the full workload (FP64 engine) on convecting CRMs: the
    bench soundings with a raining updraft and ice aloft seeded in every instance, spun up for
    --learn-spinup GCM steps; training pairs recorded from the plain scheme over the next
    --learn-record GCM steps (every stride-th point, stride chosen for --learn-pairs pairs);
    the network (seed 7) trained on them; then one GCM step from the same convecting state
    with the plain scheme and with the network, both timed, and the network's domain-mean
    heating, moistening and surface precipitation compared with the plain scheme's."""
    import math

    import crm_mlp
    from crm_hip import CrmHip, DeviceCase
    w = WORKLOAD
    nstop = int(w["dt_gl"] / w["dt"])
    ncrms = args.ncrms
    case = add_convection(make_case(ncrms, w["nx"], w["nz"], w["dx"], w["dt"], randomize=True,
                                    nx_rad=w["nx_rad"], seed_base=7 * ncrms), seed=7)
    dc = DeviceCase(case, "cuda:0")
    hip = CrmHip(0)
    m = crm_mlp.CrmMlp(0)
    for n in range(args.learn_spinup):
        hip.run_device(dc, w["dt_gl"])
        log(f"[learned] spin-up GCM step {n + 1}/{args.learn_spinup} done")
    points = ncrms * w["nx"] * w["nz"] * nstop * args.learn_record
    stride = max(1, math.ceil(points / args.learn_pairs))
    cap = points // stride + 1
    X, T = crm_mlp.record(hip, cap, stride)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for n in range(args.learn_record):
        hip.run_device(dc, w["dt_gl"])
        log(f"[learned] recording GCM step {n + 1}/{args.learn_record} done")
    npairs = crm_mlp.stop_recording(hip)
    X, T = X[:npairs], T[:npairs]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    m.set_norm(*crm_mlp.fit_norm(X, T))
    m.set_params(crm_mlp.init_params(7))
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    hist = crm_mlp.train(m, X, T, steps=args.learn_steps, batch=16384, lr=2e-3, seed=7)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    log(f"[learned] {npairs} pairs (every {stride}th point) recorded in {t1 - t0:.1f} s, "
        f"{args.learn_steps} Adam steps in {t3 - t2:.1f} s, loss {hist[0]:.3e} -> "
        f"{np.mean(hist[-50:]):.3e}")
    del X, T
    # one GCM step from the convecting state: plain, then network
    d_plain, d_net = _clone(dc), _clone(dc)
    del dc
    for c in (d_plain, d_net):  # cloud fractions accumulate onto the caller's values
        for k in ("cltot", "clhgh", "clmed", "cllow"):
            c.arrays["out"][k].zero_()
    hip.run_device(_clone(d_plain), w["dt_gl"])  # warm: the call graph of this batch
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    hip.run_device(d_plain, w["dt_gl"])
    torch.cuda.synchronize()
    sec_plain = time.perf_counter() - t4
    crm_mlp.attach(hip, m)
    torch.cuda.synchronize()
    t5 = time.perf_counter()
    hip.run_device(d_net, w["dt_gl"])
    torch.cuda.synchronize()
    sec = time.perf_counter() - t5
    crm_mlp.attach(hip, None)
    out_p, out_n = d_plain.arrays["out"], d_net.arrays["out"]
    errs = {}
    fields = {"heating": ("sltend",), "moistening": ("qltend",), "precipitation": ("precc", "precl")}
    for name, parts in fields.items():
        a = sum(out_n[p] for p in parts)
        b = sum(out_p[p] for p in parts)
        errs[name] = {"max": (a - b).abs().max().item() / b.abs().max().item(),
                      "rms": ((a - b).norm() / b.norm()).item()}
    # the evaluation step convects: cloud, precipitation production and surface rain
    activity = {"qp_src_max": out_p["qp_src"].abs().max().item(),
                "cltot_mean": out_p["cltot"][:ncrms].mean().item(),
                "precip_mean_mm_per_day": 8.64e7 * (out_p["precc"] + out_p["precl"]).mean().item(),
                "qn_max": d_plain.arrays["state"]["qn"].max().item()}
    hip.close()
    m.close()
    return {"value": ncrms * w["nx"] * nstop / sec, "unit": "column-steps/s",
            "ms_per_step": 1e3 * sec, "plain_ms_per_step_same_state": 1e3 * sec_plain,
            "dtype": "f64 state, bf16 network with f32 accumulation",
            "training": {"pairs": npairs, "stride": stride, "record_gcm_steps": args.learn_record,
                         "spinup_gcm_steps": args.learn_spinup, "adam_steps": args.learn_steps,
                         "batch": 16384, "seed": 7, "loss_first": hist[0],
                         "loss_last50": float(np.mean(hist[-50:])),
                         "record_s": t1 - t0, "train_s": t3 - t2},
            "evaluation_step_activity": activity,
            "domain_mean_error_vs_plain": errs,
            "note": "network 9-128-128-128-6 in place of micro_proc's adjustment and processes "
                    "inside its training envelope (plain scheme outside), water kept exactly by "
                    "the fixer; plain and learned GCM steps from the same convecting state"}


def state_error(d64, d32, torch):
    """Normwise relative difference max|a-b|/max|a| of the mixed-precision state and
    domain-mean outputs against the FP64 ones, computed on the device."""
    out = {}
    for group, names in (("state", ("u_wind", "w_wind", "temperature", "qt", "qp", "qn")),
                         ("out", ("sltend", "qltend", "precc", "precl", "cltot"))):
        for n in names:
            a, b = d64.arrays[group][n], d32.arrays[group][n]
            scale = a.abs().max().item()
            diff = (a - b).abs().max().item()
            out[f"{group}.{n}"] = diff / scale if scale > 0 else diff
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ncrms", type=int, default=WORKLOAD["ncrms"])
    ap.add_argument("--learn-spinup", type=int, default=2,
                    help="GCM steps of the convecting state before the training pairs are recorded")
    ap.add_argument("--learn-record", type=int, default=3,
                    help="GCM steps over which the training pairs are recorded")
    ap.add_argument("--learn-pairs", type=int, default=200_000_000,
                    help="training pairs to record (a subsample of every point of every substep)")
    ap.add_argument("--learn-steps", type=int, default=12000, help="Adam steps of 16384 pairs")
    ap.add_argument("--profile-steps", type=int, default=2,
                    help="GCM steps of the separate pass that times every kernel with HIP events")
    ap.add_argument("--cpu-per-proc", type=int, default=16,
                    help="instances per process in the CPU baseline sample (0 skips it)")
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="CPU baseline processes (default: the cores this process may use, "
                         "at most 16)")
    ap.add_argument("--fp64-only", action="store_true",
                    help="skip the mixed-precision measurement")
    ap.add_argument("--alone-steps", type=int, default=3,
                    help="GCM steps per class for the uneven batch's per-class and joint-graph times")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven-batch measurement")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-microphysics measurement")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (scale for the HBM "
                         "byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = 1 << 27
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "copies": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(instances are independent: a multi-GPU run is N copies of this one)",
              file=sys.stderr)
        sys.exit(2)

    w = WORKLOAD
    ncrms = args.ncrms
    nstop = int(w["dt_gl"] / w["dt"])
    # The CPU baseline first, before this process touches the GPU.
    cpu = cpu_native = None
    if args.cpu_per_proc > 0:
        procs = args.cpu_procs or min(16, len(os.sched_getaffinity(0)))
        log(f"CPU baseline: {procs} processes of {args.cpu_per_proc} instances")
        cpu = cpu_baseline(args.cpu_per_proc, procs)
        cpu_native = cpu_baseline(args.cpu_per_proc, procs, tag="_native")
        if cpu and cpu_native:
            cpu["o3_x86_64_v3"] = {k: cpu_native[k] for k in ("value", "sample", "build")}

    import torch

    log(f"building inputs: {ncrms} x {w['nx']}x{w['nz']}, {w['nx_rad']} radiation columns")
    case = make_case(ncrms, w["nx"], w["nz"], w["dx"], w["dt"], randomize=True, nx_rad=w["nx_rad"])

    res = {}
    for prec in ("fp64", "mixed"):
        if prec == "mixed" and args.fp64_only:
            continue
        res[prec] = measure(prec, case, args, torch)
    r64 = res["fp64"]
    value = r64["value"]
    whole = STATE_WORDS_PER_LEVEL * 8 * w["nz"] * value / 1e9
    mixed = None
    if "mixed" in res:
        r32 = res["mixed"]
        mixed = {"value": r32["value"], "ms_per_step": r32["ms_per_step"],
                 "speedup_vs_fp64": r32["value"] / value, "roofline": r32["roofline"],
                 "error_vs_fp64": state_error(r64["dcase"], r32["dcase"], torch),
                 "kernels": r32["kernels"], "autotuned": r32["schedule"],
                 "note": "FP32 bulk fields, advection, SGS and microphysics; FP64 pressure solve, "
                         "reference profiles, statistics and outputs; error after the same "
                         f"{max(1, args.warmup) + args.steps + args.profile_steps} GCM steps "
                         "from the same initial state"}

    learned = None
    if not args.no_learned:
        learned = measure_learned(args, torch)

    uneven = None
    if not args.no_uneven:
        uneven = {p: measure_uneven(p, args, torch) for p in (("fp64",) if args.fp64_only else ("fp64", "mixed"))}

    line = {
        "metric": "CRM column-steps/s on 1 MI355X, FP64 vs mixed-prec at stated error; HBM %peak",
        "value": value,
        "unit": "column-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded tropical soundings per instance",
        "config": {"workload": f"{ncrms} 2D CRMs of 64x58 (SP1V1: radiation on {w['nx_rad']} "
                               f"column groups), dx 1 km, dt 5 s, one 1200 s GCM step "
                               f"({nstop} substeps) per call, inputs resident in HBM, default "
                               f"path (one HIP graph per call), autotuned launch shapes",
                   "ncrms": ncrms, "nx": w["nx"], "nz": w["nz"], "nx_rad": w["nx_rad"],
                   "substeps": nstop, "precision": "fp64", "max_ncycle": r64["max_ncycle"],
                   "max_nprec": r64["max_nprec"],
                   "graph_substeps_timed": r64["graph_substeps_timed"],
                   "autotuned": r64["schedule"]},
        "roofline": r64["roofline"],
        "whole_step_hbm": {"achieved": whole, "frac": whole / HBM_PEAK_GBS, "unit": "GB/s",
                           "bytes_per_column_substep": STATE_WORDS_PER_LEVEL * 8 * w["nz"]},
        "kernels": r64["kernels"],
        "mixed_precision": mixed,
        "uneven_batch": uneven,
        "learned_microphysics": learned,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
