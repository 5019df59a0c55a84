#!/usr/bin/env python3
"""Throughput of the SPH step pipeline on one MI355X.

Workload: coupled thermo-mechanical Johnson-Cook cylinder compression,
R 15 mm, H 30 mm, dx 0.3 mm (759,132 particles from the reference's
AddCylinderLength), aluminium-like JC with FP64 Newton return map,
Taylor-Quinney 0.9 plastic heating with conduction, top layers at -2 m/s.
The timed region is the last --steps steps of the deck's 3,000 (default
200; the steps before them run untimed, at least --warmup), so the window
sits in the yielding regime the deck ends in, with the Newton return map and
the surrogate at work.  One online training round of the surrogate runs
every 200 steps; its cost is charged to the window amortized over those
200 steps (measured on its own), whatever number of rounds the window
happens to contain.  A step is one full pass of the Fraser step (density
rate + update, rate tensors, pressure + stress + return map + energy,
acceleration, BC + integration, conduction + leapfrog, neighbour rebuild
every 5 steps once yielding), all inputs resident in HBM.

--workload c4 runs the rigid-plate contact workload instead: the same JC
cylinder at dx 0.25 mm (1,320,424 particles) compressed by a rigid 40 x 40 mm
plate of 2 x 64 x 64 triangles at -2 m/s, penalty stiffness 0.1 E/dx,
Coulomb friction 0.2, the online-trained surrogate in the mixed run.
--workload c5 runs the friction-stir-like plate: 120 x 60 x 6 mm of the JC
aluminium at dx 0.25 mm (2,738,136 particles once the pin's volume is left
out) under a rigid tool of 12,672 triangles (shoulder R 9 mm, pin R 3 mm x
5 mm) turning at 1000 rpm and travelling at 2 mm/s, penalty 0.1 E/dx,
Coulomb 0.2 with all of the sliding-friction power as heat, bottom layer at
300 K, lists with a 0.4 dx skin rebuilt on the displacement of particles with
eps_p > 0.05.

Prints ONE JSON line: value = particle-updates/s (particles x steps / s),
plus the HBM roofline of the dominant kernel (algorithmic bytes over its
average launch in the timed steps: each pair pass records its own dispatch
start and end in HIP events on the library's stream, as a kernel trace
reports them; the other kernels of the per-kernel table are bracketed by
event pairs; avg_ms_back_to_back repeats the rates and accel passes 20
times in a row on warm caches) and the FP64 CPU oracle timed on the
host cores on a bounded sample of the same workload (on all the threads
the process may use, and on one).  error_vs_fp64 compares the mixed run's
fields with the FP64 run's at the end of the window (the deck's last step);
error_floor_fp64_perturbed is the same comparison between the FP64 run and
an FP64 run whose initial positions carry 1e-12 m of noise (how far two FP64
runs of the workload drift apart).  mixed_no_surrogate is the mixed run
without the surrogate, over the same window: the surrogate's net cost or
gain per step is the difference.
"""
import argparse
import json
import os
import sys
import time
import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-updates/s: mixed+surrogate vs plain FP64 @ stated error; HBM % peak"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per particle-update of each kernel (FP64 plain path):
# compulsory SoA reads + writes, neighbour data counted once, list excluded
ALG_BYTES_FP64 = {"dens": 80, "rates": 144, "stress": 384, "accel": 144, "integrate": 172,
                  "thermal": 120, "cond_dens": 200,
                  # the acceleration pass integrating its particles: both
                  # kernels' bytes less the acceleration's read back (24 B)
                  "accel_integrate": 144 + 172 - 24}
# mixed (FP32 pair-loop inputs and outputs), same accounting
ALG_BYTES_MIXED = {"dens": 40, "rates": 72, "stress": 204, "accel": 72, "integrate": 184,
                   "thermal": 60, "cond_dens": 100, "accel_integrate": 72 + 184 - 12}


def _cpu_info():
    """Host CPU: model name, logical CPUs visible, physical cores (from
    /proc/cpuinfo's physical id / core id pairs)."""
    model, cores, phys, core = None, set(), None, None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)), "physical_cores": len(cores) or None}


def cpu_baseline(sc, warm=5, steps=20, warm1=1, steps1=2):
    """FP64 oracle (OpenMP over particles) on the same workload from its
    initial state: `steps` timed steps after `warm` on all the host threads
    this process may use, and a shorter sample on one thread."""
    from oracle import pyoracle as po

    def timed(threads, w, k):
        po.set_threads(threads)
        st = po.State(sc.n)
        for key, v in sc.fields.items():
            st[key] = v
        st["xs"] = sc.fields["x"]
        if sc.meshes:
            st.add_contact(sc.meshes)
        st.nsearch(sc.cfg)
        st.step(sc.cfg, sc.bcs, w)
        t0 = time.perf_counter()
        st.step(sc.cfg, sc.bcs, k)
        return sc.n * k / (time.perf_counter() - t0)

    info = _cpu_info()
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = env or info["affinity"]
    rate_all = timed(threads, warm, steps)
    rate_one = timed(1, warm1, steps1)
    po.set_threads(threads)
    return {"value": rate_all, "unit": "particle-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps after {warm} warm-up steps of the same {sc.n}-particle "
                      f"workload from its initial state, FP64 oracle, OpenMP on {threads} "
                      f"threads",
            "single_thread": {"value": rate_one, "cores": 1,
                              "sample": f"{steps1} steps after {warm1} warm-up step(s), "
                                        f"OMP 1 thread"},
            "cpu_model": info["model"], "nproc": info["nproc"],
            "physical_cores": info["physical_cores"]}


def profiled_traffic(workload, kernel_id):
    """HBM bytes per launch of a kernel from the committed rocprofv3 summary of
    this workload (FETCH_SIZE + WRITE_SIZE, each scaled by what the counter
    reads for a copy kernel of known size), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"{workload}_jc_step_kernels_traffic.json")
    try:
        with open(path) as fh:
            k = json.load(fh)["kernels"].get(kernel_id)
    except (OSError, ValueError, KeyError):
        return None
    return None if k is None else k["fetch_bytes"] + k["write_bytes"]


def workload_text(sc, n):
    common = ("mixed precision (FP32 pair loops summed in FP64, FP64 stress/Newton return "
              "map/integration) with the online-trained MLP Newton guess (BF16 MFMA), "
              "conduction + Taylor-Quinney 0.9")
    if sc.name.startswith("c5"):
        ne = sum(len(m["elnode"]) for m in sc.meshes)
        how = ("local re-lists near the plastic zone"
               if sc.cfg.rebuild_policy == 2 else "plastic-zone list rebuilds")
        return (f"{sc.name}: JC plate 120x60x6mm dx0.25mm, {n} particles, rigid tool of {ne} "
                f"triangles (shoulder R9mm, pin R3mm x 5mm) at 1000 rpm and 2 mm/s, penalty "
                f"0.1 E/dx, Coulomb 0.2 with friction heat, {how}, "
                f"{common}; compared with plain FP64")
    if sc.meshes:
        ne = sum(len(m["elnode"]) for m in sc.meshes)
        return (f"{sc.name}: JC cylinder R15mm H30mm dx0.25mm, {n} particles, rigid plate "
                f"40x40mm of {ne} triangles at -2 m/s, penalty 0.1 E/dx, Coulomb 0.2, "
                f"{common}; compared with plain FP64")
    return (f"{sc.name}: JC thermo-mechanical cylinder R15mm H30mm dx0.3mm, {n} particles, "
            f"{common}, top at -2 m/s; compared with plain FP64")


DECK_STEPS = {"c3": 3000, "c4": 3000, "c5": 20000}


def calibration_copy(nbytes):
    """One device-to-device copy of a cold buffer of nbytes (hipMemcpy's copy
    kernel, which reads the source once and writes the destination once),
    after a fill of a second buffer of the same size has pushed the source
    out of L2 and the MALL: a plain copy kernel of known size for
    calibrating rocprofv3's FETCH_SIZE and WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    bufs = [C.c_void_p() for _ in range(3)]
    for b in bufs:
        if hip.hipMalloc(C.byref(b), nbytes) != 0:
            raise RuntimeError("hipMalloc failed for the calibration copy")
    src, dst, evict = bufs
    ok = (hip.hipMemset(src, 1, nbytes) == 0 and hip.hipMemset(evict, 2, nbytes) == 0 and
          hip.hipDeviceSynchronize() == 0 and hip.hipMemcpy(dst, src, nbytes, 3) == 0 and
          hip.hipDeviceSynchronize() == 0)
    for b in bufs:
        hip.hipFree(b)
    if not ok:
        raise RuntimeError("calibration copy failed")
    return {"bytes_read": nbytes, "bytes_written": nbytes,
            "kernel": "hipMemcpy device-to-device copy kernel"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200,
                    help="timed steps: the last STEPS steps of the deck")
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed steps before the window at least (the window's start is "
                         "max(warmup, deck steps - steps))")
    ap.add_argument("--deck-steps", type=int, default=None,
                    help="step count of the deck whose last steps are timed (c3/c4: 3000, "
                         "c5: 20000)")
    ap.add_argument("--timing-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--workload", choices=["c3", "c4", "c5"], default="c3")
    ap.add_argument("--rebuild", choices=["global", "local"], default="local",
                    help="c5: on the plastic-zone trigger, rebuild every tile's lists with a "
                         "reorder (global) or re-list in place the tiles near the plastic zone "
                         "(local)")
    ap.add_argument("--no-error-floor", action="store_true",
                    help="skip the FP64 run with 1e-12 m position noise that measures how "
                         "far two FP64 runs drift apart on this workload")
    ap.add_argument("--no-plain-mixed", action="store_true",
                    help="skip the mixed-precision run without the surrogate")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="after the runs, one device-to-device copy of a cold buffer of this "
                         "many MiB (read once, written once; a second buffer of the same size "
                         "is filled first so that the source is out of L2 and the MALL): a "
                         "plain copy kernel of known size for calibrating rocprofv3's "
                         "FETCH_SIZE and WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the step "
              "runs on a single GPU; sharding by space is designed, not built)",
              file=sys.stderr)
        sys.exit(2)

    from weldformgpu_amd import _abi, scenarios
    from weldformgpu_amd.domain import Domain

    deck = args.deck_steps or DECK_STEPS[args.workload]
    K = args.steps
    start = max(args.warmup, deck - K)  # untimed steps before the window

    def make():
        if args.workload == "c4":
            return scenarios.c4_contact(scale=args.scale)
        if args.workload == "c5":
            sc = scenarios.c5_friction_stir(scale=args.scale, offset=(0.137, 0.291))
            if args.rebuild == "local":
                sc.cfg.rebuild_policy = _abi.WF_REBUILD_PLASTIC_ZONE_LOCAL
            return sc
        return scenarios.c3_johnson_cook(scale=args.scale)

    def run(precision, surrogate, noise=0.0, timing_steps=0):
        sc = make()
        if noise:
            rng = np.random.default_rng(1)
            sc.fields["x"] = sc.fields["x"] + rng.normal(0.0, noise, sc.fields["x"].shape)
        sc.cfg.precision = precision
        sc.cfg.surrogate = surrogate
        dom = Domain(sc.cfg)
        scenarios.load_domain(dom, sc)
        print(f"[bench] {sc.name} precision {precision} surrogate {surrogate}: {start} "
              f"untimed steps", file=sys.stderr, flush=True)
        dom.step(start)
        dom.sync()
        s0 = dom.stats()
        t0 = time.perf_counter()
        dom.step(K)
        dom.sync()
        t1 = time.perf_counter()
        s1 = dom.stats()
        res = {"sc": sc, "secs": t1 - t0, "s0": s0, "s1": s1,
               "fields": {k: dom.get(k) for k in ["u", "sigma", "sigma_eq", "pl_strain", "T"]}}
        print(f"[bench]   window of {K} steps: {res['secs'] * 1e3 / K:.3f} ms/step",
              file=sys.stderr, flush=True)
        if timing_steps:
            # per-kernel HIP-event durations on the library stream, after the
            # window (pair passes: their own dispatch timestamps)
            dom.set_timing(True)
            dom.step(timing_steps)
            kern = {}
            alg = ALG_BYTES_FP64 if precision == _abi.WF_FP64 else ALG_BYTES_MIXED
            for k, b in alg.items():
                ms, cnt = dom.kernel_time(k)
                if cnt:
                    avg = ms / cnt
                    kern[k] = {"avg_ms": avg, "launches": cnt, "alg_bytes": b * sc.n,
                               "GBs": b * sc.n / (avg * 1e-3) / 1e9}
            for k in ("return", "mlp", "nsearch", "contact"):
                ms, cnt = dom.kernel_time(k)
                if cnt:
                    kern[k] = {"avg_ms": ms / cnt, "launches": cnt}
            dom.set_timing(False)
            # the pair passes again, 20 launches back to back (warm caches)
            for k in ("rates", "accel"):
                if k in kern:
                    avg = dom.time_pass(k, 20)
                    kern[k]["avg_ms_back_to_back"] = avg
                    kern[k]["GBs_back_to_back"] = kern[k]["alg_bytes"] / (avg * 1e-3) / 1e9
            res["kern"] = kern
            # known-size copy kernel for HBM-counter calibration: the x -> x_h
            # snapshot copies 3 planes of n doubles (24 n bytes read, 24 n written)
            for _ in range(3):
                dom.snapshot_positions()
            dom.sync()
        if surrogate:
            # one online training round (the same round wf_step runs every
            # mlp_train_interval steps), timed alone, three times
            dom.sync()
            t = time.perf_counter()
            for _ in range(3):
                dom.mlp_train()
            dom.sync()
            res["train_ms"] = (time.perf_counter() - t) * 1e3 / 3
        dom.close()
        return res

    def charged(r):
        """Window time with its training rounds replaced by the rounds' cost
        amortized over the training interval in force at the window's end
        (the cadence backs off while the network keeps fitting): a 20-step
        window that happens to end on a training step pays 20/interval of a
        round, not a whole one."""
        n_train = r["s1"]["mlp_trainings"] - r["s0"]["mlp_trainings"]
        t_train = r.get("train_ms", 0.0) * 1e-3
        interval = max(1, r["s1"]["mlp_interval"])
        amort = t_train * K / interval if r.get("train_ms") is not None else 0.0
        return r["secs"] - n_train * t_train + amort, n_train, interval

    r64 = run(_abi.WF_FP64, 0, timing_steps=args.timing_steps)
    sc = r64["sc"]
    n = sc.n
    rmx = run(_abi.WF_MIXED, 1, timing_steps=args.timing_steps)
    secs_mx, n_train_window, interval = charged(rmx)
    ms_step = secs_mx * 1e3 / K
    rate = n * K / secs_mx
    rate64 = n * K / r64["secs"]
    rplain = None if args.no_plain_mixed else run(_abi.WF_MIXED, 0,
                                                  timing_steps=args.timing_steps)

    def rl2(a, b):
        d = np.linalg.norm(b)
        return float(np.linalg.norm(a - b) / d) if d > 0 else float(np.linalg.norm(a - b))

    def field_err(r):
        e = {k: rl2(r["fields"][k], r64["fields"][k]) for k in ["u", "sigma", "sigma_eq",
                                                                "pl_strain"]}
        e["T_rise"] = rl2(r["fields"]["T"] - 300.0, r64["fields"]["T"] - 300.0)
        return e
    err = field_err(rmx)
    # the same comparison between two FP64 runs whose initial positions differ
    # by 1e-12 m noise: the drift the workload itself allows
    floor = None
    if not args.no_error_floor:
        floor = field_err(run(_abi.WF_FP64, 0, noise=1e-12))
    kern = rmx["kern"]
    stats = rmx["s1"]
    dom_k = max((k for k in kern if k in ALG_BYTES_MIXED), key=lambda k: kern[k]["avg_ms"])
    dk = kern[dom_k]
    # the dominant kernel's average launch in the timed steps
    avg_ms = dk["avg_ms"]
    achieved = dk["alg_bytes"] / (avg_ms * 1e-3) / 1e9
    roof = {"kernel": dom_k, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": profiled_traffic(args.workload, "mixed/" + dom_k),
            "alg_bytes_per_particle": ALG_BYTES_MIXED[dom_k], "avg_ms": avg_ms,
            "avg_ms_back_to_back": dk.get("avg_ms_back_to_back")}

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(sc, warm=args.cpu_warmup, steps=args.cpu_steps)

    end_step = start + K
    window = {"first_step": start, "last_step": end_step - 1, "deck_steps": deck,
              "yielding_fraction_at_end": float((rmx["fields"]["pl_strain"] > 0).mean())}
    out = {
        "metric": METRIC, "value": rate, "unit": "particle-updates/s", "n_gpus": 1,
        "steps": K, "warmup": start, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": workload_text(sc, n),
                   "particles": n, "precision": "mixed", "surrogate": True,
                   "parallelism": "single-gpu"},
        "window": window,
        "fp64": {"value": rate64, "ms_per_step": r64["secs"] * 1e3 / K,
                 "kernels": r64["kern"]},
        "speedup_vs_fp64": rate / rate64,
        "error_vs_fp64": err,
        "error_at_step": end_step,
        "error_floor_fp64_perturbed": floor,
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": kern,
        "surrogate": {
            "train_round_ms": rmx.get("train_ms"),
            "train_interval_steps": interval,
            "train_interval_base": sc.cfg.mlp_train_interval,
            "train_interval_max": sc.cfg.mlp_interval_max,
            "train_ms_per_step_amortized": rmx.get("train_ms", 0.0) / interval,
            "train_ms_per_step_run_average": rmx.get("train_ms", 0.0) *
            rmx["s1"]["mlp_trainings"] / max(1, rmx["s1"]["step"]),
            "trainings_in_window": n_train_window,
            "window_ms_per_step_measured": rmx["secs"] * 1e3 / K,
            "mlp_start_loss": rmx["s1"]["mlp_start_loss"],
        },
        "run": {"newton_calls": stats["newton_calls"], "newton_iters": stats["newton_iters"],
                "newton_iters_per_call": stats["newton_iters"] / max(1, stats["newton_calls"]),
                "fp64_plain_newton_iters_per_call": r64["s1"]["newton_iters"] /
                max(1, r64["s1"]["newton_calls"]),
                "mlp_trainings": stats["mlp_trainings"], "mlp_loss": stats["mlp_loss"],
                "ns_builds": stats["ns_builds"], "is_yielding": stats["is_yielding"],
                "ns_local_builds": stats["ns_local_builds"],
                "ns_local_tiles": stats["ns_local_tiles"],
                "newton_iters_wave": stats["newton_iters_wave"]},
    }
    if rplain is not None:
        sp = rplain["s1"]
        out["mixed_no_surrogate"] = {
            "value": n * K / rplain["secs"], "ms_per_step": rplain["secs"] * 1e3 / K,
            "speedup_vs_fp64": (n * K / rplain["secs"]) / rate64,
            "newton_iters_per_call": sp["newton_iters"] / max(1, sp["newton_calls"]),
            "error_vs_fp64": field_err(rplain),
            "surrogate_net_ms_per_step": ms_step - rplain["secs"] * 1e3 / K,
            "stress_kernel_ms": rplain["kern"].get("stress", {}).get("avg_ms")}
    if args.workload == "c4":
        out["note"] = ("on this workload two FP64 runs 1e-12 m apart already differ by ~9e-4 "
                       "in eps_p and T - T0 after 3,000 steps (error_floor_fp64_perturbed), so "
                       "the 1e-4 target is below the workload's own floor for those fields")
    if args.workload == "c5" and deck > 9000:
        out["note"] = ("past step ~8,300 this workload has two outcomes at the pin (a particle "
                       "chattering against the tool keeps or loses contact): mixed runs that "
                       "differ only by 1e-12 m of initial position noise land on either, with "
                       "u errors against FP64 of ~5e-5 or ~2e-2 at step 20,000; FP64 runs 1e-12 m "
                       "apart stay on one (error_floor_fp64_perturbed)")
    if args.calibration_copy_mib:
        out["calibration_copy"] = calibration_copy(args.calibration_copy_mib << 20)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
