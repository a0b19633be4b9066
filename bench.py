"""Throughput benchmark of the per-step PBF solve on one MI355X.

Default workload: the 1,000,000-particle dam break (100^3 jittered lattice at
spacing h/2, h = 0.1, box 12 x 6 x 6, PAPER physics, 4 constraint
iterations, FP32).  A step is one pass of the hot path: predict, cell sort,
4 x (lambda, delta), XSPH + vorticity, position update, launched as one
captured hipGraph.  Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--config 1] [--physics paper]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "position-based-fluids_amd"))

import numpy as np  # noqa: E402

import pbf  # noqa: E402
import scenarios  # noqa: E402

HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) HBM bytes per particle per launch of each kernel
# class, FP32.  PAPER passes stream a neighbour list of L = ceil(n/11)
# 16-byte groups (n neighbours, 11 bits each); "delta" is the mean over the
# I launches (the last one also reads rho and x_n and writes v).
LIST_PER_GROUP = 11
LIST_MAX = 99


def list_bytes(counts) -> float:
    """Mean list bytes per particle for neighbour counts `counts` (lists
    longer than LIST_MAX entries are not stored)."""
    c = np.asarray(counts, np.float64)
    return float(np.mean(np.where(c <= LIST_MAX,
                                  16.0 * np.ceil(c / LIST_PER_GROUP), 0.0)))


def algorithmic_bytes(cls: str, iterations: int, physics: int,
                      lb: float = 80.0) -> float:
    last_extra = 36.0 / iterations
    if physics == pbf.PHYSICS_PAPER:
        table = {
            "predict_count": 56.0,      # x, v in; x*, key, rank out
            "scatter": 24.0,
            "reorder": 88.0,
            # own record in; (x*, lambda), rho, count, 9 row starts, list out
            "search": 76.0 + lb,
            # x, count, 9 row starts, list in; (x, lambda), rho out
            "lambda": 76.0 + lb,
            # (x, lambda), count, 9 row starts, list in; x out (the last
            # pass also reads rho and x_n and writes the 32-byte x, v record)
            "delta": 72.0 + lb + last_extra,
            # own (x, rho, v, m), count, row starts, list in; v', omega, (x, |w|) out
            "xsph": 120.0 + lb,
            # (x, |w|), omega, v', count, row starts, list, id in; x, v, id out
            "vorticity": 128.0 + lb,
        }
    else:
        table = {
            "predict_count": 56.0, "scatter": 24.0, "reorder": 88.0,
            "lambda": 24.0, "delta": 36.0 + last_extra, "xsph": 84.0,
        }
    return table.get(cls, 0.0)


def step_bytes(n: int, cells: int, iterations: int, physics: int,
               lb: float = 80.0) -> float:
    """Compulsory bytes of one step: the sum of the per-launch figures above
    plus 16 bytes per cell (count, scan, table)."""
    per = sum(algorithmic_bytes(c, iterations, physics, lb)
              for c in ("predict_count", "scatter", "reorder"))
    if physics == pbf.PHYSICS_PAPER:
        per += (algorithmic_bytes("search", iterations, physics, lb)
                + (iterations - 1) * algorithmic_bytes("lambda", iterations, physics, lb)
                + iterations * algorithmic_bytes("delta", iterations, physics, lb)
                + algorithmic_bytes("xsph", iterations, physics, lb)
                + algorithmic_bytes("vorticity", iterations, physics, lb))
    else:
        per += iterations * (algorithmic_bytes("lambda", iterations, physics)
                             + algorithmic_bytes("delta", iterations, physics))
        per += algorithmic_bytes("xsph", iterations, physics)
    return n * per + 16.0 * cells


def load_traffic(workload: str):
    """Measured HBM bytes per launch of each kernel class (FETCH_SIZE +
    WRITE_SIZE per dispatch, scaled by the copy-kernel calibration), written
    by --rocprof-summary ... --traffic-out."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


# rocprof kernel name -> kernel class of pbf_profile_steps
def kernel_class(name: str):
    for prefix, cls in (("k_lambda_paper", "search"), ("k_lambda_list", "lambda"),
                        ("k_lambda_ref", "lambda"), ("k_delta", "delta"),
                        ("k_xsph", "xsph"), ("k_vort", "vorticity"),
                        ("k_predict_count", "predict_count"), ("k_scan", "scan"),
                        ("k_scatter", "scatter"), ("k_reorder", "reorder")):
        if name.startswith(prefix):
            return cls
    return None


def cpu_baseline(params, parts, steps: int, threads: int):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pbf_oracle as orc

    orc.set_threads(threads)
    pos, vel = orc.state_from_particles(parts)
    t0 = time.perf_counter()
    orc.run_f32(params, pos, vel, steps)
    dt = time.perf_counter() - t0
    return parts.shape[0] * steps / dt, dt


def _short(name: str) -> str:
    name = name.replace("void ", "").replace("pbf::", "")
    return name.split("(")[0]


def rocprof_summary(root: str, particles: int, copy_mib: int,
                    traffic_out: str = None) -> str:
    """Markdown summary of the rocprofv3 databases written by the profiling
    pass (sub-directories kt, fetch, write, valu, cfetch, cwrite); with
    traffic_out, also writes the per-class HBM bytes per launch as JSON."""
    import glob
    import sqlite3

    def db(sub):
        files = glob.glob(os.path.join(root, sub, "*.db"))
        return sqlite3.connect(files[0]) if files else None

    out = []
    kt = db("kt")
    times = {}
    if kt:
        for name, calls, avg, tot in kt.execute(
                "select name, count(*), avg(duration), sum(duration) from kernels "
                "group by name order by sum(duration) desc"):
            times[_short(name)] = (calls, avg / 1e3, tot / 1e6)

    def counter(con, cname):
        res = {}
        if con is None:
            return res
        for name, avg in con.execute(
                "select kernel_name, avg(v) from (select kernel_name, dispatch_id, "
                "sum(value) as v from counters_collection where counter_name = ? "
                "group by dispatch_id) group by kernel_name", (cname,)):
            res[_short(name)] = avg
        return res

    fetch = counter(db("fetch"), "FETCH_SIZE")
    write = counter(db("write"), "WRITE_SIZE")
    cfetch = counter(db("cfetch"), "FETCH_SIZE").get("k_copy")
    cwrite = counter(db("cwrite"), "WRITE_SIZE").get("k_copy")
    copy_kib = copy_mib * 1024.0
    fscale = copy_kib / cfetch if cfetch else None
    wscale = copy_kib / cwrite if cwrite else None
    valu_db = db("valu")
    valu = counter(valu_db, "SQ_INSTS_VALU")
    vmem = counter(valu_db, "SQ_INSTS_VMEM_RD")
    waves = counter(valu_db, "SQ_WAVES")

    out.append(f"copy calibration ({copy_mib} MiB read + write): FETCH_SIZE "
               f"{cfetch} KiB, WRITE_SIZE {cwrite} KiB -> scale factors "
               f"fetch {fscale}, write {wscale}")
    out.append("")
    out.append("| kernel | calls | avg us | total ms | HBM read B/particle | "
               "HBM write B/particle | VALU instr/particle | VMEM rd instr/particle |")
    out.append("|---|---|---|---|---|---|---|---|")
    for k, (calls, avg_us, tot_ms) in times.items():
        def per(v, s=1.0):
            return "-" if v is None or s is None else f"{v * 1024.0 * s / particles:.1f}"
        vi = valu.get(k)
        vm = vmem.get(k)
        out.append(f"| {k} | {calls} | {avg_us:.2f} | {tot_ms:.3f} | "
                   f"{per(fetch.get(k), fscale)} | {per(write.get(k), wscale)} | "
                   f"{'-' if vi is None else f'{vi * 64 / particles:.0f}'} | "
                   f"{'-' if vm is None else f'{vm * 64 / particles:.0f}'} |")
    if traffic_out and fscale and wscale:
        per_cls = {}
        for k in times:
            cls = kernel_class(k)
            if cls and k in fetch and k in write:
                b = (fetch[k] * fscale + write[k] * wscale) * 1024.0
                calls = times[k][0]
                n, tot = per_cls.get(cls, (0, 0.0))
                per_cls[cls] = (n + calls, tot + b * calls)
        with open(traffic_out, "w") as f:
            json.dump({"particles": particles,
                       "bytes_per_launch": {c: tot / n for c, (n, tot) in per_cls.items()}},
                      f, indent=1)
    return "\n".join(out)


def batch_vs_sequential(physics: int, steps: int, nscenes: int = 256):
    """256 independent scenes (N_s log-uniform in [2048, 65536]) advanced by
    one batched handle (one launch per kernel for all scenes) against the
    same scenes as separate handles stepped one after another."""
    sc = scenarios.batch_config(nscenes=nscenes, physics=physics)
    n = sc.n
    batch = pbf.BatchSimulation(sc.params, sc.particles)
    batch.step(2)
    batch.synchronize()
    ms_batch = batch.time_steps(steps)
    sims = [pbf.Simulation(p, parts) for p, parts in zip(sc.params, sc.particles)]
    for s in sims:
        s.step(2)
    for s in sims:
        s.synchronize()
    t0 = time.perf_counter()
    for s in sims:
        s.step(steps)
    for s in sims:
        s.synchronize()
    ms_seq = (time.perf_counter() - t0) * 1e3
    # both ran 2 + steps steps: the batched scenes must be the same bits
    # (hence the same density error) as the scenes run one by one
    same = True
    for k in (0, nscenes // 2, nscenes - 1):
        bp, bv = batch.dump_scene(k)
        sp, sv = sims[k].dump()
        same = same and np.array_equal(bp, sp) and np.array_equal(bv, sv)
    batch.close()
    for s in sims:
        s.close()
    return {
        "scenes": nscenes,
        "particles": n,
        "steps": steps,
        "batched_particle_steps_per_s": n * steps / (ms_batch / 1e3),
        "sequential_particle_steps_per_s": n * steps / (ms_seq / 1e3),
        "speedup": ms_seq / ms_batch,
        "batched_ms_per_step": ms_batch / steps,
        "sequential_ms_per_step": ms_seq / steps,
        "sampled_scenes_bit_identical": bool(same),
    }


def _run_to(params, parts, steps: int, timed: int, warmup: int):
    """Advance a fresh handle `steps` steps, timing `timed` of them (graph
    launches between HIP events) after `warmup`; returns (ms of the timed
    steps, diagnostics, positions) at step `steps`."""
    sim = pbf.Simulation(params, parts)
    sim.step(warmup)
    sim.synchronize()
    ms = sim.time_steps(timed)
    rest = steps - warmup - timed
    if rest > 0:
        sim.step(rest)
    diag = sim.diagnostics()
    pos, _ = sim.dump()
    sim.close()
    return ms, diag, pos


def mixed_vs_plain(timed: int, oracle_threads: int, with_oracle: bool):
    """The 4,194,304-particle sphere dam (c4), PAPER physics: FP32 against
    mixed precision (iteration 0 the FP32 neighbour search, the lambda
    passes of iterations 1-2 over FP16 neighbour positions, iteration 3
    FP32), both run
    50 steps from the same input; the speed of each over `timed` graph
    launches, the GPU FP64 density error after 50 steps, and the position
    RMS between the two.  The FP64 oracle comparison runs on a 65,536-
    particle replica of the same geometry (64 x 32 x 32 block, 6 x 2 x 2
    box, seeded sphere) for 50 steps."""
    sc = scenarios.config(4)
    p32 = sc.params[0]
    pmx = p32.copy()
    pmx.precision, pmx.fp32_iterations = pbf.PRECISION_MIXED, 1
    parts = sc.particles[0]
    warm = 5
    ms32, d32, x32 = _run_to(p32, parts, sc.steps, timed, warm)
    msmx, dmx, xmx = _run_to(pmx, parts, sc.steps, timed, warm)
    n = sc.n
    out = {
        "workload": (f"c4: sphere dam, {n} particles, {p32.iterations} "
                     f"iterations, paper physics, sphere r = "
                     f"{sc.meta['sphere'][3]:.3f}"),
        "particles": n,
        "timed_steps": timed,
        "fp32_particle_steps_per_s": n * timed / (ms32 / 1e3),
        "mixed_particle_steps_per_s": n * timed / (msmx / 1e3),
        "speedup": ms32 / msmx,
        "fp32_max_density_err": d32["max_compression"],
        "mixed_max_density_err": dmx["max_compression"],
        "fp32_mean_density_err": d32["mean_abs_density_err"],
        "mixed_mean_density_err": dmx["mean_abs_density_err"],
        "position_rms_mixed_vs_fp32": float(np.sqrt(np.mean(
            np.sum((xmx[:, :3].astype(np.float64) - x32[:, :3]) ** 2, axis=1)))),
        "steps": sc.steps,
    }
    if with_oracle:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pbf_oracle as orc

        orc.set_threads(oracle_threads)
        small = scenarios.sphere_dam((64, 32, 32), (6.0, 2.0, 2.0), seed=1005,
                                     steps=50)
        q32 = small.params[0]
        qmx = q32.copy()
        qmx.precision, qmx.fp32_iterations = pbf.PRECISION_MIXED, 1
        sp = small.particles[0]
        pos, vel = orc.state_from_particles(sp, np.float64)
        t0 = time.perf_counter()
        op, _ = orc.run_f64(q32, pos, vel, small.steps)
        secs = time.perf_counter() - t0
        rho = orc.density_f64(q32, op)
        res = {"particles": small.n, "steps": small.steps,
               "oracle_seconds": secs, "oracle_threads": oracle_threads,
               "fp64_max_density_err": float(np.abs(rho / q32.rest_density - 1).max())}
        for tag, q in (("fp32", q32), ("mixed", qmx)):
            _, d, x = _run_to(q, sp, small.steps, 1, 1)
            e = x[:, :3].astype(np.float64) - op[:, :3]
            res[f"{tag}_position_rms_vs_fp64"] = float(np.sqrt(np.mean(np.sum(e * e, axis=1))))
            res[f"{tag}_position_max_vs_fp64"] = float(np.abs(e).max())
            res[f"{tag}_max_density_err"] = d["max_compression"]
        out["vs_fp64_oracle"] = res
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", type=int, default=1, choices=[0, 1, 2, 4])
    ap.add_argument("--physics", choices=["paper", "ref"], default="paper")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=20,
                    help="steps of the 256-scene batch vs sequential "
                         "comparison (0: skip)")
    ap.add_argument("--mixed-steps", type=int, default=20,
                    help="timed steps of the c4 mixed-precision vs FP32 "
                         "comparison (0: skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain float4 copy of MIB MiB (counter "
                         "calibration) and print its time")
    ap.add_argument("--traffic-out", metavar="JSON",
                    help="with --rocprof-summary: write per-class HBM bytes "
                         "per launch")
    ap.add_argument("--rocprof-summary", metavar="DIR",
                    help="summarise the rocprofv3 databases under DIR and exit")
    args = ap.parse_args()
    if args.rocprof_summary:
        n = scenarios.config(args.config).n if args.config != 3 else 0
        print(rocprof_summary(args.rocprof_summary, n, 512, args.traffic_out))
        return

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        ms = pbf.calibrate_copy(nbytes, 10)
        print(json.dumps({"copy_bytes": nbytes, "ms": ms,
                          "GBps": 2 * nbytes / (ms / 1e3) / 1e9}), flush=True)
        return

    physics = pbf.PHYSICS_PAPER if args.physics == "paper" else pbf.PHYSICS_REF
    sc = scenarios.config(args.config, physics=physics)
    params, parts = sc.params[0], sc.particles[0]
    if args.block:
        params.block_size = args.block
    n = sc.n
    iters = params.iterations
    print(f"[bench] {sc.name}: {n} particles, {iters} iterations, "
          f"{args.physics} physics, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)

    sim = pbf.Simulation(params, parts)
    cells = sim.num_cells()
    try:
        sim.step(args.warmup)
        sim.synchronize()
        t0 = time.perf_counter()
        ms = sim.time_steps(args.steps)      # HIP events around the graph launches
        sim.synchronize()
    except pbf.PBFError as e:
        if e.code != pbf.ERR_NUMERIC:
            raise
        # REF keeps the reference's unguarded 1/rho_j in XSPH
        # (apply_vorticity_and_viscosity.cl:61); on the 1M-particle dam break
        # it turns a velocity infinite at step 82
        print(f"[bench] the state became non-finite within {args.warmup + args.steps} "
              f"steps ({args.physics} physics): {e}", file=sys.stderr, flush=True)
        sys.exit(3)
    wall = time.perf_counter() - t0
    value = n * args.steps / (ms / 1e3)
    print(f"[bench] timed {args.steps} steps: {ms:.3f} ms (wall {wall*1e3:.1f} ms)",
          file=sys.stderr, flush=True)

    diag = sim.diagnostics()
    prof = sim.profile(args.profile_steps)
    cfg = sim.config()
    # dominant kernel class by time per step among those that can be timed
    # alone; its launch duration from 20 back-to-back launches in a graph
    timeable = [k for k in prof if k in ("search", "lambda", "delta")]
    dom = max(timeable, key=lambda k: prof[k][0])
    avg_ms = sim.time_kernel(dom, 20)
    # neighbour counts at the current state give the mean list length
    lb = list_bytes(sim.probe()["nbr_count"])
    sim.close()

    total_prof = sum(v[0] for v in prof.values())
    dom_launches = prof[dom][1]
    dom_bytes = algorithmic_bytes(dom, iters, physics, lb) * n
    achieved = dom_bytes / (avg_ms / 1e3) / 1e9
    workload = f"c{args.config}"
    traffic = load_traffic(workload)

    b_step = step_bytes(n, cells, iters, physics, lb)
    step_gbs = b_step * args.steps / (ms / 1e3) / 1e9

    out = {
        "metric": ("particle-steps/s @1M (4 iters) & 256-scene batch vs plain, "
                   "equal density err"),
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{sc.name}: dam break, {n} particles, "
                         f"{iters} iterations, {args.physics} physics, FP32"),
            "particles": n,
            "cells": cells,
            "iterations": iters,
            "physics": args.physics,
            "block": cfg[0],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (traffic or {}).get("bytes_per_launch", {}).get(dom),
            "avg_launch_ms": avg_ms,
            "launches_per_step": dom_launches / args.profile_steps,
            "algorithmic_bytes_per_launch": dom_bytes,
        },
        "step_roofline": {
            "list_bytes_per_particle": lb,
            "bytes_per_step": b_step,
            "achieved_GBps": step_gbs,
            "frac": step_gbs / HBM_PEAK_GBS,
        },
        "kernel_ms_per_step": {k: v[0] / args.profile_steps for k, v in prof.items()},
        "profile_overhead": total_prof / args.profile_steps / (ms / args.steps),
        "accuracy": {
            "max_compression": diag["max_compression"],
            "mean_abs_density_err": diag["mean_abs_density_err"],
        },
    }
    if args.batch_steps > 0:
        print(f"[bench] 256-scene batch vs sequential, {args.batch_steps} steps",
              file=sys.stderr, flush=True)
        out["batch"] = batch_vs_sequential(physics, args.batch_steps)
    if args.mixed_steps > 0 and physics == pbf.PHYSICS_PAPER:
        print(f"[bench] c4 mixed precision vs FP32, {args.mixed_steps} timed steps",
              file=sys.stderr, flush=True)
        out["mixed"] = mixed_vs_plain(args.mixed_steps, min(16, os.cpu_count() or 1),
                                      not args.no_cpu)
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cps, secs = cpu_baseline(params, parts, args.cpu_steps, threads)
        out["cpu_baseline"] = {
            "value": cps,
            "unit": "particle-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": (f"oracle FP32 restatement, same {sc.name} input and "
                       f"{args.physics} physics, {args.cpu_steps} step(s), "
                       f"{secs:.1f} s on {threads} threads"),
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
