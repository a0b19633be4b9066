"""Throughput benchmark of the per-step PBF solve on one MI355X.

Default workload: the 1,000,000-particle dam break (100^3 jittered lattice at
spacing h/2, h = 0.1, box 12 x 6 x 6, PAPER physics, 4 constraint
iterations, FP32).  A step is one pass of the hot path: predict, cell sort,
4 x (lambda, delta), XSPH + vorticity, position update, launched as one
captured hipGraph.  Prints one JSON line.

Roofline: algorithmic bytes are the compulsory streams of each pass counted
once (324 + 60 I bytes per particle-step in PAPER physics, 244 + 60 I in REF,
plus 16 bytes per grid cell; 36 bytes per particle for a delta pass, 24 for
a lambda pass), over the measured time, against 8 TB/s.  The neighbour lists
this build streams are reported beside it (`with_lists`), not in it.

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

# Algorithmic (compulsory) HBM bytes per particle of the step's passes: the
# streams each pass must read and write once, neighbour re-reads counted as
# 0 (float4 = 16 B; lambda, rho, key and index 4 B each; the cell table 16 B
# per cell, written and cleared).  These are the figures `roofline` uses.
PASS_BYTES = {
    "predict_hash": 56.0,   # x, v in; x*, key, id out
    "sort": 20.0,           # keys x2, id in; sorted id out
    "reorder": 64.0,        # x, x* in; x, x* out
    "cell_table": 4.0,      # sorted keys (+ 16 B per cell below)
    "lambda": 24.0,         # x* in; lambda, rho out (per iteration)
    "delta": 36.0,          # x*, lambda in; x* out (per iteration)
    "velocity": 48.0,       # x, x* in; v out
    "xsph_paper": 68.0,     # x*, v, rho in; v', omega out
    "xsph_ref": 52.0,       # x*, v, rho in; v' out
    "vorticity": 64.0,      # x*, omega, v' in; v out
}


def step_bytes_per_particle(iterations: int, physics: int) -> float:
    """324 + 60 I bytes per particle (PAPER), 244 + 60 I (REF); the cell
    table adds 16 bytes per cell."""
    b = (PASS_BYTES["predict_hash"] + PASS_BYTES["sort"] + PASS_BYTES["reorder"]
         + PASS_BYTES["cell_table"] + PASS_BYTES["velocity"]
         + iterations * (PASS_BYTES["lambda"] + PASS_BYTES["delta"]))
    if physics == pbf.PHYSICS_PAPER:
        b += PASS_BYTES["xsph_paper"] + PASS_BYTES["vorticity"]
    else:
        b += PASS_BYTES["xsph_ref"]
    return b


def pass_bytes(cls: str) -> float:
    """Algorithmic bytes per particle of one launch of a kernel class that
    can be timed alone (the searching lambda pass, a lambda replay, a delta
    pass)."""
    return {"search": PASS_BYTES["lambda"], "lambda": PASS_BYTES["lambda"],
            "delta": PASS_BYTES["delta"]}[cls]


# The PAPER passes also stream the step's neighbour lists: L = 16 ceil(n/11)
# bytes for n neighbours (11-bit entries), plus the count and the 9 row
# starts.  Reported beside the roofline (`with_lists`), never in it.
LIST_PER_GROUP = 11
LIST_MAX = 99


def list_bytes(counts) -> float:
    """Mean list bytes per particle for neighbour counts `counts` (lists
    longer than LIST_MAX entries are not stored)."""
    c = np.asarray(counts, np.float64)
    return float(np.mean(np.where(c <= LIST_MAX,
                                  16.0 * np.ceil(c / LIST_PER_GROUP), 0.0)))


def list_pass_bytes(cls: str, lb: float) -> float:
    """Bytes per particle of one launch including the list, count and row
    starts it streams (PAPER)."""
    return {"search": 76.0 + lb, "lambda": 76.0 + lb, "delta": 72.0 + lb}[cls]


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


def cpu_info() -> dict:
    """Host core counts and CPU model of the machine the bench runs on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "usable": usable, "cpu_model": model}


def cpu_rate(params, parts, steps: int, threads: int):
    """The FP32 oracle restatement stepped `steps` times on `threads` host
    threads (step loop only); returns (particle-steps/s, seconds)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pbf_oracle as orc

    orc.set_threads(threads)
    pos, vel = orc.state_from_particles(parts)
    t0 = time.perf_counter()
    orc.run_f32(params, pos, vel, steps)
    dt = time.perf_counter() - t0
    return parts.shape[0] * steps / dt, dt


def cpu_baseline(sc, physics_name: str, steps: int, one_core_steps: int,
                 ref_steps: int) -> dict:
    """The oracle's FP32 restatement timed on the host: the benchmark's own
    workload on the host's core share (at most 16 threads: the share of one
    GPU on these machines) and on 1 core, plus the reference kernels'
    physics (REF) on the 32,768-particle dam break (c0)."""
    info = cpu_info()
    threads = max(1, min(16, info["usable"]))
    params, parts = sc.params[0], sc.particles[0]
    rate, secs = cpu_rate(params, parts, steps, threads)
    one, one_secs = cpu_rate(params, parts, one_core_steps, 1)
    c0 = scenarios.config(0, physics=pbf.PHYSICS_REF)
    ref, ref_secs = cpu_rate(c0.params[0], c0.particles[0], ref_steps, threads)
    ref1, ref1_secs = cpu_rate(c0.params[0], c0.particles[0],
                               max(1, ref_steps // 10), 1)
    return {
        "value": rate,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle FP32 restatement, the same {sc.name} input and "
                   f"{physics_name} physics, {steps} step(s), {secs:.1f} s on "
                   f"{threads} threads"),
        "nproc": info["nproc"],
        "usable_cores": info["usable"],
        "cpu_model": info["cpu_model"],
        "one_core": {"value": one, "steps": one_core_steps,
                     "seconds": one_secs},
        "ref_c0": {"value": ref, "threads": threads, "steps": ref_steps,
                   "seconds": ref_secs, "one_core_value": ref1,
                   "one_core_steps": max(1, ref_steps // 10),
                   "workload": "c0: 32,768-particle dam break, REF physics "
                               "(the reference kernels' equations), FP32"},
    }


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
    out.append("")
    out.append("| kernel class | launches | avg us |")
    out.append("|---|---|---|")
    per = {}
    for k, (calls, avg_us, tot_ms) in times.items():
        cls = kernel_class(k)
        if cls:
            c, t = per.get(cls, (0, 0.0))
            per[cls] = (c + calls, t + tot_ms)
    for cls, (c, t) in sorted(per.items(), key=lambda kv: -kv[1][1]):
        out.append(f"| {cls} | {c} | {t * 1e3 / c:.2f} |")
    if kt:
        # run phases (autotuner trials and warm-up, the timed steps, their
        # replay with kernel events ...) from the class launched once per
        # step in both physics modes (xsph): a phase ends where no xsph
        # launch began for 5 ms and ten times its median interval; every
        # class's launches are then averaged per phase
        rows = kt.execute("select name, start, end from kernels "
                          "order by start").fetchall()
        anchor = [s0 for name, s0, _ in rows
                  if kernel_class(_short(name)) == "xsph"]
        gaps = sorted(b - a for a, b in zip(anchor, anchor[1:]))
        cut = max(5e6, 10.0 * gaps[len(gaps) // 2]) if gaps else 0.0
        bounds = [(a + b) / 2 for a, b in zip(anchor, anchor[1:]) if b - a > cut]
        out.append("")
        out.append("| kernel class | phase | launches | avg us |")
        out.append("|---|---|---|---|")
        for cls in ("search", "lambda", "delta", "xsph", "vorticity",
                    "reorder", "scatter", "scan"):
            phase = {}
            for name, s0, e0 in rows:
                if kernel_class(_short(name)) != cls:
                    continue
                k = sum(1 for b in bounds if s0 > b)
                c, t = phase.get(k, (0, 0.0))
                phase[k] = (c + 1, t + (e0 - s0))
            for k in sorted(phase):
                c, t = phase[k]
                out.append(f"| {cls} | {k} | {c} | {t / c / 1e3:.2f} |")
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


def rocprof_compare(roots, particles: int) -> str:
    """Markdown table of the neighbour-pass kernels of several profiled runs
    (each root holding kt, valu and ta sub-directories of rocprofv3
    databases: a kernel trace, the SQ counters and the texture-path
    counters): launches and average duration, VALU and VMEM instructions per
    particle, the share of wave cycles issuing VALU, TA busy cycles and the
    L1 hit rate."""
    import glob
    import sqlite3

    def db(path):
        files = glob.glob(os.path.join(path, "*.db"))
        return sqlite3.connect(files[0]) if files else None

    def counters(con):
        res = {}
        if con is None:
            return res
        for name, cname, v in con.execute(
                "select kernel_name, counter_name, avg(v) from (select "
                "kernel_name, counter_name, dispatch_id, sum(value) as v from "
                "counters_collection group by dispatch_id, counter_name) "
                "group by kernel_name, counter_name"):
            res.setdefault(_short(name), {})[cname] = v
        return res

    out = ["| run | kernel | calls | avg us | VALU instr/particle | "
           "VMEM rd instr/particle | VALU active (of wave cycles) | "
           "TA busy (avg cycles per TA) | L1 hit |",
           "|---|---|---|---|---|---|---|---|---|"]
    for root in roots:
        tag = os.path.basename(os.path.normpath(root))
        kt = db(os.path.join(root, "kt"))
        sq = counters(db(os.path.join(root, "valu")))
        ta = counters(db(os.path.join(root, "ta")))
        rows = kt.execute("select name, count(*), avg(duration), "
                          "sum(duration) from kernels group by name "
                          "order by sum(duration) desc").fetchall()
        for name, calls, avg, _ in rows:
            k = _short(name)
            if kernel_class(k) not in ("search", "lambda", "delta", "xsph",
                                       "vorticity"):
                continue
            c, t = sq.get(k, {}), ta.get(k, {})
            wc = c.get("SQ_WAVE_CYCLES") or 1.0
            acc = t.get("TCP_TOTAL_CACHE_ACCESSES_sum") or 1.0
            out.append(
                f"| {tag} | {k} | {calls} | {avg / 1e3:.1f} | "
                f"{c.get('SQ_INSTS_VALU', 0) * 64 / particles:.0f} | "
                f"{c.get('SQ_INSTS_VMEM_RD', 0) * 64 / particles:.0f} | "
                f"{c.get('SQ_ACTIVE_INST_VALU', 0) / wc:.3f} | "
                f"{t.get('TA_BUSY_avr', 0):.3g} | "
                f"{100 * (1 - t.get('TCP_TCC_READ_REQ_sum', 0) / acc):.0f} % |")
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


def _state_at(params, parts, steps: int):
    """Positions, velocities and GPU diagnostics after `steps` steps of a
    fresh handle."""
    with pbf.Simulation(params, parts) as s:
        if steps:
            s.step(steps)
        pos, vel = s.dump()
        diag = s.diagnostics()
    return pos, vel, diag


def _density_errors(rho, rho0) -> dict:
    """Density error figures of pbf_diagnostics for FP64 densities: the
    largest compression over all particles, and max / mean |rho/rho0 - 1|
    over particles with a neighbour."""
    e = rho / rho0 - 1.0
    has = rho > 0.0
    return {"max_compression": float(max(e.max(), 0.0)),
            "max_abs_density_err": float(np.abs(e[has]).max()) if has.any() else 0.0,
            "mean_abs_density_err": float(np.abs(e[has]).mean()) if has.any() else 0.0}


def _vs_oracle(x, ref) -> dict:
    """Position error of a GPU state against the FP64 oracle's."""
    d = x[:, :3].astype(np.float64) - ref[:, :3]
    dist = np.sqrt(np.sum(d * d, axis=1))
    k = int(np.argmax(dist))
    return {"rms": float(np.sqrt(np.mean(dist * dist))), "max": float(dist[k]),
            "p99": float(np.percentile(dist, 99)), "p999": float(np.percentile(dist, 99.9)),
            "worst_particle": k,
            "worst_gpu_x": [float(v) for v in x[k, :3]],
            "worst_fp64_x": [float(v) for v in ref[k, :3]]}


def vs_oracle(params, parts, steps, threads: int) -> dict:
    """The workload's own input stepped by the FP64 oracle on the host and by
    a fresh GPU handle; at each step count of `steps` (ascending): position
    error of the GPU state against the oracle's, and the FP64 density error
    of both states; the FP32 oracle's own position error against FP64 beside
    it.  REF physics: also whether the GPU state equals the FP32 oracle
    restatement of the reference kernels bit for bit."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pbf_oracle as orc

    orc.set_threads(threads)
    ref = params.physics == pbf.PHYSICS_REF
    p64, v64 = orc.state_from_particles(parts, np.float64)
    p32, v32 = orc.state_from_particles(parts)
    out = {"particles": int(parts.shape[0]), "oracle_threads": threads,
           "steps": {}}
    t_oracle, done = 0.0, 0
    with pbf.Simulation(params, parts) as sim:
        for k in steps:
            if k > done:
                sim.step(k - done)
            x, v = sim.dump()
            t0 = time.perf_counter()
            p64, v64 = orc.run_f64(params, p64, v64, k - done)
            t_oracle += time.perf_counter() - t0
            at = _vs_oracle(x, p64)
            at["gpu"] = _density_errors(
                orc.density_grid_f64(params, x.astype(np.float64)),
                params.rest_density)
            at["fp64"] = _density_errors(orc.density_grid_f64(params, p64),
                                         params.rest_density)
            # the FP32 restatement on the host, against FP64 too: how far
            # FP32 rounding alone carries a trajectory by step k
            p32, v32 = orc.run_f32(params, p32, v32, k - done)
            e32 = _vs_oracle(p32, p64)
            at["fp32_oracle_vs_fp64"] = {"rms": e32["rms"], "max": e32["max"],
                                         "p99": e32["p99"]}
            if ref:
                at["fp32_oracle_bit_identical"] = bool(
                    np.array_equal(x, p32) and np.array_equal(v, v32))
            done = k
            out["steps"][str(k)] = at
            print(f"[bench] vs FP64 oracle at step {k}: rms {at['rms']:.2e} m, "
                  f"max {at['max']:.2e} m", file=sys.stderr, flush=True)
    out["oracle_seconds"] = t_oracle
    return out


def mixed_vs_plain(timed: int, oracle_threads: int, with_oracle: bool):
    """The 4,194,304-particle sphere dam (c4), PAPER physics: FP32 against
    mixed precision, both run 50 steps from the same input; the speed of
    each over `timed` graph launches, the GPU FP64 density error after 50
    steps, and the position RMS between the two.  With the oracle: the FP64
    oracle runs the same 4,194,304 particles for 50 steps on the host, and
    the FP32 and mixed runs are compared with it after 1, 10 and 50 steps
    (position RMS, maximum, and where the largest error sits) together with
    the FP64 density error of each state."""
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
        "fp32_max_abs_density_err": d32["max_abs_density_err"],
        "mixed_max_abs_density_err": dmx["max_abs_density_err"],
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
        pos, vel = orc.state_from_particles(parts, np.float64)
        res = {"particles": n, "oracle_threads": oracle_threads, "steps": {}}
        t_oracle = 0.0
        done = 0
        for k in (1, 10, sc.steps):
            t0 = time.perf_counter()
            pos, vel = orc.run_f64(p32, pos, vel, k - done)
            t_oracle += time.perf_counter() - t0
            done = k
            at = {"fp64": _density_errors(orc.density_grid_f64(p32, pos),
                                          p32.rest_density)}
            for tag, q, x_end in (("fp32", p32, x32), ("mixed", pmx, xmx)):
                x = x_end if k == sc.steps else _state_at(q, parts, k)[0]
                at[tag] = _vs_oracle(x, pos)
                at[tag].update(_density_errors(
                    orc.density_grid_f64(p32, x.astype(np.float64)),
                    p32.rest_density))
            res["steps"][str(k)] = at
            print(f"[bench] c4 vs FP64 oracle at step {k}: rms fp32 "
                  f"{at['fp32']['rms']:.2e} mixed {at['mixed']['rms']:.2e} m",
                  file=sys.stderr, flush=True)
        res["oracle_seconds"] = t_oracle
        out["vs_fp64_oracle"] = res
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", type=int, default=1, choices=[0, 1, 2, 4])
    ap.add_argument("--physics", choices=["paper", "ref"], default="paper")
    ap.add_argument("--precision", choices=["fp32", "mixed"], default="fp32",
                    help="precision of the main workload (the bench line's "
                         "value is defined on fp32)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=60,
                    help="steps of the CPU baseline on the host's core share")
    ap.add_argument("--cpu-one-core-steps", type=int, default=3,
                    help="steps of the CPU baseline on one core")
    ap.add_argument("--cpu-ref-steps", type=int, default=100,
                    help="steps of the REF-physics c0 CPU leg")
    ap.add_argument("--horizon", type=int, default=200,
                    help="steps of the fresh-handle run timed from step 0 "
                         "(0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--vs-fp64", default="", metavar="K1,K2,...",
                    help="compare the workload with the FP64 oracle after "
                         "these step counts (host cores; off by default)")
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
    ap.add_argument("--rocprof-compare", nargs="+", metavar="DIR",
                    help="compare the neighbour-pass kernels of profiled runs "
                         "(kt, valu, ta databases under each DIR) and exit")
    ap.add_argument("--rocprof-summary", metavar="DIR",
                    help="summarise the rocprofv3 databases under DIR and exit")
    args = ap.parse_args()
    if args.rocprof_compare:
        print(rocprof_compare(args.rocprof_compare,
                              scenarios.config(args.config).n))
        return
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
    if args.precision == "mixed":
        params.precision, params.fp32_iterations = pbf.PRECISION_MIXED, 1
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
        # the state the timed steps start from (a dump is a checkpoint)
        start_pos, start_vel = sim.dump()
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
    cfg = sim.config()
    # neighbour counts at the final state give the mean list length
    lb = list_bytes(sim.probe()["nbr_count"])
    sim.close()

    # The timed steps once more, from the state they started from (a handle
    # created from a dump continues bit for bit: every step re-sorts by
    # cell and id), each kernel launched between HIP events on the handle's
    # stream: every kernel class's time and launch count over exactly the
    # timed region, hence the dominant kernel's average launch duration.
    resumed = np.zeros_like(parts)
    resumed[:, 0] = start_vel[:, 3]
    resumed[:, 1:4] = start_pos[:, :3]
    resumed[:, 4:7] = start_vel[:, :3]
    with pbf.Simulation(params, resumed) as again:
        prof = again.profile(args.steps)
    # dominant kernel class by time per step among those that can be timed
    # alone (search, lambda replay, delta)
    timeable = [k for k in prof if k in ("search", "lambda", "delta")]
    dom = max(timeable, key=lambda k: prof[k][0])
    dom_launches = prof[dom][1]
    avg_ms = prof[dom][0] / dom_launches
    total_prof = sum(v[0] for v in prof.values())
    alg = pass_bytes(dom) * n
    achieved = alg / (avg_ms / 1e3) / 1e9
    traffic = (load_traffic(f"c{args.config}") or {}).get("bytes_per_launch", {}).get(dom)
    step_b = step_bytes_per_particle(iters, physics)
    # the cell table of the configuration's grid (x_n y_n z_n cells; the
    # PAPER search bins x four times finer, which is this build's choice)
    grid_cells = int(params.x_n * params.y_n * params.z_n)
    b_step = step_b * n + 16.0 * grid_cells
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
        "dtype": "f32" if args.precision == "fp32" else "f32+f16",
        "data": "synthetic",
        "config": {
            "workload": (f"{sc.name}: dam break, {n} particles, "
                         f"{iters} iterations, {args.physics} physics, "
                         f"{'FP32' if args.precision == 'fp32' else 'mixed precision'}"),
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
            "traffic": traffic,
            "traffic_over_algorithmic": (traffic / alg) if traffic else None,
            "avg_launch_ms": avg_ms,
            "launches_per_step": dom_launches / args.steps,
            "algorithmic_bytes_per_particle": pass_bytes(dom),
            "algorithmic_bytes_per_launch": alg,
            "with_lists": {
                "bytes_per_particle": list_pass_bytes(dom, lb),
                "list_bytes_per_particle": lb,
                "achieved_GBps": list_pass_bytes(dom, lb) * n / (avg_ms / 1e3) / 1e9,
                "frac": list_pass_bytes(dom, lb) * n / (avg_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
            },
        },
        "step_roofline": {
            "bytes_per_particle": step_b + 16.0 * grid_cells / n,
            "bytes_per_step": b_step,
            "achieved_GBps": step_gbs,
            "frac": step_gbs / HBM_PEAK_GBS,
        },
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in prof.items()},
        # kernel time of the replayed steps over the graph-timed step time
        "kernels_over_step": total_prof / ms,
        "accuracy": {
            "max_compression": diag["max_compression"],
            "max_abs_density_err": diag["max_abs_density_err"],
            "mean_abs_density_err": diag["mean_abs_density_err"],
            "at_step": args.warmup + args.steps,
        },
    }
    if args.horizon > 0:
        # the configuration's own horizon, timed from step 0 of a fresh
        # handle (the autotuner's trial steps restore the state first)
        with pbf.Simulation(params, parts) as h:
            hms = h.time_steps(args.horizon)
            hd = h.diagnostics()
        out["horizon"] = {
            "steps": args.horizon,
            "ms_per_step": hms / args.horizon,
            "particle_steps_per_s": n * args.horizon / (hms / 1e3),
            "step_roofline_frac": b_step * args.horizon / (hms / 1e3) / 1e9 / HBM_PEAK_GBS,
            "max_compression": hd["max_compression"],
            "mean_abs_density_err": hd["mean_abs_density_err"],
        }
        print(f"[bench] steps 0-{args.horizon} of a fresh handle: "
              f"{hms / args.horizon:.4f} ms/step", file=sys.stderr, flush=True)
    if args.batch_steps > 0:
        print(f"[bench] 256-scene batch vs sequential, {args.batch_steps} steps",
              file=sys.stderr, flush=True)
        out["batch"] = batch_vs_sequential(physics, args.batch_steps)
    threads = max(1, min(16, cpu_info()["usable"]))
    if args.vs_fp64:
        ks = sorted({int(k) for k in args.vs_fp64.split(",") if k.strip()})
        print(f"[bench] {sc.name} against the FP64 oracle at steps {ks}",
              file=sys.stderr, flush=True)
        out["vs_fp64_oracle"] = vs_oracle(params, parts, ks, threads)
    if args.mixed_steps > 0 and physics == pbf.PHYSICS_PAPER:
        print(f"[bench] c4 mixed precision vs FP32, {args.mixed_steps} timed steps",
              file=sys.stderr, flush=True)
        out["mixed"] = mixed_vs_plain(args.mixed_steps, threads, not args.no_cpu)
    if not args.no_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(sc, args.physics, args.cpu_steps,
                                           args.cpu_one_core_steps,
                                           args.cpu_ref_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
