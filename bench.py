#!/usr/bin/env python3
"""Benchmark: zone-updates/s of the hydro step at 3D 256^3 on one MI355X.

Workload (the 3D point blast): 256^3 cells on [0,1]^3, reflective faces,
superbee limiter, gamma = 5/3, rho = 1 +- 1% seeded jitter, p = 1e-5 at
rest, unit energy deposited as pressure in the cells within 3 cells of the
centre; exact Riemann solver with tol 1e-6 and 100 iterations, Courant
number 0.2 (the reference's own 3D example value; at 0.8 the unsplit 3D
update turns pressures negative on the first step in the reference
algorithm itself).

A step is one main-loop step (dt from the CFL reduction, ghost fill,
MUSCL-Hancock update) of the whole box, state resident in HBM.  W warmup
steps run untimed from the initial condition, then K steps are timed between
two stream synchronisations; zone-updates/s = 256^3 * K / elapsed.

Also reported:
  roofline      algorithmic HBM bytes of the step (2 * 5 * 8 = 80 B per
                zone-update: one read and one write of the FP64 state) over
                the measured time of one step's kernels (achieved) and over
                the wall time of a step (achieved_wall), against 8 TB/s;
                `traffic` is what the step's kernels actually move per step,
                from calibrated rocprofv3 FETCH_SIZE/WRITE_SIZE passes
                summarised by --pmc-summary, flagged stale when the library
                sources changed since.
  cpu_baseline  the CPU oracle (a restatement of the reference Fortran path,
                both loop structures) on the host cores, on a bounded sample
                of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-updates/s at 3D 256^3 (FP64 vs mixed), % HBM roofline; serve p99 latency"
HBM_PEAK_GBS = 8000.0
EVENT_EVERY = 5  # HIP events around the launches of every 5th timed step
# Counter figures of the step (HBM bytes per zone-update from calibrated
# FETCH_SIZE/WRITE_SIZE passes, VALU issue of the walk kernels from SQ
# passes), written by `bench.py --pmc-summary` from rocprofv3 --pmc runs of
# this bench and stamped with the hash of the sources they were taken with;
# a line whose sources hash differently marks them stale.
PMC_SUMMARY = "profiles/blast256_counters.json"
# gfx950 issue model: a wave64 VALU instruction holds its SIMD 4 cycles, a
# transcendental 16; 256 CUs x 4 SIMDs at the 2.4 GHz peak clock
SIMDS, CLOCK_GHZ = 1024, 2.4


def source_hash():
    """sha256 over the library's sources and build recipe (what the counters
    depend on): rhyme_amd/csrc/*.hip|hpp, include/rhyme_hip.h, rhyme_amd/build.py."""
    import hashlib
    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "rhyme_amd", "csrc")
    files = sorted(os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith((".hip", ".hpp")))
    files += [os.path.join(ROOT, "include", "rhyme_hip.h"),
              os.path.join(ROOT, "rhyme_amd", "build.py")]
    for f in files:
        h.update(os.path.relpath(f, ROOT).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def load_pmc_summary():
    path = os.path.join(ROOT, PMC_SUMMARY)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    d["stale"] = d.get("source_hash") != source_hash()
    return d


# Kernels of one step of each leg, for the counter summary (names as
# rocprofv3 prints them; kernels shared by the legs of both precisions,
# such as dt_kernel, enter at their mean over all dispatches).
STEP_KERNELS = {
    "fp64": r"^void rhy::(flux_kernel<3, 0, \d, double|acc_walk_kernel<3, \d, \d, double|"
            r"acc_fixup_kernel<3, \d, double|flux_fixup_kernel<3, double|update_kernel<3, \d, \w+, "
            r"double|bc_kernel<3, double)|^rhy::(dt_kernel|step_reset_kernel)",
    "mixed": r"^void rhy::(flux_kernel<3, 0, \d, float|acc_walk_kernel<3, \d, \d, float|"
             r"acc_fixup_kernel<3, \d, float|flux_fixup_kernel<3, float|update_kernel<3, \d, \w+, "
             r"float|bc_kernel<3, float)|^rhy::(dt_kernel|step_reset_kernel)",
    "batched_fp64": r"batch_(flux|fixup|update)_kernel<3, [^>]*double|batch_dt_kernel|"
                    r"batch_step_counter_kernel|^rhy::set_u32_kernel",
    "batched_mixed": r"batch_(flux|fixup|update)_kernel<3, [^>]*float|batch_dt_kernel|"
                     r"batch_step_counter_kernel|^rhy::set_u32_kernel",
}
# the kernels that walk faces (VALU-issue bound) in each leg
WALK_KERNELS = r"flux_kernel<|acc_walk_kernel<|batch_flux_kernel<"
# kernels the run also launches for other (smaller) boxes: only their
# dispatches of at least half the largest size enter the summary
LARGEST_ONLY = r"bc_kernel<"


def pmc_summary(root):
    """Summarise the rocprofv3 --pmc passes under `root` (valu/, misc/,
    fetch/, write/, each holding one *_counter_collection.csv, and the bench
    line each pass printed as <pass>.json) into the figures the bench line
    reports.  FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by
    what they read for the --copy-probe kernels of known traffic (a 1 GiB
    fill and a 1 GiB scaled copy)."""
    import collections
    import csv
    import glob
    import re

    def load(name):
        path = glob.glob(os.path.join(root, name, "*_counter_collection.csv"))[0]
        vals = collections.defaultdict(lambda: collections.defaultdict(
            lambda: collections.defaultdict(float)))
        disp = collections.defaultdict(dict)
        for r in csv.DictReader(open(path)):
            k, d = r["Kernel_Name"], r["Dispatch_Id"]
            vals[k][d][r["Counter_Name"]] += float(r["Counter_Value"])
            disp[k][d] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        for k in vals:
            if re.search(LARGEST_ONLY, k):
                size = {d: max(c.values()) for d, c in vals[k].items()}
                big = max(size.values())
                for d in [d for d in size if size[d] < 0.5 * big]:
                    del vals[k][d], disp[k][d]
        per = collections.defaultdict(lambda: collections.defaultdict(float))
        for k in vals:
            for c in vals[k].values():
                for n, v in c.items():
                    per[k][n] += v
        return per, disp

    passes = {n: load(n) for n in ("valu", "misc", "fetch", "write")}
    with open(os.path.join(root, "fetch.json")) as f:
        line = json.loads(f.read().strip().splitlines()[-1])

    def only(per, pat):
        ks = [k for k in per if re.search(pat, k)]
        assert len(ks) == 1, (pat, ks)
        return ks[0]
    fetch, _ = passes["fetch"]
    write, _ = passes["write"]
    gib_kib = float(1 << 20)
    f_cal = fetch[only(fetch, r"AUnaryFunctor<double")]["FETCH_SIZE"] / gib_kib
    w_cal = write[only(write, r"FillFunctor<double")]["WRITE_SIZE"] / gib_kib

    def kernel_row(k):
        vp, vd = passes["valu"]
        mp, md = passes["misc"]
        n = len(vd[k])
        waves = vp[k]["SQ_WAVES"] / n
        per_wave = {c.replace("SQ_INSTS_", "").lower(): v / vp[k]["SQ_WAVES"]
                    for c, v in list(vp[k].items()) + list(mp[k].items())
                    if c.startswith("SQ_INSTS")}
        dur_us = sum(vd[k].values()) / n / 1e3
        trans = per_wave.get("valu_trans_f64", 0.0) + per_wave.get("valu_trans_f32", 0.0)
        issue = waves * (4 * per_wave["valu"] + 12 * trans) / (dur_us * 1e-6 * CLOCK_GHZ * 1e9
                                                               * SIMDS)
        f64 = sum(per_wave.get("valu_%s_f64" % o, 0.0) for o in ("add", "mul", "fma", "trans"))
        return dict(dispatches=n, avg_us=dur_us, waves=waves,
                    read_bytes=fetch[k]["FETCH_SIZE"] / len(passes["fetch"][1][k]) / f_cal * 1024,
                    written_bytes=write[k]["WRITE_SIZE"] / len(passes["write"][1][k]) / w_cal
                    * 1024,
                    valu_per_wave=per_wave["valu"], f64_arith_per_wave=f64,
                    f64_arith_share=f64 / per_wave["valu"] if per_wave["valu"] else 0.0,
                    int32_per_wave=per_wave.get("valu_int32"),
                    int64_per_wave=per_wave.get("valu_int64"),
                    cvt_per_wave=per_wave.get("valu_cvt"), trans_per_wave=trans,
                    valu_issue=issue)

    out = dict(source_hash=line.get("source_hash"),
               calibration=dict(fetch_kib_per_kib_read=f_cal, write_kib_per_kib_written=w_cal))
    zones = {"fp64": line["config"]["grid"][0] ** 3, "mixed": line["config"]["grid"][0] ** 3}
    if "batched" in line:
        zones["batched_fp64"] = line["batched"]["fp64"]["zones_per_step"]
        zones["batched_mixed"] = line["batched"]["mixed"]["zones_per_step"]
    for leg, z in zones.items():
        ks = [k for k in passes["valu"][0] if re.search(STEP_KERNELS[leg], k)]
        rows = {k: kernel_row(k) for k in ks}
        step_bytes = sum(r["read_bytes"] + r["written_bytes"] for r in rows.values())
        walks = [r for k, r in rows.items() if re.search(WALK_KERNELS, k)]
        wt = sum(r["avg_us"] for r in walks)
        out[leg] = dict(zones=z, bytes_per_step=step_bytes, bytes_per_zone=step_bytes / z,
                        step_kernel_us=sum(r["avg_us"] for r in rows.values()),
                        walk_valu_issue=sum(r["valu_issue"] * r["avg_us"] for r in walks) / wt,
                        walk_f64_arith_share=sum(r["f64_arith_share"] * r["avg_us"]
                                                 for r in walks) / wt,
                        kernels=rows)
    return out


def blast_params(hydro, solver):
    return hydro.make_params(
        3, thermo=hydro.Thermo(state_of_matter=hydro.MONATOMIC),
        rp=hydro.RiemannProblem(irs=hydro.Irs(n_iteration=100, tolerance=1e-6)),
        sl=hydro.SlopeLimiter(type=hydro.SUPERBEE, w=0.0),
        mh=hydro.MusclHancock(solver_type=solver),
        cfl=hydro.Cfl(0.2), bc=hydro.SamrBc(types=[hydro.REFLECTIVE] * 6))


def gpu_leg(n, steps, warmup, seed, precision=0, event_every=EVENT_EVERY):
    import numpy as np
    from rhyme_amd import hydro, ics

    cells = ics.blast_3d(n=n, seed=seed)
    params = blast_params(hydro, hydro.MEMORY_INTENSIVE)
    dx = [1.0 / n] * 3
    box = hydro.SamrBox(3, [n, n, n], cells=cells, precision=precision)
    box.upload()
    if warmup:
        hydro.run(params, box, dx, warmup)
    # timed region: K steps enqueued, bracketed by stream syncs; HIP events
    # recorded on the box's stream around every launch of the steps 0, E,
    # 2E, ... of the K (E = event_every; events around every launch of
    # every step cost ~4% of the step)
    box.set_profiling(event_every)
    box.sync()
    t0 = time.perf_counter()
    hydro.run_async(params, box, dx, steps)
    box.sync()
    t1 = time.perf_counter()
    stats = {k: box.kernel_stats(k) for k in range(6)}
    dts = hydro.last_dts(box, steps)
    out = box.download()
    interior = out[:5, 2:-2, 2:-2, 2:-2].copy()
    finite = bool(np.isfinite(interior).all())
    box.close()
    return dict(elapsed=t1 - t0, stats=stats, dts=dts, finite=finite, state=interior)


def batched_leg(steps, warmup, seed):
    """Batched AMR boxes: a 128^3 base box plus 200 refined boxes with
    per-axis extents U{8..64} (dx halved), seeded two-state/blast initial
    conditions, van Albada, outflow faces, each box with its own dt; every
    phase of a step is one launch over all boxes, the step is replayed from a
    hipGraph, tile shape from the autotuner.  FP64 and mixed."""
    import numpy as np
    from rhyme_amd import autotune, hydro, ics
    from rhyme_amd.batch import SamrBatch
    boxes = ics.amr_boxes(seed=seed)
    p = hydro.make_params(3, rp=hydro.RiemannProblem(irs=hydro.Irs(100, 1e-6)),
                          sl=hydro.SlopeLimiter(type=hydro.VAN_ALBADA),
                          cfl=hydro.Cfl(0.2), bc=hydro.SamrBc(types=[hydro.OUTFLOW] * 6))
    out = dict(workload="128^3 base + 200 boxes U{8..64}^3, van Albada, outflow, "
                        "per-box dt, graph-replayed steps", nboxes=len(boxes))
    for name, prec in (("fp64", hydro.FP64), ("mixed", hydro.MIXED)):
        b = SamrBatch(3, [d for d, _, _ in boxes], [x for _, x, _ in boxes], precision=prec)
        init = [c for _, _, c in boxes]
        best, timings = autotune.tune(b, p, init, steps=3)
        if warmup:
            b.run(p, warmup, graph=True)
        b.sync()
        t0 = time.perf_counter()
        b.run_async(p, steps, graph=True)
        b.sync()
        el = time.perf_counter() - t0
        zones = b.zones()
        finite = all(bool(np.isfinite(b.download(i)).all()) for i in range(b.n))
        out[name] = dict(value=zones * steps / el, unit="zone-updates/s",
                         ms_per_step=1e3 * el / steps, zones_per_step=zones,
                         tiles=list(best), device_bytes=b.nbytes(), state_finite=finite)
        b.close()
    out["mixed_speedup_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    return out


def amr_leg(n, steps, warmup, seed):
    """Two-level hierarchy: the n^3 point blast as the coarse level, the
    central (n/4)^3 coarse cells refined by 2x2x2 fine boxes of (n/4)^3 cells
    each (dx / 2) that start as their coarse parents.  One coarse step is the
    coarse CFL/fine CFL dt, the coarse step, two subcycled fine substeps with
    ghosts prolonged from the coarse level, and restriction; zone-updates
    count the coarse cells once and the fine cells twice.  FP64 and mixed."""
    import numpy as np
    from rhyme_amd import amr, hydro, ics
    m = n // 4
    c0 = n // 2 - m // 2 + 1
    lo = [[c0 + m // 2 * a, c0 + m // 2 * b, c0 + m // 2 * c]
          for c in (0, 1) for b in (0, 1) for a in (0, 1)]
    fdims = [[m, m, m]] * 8
    cells = ics.blast_3d(n=n, seed=seed)
    fines = [amr.fine_from_coarse(3, cells, l, d) for l, d in zip(lo, fdims)]
    p = blast_params(hydro, hydro.MEMORY_INTENSIVE)
    out = dict(workload="3D %d^3 point blast + 8 fine boxes %d^3 (ratio 2) over the central "
                        "%d^3 coarse cells, subcycled, superbee, reflective" % (n, m, m))
    for name, prec in (("fp64", hydro.FP64), ("mixed", hydro.MIXED)):
        h = amr.AmrHierarchy(3, [n] * 3, [1.0 / n] * 3, lo, fdims, precision=prec)
        h.upload(cells, fines)
        if warmup:
            h.run(p, warmup)
        t0 = time.perf_counter()
        dts = h.run(p, steps)
        el = time.perf_counter() - t0
        zones = h.zones_per_step()
        dc, dfs = h.download()
        finite = bool(np.isfinite(dc).all()) and all(bool(np.isfinite(f).all()) for f in dfs)
        out[name] = dict(value=zones * steps / el, unit="zone-updates/s",
                         ms_per_coarse_step=1e3 * el / steps, zones_per_coarse_step=zones,
                         dt_first=float(dts[0]), dt_last=float(dts[-1]), state_finite=finite)
        h.close()
    return out


def snapshot_leg(n, seed, reps=5):
    """Snapshot payload of the n^3 blast (FP64 state): the FP32 interior,
    component after component, packed on the device and copied to pageable
    host memory (rhy_box_snapshot_pack), then unpacked into a box again.
    Whole calls are timed (kernel, PCIe copy and waits); the kernel alone is
    in the rocprofv3 summary."""
    import numpy as np
    from rhyme_amd import hydro, ics
    box = hydro.SamrBox(3, [n] * 3, cells=ics.blast_3d(n=n, seed=seed))
    box.upload()
    data = box.snapshot()
    pack, unpack = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        data = box.snapshot()
        pack.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        box.load_snapshot(data)
        unpack.append(time.perf_counter() - t0)
    box.close()
    mb = data.nbytes / 1e6
    return dict(payload_mb=mb, components=int(data.size // n ** 3),
                pack_call_ms=1e3 * float(np.median(pack)),
                unpack_call_ms=1e3 * float(np.median(unpack)),
                pack_call_gbs=mb / 1e3 / float(np.median(pack)),
                note="median of %d calls, pageable host buffer, PCIe-inclusive" % reps)


def cpu_leg(n_sample, steps_sample, seed, threads_all):
    """The oracle on the host cores: both reference loop structures with all
    threads, plus memory_intensive on one thread (at a quarter of the edge),
    on the first steps of the same blast workload.  Each configuration runs
    in its own process so that OMP_NUM_THREADS takes effect."""
    from rhyme_amd import hydro
    return {
        "memory_intensive": _cpu_subprocess(n_sample, steps_sample, seed,
                                            hydro.MEMORY_INTENSIVE, threads_all),
        "cpu_intensive": _cpu_subprocess(n_sample, steps_sample, seed,
                                         hydro.CPU_INTENSIVE, threads_all),
        "memory_intensive_1thread": _cpu_subprocess(max(32, n_sample // 4), steps_sample, seed,
                                                    hydro.MEMORY_INTENSIVE, 1),
    }


def _host_cpu():
    """nproc and the CPU model of the host the baseline ran on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(nproc=os.cpu_count(), model=model,
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"))


def _cpu_subprocess(n, steps, seed, solver, threads):
    """One warm-up main-loop step of the blast from its initial condition,
    then `steps` single steps each timed on its own; the rate is the median
    step's (zone-updates/s), with the per-step times reported."""
    import subprocess
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    code = (
        "import sys, time, json; sys.path.insert(0, %r)\n"
        "import oracle\n"
        "from rhyme_amd import hydro, ics\n"
        "import bench\n"
        "cells = ics.blast_3d(n=%d, seed=%d)\n"
        "p = oracle.params_from(bench.blast_params(hydro, %d))\n"
        "dx = [1.0 / %d] * 3\n"
        "oracle.run(p, [%d] * 3, cells, dx, 1)\n"
        "ts = []\n"
        "for _ in range(%d):\n"
        "    t0 = time.perf_counter()\n"
        "    oracle.run(p, [%d] * 3, cells, dx, 1)\n"
        "    ts.append(time.perf_counter() - t0)\n"
        "print(json.dumps(dict(seconds=ts, threads=oracle.num_threads())))\n"
    ) % (ROOT, n, seed, solver, n, n, steps, n)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("cpu baseline failed:\n" + r.stderr)
    d = json.loads(r.stdout.strip().splitlines()[-1])
    ts = sorted(d["seconds"])
    med = ts[len(ts) // 2]
    return dict(zone_updates_per_s=n ** 3 / med, seconds_per_step=d["seconds"],
                threads=d["threads"],
                sample="3D blast %d^3: one warm-up step, then the median of %d timed steps"
                       % (n, steps))


def serving_leg(njobs, rate, budget_gib, quantum, seed, precision=0):
    """Serving stream: `njobs` independent jobs with Poisson arrivals at
    `rate` jobs/s, half 2D n^2 (n log-uniform in [64, 512]), half 3D n^3 (n
    log-uniform in [32, 128]), 100 steps each, seeded two-state or blast
    state, limiter uniform over the four; admitted first come first served
    under an HBM budget, state resident, continuously batched in quanta.
    A job's initial state is painted on the device at admission (inside the
    measured latency); a job completes at the end of the quantum in which it
    reaches its step count, when its slot is freed (final states are not
    copied back here)."""
    from rhyme_amd import serve
    jobs = serve.serving_jobs(njobs=njobs, rate=rate, seed=seed)
    srv = serve.HydroServer(int(budget_gib * (1 << 30)), precision=precision, quantum=quantum,
                            keep_results=False)
    t0 = time.perf_counter()
    srv.serve(jobs)
    wall = time.perf_counter() - t0
    rep = serve.latency_report(jobs)
    srv.close()
    rep.update(rate_jobs_per_s=rate, budget_gib=budget_gib, quantum_steps=quantum,
               peak_resident_gib=srv.peak_bytes / float(1 << 30), admission_deferrals=srv.deferrals,
               admission=srv.admission, admitted_past_head=srv.bypassed,
               quanta=srv.quanta, wall_s=wall, host_seconds=srv.seconds, unit="zone-updates/s (aggregate), ms latency",
               precision="mixed" if precision else "fp64")
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--skip-cpu", action="store_true",
                    help="omit the CPU baseline leg (profiling runs)")
    ap.add_argument("--skip-mixed", action="store_true",
                    help="omit the mixed-precision leg")
    ap.add_argument("--skip-batched", action="store_true",
                    help="omit the batched AMR-boxes leg")
    ap.add_argument("--batched-steps", type=int, default=50)
    ap.add_argument("--skip-amr", action="store_true", help="omit the two-level hierarchy leg")
    ap.add_argument("--amr-steps", type=int, default=20)
    ap.add_argument("--skip-snapshot", action="store_true",
                    help="omit the snapshot payload pack/unpack leg")
    ap.add_argument("--skip-serve", action="store_true", help="omit the serving-stream leg")
    ap.add_argument("--serve-jobs", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=40.0, help="Poisson arrivals, jobs/s")
    ap.add_argument("--serve-budget-gib", type=float, default=16.0)
    ap.add_argument("--serve-quantum", type=int, default=10)
    ap.add_argument("--serve-saturated-rate", type=float, default=1000.0,
                    help="second pass with arrivals faster than service (capacity); 0: skip")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run two torch kernels of known traffic, a 1 GiB "
                         "fill (writes 1 GiB) and a 1 GiB scaled copy (reads 1 GiB, "
                         "writes 1 GiB), that scale rocprof FETCH/WRITE_SIZE")
    ap.add_argument("--event-every", type=int, default=EVENT_EVERY,
                    help="HIP events around the launches of every E-th timed step "
                         "(1: every step)")
    ap.add_argument("--pmc-summary", metavar="DIR",
                    help="no run: summarise the rocprofv3 --pmc passes of this bench under DIR "
                         "(valu/, misc/, fetch/, write/ and their printed lines) into "
                         + PMC_SUMMARY)
    a = ap.parse_args()
    if a.pmc_summary:
        d = pmc_summary(a.pmc_summary)
        with open(os.path.join(ROOT, PMC_SUMMARY), "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
            f.write("\n")
        print(json.dumps({leg: {k: v for k, v in d[leg].items() if k != "kernels"}
                          for leg in d if isinstance(d[leg], dict) and "zones" in d[leg]}))
        return 0
    if a.gpus != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        return 2

    if a.copy_probe:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    g = gpu_leg(a.n, a.steps, a.warmup, a.seed, event_every=a.event_every)
    zones = a.n ** 3
    zu = zones * a.steps / g["elapsed"]
    ms_step = 1e3 * g["elapsed"] / a.steps
    st = g["stats"]
    # kernels of one step: dt (4), deferred-face list reset and fixup (5),
    # x flux walk and y/z accumulating walks (0), update (1, 1D boxes only),
    # ghost fill of the new state (2)
    kern_ms = sum(st[k][1] for k in (0, 1, 2, 4, 5)) / max(st[4][0], 1)  # per timed step
    flux_ms = st[0][1] / max(st[0][0], 1)
    bytes_alg = 80.0 * zones
    achieved = bytes_alg / (kern_ms * 1e-3) / 1e9
    pmc = load_pmc_summary()
    traffic = pmc["fp64"]["bytes_per_zone"] * zones if pmc and a.n ** 3 == pmc["fp64"]["zones"] \
        else None
    achieved_wall = bytes_alg / (ms_step * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=achieved / HBM_PEAK_GBS, traffic=traffic,
                    traffic_unit="HBM bytes per step (FETCH_SIZE, WRITE_SIZE, calibrated)",
                    counters_source=PMC_SUMMARY if pmc else None,
                    counters_stale=pmc["stale"] if pmc else None,
                    achieved_wall=achieved_wall, frac_wall=achieved_wall / HBM_PEAK_GBS,
                    flux_kernel_bound="valu issue",
                    flux_kernel_valu_issue_frac=pmc["fp64"]["walk_valu_issue"] if pmc else None,
                    kernel="one step: dt_kernel + deferred-list reset + x flux walk + its fixup "
                           "+ y and z accumulating walks + their fixups + ghost fill (achieved: "
                           "over the HIP-event sum of these kernels; achieved_wall: over "
                           "ms_per_step)",
                    events="HIP events on the box stream around every launch of timed steps "
                           "0, %d, %d, ... (%d of the %d)" % (a.event_every, 2 * a.event_every,
                                                              st[4][0], a.steps),
                    step_kernel_ms=kern_ms, walk_kernel_avg_ms=flux_ms,
                    bytes_per_zone_update=80)
    line = dict(metric=METRIC, value=zu, unit="zone-updates/s", n_gpus=1, steps=a.steps,
                warmup=a.warmup, ms_per_step=ms_step, higher_is_better=True, scaling="weak",
                vs_baseline=None, dtype="f64", data="synthetic",
                config=dict(workload="3D %d^3 point blast, FP64 drop-in path "
                            "(memory_intensive order), superbee, reflective, CFL 0.2, "
                            "exact RS tol 1e-6 / 100 it" % a.n,
                            grid=[a.n] * 3, precision="fp64", state_finite=g["finite"],
                            dt_first=float(g["dts"][0]), dt_last=float(g["dts"][-1])),
                roofline=roofline, source_hash=source_hash())
    if not a.skip_mixed:
        import numpy as np
        m = gpu_leg(a.n, a.steps, a.warmup, a.seed, precision=1, event_every=a.event_every)
        zu_m = zones * a.steps / m["elapsed"]
        ref = g["state"]
        rel = [float(np.abs(m["state"][c] - ref[c]).sum() / np.abs(ref[c]).sum())
               for c in range(5)]
        line["mixed"] = dict(
            value=zu_m, unit="zone-updates/s", ms_per_step=1e3 * m["elapsed"] / a.steps,
            speedup_vs_fp64=zu_m / zu, dtype="f32 state/bulk, f64 Riemann p*/sampling/dt",
            rel_l1_vs_fp64=dict(zip(["rho", "rho_u", "rho_v", "rho_w", "e_tot"], rel)),
            rel_l1_note="against the FP64 device path after the same warmup+steps; the FP64 "
                        "path matches the FP64 oracle to <= 1e-12 (tests/test_gpu_parity.py)",
            walk_kernel_avg_ms=m["stats"][0][1] / max(m["stats"][0][0], 1),
            roofline_achieved_gbs=40.0 * zones / (sum(m["stats"][k][1] for k in (0, 1, 2, 4, 5))
                                                   / max(m["stats"][4][0], 1) * 1e-3) / 1e9,
            bytes_per_zone_update=40, state_finite=m["finite"],
            traffic=pmc["mixed"]["bytes_per_zone"] * zones if traffic else None,
            flux_kernel_valu_issue_frac=pmc["mixed"]["walk_valu_issue"] if pmc else None)
    if not a.skip_batched:
        line["batched"] = batched_leg(a.batched_steps, a.warmup, a.seed)
    if not a.skip_amr:
        line["amr"] = amr_leg(a.n, a.amr_steps, min(a.warmup, 2), a.seed)
    if not a.skip_snapshot:
        line["snapshot"] = snapshot_leg(a.n, a.seed)
    if not a.skip_serve:
        line["serving"] = serving_leg(a.serve_jobs, a.serve_rate, a.serve_budget_gib,
                                      a.serve_quantum, a.seed)
        if a.serve_saturated_rate > 0:
            line["serving"]["saturated"] = serving_leg(
                a.serve_jobs, a.serve_saturated_rate, a.serve_budget_gib, a.serve_quantum,
                a.seed)
            line["serving"]["mixed_saturated"] = serving_leg(
                a.serve_jobs, a.serve_saturated_rate, a.serve_budget_gib, a.serve_quantum,
                a.seed, precision=1)
            line["serving"]["mixed_saturated_vs_plain_fp64"] = (
                line["serving"]["mixed_saturated"]["zone_updates_per_s"] / zu)
    if not a.skip_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        c = cpu_leg(a.cpu_n, a.cpu_steps, a.seed, threads)
        m = c["memory_intensive"]
        line["cpu_baseline"] = dict(
            value=m["zone_updates_per_s"], unit="zone-updates/s", cores=m["threads"],
            kind="port", sample=m["sample"] + " (memory_intensive, all threads)",
            cpu_intensive=c["cpu_intensive"],
            one_thread=c["memory_intensive_1thread"], host=_host_cpu())
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
