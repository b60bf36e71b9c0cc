#!/usr/bin/env python3
"""Benchmark of the res_calc edge loop on one MI355X.

Workload: the synthetic airfoil quad mesh of 1200 x 600 cells (721 801 nodes,
720 000 cells, 1 438 200 interior edges), nodes / cells / edges shuffled with
seed 0, FP64 q = freestream * (1 + 1e-2 N(0,1)), adt ~ U[0.5, 1.5].

A step is one res_calc iteration over all edges: the increments of every edge
added into res on the device (the reference's timed region, problem.hpp:
269-304; its out->in copy is timed separately there and is a no-op for
res_calc, whose user function never reads res).  Inputs are resident in HBM.
The configuration is the learned autotuner's pick for this mesh (--tune
learned, the default; see tune_res_calc), recorded in config.tune.

Prints ONE JSON line (rank 0).  --gpus above 1 is refused: the loops run on
one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# reference unique-bytes formula (scripts/visualise_paper.py:19-22):
# x 16 B/node + q,adt 40 B/cell + res 32 B read + 32 B write per cell +
# 2 maps x 2 x 4 B per edge
PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s


def unique_bytes(num_nodes: int, num_cells: int, num_edges: int) -> int:
    return 16 * num_nodes + 40 * num_cells + 64 * num_cells + 16 * num_edges


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="res_calc",
                    choices=["res_calc", "volna-batch", "mini_aero", "hex", "lulesh", "mine",
                             "getacc", "copy"],
                    help="res_calc: the airfoil edge loop (the headline line); the others "
                         "print their own line: 64 batched Volna meshes, mini_aero 96^3, "
                         "hex scatter 128^3, lulesh element forces 128^3, getacc corner "
                         "scatter on 1024^2 quads, mine on the "
                         "1024^2 grid with its serial and OpenMP CPU loops")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=1200)
    ap.add_argument("--ny", type=int, default=600)
    ap.add_argument("--strategy", default="hierarchical",
                    choices=["atomic", "global", "hierarchical"])
    ap.add_argument("--block-size", type=int, default=None,
                    help="default 512 (mini_aero 128)")
    ap.add_argument("--partition", default=None,
                    choices=["none", "kway", "bisect", "ours", "gps", "points"],
                    help="points: owner partition of mapping-0 points (owner compute); "
                         "default points (res_calc, volna-batch), else bisect")
    ap.add_argument("--points-per-block", type=int, default=240)
    ap.add_argument("--numbering", default="cm", choices=["id", "cm"],
                    help="owner partition: points inside a part in their previous order "
                         "(id) or in Cuthill-McKee order (cm)")
    ap.add_argument("--tune", default=None, choices=["none", "exhaustive", "learned"],
                    help="res_calc (default learned): the configuration the autotuner "
                         "picks for the airfoil as generated; volna-batch (default none): "
                         "each mesh's partition size")
    ap.add_argument("--ept", type=int, default=2, choices=[1, 2],
                    help="owner compute: cells per thread")
    ap.add_argument("--no-stage", action="store_true")
    ap.add_argument("--accumulate", default="gather", choices=["gather", "colours"],
                    help="owner compute: owned cells gather their edges' fluxes from LDS "
                         "(three barriers), or add them colour by colour (a barrier each)")
    ap.add_argument("--persistent", action="store_true",
                    help="owner compute, gather: persistent workgroups that load the next "
                         "block's plan entries while computing the current one")
    ap.add_argument("--hbm-lists", action="store_true",
                    help="owner compute, gather: load the owned cells' edge lists from the "
                         "plan in HBM instead of building them in LDS")
    ap.add_argument("--nontemporal", default="auto", choices=["auto", "on", "off"],
                    help="nontemporal hints on read-once data (res_calc owner gather: the "
                         "owned res; hex, mini_aero and lulesh two-phase: the cell data); auto: when "
                         "the step's data exceed the 256 MB Infinity Cache")
    ap.add_argument("--point-terms", action="store_true",
                    help="owner compute: 1/rho and pressure once per staged cell")
    ap.add_argument("--nodes", default="gather", choices=["gather", "stage"],
                    help="owner compute: gather node coordinates per edge, or stage each "
                         "block's nodes in LDS first")
    ap.add_argument("--order", default=None, choices=["owner", "dep", "launch", "two-phase"],
                    help="hierarchical block scheduling: owner compute in one launch (owner), "
                         "block colours ordered by completion flags in one launch (dep), "
                         "one launch per block colour (launch) or every block in one launch "
                         "with shared points merged by a second (two-phase); default owner "
                         "(res_calc, volna-batch), else two-phase")
    ap.add_argument("--graph", action="store_true", help="replay each step from a hipGraph")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-atomic-ref", action="store_true")
    ap.add_argument("--sweep-block-sizes", action="store_true",
                    help="mini_aero / hex / lulesh / getacc / mine: also time plain atomics "
                         "and the "
                         "reference-shaped hierarchical loop (GPS-ordered mesh, one launch per "
                         "block colour) at block sizes 128-1024")
    ap.add_argument("--soa", action="store_true",
                    help="res_calc: point data in SOA layout (d * n + i) instead of AOS")
    ap.add_argument("--share", type=int, default=0,
                    help="res_calc: time, on this GPU, each share of a SHARE-way partition "
                         "of the cells (owned cells, the edges touching them, their halo) "
                         "as one GPU of a sharded run would run it")
    args = ap.parse_args()
    # the 3D element loops, getacc and mine default to the two-phase
    # write-back on recursive-bisection blocks (the autotuner's pick on hex
    # and mini_aero; measured fastest on lulesh, getacc and mine), res_calc
    # and the volna batch to owner compute on point parts
    part, bs, order = WORKLOAD_DEFAULTS.get(args.workload, ("points", 512, "owner"))
    args.partition = part if args.partition is None else args.partition
    args.block_size = bs if args.block_size is None else args.block_size
    args.order = order if args.order is None else args.order
    return args


WORKLOAD_DEFAULTS = {"hex": ("bisect", 512, "two-phase"), "mini_aero": ("bisect", 128, "two-phase"),
                     "lulesh": ("bisect", 512, "two-phase"), "getacc": ("bisect", 512, "two-phase"),
                     "mine": ("bisect", 512, "two-phase")}


def build_problem(args, partition_kind):
    from cuda_fot_amd import meshgen, res_calc_problem
    mesh = meshgen.airfoil(args.nx, args.ny, seed=0, shuffle=7)
    p = res_calc_problem(mesh, block_size=args.block_size, soa=getattr(args, "soa", False))
    if partition_kind == "gps":
        p.reorder("gps")
    elif partition_kind in ("kway", "bisect"):
        p.reorder("gps")
        p.partition(1.001, method=partition_kind)
        p.reorder_to_partition()
        p.renumber_points()
    elif partition_kind == "ours":
        p.partition_ours(1.001, 8)
        p.reorder_to_partition()
        p.renumber_points()
    elif partition_kind == "points":
        p.partition_points(args.points_per_block, args.numbering)
    return mesh, p


def make_loop(p, strategy, block_size, stage, graph, order="launch", ept=1, gather=False,
              point_terms=False, owner_gather=False, persistent=False, hbm_lists=False,
              nontemporal="auto"):
    from cuda_fot_amd import DeviceArray, Loop, capi
    strat = {"atomic": capi.STRATEGY_ATOMIC, "global": capi.STRATEGY_GLOBAL_COLOUR,
             "hierarchical": capi.STRATEGY_HIERARCHICAL}[strategy]
    dev = [DeviceArray.from_host(d) for d in p.point_data]
    out = DeviceArray.from_host(p.point_data[0])
    flags = (capi.LOOP_STAGE_INPUTS if stage else 0) | \
        {"launch": 0, "dep": capi.LOOP_DEPENDENCY_ORDER, "owner": capi.LOOP_OWNER_COMPUTE,
         "two-phase": capi.LOOP_TWO_PHASE}[order]
    if ept == 2:
        flags |= capi.LOOP_TWO_PER_THREAD
    if gather:
        flags |= capi.LOOP_GATHER_NODES
    if point_terms:
        flags |= capi.LOOP_POINT_TERMS
    if owner_gather:
        flags |= capi.LOOP_OWNER_GATHER
    if persistent:
        flags |= capi.LOOP_PERSISTENT
    if hbm_lists:
        flags |= capi.LOOP_LISTS_IN_HBM
    flags |= nt_flags(nontemporal)
    part = p.partition_vector if strat == capi.STRATEGY_HIERARCHICAL else None
    if strat == capi.STRATEGY_HIERARCHICAL and order == "owner" and p.point_partition is not None:
        part = p.point_partition
        flags |= capi.LOOP_POINT_PARTITION
    loop = Loop(capi.functor(capi.KERNEL_RES_CALC), p.maps, p.num_points, dev, out,
                soa=p.soa, strategy=strat, block_size=block_size, partition=part,
                use_graph=graph, flags=flags)
    return loop, dev, out


def time_loop(loop, steps, warmup):
    from cuda_fot_amd import Timer, synchronize
    for _ in range(warmup):
        loop.step()
    synchronize()
    t = Timer()
    synchronize()
    t0 = time.perf_counter()
    t.start()
    for _ in range(steps):
        loop.step()
    t.stop()
    synchronize()
    wall = time.perf_counter() - t0
    return wall, t.elapsed_ms() / 1e3


def launch_durations(loop, steps):
    """Mean and median duration of one launch: a HIP event pair recorded
    around each of `steps` launches on the loop's stream (the kernel's own
    time, without the gaps between back-to-back launches)."""
    from cuda_fot_amd import Timer, synchronize
    timers = [Timer() for _ in range(steps)]
    synchronize()
    for t in timers:
        t.start()
        loop.step()
        t.stop()
    synchronize()
    ms = np.array([t.elapsed_ms() for t in timers])
    return float(ms.mean()), float(np.median(ms))


def copy_bandwidth(nbytes=256 << 20, reps=10, how="kernel"):
    """Bytes/s a device copy moves, read + write: fot_copy (the streaming
    copy kernel, 16-byte nontemporal loads and stores, four in flight per
    lane) or hipMemcpyAsync device to device ("runtime").  Beyond the 256 MB
    Infinity Cache this is what DRAM delivers to a streaming kernel on this
    GPU, for context beside the nominal peak."""
    from cuda_fot_amd import DeviceArray, Timer, capi, synchronize
    fn = {"kernel": "fot_copy", "runtime": "fot_memcpy_dtod"}[how]
    a = DeviceArray((nbytes // 8,), np.float64)
    b = DeviceArray((nbytes // 8,), np.float64)
    capi.call(fn, b.ptr, a.ptr, nbytes, None)
    synchronize()
    t = Timer()
    t.start()
    for _ in range(reps):
        capi.call(fn, b.ptr, a.ptr, nbytes, None)
    t.stop()
    synchronize()
    return 2.0 * nbytes * reps / (t.elapsed_ms() / 1e3)


def bench_copy(args):
    """The DRAM ceiling the loops are compared with: device copies of 32 MiB
    to 1 GiB each way by the streaming copy kernel and by hipMemcpyAsync,
    best of three timings of 10 copies each."""
    sizes = [32 << 20, 64 << 20, 128 << 20, 256 << 20, 512 << 20, 1 << 30]
    rows = {}
    for n in sizes:
        rows[str(n >> 20)] = {how: round(max(copy_bandwidth(n, 10, how) for _ in range(3)) / 1e9, 1)
                              for how in ("kernel", "runtime")}
    best_dram = max(max(r.values()) for k, r in rows.items() if int(k) >= 512)
    return _line("device copy GB/s (read + write) by size each way", best_dram * 1e9, "B/s",
                 10, 1, 2.0 * (1 << 30) / (best_dram * 1e9) * 1e3,
                 "device copy, 32 MiB - 1 GiB each way",
                 copy_gbps_by_mib=rows,
                 note="value: the best rate at 512 MiB - 1 GiB each way (beyond the 256 MB "
                      "Infinity Cache)")


def load_traffic(args, stats):
    """HBM bytes per launch from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE, scaled by a copy-kernel calibration) when it
    was taken on this configuration; else None."""
    path = os.path.join(ROOT, "profiles", f"res_calc_owner_airfoil_{args.nx}x{args.ny}_"
                        f"ppb{args.points_per_block}_traffic.json")
    if not os.path.exists(path):
        return None
    t = json.load(open(path))
    c = t["config"]
    same = (args.strategy == "hierarchical" and args.order == "owner" and
            args.partition == c["partition"] and args.nx == c["nx"] and args.ny == c["ny"] and
            args.points_per_block == c["points_per_block"] and
            args.numbering == c.get("numbering", "id") and
            args.ept == c["cells_per_thread"] and args.nodes == c["nodes"] and
            args.point_terms == c.get("point_terms", False) and
            args.accumulate == c.get("accumulate", "colours") and
            args.persistent == c.get("persistent", False) and
            ("hbm" if args.hbm_lists else "lds") == c.get("lists", "hbm") and
            stats["nontemporal"] == c.get("nontemporal", 0) and
            args.soa == c.get("soa", False))
    return t["traffic_bytes"] if same else None


def nt_flags(mode: str) -> int:
    """The loop flags of --nontemporal auto / on / off."""
    from cuda_fot_amd import capi
    return {"auto": 0, "on": capi.LOOP_NONTEMPORAL, "off": capi.LOOP_CACHED}[mode]


def load_workload_traffic(workload: str, variant: str, nontemporal: int = 0):
    """HBM bytes per step of a secondary workload's line from its committed
    rocprofv3 counter summary (the L2-to-fabric request counters summed over
    the step's launches), or None when no summary was taken for this variant
    with the same nontemporal setting."""
    path = os.path.join(ROOT, "profiles", f"{workload}_{variant}_traffic.json")
    if not os.path.exists(path):
        return None
    t = json.load(open(path))
    return t["traffic_bytes"] if t.get("nontemporal", 0) == nontemporal else None


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), as lscpu reports it."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cpu() -> dict:
    """The host's CPUs (lscpu's model name, sockets, cores, threads) and the
    threads the CPU legs run: OMP_NUM_THREADS when the environment sets it
    (a shared GPU host gives each GPU's jobs that share of its cores), else
    every CPU this process may run on."""
    import subprocess
    info = {"cpu_model": cpu_model(), "host_logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info["lscpu_" + k.lower().replace("(s)", "s").replace(" ", "_")] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else \
        (os.cpu_count() or 1)
    info["affinity_cpus"] = affinity
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        info["threads"] = min(int(env), affinity)
        info["threads_from"] = "OMP_NUM_THREADS (this host's CPU share for the job)"
    else:
        info["threads"] = affinity
        info["threads_from"] = "every CPU in the process's affinity mask"
    return info


def cpu_baseline(mesh, seconds):
    """The oracle's colour-parallel loop (problem.hpp:405-437 restated in C,
    OpenMP over all host cores) timed on a bounded number of iterations of
    the same mesh; iteration-loop time only, colouring excluded."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from cuda_fot_amd import res_calc_problem
    p = res_calc_problem(mesh)
    o = O.OracleProblem(O.RES_CALC, p.maps, p.num_points, p.point_data)
    host = host_cpu()
    cores = host["threads"]
    # size the sample from one timed iteration so the leg takes ~`seconds`
    per = o.loop_colour(1, cores)
    iters = int(max(1, min(2000, seconds / max(per, 1e-6))))
    secs = o.loop_colour(iters, cores)
    serial = O.OracleProblem(O.RES_CALC, p.maps, p.num_points, p.point_data)
    serial.loop_serial(2)
    serial_per = serial.last_loop_seconds / 2
    return {"value": mesh["num_edges"] * iters / secs, "unit": "edges/s", "cores": cores,
            "kind": "port", "host": host, "build": "gcc -O3 -fopenmp -ffp-contract=off",
            "sample": f"{iters} iterations ({secs:.1f} s) of the colour-parallel OpenMP oracle "
                      f"loop (problem.hpp:382-437 restated) on {cores} threads, on the same "
                      f"{mesh['nx']}x{mesh['ny']} mesh, iteration loop only",
            "serial_edges_per_s": mesh["num_edges"] / serial_per}


def cpu_baseline_mine(p, f, seconds):
    """mine's CPU loops (loopCPUCellCentred / loopCPUCellCentredOMP,
    problem.hpp:342-437, restated by the oracle) on the same grid: a bounded
    number of iterations each, iteration-loop time only (colouring and the
    per-iteration copy excluded, as the reference times them)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    host = host_cpu()
    cores = host["threads"]
    ne = p.maps[0].shape[0]
    o = O.OracleProblem(O.MINE2, p.maps, p.num_points, p.point_data, p.cell_data)
    per = o.loop_colour(1, cores)
    iters = int(max(1, min(100, seconds / 2 / max(per, 1e-6))))
    secs = o.loop_colour(iters, cores)
    s = O.OracleProblem(O.MINE2, p.maps, p.num_points, p.point_data, p.cell_data)
    s.loop_serial(1)
    per_s = s.last_loop_seconds
    it_s = int(max(1, min(100, seconds / 2 / max(per_s, 1e-6))))
    s.loop_serial(it_s)
    secs_s = s.last_loop_seconds
    return {"value": ne * iters / secs, "unit": "edges/s", "cores": cores, "kind": "port",
            "host": host, "build": "gcc -O3 -fopenmp -ffp-contract=off",
            "sample": f"{iters} iterations ({secs:.1f} s) of the colour-parallel OpenMP oracle "
                      f"loop and {it_s} ({secs_s:.1f} s) of the serial one on the same grid, "
                      f"iteration loop only",
            "serial_edges_per_s": ne * it_s / secs_s}


def _line(metric, value, unit, steps, warmup, ms, workload, **extra):
    d = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
         "warmup": warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": "f64", "data": "synthetic",
         "config": {"workload": workload}}
    d.update(extra)
    return d


def bench_volna_batch(args):
    """64 shallow-water triangle meshes, cells log-uniform in [2e4, 4e5], one
    persistent owner-compute launch per step on meshes renumbered by their
    point partition; the part size is the autotuner's (the batch timed at
    each size; --tune none takes --points-per-block); fot_batch_run timed
    with eager launches and with hipGraph replays, beside one owner loop per
    mesh."""
    from cuda_fot_amd import Batch, DeviceArray, Timer, capi, meshgen, synchronize, volna_problem
    from cuda_fot_amd.device import Loop
    rng = np.random.default_rng(0)
    probs = []
    for i in range(64):
        cells = float(np.exp(rng.uniform(np.log(2e4), np.log(4e5))))
        aspect = float(rng.uniform(0.5, 2.0))
        nx = max(4, int(round(np.sqrt(cells / 2 * aspect))))
        ny = max(4, int(round(cells / 2 / nx)))
        probs.append(volna_problem(meshgen.tri_mesh(nx, ny, seed=100 + i),
                                   block_size=args.points_per_block))
    f = capi.functor(capi.KERNEL_VOLNA)
    tune = {"mode": args.tune}
    # every mesh runs owner compute on its point partition, renumbered so a
    # part's cells are contiguous (Problem.partition_points): a block's state
    # and volumes are then contiguous runs instead of scattered lines
    import copy

    def sized(size):
        out = []
        for p in probs:
            q = copy.deepcopy(p)
            q.partition_points(size)
            q.block_size = size
            out.append(q)
        return out

    def upload(ps):
        specs, keep = [], []
        for p in ps:
            dev = [DeviceArray.from_host(d) for d in p.point_data]
            out = DeviceArray(p.point_data[0].shape, np.float64)
            cdev = [DeviceArray.from_host(c) for c in p.cell_data]
            keep.append((dev, out, cdev))
            specs.append(dict(maps=p.maps, num_points=p.num_points, point_data=dev, out=out,
                              cell_data=cdev, block_size=p.block_size,
                              point_partition=p.point_partition))
        return specs, keep

    ppb = args.points_per_block
    if args.tune != "none":
        # the batch's part size: the whole batch timed at each owner part
        # size (128 / 256 / 512 points; the meshes fill the GPU together, so
        # a size measured on one mesh alone would mislead), the fastest kept
        from cuda_fot_amd import tuner
        times = {}
        for size in tuner.owner_points(f):
            sp, kp = upload(sized(size))
            b = Batch(f, sp)
            b.run(2)
            synchronize()
            t = Timer()
            t.start()
            b.run(10)
            t.stop()
            synchronize()
            times[str(size)] = round(t.elapsed_ms() / 10, 4)
            b.close()
            del sp, kp
        ppb = int(min(times, key=times.get))
        tune.update(mode="measured", measured_ms_per_step=times, pick=ppb)
    probs = sized(ppb)
    specs, keep = upload(probs)
    edges = sum(p.num_cells for p in probs)
    cells = sum(p.num_points[0] for p in probs)
    ubytes = 72 * cells + 72 * edges  # state r+w 64 + volume 8 / cell; 64 data + 8 map / edge
    res = {}
    for graph in (False, True):
        b = Batch(f, specs, use_graph=graph)
        b.run(args.warmup)
        synchronize()
        t = Timer()
        t0 = time.perf_counter()
        t.start()
        # copy in, the steps (eager launches or graph replays), copy back
        b.run(args.steps)
        t.stop()
        synchronize()
        res[graph] = (time.perf_counter() - t0, t.elapsed_ms() / 1e3, b.stats())
        b.close()
    # the same meshes one owner loop (one launch) each
    loops = [Loop(f, sp["maps"], p.num_points, sp["point_data"], sp["out"], sp["cell_data"],
                  strategy=capi.STRATEGY_HIERARCHICAL, block_size=p.block_size,
                  partition=p.point_partition,
                  flags=capi.LOOP_OWNER_COMPUTE | capi.LOOP_POINT_PARTITION)
             for sp, p in zip(specs, probs)]
    for _ in range(args.warmup):
        for l in loops:
            l.step()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for l in loops:
            l.step()
    synchronize()
    seq = time.perf_counter() - t0
    for l in loops:
        l.close()
    wall, ev, st = res[False]
    gwall, gev, gst = res[True]
    return _line("volna edges/s (64 batched meshes, one persistent launch per step)",
                 edges * args.steps / wall, "edges/s", args.steps, args.warmup,
                 wall / args.steps * 1e3,
                 f"volna batch: 64 triangle meshes, {cells} cells, {edges} edges, seed 0",
                 roofline={"bound": "hbm", "achieved": ubytes * args.steps / ev / 1e9,
                           "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                           "frac": ubytes * args.steps / ev / PEAK_HBM,
                           "traffic": load_workload_traffic("volna-batch", f"owner_ppb{ppb}"),
                           "algorithmic_bytes_per_step": ubytes},
                 graph_edges_per_s=edges * args.steps / gwall,
                 graph_ms_per_step=gwall / args.steps * 1e3,
                 per_mesh_launch_edges_per_s=edges * args.steps / seq,
                 tuning=tune,
                 batch={"points_per_block": ppb, "grid": st["grid"],
                        "work_items": st["num_work_items"], "threads": st["threads"],
                        "lds_bytes": st["lds_bytes"]})


def share_mesh(mesh: dict, part: np.ndarray, s: int):
    """The submesh one GPU runs for share s of a partition of the cells:
    the owned cells (numbered first), the edges touching them, and the halo
    cells and nodes those edges read.  Returns (mesh dict, owned, halo)."""
    pecell, pedge = mesh["pecell"], mesh["pedge"]
    own = part == s
    emask = own[pecell[:, 0]] | own[pecell[:, 1]]
    e_cells, e_nodes = pecell[emask], pedge[emask]
    owned = np.nonzero(own)[0]
    halo = np.setdiff1d(np.unique(e_cells), owned)
    cells = np.concatenate([owned, halo])
    cmap = np.full(mesh["num_cells"], -1, np.int64)
    cmap[cells] = np.arange(cells.size)
    nodes = np.unique(e_nodes)
    nmap = np.full(mesh["num_nodes"], -1, np.int64)
    nmap[nodes] = np.arange(nodes.size)
    sub = {"pecell": cmap[e_cells].astype(np.uint32), "pedge": nmap[e_nodes].astype(np.uint32),
           "x": mesh["x"][nodes], "q": mesh["q"][cells], "adt": mesh["adt"][cells],
           "num_cells": int(cells.size), "num_nodes": int(nodes.size),
           "num_edges": int(emask.sum())}
    return sub, int(owned.size), int(halo.size)


def bench_share(args) -> dict:
    """Each share of a G-way partition of the airfoil's cells (recursive
    bisection of the cell graph into equal parts), run on this GPU as one
    GPU of a sharded run would run it: its owned cells, every edge touching
    one of them (edges across the cut are computed on both sides) and the
    halo cells and nodes they read, with the default plan (owner compute on
    240-point parts, the gather kernel).  The halo cells are owned by blocks
    of the share's plan too, so their (discarded) sums are written: O(cut)
    extra work.  Reports each share's time per step and its halo, the
    q / adt values a timestep would refresh from the neighbours."""
    from cuda_fot_amd import capi, meshgen, plan, res_calc_problem
    g = args.share
    mesh = meshgen.airfoil(args.nx, args.ny, seed=0, shuffle=7)
    ind, end = plan.graph_csr(mesh["pecell"], mesh["num_cells"])
    part, nparts = plan.bisect_partition(ind, end, -(-mesh["num_cells"] // g))
    shares = []
    for s in range(nparts):
        sub, owned, halo = share_mesh(mesh, part, s)
        p = res_calc_problem(sub)
        p.partition_points(240)
        loop, _, _ = make_loop(p, "hierarchical", 512, True, False, "owner", 2, True, False,
                               True, False, False, args.nontemporal)
        wall, ev = time_loop(loop, args.steps, args.warmup)
        loop.close()
        shares.append({"owned_cells": owned, "halo_cells": halo, "edges": sub["num_edges"],
                       "us_per_step": round(ev / args.steps * 1e6, 2),
                       "halo_refresh_bytes": 40 * halo})
    slowest = max(sh["us_per_step"] for sh in shares)
    return {"metric": f"res_calc: the {nparts} shares of a {nparts}-way cell partition, each "
                      "timed alone on one GPU",
            "workload": f"res_calc airfoil {args.nx}x{args.ny}, shuffled, seed 0",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "shares": shares,
            "slowest_share_us_per_step": slowest,
            "edges_total": mesh["num_edges"],
            "edges_computed": int(sum(sh["edges"] for sh in shares)),
            "note": "one GPU runs each share in turn; no multi-GPU run is measured"}


def tune_res_calc(args) -> dict:
    """The autotuner's choice for the airfoil as generated (shuffled, not
    reordered): the learned cost model's shortlist of strategy classes,
    their configurations measured; when the fastest misses its class's
    prediction by more than 5 %, further classes measured one at a time, or
    all 21 past a miss of 30 % (tuner.choose); --tune exhaustive measures
    all 21.  Sets args to the
    chosen configuration (its ordering, block or part size and launch
    mode); returns the tuning record."""
    from cuda_fot_amd import capi, tuner
    _, p = build_problem(args, "none")
    f = capi.functor(capi.KERNEL_RES_CALC)
    rec = {"mode": args.tune}
    if args.tune == "learned":
        rows, X, T = tuner.load_corpus(tuner.CORPUS_PATH)
        model = tuner.CostModel(seed=0)
        model.fit(X, tuner.class_times(T))
        r = tuner.Tuner(model).learned(p, f, warmup=2, steps=5)
        pick, t = r["pick"], r["times"]
        rec.update(corpus_rows=len(rows), fell_back=bool(r["fell_back"]),
                   measured_candidates=int(r["measured"]),
                   shortlist=[tuner.CLASSES[c] for c in r["shortlist"]],
                   predicted_class_ms={tuner.CLASSES[c]: round(float(v), 4)
                                       for c, v in enumerate(r["predicted"])},
                   measured_pick_ms=round(float(t[pick]), 4))
    else:
        pick, t = tuner.Tuner.exhaustive(p, f, warmup=2, steps=5)
        rec.update(measured_candidates=int(np.isfinite(t).sum()))
    rec["measured_ms"] = {tuner.candidate_name(f, i): round(float(v), 4)
                          for i, v in enumerate(t) if np.isfinite(v)}
    k = tuner.candidates(f)[pick]
    rec["pick"] = k["name"]
    if k["strategy"] == capi.STRATEGY_ATOMIC:
        args.strategy, args.partition = "atomic", "none"
    elif k["strategy"] == capi.STRATEGY_GLOBAL_COLOUR:
        args.strategy, args.partition, args.block_size = "global", "none", 256
    elif k["flags"] & capi.LOOP_OWNER_COMPUTE:
        args.strategy, args.order, args.partition = "hierarchical", "owner", "points"
        args.points_per_block = k["bs"]
        args.ept, args.accumulate, args.nodes = 2, "gather", "gather"
    else:
        args.strategy, args.block_size = "hierarchical", k["bs"]
        args.order = "two-phase" if k["flags"] & capi.LOOP_TWO_PHASE else "launch"
        args.partition = {"given": "none", "gps": "gps", "bisect": "bisect"}[k["ordering"]]
        args.no_stage = not k["flags"] & capi.LOOP_STAGE_INPUTS
    return rec


def bench_single(args, workload):
    """mini_aero 96^3 faces, hex scatter or lulesh forces on 128^3 elements,
    getacc on 1024^2 quads, mine on the 1024^2 grid: owner compute on a
    point-partitioned mesh beside plain atomics."""
    from cuda_fot_amd import Problem, capi, getacc_problem, lulesh_problem, meshgen, mini_aero_problem
    if workload == "getacc":
        n = 1024
        p = getacc_problem(n + 1, n + 1, seed=0, block_size=512)
        f = capi.functor(capi.KERNEL_GETACC)
        nn, nc = (n + 1) ** 2, n * n
        # ndmass/ndarea/ndub/ndvb read and written (64 B per node); cnmass,
        # rho, cnwt, cnfx, cnfy (17 doubles) and 4 map ids (152 B per quad)
        units, ubytes = nc, 64 * nn + 152 * nc
        name = f"getacc corner scatter {n}x{n} quads, shuffled, seed 0"
    elif workload == "mine":
        # the fillEdgeList grid as generate_grid lays it out (not shuffled),
        # point data and edge weights dim 4 U[0,1), seed 0
        n = 1024
        e = meshgen.grid_edges(n, n)
        pd = meshgen.uniform(n * n * 4, 0).reshape(n * n, 4)
        w = meshgen.uniform(e.shape[0] * 4, 1).reshape(4, -1)
        p = Problem([e], [n * n], [pd], [w], 512)
        f = capi.functor(capi.KERNEL_MINE, 2, 4, 4)
        # unique bytes: point data read (32 B per point), weights (32 B) and
        # the map (8 B) per edge, out read and written (64 B per point):
        # 184 467 456 B, 88 B per edge, at 1024^2
        units, ubytes = e.shape[0], 32 * n * n + 32 * e.shape[0] + 8 * e.shape[0] + 64 * n * n
        name = f"mine {n}x{n} grid, {e.shape[0]} edges, dim 4 point and edge data, seed 0"
    elif workload == "lulesh":
        n = 128
        p = lulesh_problem(n + 1, n + 1, n + 1, seed=0, block_size=512)
        f = capi.functor(capi.KERNEL_LULESH)
        nn, nc = (n + 1) ** 3, n ** 3
        # xyz 24 B + f 24 B read + 24 B written per node; sigma 24 B, volume
        # 8 B written and 8 map ids (32 B) per element
        units, ubytes = nc, 72 * nn + 64 * nc
        name = f"lulesh element forces {n}^3 elements, shuffled, seed 0"
    elif workload == "mini_aero":
        n = 96
        m = meshgen.hex_faces(n, n, n, seed=0)
        p = mini_aero_problem(m, block_size=512)
        f = capi.functor(capi.KERNEL_MINI_AERO)
        units, ubytes = m["num_faces"], 304 * m["num_cells"] + 104 * m["num_faces"]
        name = f"mini_aero {n}^3 hex cells, {m['num_faces']} interior faces, shuffled, seed 0"
    else:
        n = 128
        cells = meshgen.hex_grid(n + 1, n + 1, n + 1)
        cells = cells[meshgen.permutation(cells.shape[0], 5)]
        nn = (n + 1) ** 3
        forces = meshgen.normal(cells.shape[0] * 24, 6).reshape(24, -1)
        p = Problem([cells], [nn], [np.zeros((nn, 3))], [forces], 512)
        f = capi.functor(capi.KERNEL_HEX_SCATTER)
        units, ubytes = cells.shape[0], 48 * nn + 224 * cells.shape[0]
        name = f"hex scatter {n}^3 elements, 8 corners x 3 forces, shuffled, seed 0"
    from cuda_fot_amd import DeviceArray
    from cuda_fot_amd.device import Loop
    import copy
    # mine's CPU loops run on the grid as generated, before any reordering
    p_cpu = copy.deepcopy(p) if workload == "mine" and not args.no_cpu else None

    def run(prob, strategy, flags, bs, with_stats=False):
        flags |= nt_flags(args.nontemporal)
        dev = [DeviceArray.from_host(d) for d in prob.point_data]
        out = DeviceArray.from_host(prob.point_data[0])
        cdev = [DeviceArray.from_host(c) for c in prob.cell_data]
        if flags & capi.LOOP_OWNER_COMPUTE:
            part = prob.point_partition
            if part is not None:
                flags |= capi.LOOP_POINT_PARTITION
        else:
            part = prob.partition_vector if strategy == capi.STRATEGY_HIERARCHICAL else None
        l = Loop(f, prob.maps, prob.num_points, dev, out, cdev, soa=prob.soa, strategy=strategy,
                 block_size=bs, partition=part, flags=flags)
        wall, ev = time_loop(l, args.steps, args.warmup)
        st = l.stats()
        l.close()
        return (wall, ev, st) if with_stats else (wall, ev)

    wa = run(p, capi.STRATEGY_ATOMIC, 0, 256)[0] if not args.no_atomic_ref else None
    sweep = None
    if args.sweep_block_sizes:
        # atomics on the mesh as generated; the hierarchical loop (LDS-cached
        # point data, thread colours, one launch per block colour) on the
        # GPS-ordered mesh with consecutive blocks
        g = copy.deepcopy(p)
        g.reorder("gps")
        sweep = {}
        for bs in (128, 256, 512, 1024):
            _, ea = run(p, capi.STRATEGY_ATOMIC, 0, bs)
            row = {"atomic_ms": ea / args.steps * 1e3}
            try:
                _, eh, st = run(g, capi.STRATEGY_HIERARCHICAL, 0, bs, with_stats=True)
                row.update(hierarchical_gps_ms=eh / args.steps * 1e3,
                           hierarchical_launches=st["num_launches_per_iter"])
            except capi.ColourOverflow as e:
                row["hierarchical_gps_ms"] = None
                row["hierarchical_error"] = str(e)
            sweep[str(bs)] = row
        del g
    if args.tune in ("learned", "exhaustive"):
        # the autotuner's choice on the mesh as generated: the learned
        # shortlist of strategy classes measured (with the fallback), or all
        # 21 configurations; the pick runs as deployed (reordered copy for
        # the ordered configurations)
        from cuda_fot_amd import tuner
        t0 = time.perf_counter()
        rec = {"mode": args.tune}
        if args.tune == "learned":
            rows, X, T = tuner.load_corpus(tuner.CORPUS_PATH)
            model = tuner.CostModel(seed=0)
            model.fit(X, tuner.class_times(T))
            r = tuner.Tuner(model).learned(p, f, warmup=2, steps=5)
            pick, t = r["pick"], r["times"]
            rec.update(fell_back=bool(r["fell_back"]), shortlist=[tuner.CLASSES[c] for c in
                                                                   r["shortlist"]],
                       predicted_class_ms={tuner.CLASSES[c]: round(float(v), 4)
                                           for c, v in enumerate(r["predicted"])},
                       measured_candidates=int(r["measured"]))
        else:
            pick, t = tuner.Tuner.exhaustive(p, f, warmup=2, steps=5)
            rec["measured_candidates"] = int(np.isfinite(t).sum())
        rec["measured_ms"] = {tuner.candidate_name(f, i): round(float(v), 4)
                              for i, v in enumerate(t) if np.isfinite(v)}
        rec["tune_seconds"] = round(time.perf_counter() - t0, 1)
        k = tuner.candidates(f)[pick]
        rec["pick"] = k["name"]
        q, _ = tuner.deploy(p, f, pick)
        wall, ev, st = run(q, k["strategy"], k["flags"], k["bs"], with_stats=True)
        # the same name as the explicit options give this configuration
        if k["flags"] & capi.LOOP_OWNER_COMPUTE:
            variant = f"owner_ppb{k['bs']}"
        elif k["strategy"] == capi.STRATEGY_HIERARCHICAL:
            variant = f"{ {'given': 'none'}.get(k['ordering'], k['ordering'])}_b{k['bs']}" + \
                ("_two-phase" if k["flags"] & capi.LOOP_TWO_PHASE else "")
        else:
            variant = k["name"]
        cfg = {"strategy": tuner.CLASSES[k["cls"]], "block_size": k["bs"],
               "ordering": k["ordering"], "launches_per_step": st["num_launches_per_iter"],
               "block_colours": st["num_colours"], "blocks": st["num_blocks"],
               "tune": rec}
    elif args.partition in ("gps", "kway", "bisect"):
        # the hierarchical loop on the GPS-ordered mesh: consecutive blocks
        # of --block-size cells (gps), or k-way parts of --block-size cells
        # with the mesh reordered to them and the points renumbered (kway);
        # one launch per block colour (--order launch), the two-phase
        # writeback (two-phase) or the dependency-ordered launch (dep)
        p.reorder("gps")
        if args.partition in ("kway", "bisect"):
            p.block_size = args.block_size
            p.partition(1.001, method=args.partition)
            p.reorder_to_partition()
            p.renumber_points()
        order = args.order if args.order != "owner" else "launch"
        flags = {"launch": 0, "dep": capi.LOOP_DEPENDENCY_ORDER,
                 "two-phase": capi.LOOP_TWO_PHASE}[order]
        wall, ev, st = run(p, capi.STRATEGY_HIERARCHICAL, flags, args.block_size,
                           with_stats=True)
        variant = f"{args.partition}_b{args.block_size}" + \
            ("" if order == "launch" else "_" + order)
        cfg = {"strategy": "hierarchical", "partition": args.partition,
               "block_size": args.block_size, "block_order": order,
               "launches_per_step": st["num_launches_per_iter"],
               "block_colours": st["num_colours"], "blocks": st["num_blocks"],
               "avg_thread_colours": round(st["avg_thread_colours"], 3),
               "reuse_factor": round(st["reuse_factor"], 4)}
    else:
        p.partition_points(args.points_per_block, args.numbering)
        wall, ev, st = run(p, capi.STRATEGY_HIERARCHICAL, capi.LOOP_OWNER_COMPUTE,
                           args.points_per_block, with_stats=True)
        variant = f"owner_ppb{args.points_per_block}"
        cfg = {"strategy": "owner compute", "points_per_block": args.points_per_block,
               "redundancy": round(st["redundancy"], 4)}
    unit = {"mini_aero": "faces/s", "mine": "edges/s"}.get(workload, "elements/s")
    extra = {}
    if wa is not None:
        extra.update(atomic_per_s=units * args.steps / wa, speedup_vs_atomic=wa / wall)
    if workload == "mine" and not args.no_cpu:
        extra["cpu_baseline"] = cpu_baseline_mine(p_cpu, f, args.cpu_seconds)
    line = _line(f"{workload} {unit}", units * args.steps / wall, unit, args.steps, args.warmup,
                 wall / args.steps * 1e3, name,
                 roofline={"bound": "hbm", "achieved": ubytes * args.steps / ev / 1e9,
                           "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                           "frac": ubytes * args.steps / ev / PEAK_HBM,
                           "traffic": load_workload_traffic(workload, variant,
                                                            st["nontemporal"]),
                           "algorithmic_bytes_per_step": ubytes},
                 **extra)
    line["config"].update(cfg, nontemporal=st["nontemporal"])
    if sweep is not None:
        line["block_size_sweep"] = sweep
    return line


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the engine has no multi-GPU path; an "
              "owner-compute plan sharded by point partition would need no result "
              "exchange, only a halo refresh of the read-only inputs (--share G times "
              "each share of a G-way partition on this GPU)", file=sys.stderr)
        sys.exit(2)
    from cuda_fot_amd import capi
    if capi.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    if args.workload == "copy":
        print(json.dumps(bench_copy(args)), flush=True)
        return
    if args.workload == "volna-batch":
        args.tune = args.tune or "exhaustive"
        print(json.dumps(bench_volna_batch(args)), flush=True)
        return
    if args.workload in ("mini_aero", "hex", "lulesh", "mine", "getacc"):
        print(json.dumps(bench_single(args, args.workload)), flush=True)
        return
    if args.share > 1:
        print(json.dumps(bench_share(args)), flush=True)
        return
    tune = None
    if (args.tune or "learned") != "none":
        args.tune = args.tune or "learned"
        tune = tune_res_calc(args)
    t_plan = time.perf_counter()
    mesh, p = build_problem(args, args.partition)
    loop, dev, out = make_loop(p, args.strategy, args.block_size, not args.no_stage,
                               args.graph, args.order, args.ept, args.nodes == "gather", args.point_terms,
                               args.accumulate == "gather", args.persistent, args.hbm_lists,
                               args.nontemporal)
    plan_s = time.perf_counter() - t_plan
    stats = loop.stats()
    wall, ev = time_loop(loop, args.steps, args.warmup)
    launch_ms, launch_med = launch_durations(loop, args.steps)
    capi.call("fot_loop_check", loop.handle)
    ne = mesh["num_edges"]
    ubytes = unique_bytes(mesh["num_nodes"], mesh["num_cells"], ne)
    # the dominant kernel's bandwidth: algorithmic bytes over its average
    # launch duration, from HIP events bracketing the K timed launches (the
    # per-launch event pairs are reported too; their own packets add ~1 us)
    achieved = ubytes * args.steps / ev
    traffic = load_traffic(args, stats)
    result = {
        "metric": "res_calc edges/s, achieved HBM GB/s (% of peak); learned-tuned vs plain atomics",
        "value": ne * args.steps / wall,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"res_calc airfoil {args.nx}x{args.ny} quad cells, "
                               f"{ne} interior edges, shuffled, seed 0",
                   "strategy": args.strategy, "block_size": args.block_size,
                   "partition": args.partition,
                   "points_per_block": args.points_per_block if args.partition == "points" else None,
                   "numbering": args.numbering if args.partition == "points" else None,
                   "stage_inputs": not args.no_stage, "graph": args.graph,
                   "layout": "SOA" if args.soa else "AOS",
                   "block_order": args.order, "cells_per_thread": args.ept,
                   "nodes": args.nodes, "point_terms": args.point_terms,
                   "accumulate": args.accumulate,
                   "launches_per_step": stats["num_launches_per_iter"],
                   "block_colours": stats["num_colours"],
                   "avg_thread_colours": round(stats["avg_thread_colours"], 3),
                   "reuse_factor": round(stats["reuse_factor"], 4),
                   "redundancy": round(stats["redundancy"], 4),
                   "threads_per_block": stats["threads_per_block"],
                   "stage_per_thread": stats["stage_per_thread"],
                   "persistent": args.persistent, "launch_grid": stats["launch_grid"],
                   "lists": "hbm" if args.hbm_lists else "lds",
                   "nontemporal": stats["nontemporal"],
                   "plan_seconds": round(plan_s, 2), "tune": tune},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": traffic,
                     "algorithmic_bytes_per_launch": ubytes,
                     "launch_ms_mean": launch_ms, "launch_ms_median": launch_med,
                     "event_ms_per_step": ev / args.steps * 1e3},
    }
    loop.close()
    result["roofline"]["copy_gbps"] = copy_bandwidth() / 1e9
    if args.strategy != "atomic" and not args.no_atomic_ref:
        # the plain atomics loop on the same (reordered) mesh, and on the
        # shuffled mesh as generated
        la, _, _ = make_loop(p, "atomic", 256, False, args.graph)
        wa, _ = time_loop(la, args.steps, args.warmup)
        la.close()
        _, ps = build_problem(args, "none")
        ls, _, _ = make_loop(ps, "atomic", 256, False, args.graph)
        wsh, _ = time_loop(ls, args.steps, args.warmup)
        ls.close()
        result["atomic_edges_per_s"] = ne * args.steps / wa
        result["atomic_shuffled_edges_per_s"] = ne * args.steps / wsh
        result["speedup_vs_atomic"] = result["value"] / max(result["atomic_edges_per_s"],
                                                            result["atomic_shuffled_edges_per_s"])
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(mesh, args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
