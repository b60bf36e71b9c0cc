"""Autotuner and learned cost model for the loop engine (no reference
counterpart).

Six candidate strategies per loop (the choices the engine offers):
  0 atomic       FP64 global atomics, 256 threads
  1 global       global colouring, 256 threads
  2 hier256      hierarchical, per-block-colour launches, 256 cells per block
  3 owner small  owner-compute hierarchical, parts of 128 points
  4 owner medium owner-compute hierarchical, parts of 256 points
  5 owner large  owner-compute hierarchical, parts of 512 points
For res_calc the owner candidates run the gather kernel with two edges per
thread on parts of 120 / 240 / 480 points: about 2.13 edges per point, so
those parts fill 128 / 256 / 512 threads exactly (256 points would round up
to 320 threads).  They are measured as they are deployed: on a copy of the
mesh partitioned into parts of that size and renumbered so that each part's
points are contiguous (Problem.partition_points), which that kernel needs to
address its owned cells directly.

`measure` is the exhaustive search: every candidate planned and timed on the
GPU (fot_autotune_measure).  `CostModel` is an FP32 MLP 24 -> 64 -> 64 -> 6
(libfot's MFMA kernels) trained on the autotuner's own timings to predict
log(ms per step) of every candidate from 24 cheap mesh features
(fot_loop_features).  `Tuner.learned` measures only the candidates the model
predicts within the tolerance (5 %) of its best prediction (the shortlist:
one candidate unless the model sees a near tie it cannot resolve) and takes
the fastest of them; when that one's measurement exceeds the model's
prediction for it by more than the tolerance, or nothing on the shortlist can
be planned, it falls back to the measured search.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

from . import capi
from .capi import call, ptr
from .device import DeviceArray, make_desc
from .mlp import MLP

NAMES = ["atomic", "global", "hier256", "owner128", "owner256", "owner512"]
# the autotuner's own timings, the cost model's training corpus
CORPUS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                           "tuner_corpus.jsonl")
# held-out evaluation set: mini_aero 48^3 .. 96^3 (heldout_specs), measured
HELDOUT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                            "tuner_heldout_mini_aero.jsonl")
NUM_FEATURES = 24


def owner_points(func: capi.Functor) -> tuple[int, int, int]:
    """Points per part of the three owner-compute candidates."""
    return (120, 240, 480) if func.kernel == capi.KERNEL_RES_CALC else (128, 256, 512)


def owner_flags(func: capi.Functor) -> int:
    """Loop flags of the owner-compute candidates."""
    if func.kernel == capi.KERNEL_RES_CALC:
        return capi.LOOP_OWNER_COMPUTE | capi.LOOP_TWO_PER_THREAD | capi.LOOP_OWNER_GATHER
    return capi.LOOP_OWNER_COMPUTE


def candidate_name(func: capi.Functor, i: int) -> str:
    if i < 3:
        return NAMES[i]
    return f"owner{owner_points(func)[i - 3]}"


def candidate_configs(func: capi.Functor):
    """The six fot_loop_config candidates for a functor."""
    res_calc = func.kernel == capi.KERNEL_RES_CALC
    own, pts = owner_flags(func), owner_points(func)
    spec = [(capi.STRATEGY_ATOMIC, 256, 0), (capi.STRATEGY_GLOBAL_COLOUR, 256, 0),
            (capi.STRATEGY_HIERARCHICAL, 256, capi.LOOP_STAGE_INPUTS if res_calc else 0),
            (capi.STRATEGY_HIERARCHICAL, pts[0], own), (capi.STRATEGY_HIERARCHICAL, pts[1], own),
            (capi.STRATEGY_HIERARCHICAL, pts[2], own)]
    cfgs = (capi.LoopConfig * len(spec))()
    for i, (s, bs, fl) in enumerate(spec):
        cfgs[i].strategy, cfgs[i].block_size, cfgs[i].flags = s, bs, fl
    return cfgs


def features(problem, func: capi.Functor) -> np.ndarray:
    """fot_loop_features of a Problem (host arrays only)."""
    maps = [capi.u32(m) for m in problem.maps]
    d = capi.ProblemDesc()
    d.num_cells = maps[0].shape[0]
    d.num_mappings = len(maps)
    for k, m in enumerate(maps):
        d.cell_to_node[k] = ptr(m)
        d.mesh_dim[k] = m.shape[1]
        d.num_points[k] = int(problem.num_points[k])
        pd = problem.point_data[k]
        d.point_dim[k] = 1 if pd is None or pd.ndim == 1 else (pd.shape[0] if problem.soa
                                                               else pd.shape[1])
    d.num_cell_data = len(problem.cell_data)
    for k, c in enumerate(problem.cell_data):
        d.cell_dim[k] = c.shape[0] if c.ndim > 1 else 1
    d.soa = int(problem.soa)
    out = np.empty(NUM_FEATURES, np.float32)
    call("fot_loop_features", C.byref(func), C.byref(d), ptr(out))
    return out


class _Uploaded:
    """A Problem's data on the device and its fot_problem_desc."""

    def __init__(self, problem):
        self.dev = [None if x is None else DeviceArray.from_host(x) for x in problem.point_data]
        self.out = DeviceArray.from_host(problem.point_data[0])
        self.cdev = [DeviceArray.from_host(c) for c in problem.cell_data]
        self.desc, self._keep = make_desc(problem.maps, problem.num_points, self.dev, self.out,
                                          self.cdev, problem.soa)


def renumbered_owner(func: capi.Functor) -> bool:
    """Whether the owner candidates run on the mesh renumbered for their
    partition (res_calc's gather kernel)."""
    return func.kernel == capi.KERNEL_RES_CALC


def _measure_cfgs(problem, func, cfgs_list, warmup, steps, partition=None):
    up = _Uploaded(problem)
    cfgs = (capi.LoopConfig * len(cfgs_list))(*cfgs_list)
    keep = None
    if partition is not None:
        keep = capi.u32(partition)
        for c in cfgs:
            c.partition = ptr(keep)
            c.flags |= capi.LOOP_POINT_PARTITION
    t = np.empty(len(cfgs_list), np.float32)
    call("fot_autotune_measure", C.byref(func), C.byref(up.desc), cfgs, len(cfgs_list), warmup,
         steps, ptr(t))
    return t


def measure(problem, func: capi.Functor, which=None, warmup: int = 3,
            steps: int = 10) -> np.ndarray:
    """ms per step of the candidates `which` (default all); -1 where a
    candidate cannot be planned, NaN where it was not measured."""
    import copy
    all_cfgs = candidate_configs(func)
    idx = list(range(len(NAMES))) if which is None else list(which)
    out = np.full(len(NAMES), np.nan, np.float32)
    own = [i for i in idx if i >= 3] if renumbered_owner(func) else []
    plain = [i for i in idx if i not in own]
    if plain:
        out[plain] = _measure_cfgs(problem, func, [all_cfgs[i] for i in plain], warmup, steps)
    for i in own:
        q = copy.deepcopy(problem)
        q.partition_points(owner_points(func)[i - 3])
        out[i] = _measure_cfgs(q, func, [all_cfgs[i]], warmup, steps, q.point_partition)[0]
    return out


class CostModel:
    """MLP on standardised features -> standardised log(ms) per candidate."""

    def __init__(self, seed: int = 0, hidden: int = 64):
        self.net = MLP(NUM_FEATURES, hidden, len(NAMES), seed=seed)
        self.fmu = self.fsd = self.tmu = self.tsd = None

    def fit(self, X, T, epochs: int = 3000, lr: float = 3e-3, weight_decay: float = 1e-4):
        X = np.asarray(X, np.float64)
        T = np.asarray(T, np.float64).copy()
        # a candidate that cannot be planned (-1) trains as 10x the mesh's
        # slowest measured time, so the model learns to avoid it
        for i in range(len(T)):
            ok = np.isfinite(T[i]) & (T[i] > 0)
            if ok.any():
                T[i, np.isfinite(T[i]) & (T[i] < 0)] = 10.0 * T[i, ok].max()
        mask = np.isfinite(T) & (T > 0)
        logT = np.where(mask, np.log(np.where(mask, T, 1.0)), 0.0)
        self.fmu, self.fsd = X.mean(0), X.std(0) + 1e-6
        self.tmu = (logT * mask).sum() / max(mask.sum(), 1)
        self.tsd = math.sqrt(((logT - self.tmu) ** 2 * mask).sum() / max(mask.sum(), 1)) + 1e-6
        xs = (X - self.fmu) / self.fsd
        ts = np.where(mask, (logT - self.tmu) / self.tsd, 0.0)
        return self.net.train(xs, ts, mask.astype(np.float32), epochs=epochs, lr=lr,
                              weight_decay=weight_decay)

    def predict(self, X) -> np.ndarray:
        xs = (np.atleast_2d(np.asarray(X, np.float64)) - self.fmu) / self.fsd
        return np.exp(self.net.forward(xs).astype(np.float64) * self.tsd + self.tmu)


class Tuner:
    def __init__(self, model: CostModel | None = None, tolerance: float = 0.05):
        self.model, self.tolerance = model, tolerance

    @staticmethod
    def exhaustive(problem, func, **kw):
        t = measure(problem, func, **kw)
        ok = np.where(t > 0, t, np.inf)
        return int(np.argmin(ok)), t

    def learned(self, problem, func, **kw) -> dict:
        pred = self.model.predict(features(problem, func))[0]
        short = shortlist(pred, self.tolerance)
        t = measure(problem, func, which=short, **kw)
        ok = np.where(t[short] > 0, t[short], np.inf)
        pick = short[int(np.argmin(ok))]
        res = {"pick": pick, "predicted": pred, "shortlist": short,
               "measured_pick": float(t[pick]), "fell_back": False, "measured": len(short)}
        if not (t[pick] > 0) or t[pick] > (1.0 + self.tolerance) * pred[pick]:
            best, times = self.exhaustive(problem, func, **kw)
            res.update(pick=best, fell_back=True, measured=len(short) + len(NAMES), times=times)
        return res


def shortlist(pred_ms, tolerance: float = 0.05) -> list[int]:
    """Candidates predicted within `tolerance` of the best prediction."""
    pred_ms = np.asarray(pred_ms, np.float64)
    return [int(i) for i in np.nonzero(pred_ms <= (1.0 + tolerance) * pred_ms.min())[0]]


def shortlist_pick(pred_ms, measured_ms, tolerance: float = 0.05):
    """The learned choice on known measurements: the fastest measured of the
    shortlist; if that one exceeds the model's prediction for it by more
    than `tolerance` (or nothing on the shortlist can be planned), the
    measured best.  Returns (choice, fell_back, candidates measured)."""
    short = shortlist(pred_ms, tolerance)
    m = np.asarray(measured_ms, np.float64)
    ok = np.where(m > 0, m, np.inf)
    pick = short[int(np.argmin(ok[short]))]
    if not (m[pick] > 0) or m[pick] > (1.0 + tolerance) * pred_ms[pick]:
        return int(np.argmin(ok)), True, len(short) + len(NAMES)
    return pick, False, len(short)


def fallback_pick(pred_ms, measured_ms, tolerance: float = 0.05):
    """The learned choice with the fallback rule, on known measurements:
    take the predicted-best candidate; if its measured time exceeds the
    model's prediction for it by more than `tolerance` (or it cannot be
    planned), take the measured best instead.  Returns (choice, fell_back)."""
    pick = int(np.argmin(pred_ms))
    t = measured_ms[pick]
    ok = np.where(measured_ms > 0, measured_ms, np.inf)
    if not (t > 0) or t > (1.0 + tolerance) * pred_ms[pick]:
        return int(np.argmin(ok)), True
    return pick, False


# ---- evaluation ---------------------------------------------------------------
DEFAULT_CHOICE = 4  # the method without the learned part: the medium owner parts everywhere


def evaluate(model: CostModel, X, T, tolerance: float = 0.05) -> dict:
    """The learned choice with its fallback against the exhaustive search on
    rows whose six times are all measured (T, ms; -1 = cannot be planned).
    Reports, over the rows: how often the predicted-best pick is within the
    tolerance of the best; the mean slowdown of the pick, of the choice after
    the fallback rule and of the fixed default; how many candidates the
    learned tuner measured against the six of the exhaustive search."""
    X, T = np.asarray(X, np.float64), np.asarray(T, np.float64)
    pred = model.predict(X)
    best = np.where(T > 0, T, np.inf).min(1)
    pick_t, final_t, measured, within, fell = [], [], 0, 0, 0
    sl_t, sl_measured, sl_fell, sl_within = [], 0, 0, 0
    for i in range(len(X)):
        pick = int(np.argmin(pred[i]))
        choice, fb = fallback_pick(pred[i], T[i], tolerance)
        t_pick = T[i, pick] if T[i, pick] > 0 else np.inf
        pick_t.append(t_pick / best[i])
        final_t.append(T[i, choice] / best[i])
        within += t_pick <= (1 + tolerance) * best[i]
        fell += fb
        measured += 1 + (len(NAMES) if fb else 0)
        c, f2, k = shortlist_pick(pred[i], T[i], tolerance)
        sl_t.append(T[i, c] / best[i])
        sl_within += T[i, c] <= (1 + tolerance) * best[i]
        sl_fell += f2
        sl_measured += k
    d = T[:, DEFAULT_CHOICE]
    default = np.where(d > 0, d, np.inf) / best
    n = len(X)
    return {"rows": n, "pick_within_tolerance": within / n,
            "pick_mean_slowdown": float(np.mean(pick_t)),
            "final_mean_slowdown": float(np.mean(final_t)),
            "final_max_slowdown": float(np.max(final_t)),
            "fallbacks": fell / n,
            "measured_per_mesh": measured / n, "exhaustive_measured_per_mesh": len(NAMES),
            "shortlist_within_tolerance": sl_within / n,
            "shortlist_mean_slowdown": float(np.mean(sl_t)),
            "shortlist_max_slowdown": float(np.max(sl_t)),
            "shortlist_fallbacks": sl_fell / n,
            "shortlist_measured_per_mesh": sl_measured / n,
            "default_mean_slowdown": float(np.mean(default)),
            "default_within_tolerance": float(np.mean(default <= 1 + tolerance))}


def load_corpus(path):
    rows = [json.loads(l) for l in open(path) if l.strip()]
    X = np.array([r["features"] for r in rows], np.float64)
    T = np.array([r["ms"] for r in rows], np.float64)
    return rows, X, T


# ---- corpus of autotuner timings ---------------------------------------------
def corpus_meshes(n: int, seed: int = 0, min_edges: float = 5e4, max_edges: float = 2e6):
    """(family, ordering, size) specs: 2D quad (mine), airfoil quads
    (res_calc), triangles (volna), 3D hexes (scatter), each in random, RCM
    and point-partitioned orderings, edges log-uniform in [min, max]."""
    rng = np.random.default_rng(seed)
    fams = ["mine", "res_calc", "volna", "hex"]
    orders = ["random", "rcm", "partitioned"]
    out = []
    for i in range(n):
        e = float(np.exp(rng.uniform(np.log(min_edges), np.log(max_edges))))
        out.append({"id": i, "family": fams[i % 4], "ordering": orders[(i // 4) % 3],
                    "edges": int(e), "seed": int(rng.integers(1 << 30))})
    return out


def build_corpus_problem(spec):
    """The Problem and functor of a corpus spec."""
    from . import meshgen
    from .problem import Problem, res_calc_problem, volna_problem
    fam, e, s = spec["family"], spec["edges"], spec["seed"]
    rng = np.random.default_rng(s)
    aspect = float(rng.uniform(0.5, 2.0))
    if fam == "mine":
        # ~2 edges per point
        N = max(8, int(math.sqrt(e / 2 * aspect)))
        M = max(8, int(e / 2 / N))
        edges = meshgen.grid_edges(N, M)
        edges = edges[meshgen.permutation(edges.shape[0], s + 1)]
        n = N * M
        pd = meshgen.uniform(n * 4, s + 2).reshape(n, 4)
        w = meshgen.uniform(edges.shape[0] * 4, s + 3).reshape(4, -1)
        p, func = Problem([edges], [n], [pd], [w], 256), capi.functor(capi.KERNEL_MINE, 2, 4, 4)
    elif fam == "res_calc":
        nx = max(8, int(math.sqrt(e / 2 * aspect)))
        ny = max(8, int(e / 2 / nx))
        p, func = res_calc_problem(meshgen.airfoil(nx, ny, seed=s)), \
            capi.functor(capi.KERNEL_RES_CALC)
    elif fam == "volna":
        nx = max(8, int(math.sqrt(e / 3 * aspect)))
        ny = max(8, int(e / 3 / nx))
        p, func = volna_problem(meshgen.tri_mesh(nx, ny, seed=s)), \
            capi.functor(capi.KERNEL_VOLNA)
    elif fam == "mini_aero":
        from .problem import mini_aero_problem
        n = max(4, int(round((e / 3) ** (1 / 3))))   # ~3 n^3 interior faces
        p, func = mini_aero_problem(meshgen.hex_faces(n, n, n, seed=s)), \
            capi.functor(capi.KERNEL_MINI_AERO)
    else:
        m = max(4, int(round(e ** (1 / 3))))
        cells = meshgen.hex_grid(m + 1, m + 1, m + 1)
        cells = cells[meshgen.permutation(cells.shape[0], s + 5)]
        nn = (m + 1) ** 3
        forces = meshgen.normal(cells.shape[0] * 24, s + 6).reshape(24, -1)
        p, func = Problem([cells], [nn], [np.zeros((nn, 3))], [forces], 256), \
            capi.functor(capi.KERNEL_HEX_SCATTER)
    if spec["ordering"] == "rcm":
        p.reorder("rcm")
    elif spec["ordering"] == "partitioned":
        p.partition_points(256)
        p.point_partition = None  # the ordering stays; candidates plan their own
    return p, func


def corpus_extra_specs(n: int = 60, seed: int = 1):
    """mini_aero-style hex face meshes of 16^3 to 44^3 cells in the three
    orderings, appended to the corpus (ids 500 on); the evaluation set
    (heldout_specs, 48^3 to 96^3) stays out of it."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        c = int(rng.integers(16, 45))
        out.append({"id": 500 + i, "family": "mini_aero", "ordering":
                    ["random", "rcm", "partitioned"][i % 3], "edges": 3 * c ** 3,
                    "seed": int(rng.integers(1 << 30))})
    return out


def heldout_specs():
    """The held-out evaluation set: mini_aero-style hex face meshes (a kernel
    the corpus never holds) of 48^3 to 96^3 cells in the three orderings."""
    out = []
    for i, n in enumerate([48, 64, 80, 96]):
        for j, order in enumerate(["random", "rcm", "partitioned"]):
            out.append({"id": 1000 + 3 * i + j, "family": "mini_aero", "ordering": order,
                        "edges": 3 * n ** 3, "seed": 7 + i})
    return out


def measure_specs(specs, out):
    """Measures every candidate on each spec, one JSON line per mesh."""
    for spec in specs:
        t0 = time.perf_counter()
        p, func = build_corpus_problem(spec)
        x = features(p, func)
        t = measure(p, func)
        rec = dict(spec, cells=int(p.num_cells), features=[float(v) for v in x],
                   ms=[float(v) for v in t], seconds=round(time.perf_counter() - t0, 2))
        with open(out, "a") as f:
            f.write(json.dumps(rec) + "\n")
        print(f"mesh {spec['id']} {spec['family']}/{spec['ordering']} cells {p.num_cells}: "
              + " ".join(f"{v:.3f}" for v in t) + f" ({rec['seconds']} s)", flush=True)


def corpus_main(argv=None):
    """python -m cuda_fot_amd.tuner corpus OUT.jsonl N [SEED] [FIRST]:
    measures every candidate on corpus meshes FIRST..N-1, appending one JSON
    line per mesh (features, ms per step) and printing progress."""
    argv = sys.argv[1:] if argv is None else argv
    out, n = argv[1], int(argv[2])
    seed = int(argv[3]) if len(argv) > 3 else 0
    first = int(argv[4]) if len(argv) > 4 else 0
    done = set()
    if os.path.exists(out):
        with open(out) as f:
            done = {json.loads(l)["id"] for l in f if l.strip()}
    for spec in corpus_meshes(n, seed)[first:]:
        if spec["id"] in done:
            continue
        t0 = time.perf_counter()
        p, func = build_corpus_problem(spec)
        x = features(p, func)
        t = measure(p, func)
        rec = dict(spec, cells=int(p.num_cells), features=[float(v) for v in x],
                   ms=[float(v) for v in t], seconds=round(time.perf_counter() - t0, 2))
        with open(out, "a") as f:
            f.write(json.dumps(rec) + "\n")
        print(f"mesh {spec['id']} {spec['family']}/{spec['ordering']} cells {p.num_cells}: "
              + " ".join(f"{v:.3f}" for v in t) + f" ({rec['seconds']} s)", flush=True)


def eval_main(argv=None):
    """python -m cuda_fot_amd.tuner eval CORPUS.jsonl [HELDOUT_FAMILY|-] [SEED]:
    trains the cost model on the GPU on every mesh but the held-out family
    (or a seeded 20 % split when '-') and prints the evaluation as JSON."""
    argv = sys.argv[1:] if argv is None else argv
    rows, X, T = load_corpus(argv[1])
    held = argv[2] if len(argv) > 2 else "-"
    seed = int(argv[3]) if len(argv) > 3 else 0
    if os.path.exists(held):
        # a separate held-out file: train on all of the corpus
        _, Xh, Th = load_corpus(held)
        model = CostModel(seed=seed)
        losses = model.fit(X, T)
        out = evaluate(model, Xh[np.all(np.isfinite(Th), 1)], Th[np.all(np.isfinite(Th), 1)])
        out.update(train_rows=len(rows), held_out=os.path.basename(held),
                   final_train_loss=float(losses[-1]), first_train_loss=float(losses[0]))
        print(json.dumps(out))
        return out
    if held == "-":
        test = np.random.default_rng(seed).random(len(rows)) < 0.2
    else:
        test = np.array([r["family"] == held for r in rows])
    full = np.all(np.isfinite(T), 1)
    model = CostModel(seed=seed)
    losses = model.fit(X[~test], T[~test])
    out = evaluate(model, X[test & full], T[test & full])
    out.update(train_rows=int((~test).sum()), held_out=held,
               final_train_loss=float(losses[-1]), first_train_loss=float(losses[0]))
    print(json.dumps(out))
    return out


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "corpus":
        corpus_main()
    elif len(sys.argv) > 1 and sys.argv[1] == "eval":
        eval_main()
    elif len(sys.argv) > 1 and sys.argv[1] == "corpus-extra":
        # python -m cuda_fot_amd.tuner corpus-extra OUT.jsonl
        measure_specs(corpus_extra_specs(), sys.argv[2])
    elif len(sys.argv) > 1 and sys.argv[1] == "remeasure":
        # python -m cuda_fot_amd.tuner remeasure FAMILY OUT.jsonl: measure the
        # seed-0 corpus meshes of one family again (after a candidate change)
        measure_specs([sp for sp in corpus_meshes(500, 0) if sp["family"] == sys.argv[2]],
                      sys.argv[3])
    elif len(sys.argv) > 1 and sys.argv[1] == "heldout":
        # python -m cuda_fot_amd.tuner heldout OUT.jsonl: measure the held-out set
        measure_specs(heldout_specs(), sys.argv[2])
