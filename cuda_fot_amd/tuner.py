"""Autotuner and learned cost model for the loop engine (no reference
counterpart).

The autotuner searches 21 configurations per loop, grouped in six strategy
classes (the reference's measurement matrix varies the ordering and the
block size the same way, res_calc.cu:183-264):
  0 atomic             FP64 global atomics, 256 threads
  1 global             global colouring, 256 threads
  2 hier               hierarchical, one launch per block colour, blocks of
                       B = 128, 256, 512, 1024 cells on the mesh as given
  3 two-phase          hierarchical, every block in one launch with the
                       shared points merged by a second (FOT_LOOP_TWO_PHASE),
                       B = 128 .. 1024 on the mesh as given
  4 two-phase ordered  the same on the GPS-ordered mesh (consecutive blocks)
                       and on recursive-bisection parts of B cells with the
                       mesh reordered to them and its points renumbered,
                       B = 128 .. 1024 each
  5 owner              owner-compute hierarchical, parts of 128 / 256 / 512
                       points
res_calc's hierarchical configurations stage q and adt in LDS; its owner
configurations run the gather kernel with two edges per thread on parts of
120 / 240 / 480 points (about 2.13 edges per point fill 128 / 256 / 512
threads), on a copy of the mesh partitioned into parts of that size and
renumbered so each part's points are contiguous (Problem.partition_points).
Every configuration is measured as it is deployed: the ordered ones on a
reordered copy of the mesh.

`measure` times configurations on the GPU (fot_autotune_measure); the
exhaustive search measures all 21.  `CostModel` is an FP32 MLP 24 -> 64 ->
64 -> 6 (libfot's MFMA kernels) trained on the autotuner's own timings to
predict log(ms per step) of each class's best configuration from 24 cheap
mesh features (fot_loop_features).  `Tuner.learned` measures only the
configurations of the classes the model predicts within the tolerance (5 %)
of its best prediction (the shortlist) and takes the fastest; when that
one's measurement exceeds the model's prediction for its class by more than
the tolerance, it measures further classes one at a time (those predicted
below the best time measured so far), or, when the miss exceeds
WIDEN_LIMIT or nothing on the shortlist can be planned, all the rest
(`choose`).
"""
from __future__ import annotations

import copy
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

CLASSES = ["atomic", "global", "hier", "two-phase", "two-phase-ordered", "owner"]
BLOCK_SIZES = (128, 256, 512, 1024)
# the autotuner's own timings, the cost model's training corpus
CORPUS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                           "tuner_corpus.jsonl")
# held-out evaluation set: mini_aero 48^3 .. 96^3 (heldout_specs), measured
HELDOUT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                            "tuner_heldout_mini_aero.jsonl")
NUM_FEATURES = 24


def owner_points(func: capi.Functor) -> tuple[int, int, int]:
    """Points per part of the three owner-compute configurations."""
    return (120, 240, 480) if func.kernel == capi.KERNEL_RES_CALC else (128, 256, 512)


def owner_flags(func: capi.Functor) -> int:
    """Loop flags of the owner-compute configurations."""
    if func.kernel == capi.KERNEL_RES_CALC:
        return capi.LOOP_OWNER_COMPUTE | capi.LOOP_TWO_PER_THREAD | capi.LOOP_OWNER_GATHER
    return capi.LOOP_OWNER_COMPUTE


def renumbered_owner(func: capi.Functor) -> bool:
    """Whether the owner configurations run on the mesh renumbered for their
    partition (res_calc's gather kernel)."""
    return func.kernel == capi.KERNEL_RES_CALC


def candidates(func: capi.Functor) -> list[dict]:
    """Every configuration the autotuner measures: class, name, strategy,
    block size (cells per block, or points per owner part), flags and the
    ordering it runs on ("given", "gps", "bisect" or "points")."""
    stage = capi.LOOP_STAGE_INPUTS if func.kernel == capi.KERNEL_RES_CALC else 0
    tp = capi.LOOP_TWO_PHASE
    H = capi.STRATEGY_HIERARCHICAL
    out = [dict(cls=0, name="atomic", strategy=capi.STRATEGY_ATOMIC, bs=256, flags=0,
                ordering="given"),
           dict(cls=1, name="global", strategy=capi.STRATEGY_GLOBAL_COLOUR, bs=256, flags=0,
                ordering="given")]
    out += [dict(cls=2, name=f"hier{b}", strategy=H, bs=b, flags=stage, ordering="given")
            for b in BLOCK_SIZES]
    out += [dict(cls=3, name=f"2phase{b}", strategy=H, bs=b, flags=stage | tp,
                 ordering="given") for b in BLOCK_SIZES]
    out += [dict(cls=4, name=f"2phase{b}-gps", strategy=H, bs=b, flags=stage | tp,
                 ordering="gps") for b in BLOCK_SIZES]
    out += [dict(cls=4, name=f"2phase{b}-bisect", strategy=H, bs=b, flags=stage | tp,
                 ordering="bisect") for b in BLOCK_SIZES]
    own = "points" if renumbered_owner(func) else "given"
    out += [dict(cls=5, name=f"owner{n}", strategy=H, bs=n, flags=owner_flags(func),
                 ordering=own) for n in owner_points(func)]
    return out


NUM_CANDIDATES = 21


def candidate_name(func: capi.Functor, i: int) -> str:
    return candidates(func)[i]["name"]


def class_members(func: capi.Functor, c) -> list[int]:
    """Indices of the configurations of class c (index or name)."""
    c = CLASSES.index(c) if isinstance(c, str) else c
    return [i for i, k in enumerate(candidates(func)) if k["cls"] == c]


# class of each configuration index (the same for every functor)
CLASS_OF = np.array([0, 1] + [2] * 4 + [3] * 4 + [4] * 8 + [5] * 3)


def class_times(T) -> np.ndarray:
    """Per class, the best measured time of its configurations (rows x 21 ->
    rows x 6): -1 when none of them can be planned, NaN when none was
    measured."""
    T = np.atleast_2d(np.asarray(T, np.float64))
    out = np.full((T.shape[0], len(CLASSES)), np.nan)
    for c in range(len(CLASSES)):
        sub = T[:, CLASS_OF == c]
        pos = np.where(np.isfinite(sub) & (sub > 0), sub, np.inf).min(1)
        unplannable = np.all(np.isfinite(sub) & (sub < 0), 1)
        out[:, c] = np.where(np.isfinite(pos), pos, np.where(unplannable, -1.0, np.nan))
    return out


def deploy(problem, func: capi.Functor, i: int, gps=None):
    """The problem as configuration i runs on it (a reordered copy for the
    ordered configurations, the problem itself otherwise) and the partition
    its config takes (None: the engine's own blocks).  `gps`: the problem's
    GPS-ordered copy, when one is at hand."""
    k = candidates(func)[i]
    if k["ordering"] == "given":
        return problem, None
    if k["ordering"] == "points":
        q = copy.deepcopy(problem)
        q.partition_points(k["bs"])
        return q, q.point_partition
    if gps is None:
        gps = copy.deepcopy(problem)
        gps.reorder("gps")
    if k["ordering"] == "gps":
        return gps, None
    q = copy.deepcopy(gps)
    q.block_size = k["bs"]
    q.partition(1.001, method="bisect")
    q.reorder_to_partition()
    q.renumber_points()
    return q, q.partition_vector


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


def measure_configs(problem, func, cands, warmup, steps, partition=None):
    """ms per step of the configurations `cands` (candidates() entries) on
    `problem` as given, `partition` passed to each (a point partition for the
    owner ones); -1 where one cannot be planned."""
    up = _Uploaded(problem)
    cfgs = (capi.LoopConfig * len(cands))()
    keep = None if partition is None else capi.u32(partition)
    for c, k in zip(cfgs, cands):
        c.strategy, c.block_size, c.flags = k["strategy"], k["bs"], k["flags"]
        if keep is not None:
            c.partition = ptr(keep)
            if k["ordering"] == "points":
                c.flags |= capi.LOOP_POINT_PARTITION
    t = np.empty(len(cands), np.float32)
    call("fot_autotune_measure", C.byref(func), C.byref(up.desc), cfgs, len(cands), warmup,
         steps, ptr(t))
    return t


def measure(problem, func: capi.Functor, which=None, warmup: int = 3,
            steps: int = 10) -> np.ndarray:
    """ms per step of the configurations `which` (default all 21); -1 where
    one cannot be planned, NaN where it was not measured.  Configurations on
    the mesh as given share one upload; the GPS-ordered ones share one
    reordered copy; each bisection and owner part size gets its own."""
    cands = candidates(func)
    idx = list(range(len(cands))) if which is None else sorted(set(int(i) for i in which))
    out = np.full(len(cands), np.nan, np.float32)
    given = [i for i in idx if cands[i]["ordering"] == "given"]
    if given:
        out[given] = measure_configs(problem, func, [cands[i] for i in given], warmup, steps)
    ordered = None
    if any(cands[i]["ordering"] in ("gps", "bisect") for i in idx):
        ordered = copy.deepcopy(problem)
        ordered.reorder("gps")
    gps = [i for i in idx if cands[i]["ordering"] == "gps"]
    if gps:
        out[gps] = measure_configs(ordered, func, [cands[i] for i in gps], warmup, steps)
    for i in idx:
        if cands[i]["ordering"] in ("bisect", "points"):
            q, part = deploy(problem, func, i, ordered)
            out[i] = measure_configs(q, func, [cands[i]], warmup, steps, part)[0]
    return out


class CostModel:
    """MLP on standardised features -> standardised log(ms) of each class's
    best configuration."""

    def __init__(self, seed: int = 0, hidden: int = 64):
        self.net = MLP(NUM_FEATURES, hidden, len(CLASSES), seed=seed)
        self.fmu = self.fsd = self.tmu = self.tsd = None

    def fit(self, X, T, epochs: int = 3000, lr: float = 3e-3, weight_decay: float = 1e-4):
        """X: rows x 24 features; T: rows x 6 class times (class_times), ms,
        -1 = cannot be planned, NaN = not measured."""
        X = np.asarray(X, np.float64)
        T = np.asarray(T, np.float64).copy()
        # a class that cannot be planned (-1) trains as 10x the mesh's
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


def shortlist(pred_ms, tolerance: float = 0.05) -> list[int]:
    """Classes predicted within `tolerance` of the best prediction."""
    pred_ms = np.asarray(pred_ms, np.float64)
    return [int(i) for i in np.nonzero(pred_ms <= (1.0 + tolerance) * pred_ms.min())[0]]


def members(classes) -> list[int]:
    """Configuration indices of the given classes."""
    return [int(i) for i in np.nonzero(np.isin(CLASS_OF, list(classes)))[0]]


# A miss of the picked class's prediction by more than this factor means the
# model does not know the mesh: the measured search over all 21.  A smaller
# miss widens the shortlist one class at a time instead.
WIDEN_LIMIT = 1.3


def choose(pred_cls, measure, tolerance: float = 0.05):
    """The learned choice, measuring through `measure(classes)`, which
    returns the 21 configuration times with those classes' configurations
    filled in (ms; -1 = cannot be planned).

    The shortlisted classes are measured.  The fastest of them is kept when
    it is within `tolerance` of its class's prediction.  When it misses by
    more than WIDEN_LIMIT (or nothing on the shortlist can be planned), every
    configuration is measured and the fastest kept.  Otherwise the class
    predicted fastest among those not yet measured and predicted below the
    best time measured so far (within the tolerance) is measured, and the
    test repeats, until the pick is within its prediction or no such class
    is left.  Returns (configuration, measured beyond the shortlist,
    configurations measured, times)."""
    pred = np.asarray(pred_cls, np.float64)
    short = shortlist(pred, tolerance)
    done = list(short)
    t = np.asarray(measure(short), np.float64)
    count = len(members(short))
    while True:
        subs = members(done)
        ok = np.where(t[subs] > 0, t[subs], np.inf)
        pick = subs[int(np.argmin(ok))]
        best = float(ok.min())
        limit = pred[CLASS_OF[pick]]
        if np.isfinite(best) and best <= (1.0 + tolerance) * limit:
            return pick, len(done) > len(short), count, t
        if len(done) == len(short) and not (np.isfinite(best) and best <= WIDEN_LIMIT * limit):
            t = np.asarray(measure([c for c in range(len(CLASSES)) if c not in done]), np.float64)
            return int(np.argmin(np.where(t > 0, t, np.inf))), True, NUM_CANDIDATES, t
        rest = [int(c) for c in np.argsort(pred, kind="stable")
                if int(c) not in done and pred[c] < (1.0 + tolerance) * best]
        if not rest:
            return pick, len(done) > len(short), count, t
        t = np.asarray(measure([rest[0]]), np.float64)
        done.append(rest[0])
        count += len(members([rest[0]]))


def learned_choice(pred_cls, measured_ms, tolerance: float = 0.05):
    """choose() on known measurements (the 21 configuration times): returns
    (configuration, measured beyond the shortlist, configurations
    measured)."""
    m = np.asarray(measured_ms, np.float64)
    seen = np.full(m.shape, np.nan)

    def reveal(classes):
        idx = members(classes)
        seen[idx] = m[idx]
        return seen

    pick, widened, count, _ = choose(pred_cls, reveal, tolerance)
    return pick, widened, count


class Tuner:
    def __init__(self, model: CostModel | None = None, tolerance: float = 0.05):
        self.model, self.tolerance = model, tolerance

    @staticmethod
    def exhaustive(problem, func, **kw):
        t = measure(problem, func, **kw)
        ok = np.where(t > 0, t, np.inf)
        return int(np.argmin(ok)), t

    def learned(self, problem, func, **kw) -> dict:
        """choose() on the problem: the configurations measured as deployed.
        The record: pick, predicted class times, shortlist, whether classes
        beyond the shortlist were measured (fell_back), how many
        configurations were measured and their times (NaN: not measured)."""
        pred = self.model.predict(features(problem, func))[0]
        times = np.full(NUM_CANDIDATES, np.nan)

        def run(classes):
            idx = members(classes)
            t = measure(problem, func, which=idx, **kw)
            times[idx] = t[idx]
            return times

        pick, widened, count, t = choose(pred, run, self.tolerance)
        return {"pick": pick, "predicted": pred, "shortlist": shortlist(pred, self.tolerance),
                "measured_pick": float(t[pick]), "fell_back": bool(widened),
                "measured": int(count), "times": t}


# ---- evaluation ---------------------------------------------------------------
# the method without the learned part: one fixed configuration everywhere,
# the medium owner-compute parts
DEFAULT_CHOICE = 19


def evaluate(model: CostModel, X, T, tolerance: float = 0.05) -> dict:
    """The learned choice with its fallback against the exhaustive search on
    rows whose 21 configuration times are all measured (T, ms; -1 = cannot
    be planned).  Reports, over the rows: how often the choice is within the
    tolerance of the best, its mean and largest slowdown, how often it fell
    back, how many configurations it measured against the 21 of the
    exhaustive search, the same for the predicted-best class alone (its
    configurations measured, no fallback), and the fixed default's
    slowdown."""
    X, T = np.asarray(X, np.float64), np.asarray(T, np.float64)
    pred = model.predict(X)
    ok = np.where(T > 0, T, np.inf)
    best = ok.min(1)
    final_t, measured, within, fell = [], 0, 0, 0
    top_t, top_within, top_measured = [], 0, 0
    for i in range(len(X)):
        c, fb, k = learned_choice(pred[i], T[i], tolerance)
        final_t.append(T[i, c] / best[i])
        within += T[i, c] <= (1 + tolerance) * best[i]
        fell += fb
        measured += k
        subs = members([int(np.argmin(pred[i]))])
        t_top = ok[i, subs].min()
        top_t.append(t_top / best[i])
        top_within += t_top <= (1 + tolerance) * best[i]
        top_measured += len(subs)
    default = ok[:, DEFAULT_CHOICE] / best
    n = len(X)
    return {"rows": n, "final_within_tolerance": within / n,
            "final_mean_slowdown": float(np.mean(final_t)),
            "final_max_slowdown": float(np.max(final_t)),
            "fallbacks": fell / n,
            "measured_per_mesh": measured / n,
            "exhaustive_measured_per_mesh": NUM_CANDIDATES,
            "top_class_within_tolerance": top_within / n,
            "top_class_mean_slowdown": float(np.mean(top_t)),
            "top_class_measured_per_mesh": top_measured / n,
            "default_mean_slowdown": float(np.mean(default)),
            "default_within_tolerance": float(np.mean(default <= 1 + tolerance))}


def load_corpus(path):
    """Rows, features (rows x 24) and configuration times (rows x 21, ms)."""
    rows = [json.loads(l) for l in open(path) if l.strip()]
    X = np.array([r["features"] for r in rows], np.float64)
    T = np.array([r["ms"] for r in rows], np.float64)
    return rows, X, T


# ---- corpus of autotuner timings ---------------------------------------------
FAMILIES = ["mine", "res_calc", "volna", "hex", "mini_aero"]
ORDERINGS = ["random", "rcm", "partitioned"]
# mini_aero meshes of the first 500 corpus meshes stay at 44^3 cells or
# fewer (corpus_mini_aero_specs adds 45^3 .. 62^3; the held-out set is 64^3
# .. 96^3)
MINI_AERO_MAX_EDGES = 3 * 44 ** 3


def corpus_meshes(n: int, seed: int = 0, min_edges: float = 2e4, max_edges: float = 5e5):
    """(family, ordering, size) specs: 2D quads (mine), airfoil quads
    (res_calc), triangles (volna), 3D hexes (scatter) and hex faces
    (mini_aero, at most 44^3 cells), each in random, RCM and
    point-partitioned orderings, edges log-uniform in [min, max]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        fam = FAMILIES[i % len(FAMILIES)]
        hi = min(max_edges, MINI_AERO_MAX_EDGES) if fam == "mini_aero" else max_edges
        e = float(np.exp(rng.uniform(np.log(min_edges), np.log(hi))))
        out.append({"id": i, "family": fam, "ordering": ORDERINGS[(i // len(FAMILIES)) % 3],
                    "edges": int(e), "seed": int(rng.integers(1 << 30))})
    return out


def corpus_large_specs(n: int = 150, seed: int = 1, min_edges: float = 5e5,
                       max_edges: float = 2e6):
    """Larger meshes appended to the corpus (ids 500 on): the four families
    other than mini_aero (which stays below the held-out sizes) in the three
    orderings, edges log-uniform in [5e5, 2e6]."""
    rng = np.random.default_rng(seed)
    fams = [f for f in FAMILIES if f != "mini_aero"]
    out = []
    for i in range(n):
        e = float(np.exp(rng.uniform(np.log(min_edges), np.log(max_edges))))
        out.append({"id": 500 + i, "family": fams[i % len(fams)],
                    "ordering": ORDERINGS[(i // len(fams)) % 3], "edges": int(e),
                    "seed": int(rng.integers(1 << 30))})
    return out


def corpus_mini_aero_specs(n: int = 45, seed: int = 2):
    """mini_aero hex face meshes of 45^3 to 62^3 cells appended to the corpus
    (ids 650 on), below the held-out set's 64^3 .. 96^3, in the three
    orderings."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        c = int(rng.integers(45, 63))
        out.append({"id": 650 + i, "family": "mini_aero", "ordering": ORDERINGS[i % 3],
                    "edges": 3 * c ** 3, "seed": int(rng.integers(1 << 30))})
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


def heldout_specs():
    """The held-out evaluation set: mini_aero-style hex face meshes of 64^3
    to 96^3 cells (the corpus holds mini_aero up to 44^3 only) in the three
    orderings."""
    out = []
    for i, n in enumerate([64, 80, 96]):
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
    measures every configuration on corpus meshes FIRST..N-1 not yet in OUT,
    appending one JSON line per mesh (features, ms per step) and printing
    progress."""
    argv = sys.argv[1:] if argv is None else argv
    out, n = argv[1], int(argv[2])
    seed = int(argv[3]) if len(argv) > 3 else 0
    first = int(argv[4]) if len(argv) > 4 else 0
    done = set()
    if os.path.exists(out):
        with open(out) as f:
            done = {json.loads(l)["id"] for l in f if l.strip()}
    measure_specs([sp for sp in corpus_meshes(n, seed)[first:] if sp["id"] not in done], out)


def eval_main(argv=None):
    """python -m cuda_fot_amd.tuner eval CORPUS.jsonl [HELDOUT_FAMILY|HELDOUT.jsonl|-] [SEED]:
    trains the cost model on the GPU on every mesh but the held-out family
    (or on all of it against a held-out file, or on a seeded 80 % split when
    '-') and prints the evaluation as JSON."""
    argv = sys.argv[1:] if argv is None else argv
    rows, X, T = load_corpus(argv[1])
    held = argv[2] if len(argv) > 2 else "-"
    seed = int(argv[3]) if len(argv) > 3 else 0
    if os.path.exists(held):
        # a separate held-out file: train on all of the corpus
        _, Xh, Th = load_corpus(held)
        model = CostModel(seed=seed)
        losses = model.fit(X, class_times(T))
        full = np.all(np.isfinite(Th), 1)
        out = evaluate(model, Xh[full], Th[full])
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
    losses = model.fit(X[~test], class_times(T[~test]))
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
    elif len(sys.argv) > 1 and sys.argv[1] == "corpus-large":
        # python -m cuda_fot_amd.tuner corpus-large OUT.jsonl [FIRST] [LAST]
        specs = corpus_large_specs()
        first = int(sys.argv[3]) if len(sys.argv) > 3 else 0
        last = int(sys.argv[4]) if len(sys.argv) > 4 else len(specs)
        measure_specs(specs[first:last], sys.argv[2])
    elif len(sys.argv) > 1 and sys.argv[1] == "corpus-mini-aero":
        # python -m cuda_fot_amd.tuner corpus-mini-aero OUT.jsonl
        measure_specs(corpus_mini_aero_specs(), sys.argv[2])
    elif len(sys.argv) > 1 and sys.argv[1] == "heldout":
        # python -m cuda_fot_amd.tuner heldout OUT.jsonl: measure the held-out set
        measure_specs(heldout_specs(), sys.argv[2])
