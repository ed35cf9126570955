#!/usr/bin/env python3
"""Benchmark of the masked-product path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs ktruss,tc,bc,batch,cmask] [--no-cpu]

Headline (`value`): full truss decomposition (k = 3..max) of the seeded
R-MAT scale-20 graph (edge factor 16, a/b/c/d = .57/.19/.19/.05,
symmetrised, self loops and duplicates removed, degree-relabelled; int32
CSR), in undirected edges per second: |E| / wall time of one decomposition
with the graph resident in HBM.  A step is one whole decomposition.

Also reported, each in its own object:
  tc  triangle counting C<L> = L*L on R-MAT scale 22 (same generator),
      edges/s: plain fixed thresholds (every B row streamed), the library
      default (plus hub bitmaps), autotuned, learned, learned + autotuned;
      the same variants and the k-truss on a 2048^2 grid and a 4M-vertex
      Chung-Lu graph;
  bc  batched Brandes BC, 4 batches of 256 seeded sources on the scale-20
      graph, FP64, source-edges/s;
  batch  triangle count + 4-truss of each of 4096 seeded uneven graphs in one
      hipGraph-captured step, edges/s, knobs autotuned;
  cmask  complement-mask product C<not A> = A*A on R-MAT scale 16, product
      terms/s.
`roofline` is for the dominant kernel of the headline run (the accumulator
kernel with the most device time): algorithmic bytes per launch over its
average launch duration (HIP events on the library's stream).
`cpu_baseline` times the oracle (the C++/OpenMP restatement of the
reference's CPU algorithm) on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "TC/k-truss edges/s, BC source-edges/s; learned vs plain, % HBM peak"
RMAT = dict(a=0.57, b=0.19, c=0.19)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def dominant(t: dict) -> dict:
    from graphkokkos_amd._lib import KERNEL_NAMES as ACC_NAMES
    a = int(np.argmax(t["acc_ms"]))
    launches = max(int(t["acc_launches"][a]), 1)
    ms = t["acc_ms"][a]
    byts = t["acc_bytes"][a]
    gbs = (byts / 1e9) / (ms / 1e3) if ms > 0 else 0.0
    return {"kernel": ACC_NAMES[a], "index": a, "launches": launches, "avg_launch_ms": ms / launches,
            "bytes_per_launch": byts / launches, "achieved_gbs": gbs}


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/hbm_traffic.txt),
    or None when that pair was not measured."""
    path = ROOT / "profiles" / "hbm_traffic.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        f = line.split()
        if len(f) == 4 and not line.startswith("#") and f[0] == kernel and f[1] == workload:
            return float(f[2]) / float(f[3])
    return None


def roofline_obj(d: dict, workload: str | None = None) -> dict:
    traffic = measured_traffic(d["kernel"], workload) if workload else None
    return {"bound": "hbm", "achieved": round(d["achieved_gbs"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(d["achieved_gbs"] / HBM_PEAK_GBS, 5),
            "traffic": round(traffic) if traffic is not None else None,
            "kernel": d["kernel"], "avg_launch_ms": round(d["avg_launch_ms"], 6),
            "bytes_per_launch": round(d["bytes_per_launch"]), "launches": d["launches"]}


def run_ktruss(ctx, A, steps: int, warmup: int) -> dict:
    from graphkokkos_amd import DeviceCsr  # noqa: F401
    dA = ctx.upload(A)
    buf = ctx.alloc(4 * A.nz)
    from graphkokkos_amd import _lib
    import ctypes as C
    kmax = C.c_int32(0)
    o = _lib.default_opts()

    def once():
        _lib.check(_lib.lib().gk_truss_decompose_dev(ctx.handle, dA.handle, buf.ptr, C.byref(kmax), C.byref(o)),
                   "gk_truss_decompose_dev")

    for i in range(warmup):
        once()
        log(f"ktruss warmup {i + 1}/{warmup}: {ctx.last_timing()['total_ms']:.1f} ms")
    ctx.sync()
    t0 = time.perf_counter()
    dev_ms = []
    for i in range(steps):
        once()
        dev_ms.append(ctx.last_timing()["total_ms"])
        log(f"ktruss step {i + 1}/{steps}: {dev_ms[-1]:.1f} ms")
    ctx.sync()
    wall = (time.perf_counter() - t0) / steps
    t = ctx.last_timing()
    truss = buf.to_host(np.uint32, A.nz)
    return {"wall_s": wall, "dev_ms": dev_ms, "timing": t, "kmax": int(kmax.value),
            "rounds": int(t["iterations"]), "truss_hist_head": np.bincount(truss)[:8].tolist()}


def run_tc(ctx, L, steps: int, warmup: int, opts=None) -> dict:
    dL = ctx.upload(L)
    count = None
    for _ in range(warmup):
        count = ctx.tc_count(dL, opts)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        count = ctx.tc_count(dL, opts)
    ctx.sync()
    wall = (time.perf_counter() - t0) / steps
    return {"wall_s": wall, "count": count, "timing": ctx.last_timing()}


def run_graph_build(ctx, scale: int, seed: int, reps: int = 3) -> float:
    """Device time (ms) of building the R-MAT graph in HBM (generate,
    symmetrise, deduplicate, degree-relabel, CSR and its strict lower part)
    with gk_dgraph_rmat; the tests check it against the host builder."""
    ctx.dgraph_rmat(scale, 16, seed=seed, **RMAT)  # warm-up
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        A, L = ctx.dgraph_rmat(scale, 16, seed=seed, **RMAT)
        A.free()
        L.free()
    ctx.sync()
    return (time.perf_counter() - t0) / reps * 1e3


def run_bc(ctx, A, sources, steps: int) -> dict:
    dA = ctx.upload(A)
    out = ctx.alloc(8 * A.m)
    ctx.bc_batch_dev(dA, sources, out)  # warm-up
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.bc_batch_dev(dA, sources, out)
    ctx.sync()
    wall = (time.perf_counter() - t0) / steps
    t = ctx.last_timing()
    bc = out.to_host(np.float64, A.m)
    return {"wall_s": wall, "timing": t, "levels": int(t["iterations"]), "bc_sum": float(bc.sum())}


def run_cmask(ctx, scale: int, seed: int, steps: int, no_cpu: bool, cpu_reps: int = 5) -> dict:
    """C<not A> = A*A on the symmetric R-MAT graph: every pair two hops
    apart that is not an edge, with its number of common neighbours
    (the diagonal included), through gk_masked_spgemm_cmask_dev."""
    from graphkokkos_amd import HostGraph
    G = HostGraph.rmat(scale, 16, seed=seed, **RMAT)
    A = G.A
    dA = ctx.upload(A)

    def timed(dense_rows, keep=False):
        # one result refilled step after step, as a caller that keeps C would
        C_ = ctx.masked_spgemm_cmask_dev(dA, dA, dA, dense_rows)  # warm-up
        nz = C_.nz
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.masked_spgemm_cmask_dev(dA, dA, dA, dense_rows, out=C_)
        ctx.sync()
        wall = (time.perf_counter() - t0) / steps
        t_ = ctx.last_timing()
        host = C_.to_host(values=True) if keep else None
        C_.free()
        return wall, nz, t_, host

    wall_sorted, nz_sorted, t_sorted, _ = timed(0)  # every row expanded and sorted
    log(f"cmask s{scale} sorted rows only: {wall_sorted * 1e3:.1f} ms")
    wall, nz, t, C_host = timed(-1, keep=not no_cpu)
    assert nz == nz_sorted, "the two ways of the rows disagree on C's size"
    terms = t["flop"]
    exp_gbs = t["bytes_alg"] / 1e9 / (t["kernel_ms"] / 1e3) if t["kernel_ms"] > 0 else 0.0
    dense_gbs = t["acc_bytes"][0] / 1e9 / (t["acc_ms"][0] / 1e3) if t["acc_ms"][0] > 0 else 0.0
    out = {"workload": f"complement-mask product C<not A> = A*A (two-hop non-adjacent pairs with their common "
                       f"neighbour counts), R-MAT scale {scale} ef16, pattern values",
           "vertices": G.n, "edges": G.edges, "terms": terms, "entries": nz,
           "value": terms / wall, "unit": "product terms/s", "ms": wall * 1e3, "device_ms": t["total_ms"],
           "expand_ms": t["kernel_ms"], "sort_reduce_ms": t["select_ms"], "chunks": t["launches"],
           "dense_rows": t["iterations"], "dense_ms": t["acc_ms"][0],
           "sorted_rows_only": {"ms": wall_sorted * 1e3, "value": terms / wall_sorted,
                                "expand_ms": t_sorted["kernel_ms"], "sort_reduce_ms": t_sorted["select_ms"]},
           "roofline": {"bound": "hbm", "achieved": round(dense_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(dense_gbs / HBM_PEAK_GBS, 5),
                        "traffic": measured_traffic("k_cm_dense", f"cmask-rmat{scale}"),
                        "kernel": "k_cm_dense (pool pass) + k_cm_dense_place", "bytes_per_launch": t["acc_bytes"][0] / 2,
                        "avg_launch_ms": t["acc_ms"][0] / 2, "launches": 2},
           "expand_roofline": {"bound": "hbm", "achieved": round(exp_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(exp_gbs / HBM_PEAK_GBS, 5), "traffic": None,
                               "kernel": "k_cm_expand", "bytes": t["bytes_alg"]}}
    log(f"cmask s{scale}: {wall * 1e3:.1f} ms, {terms:.3g} terms, {nz} entries, {t['iterations']} dense rows "
        f"({t['acc_ms'][0]:.1f} ms), sorted rows: expand {t['kernel_ms']:.1f} ms at {exp_gbs:.0f} GB/s, "
        f"sort+reduce {t['select_ms']:.1f} ms")
    if not no_cpu:
        from oracle import oracle as O
        dt, (rst, col, val) = best_of(lambda: O.mxmm_cmask(A, A, A), cpu_reps)
        assert int(rst[-1]) == nz, "oracle and GPU entry counts differ"
        assert np.array_equal(rst, C_host.Rst) and np.array_equal(col, C_host.Col) and \
            np.array_equal(val, C_host.Val), "oracle and GPU complement-mask products differ"
        del C_host
        out["cpu_baseline"] = {"value": terms / dt, "unit": "product terms/s", "cores": O.threads(), "kind": "port",
                               "sample": f"the same product (R-MAT scale {scale}, {terms:.3g} terms) by the "
                                         f"parallel restatement of the reference's two-pass complement-mask MSA, "
                                         f"best of {cpu_reps} after a warm-up: {dt:.2f}s; pattern and values equal "
                                         "the GPU's entry for entry"}
        log(f"cpu baseline cmask: {out['cpu_baseline']['value']:.3g} terms/s")
    dA.free()
    return out


def bc_bytes(n: int, nnz: int, levels: int, nsrc: int) -> float:
    # 2D (4(n+1) + 4 nnz) structure per level (forward + backward), 24 B per
    # source-edge (sigma read forward; W read backward; 8 B state each), 24 B
    # per source-vertex of state and reduction.
    return 2 * levels * (4 * (n + 1) + 4 * nnz) + 24.0 * nsrc * nnz + 24.0 * nsrc * n


# The plain fixed-heuristic kernel: the fixed mask-length thresholds as
# first set (MCA-wave up to 8, the short-row window kernel up to 128) pick
# each row's accumulator and every B row is streamed (hub bitmaps off).  The
# library default ("heuristic_hub") adds hub bitmaps with bit tests, sends
# masks up to 48 to MCA-wave and up to 32 to the short-row window kernel, and
# runs 16 long-row and 16 MCA workgroups per CU (triangle counting's own
# defaults, found by the autotuner on this graph).
PLAIN = {"hub_window": 0, "mca_max_mask": 8, "small_max_mask": 128}


# Triangle-counting options of the committed HBM-traffic profile of the
# tuned product (profiles/tc_rmat22_tuned_counters.md).
PROFILED_TC_OPTS = {"mca_max_mask": 48, "small_max_mask": 32, "hub_probe": 2, "blocks_long": 16, "blocks_mca": 16}


def tc_variants(tuned_h: dict, tuned_l: dict) -> list:
    from graphkokkos_amd import SEL_LEARNED
    return [("plain", PLAIN), ("heuristic_hub", {}), ("tuned_heuristic", tuned_h),
            ("learned", {"selector": SEL_LEARNED}), ("learned_tuned", tuned_l)]


def run_learned_graphs(ctx) -> dict:
    """The learned selector (already trained) against the fixed-threshold
    heuristic on the other two graphs of the learned-selector evaluation:
    a triangulated 2048 x 2048 grid (one diagonal per cell) and a Chung-Lu
    graph with exponent 2.0, 4M vertices, mean degree 16.  Triangle counting
    in the five variants, the knobs autotuned on each graph itself (outside
    the timed runs), and the full truss decomposition with each selector;
    counts and truss numbers must agree."""
    from graphkokkos_amd import HostGraph, SEL_LEARNED
    from graphkokkos_amd import _lib
    from graphkokkos_amd import selector as SEL
    import ctypes as C
    out = {}
    for name, make in (("grid-2048x2048-diag", lambda: HostGraph.grid(2048)),
                       ("chung-lu-4M-exp2.0-d16", lambda: HostGraph.chung_lu(1 << 22, 2.0, 16.0, seed=1))):
        t0 = time.perf_counter()
        G = make()
        log(f"{name}: n={G.n} |E|={G.edges} built in {time.perf_counter() - t0:.1f}s")
        res = {"edges": G.edges}
        dL = ctx.upload(G.L)
        tuned_h, _, _ = SEL.autotune_tc(ctx, dL, rounds=1, reps=3)
        tuned_l, _, _ = SEL.autotune_tc(ctx, dL, base={"selector": SEL_LEARNED}, rounds=1, reps=3,
                                        space=SEL.LEARNED_SPACE)
        dL.free()
        res["tuned_heuristic_opts"], res["learned_tuned_opts"] = tuned_h, tuned_l
        base = None
        for var, opts in tc_variants(tuned_h, tuned_l):
            r = run_tc(ctx, G.L, 5, 1, opts)
            if base is None:
                base = r
            assert r["count"] == base["count"], f"{name}: {var} changed the triangle count"
            res[f"tc_{var}_ms"] = r["wall_s"] * 1e3
        res["triangles"] = base["count"]
        for var in ("heuristic_hub", "tuned_heuristic", "learned", "learned_tuned"):
            res[f"tc_{var}_speedup"] = res["tc_plain_ms"] / res[f"tc_{var}_ms"]
        for var in ("learned", "learned_tuned"):
            res[f"tc_{var}_vs_default"] = res["tc_heuristic_hub_ms"] / res[f"tc_{var}_ms"]
            res[f"tc_{var}_vs_tuned_heuristic"] = res["tc_tuned_heuristic_ms"] / res[f"tc_{var}_ms"]
        dA = ctx.upload(G.A)
        buf = ctx.alloc(4 * max(G.A.nz, 1))
        kmax = C.c_int32(0)
        truss = {}
        for var, opts in (("heuristic", None), ("learned", {"selector": SEL_LEARNED})):
            o = _lib.default_opts(**(opts or {}))
            times = []
            for _ in range(3):
                t1 = time.perf_counter()
                _lib.check(_lib.lib().gk_truss_decompose_dev(ctx.handle, dA.handle, buf.ptr, C.byref(kmax),
                                                             C.byref(o)), "gk_truss_decompose_dev")
                ctx.sync()
                times.append(time.perf_counter() - t1)
            truss[var] = buf.to_host(np.uint32, G.A.nz)
            res[f"ktruss_{var}_ms"] = min(times) * 1e3
        assert np.array_equal(truss["heuristic"], truss["learned"]), f"{name}: selectors disagree on truss numbers"
        res["kmax"] = int(kmax.value)
        res["ktruss_learned_speedup"] = res["ktruss_heuristic_ms"] / res["ktruss_learned_ms"]
        buf.free()
        dA.free()
        log(f"{name}: tc plain {res['tc_plain_ms']:.2f} ms, heuristic+hub {res['tc_heuristic_hub_ms']:.2f}, "
            f"tuned {res['tc_tuned_heuristic_ms']:.2f}, learned {res['tc_learned_ms']:.2f}, "
            f"learned+tuned {res['tc_learned_tuned_ms']:.2f}; k-truss heuristic {res['ktruss_heuristic_ms']:.1f} ms, "
            f"learned {res['ktruss_learned_ms']:.1f} (kmax {res['kmax']})")
        out[name] = res
    return out


def run_batch(ctx, a) -> dict:
    """Triangle count + 4-truss of every graph of a seeded batch of uneven
    graphs (vertex counts log-uniform in [1024, 65536], half Erdos-Renyi with
    mean degree U[8,32], half Chung-Lu with exponent U[2.1,2.8] and mean
    degree 16), as one hipGraph-captured step; knobs autotuned."""
    from graphkokkos_amd import HostGraph
    from graphkokkos_amd import selector as SEL
    from graphkokkos_amd.graph import batch_config
    t0 = time.perf_counter()
    cfg = batch_config(a.batch_graphs, seed=a.seed)
    U = HostGraph.batch(**cfg)
    E = U.edges
    log(f"batch: {a.batch_graphs} graphs, {U.n} vertices, {E} edges built in {time.perf_counter() - t0:.1f}s")
    dA = ctx.upload(U.A)
    voff = U.voff
    log(f"batch: union uploaded ({U.A.nz} entries)")
    eager = ctx.batch_tc_ktruss(dA, voff, k=4, use_graph=False, reps=2)
    log(f"batch: eager step {eager['ms']:.2f} ms")
    plain = ctx.batch_tc_ktruss(dA, voff, k=4, use_graph=True, reps=max(a.steps, 3))
    assert np.array_equal(eager["triangles"], plain["triangles"]) and \
        np.array_equal(eager["truss_edges"], plain["truss_edges"]), "graph replay differs from the eager run"
    log(f"batch: eager {eager['ms']:.2f} ms, graph {plain['ms']:.2f} ms per step")
    t0 = time.perf_counter()
    tuned, _, tlog = SEL.autotune_batch(ctx, dA, voff, k=4, rounds=1, reps=2,
                                        progress=lambda o, t: log(f"batch: trial {o}: {t:.2f} ms"))
    best = ctx.batch_tc_ktruss(dA, voff, k=4, use_graph=True, reps=max(a.steps, 3), opts=tuned)
    assert np.array_equal(best["triangles"], plain["triangles"]) and \
        np.array_equal(best["truss_edges"], plain["truss_edges"]), "tuned options changed the results"
    log(f"batch: autotuned {tuned} ({len(tlog)} trials, {time.perf_counter() - t0:.1f}s): {best['ms']:.2f} ms")
    dA.free()
    return {"workload": f"{a.batch_graphs} graphs (log-uniform n in [1024, 65536], half ER d~U[8,32], half "
                        "Chung-Lu exponent~U[2.1,2.8] d=16): triangle count + 4-truss per graph, one "
                        "hipGraph-captured step",
            "graphs": a.batch_graphs, "vertices": int(U.n), "edges": int(E),
            "value": E / (best["ms"] / 1e3), "unit": "edges/s", "ms": best["ms"], "ms_eager": eager["ms"],
            "ms_graph_default": plain["ms"], "tuned_opts": tuned,
            "triangles_total": int(plain["triangles"].sum()),
            "truss4_edges_total": int(plain["truss_edges"].sum()),
            "_per_graph": (plain["triangles"], plain["truss_edges"])}


def batch_cpu_baseline(a, threads: int, gpu_triangles, gpu_truss_edges, reps: int) -> dict:
    """The oracle's triangle count and 4-truss loop on the first graphs of the
    same batch, up to about 3 s of work per pass; every graph's results must
    equal the batch step's, and the time is the best of `reps` passes after a
    warm-up pass."""
    from graphkokkos_amd import HostGraph
    from graphkokkos_amd.graph import batch_config
    from oracle import oracle as O
    cfg = batch_config(a.batch_graphs, seed=a.seed)
    graphs, edges, dt = [], 0, 0.0
    for g in range(a.batch_graphs):
        if cfg["kind"][g] == 0:
            G = HostGraph.er(int(cfg["n"][g]), float(cfg["mean_degree"][g]), seed=int(cfg["seed"][g]))
        else:
            G = HostGraph.chung_lu(int(cfg["n"][g]), float(cfg["exponent"][g]), float(cfg["mean_degree"][g]),
                                   seed=int(cfg["seed"][g]))
        t0 = time.perf_counter()
        tri = O.tc_sandia(G.L)
        keep, _ = O.ktruss(G.A, 4)
        dt += time.perf_counter() - t0
        assert tri == int(gpu_triangles[g]), f"batch graph {g}: oracle and GPU triangle counts differ"
        assert int(keep.sum()) // 2 == int(gpu_truss_edges[g]), f"batch graph {g}: oracle and GPU 4-truss sizes differ"
        graphs.append(G)
        edges += G.edges
        if dt > 3.0:
            break

    def one_pass():
        for G in graphs:
            O.tc_sandia(G.L)
            O.ktruss(G.A, 4)

    best, _ = best_of(one_pass, reps)
    return {"value": edges / best, "unit": "edges/s", "cores": threads, "kind": "port",
            "sample": f"the first {len(graphs)} graphs of the same batch ({edges} edges): triangle count (parallel "
                      f"MSA) and the reference's 4-truss loop, best of {reps} passes after a warm-up: {best:.2f}s; "
                      f"every graph's triangles and 4-truss edges equal the GPU step's"}


def calibrate_copy(ctx, nbytes: int = 256 << 20) -> None:
    """One device-to-device copy of nbytes (read nbytes, write nbytes): the
    reference kernel for converting FETCH_SIZE / WRITE_SIZE counter readings
    into bytes."""
    from graphkokkos_amd import _lib
    src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
    _lib.check(_lib.lib().gk_memcpy_d2d(ctx.handle, dst.ptr, src.ptr, nbytes), "gk_memcpy_d2d")
    src.free()
    dst.free()
    log(f"calibration copy: {nbytes} bytes read and written")


def best_of(fn, reps: int):
    """One untimed warm-up call of fn, then the best wall time of `reps`
    calls; returns (seconds, the last call's result)."""
    out = fn()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        best = min(best, time.perf_counter() - t0)
    return best, out


def cpu_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--legs", default="ktruss,tc,bc,batch,cmask")
    ap.add_argument("--cmask-scale", type=int, default=16)
    ap.add_argument("--ktruss-scale", type=int, default=20)
    ap.add_argument("--tc-scale", type=int, default=22)
    ap.add_argument("--bc-batches", type=int, default=4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=5,
                    help="CPU baselines: best of this many timed runs after one warm-up")
    ap.add_argument("--tc-variants", default="all", choices=["all", "plain", "default", "opts"],
                    help="plain: only the fixed-heuristic triangle count (no selector training or autotuning); "
                         "default: plain and the library default (hub bitmaps); opts: plain and the options "
                         "of --tc-opts (a fixed configuration, for profiling it)")
    ap.add_argument("--tc-opts", default="{}",
                    help='options of the "opts" triangle-counting variant, as JSON (gk_opts field names)')
    ap.add_argument("--batch-graphs", type=int, default=4096)
    ap.add_argument("--no-learned-graphs", action="store_true",
                    help="skip the grid and Chung-Lu graphs of the learned-selector evaluation")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with a 256 MiB device-to-device copy, the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        return 2
    legs = set(a.legs.split(","))

    from graphkokkos_amd import Context, HostGraph, SEL_LEARNED

    ctx = Context(0)
    result: dict = {}
    t0 = time.perf_counter()
    G20 = HostGraph.rmat(a.ktruss_scale, 16, seed=a.seed, **RMAT)
    E20 = G20.edges
    log(f"R-MAT s{a.ktruss_scale}: n={G20.n} |E|={E20} built in {time.perf_counter() - t0:.1f}s")

    kt = None
    if "ktruss" in legs:
        kt = run_ktruss(ctx, G20.A, a.steps, a.warmup)
        d = dominant(kt["timing"])
        kt["dominant"] = d
        log(f"ktruss: {kt['wall_s'] * 1e3:.1f} ms/decomposition, kmax={kt['kmax']}, rounds={kt['rounds']}, "
            f"dominant {d['kernel']} {d['achieved_gbs']:.0f} GB/s")

    tc = None
    if "tc" in legs:
        t0 = time.perf_counter()
        G22 = HostGraph.rmat(a.tc_scale, 16, seed=a.seed, **RMAT)
        host_build_s = time.perf_counter() - t0
        log(f"R-MAT s{a.tc_scale}: n={G22.n} |E|={G22.edges} built in {host_build_s:.1f}s")
        dev_build_ms = run_graph_build(ctx, a.tc_scale, a.seed)
        log(f"R-MAT s{a.tc_scale} built on the device in {dev_build_ms:.1f} ms")
        plain = run_tc(ctx, G22.L, a.steps, a.warmup, PLAIN)
        pd = dominant(plain["timing"])
        tc = {"workload": f"triangle counting C<L>=L*L, R-MAT scale {a.tc_scale} ef16", "edges": G22.edges,
              "triangles": plain["count"],
              "plain": {"value": G22.edges / plain["wall_s"], "ms": plain["wall_s"] * 1e3, "opts": PLAIN,
                        "kernel_ms": plain["timing"]["kernel_ms"], "select_ms": plain["timing"]["select_ms"],
                        "roofline": roofline_obj(pd),
                        "bytes_alg_all": plain["timing"]["bytes_alg"],
                        "traffic_all_kernels": measured_traffic("tc-product", f"tc-rmat{a.tc_scale}"),
                        "hbm_frac_measured": (measured_traffic("tc-product", f"tc-rmat{a.tc_scale}") or 0) / 1e9 /
                        plain["wall_s"] / HBM_PEAK_GBS or None,
                        "frac_all_kernels": (plain["timing"]["bytes_alg"] / 1e9) /
                        (plain["timing"]["total_ms"] / 1e3) / HBM_PEAK_GBS}}
        log(f"tc plain: {plain['wall_s'] * 1e3:.2f} ms, T={plain['count']}, dominant {pd['kernel']} "
            f"{pd['achieved_gbs']:.0f} GB/s")
        tc["plain"]["rows_per_acc"] = plain["timing"]["rows_per_acc"]
        tc["graph_build"] = {"what": f"R-MAT scale {a.tc_scale} ef16 generated, symmetrised, deduplicated, "
                                     "degree-relabelled, CSR + strict lower part",
                             "device_ms": dev_build_ms, "host_s": host_build_s,
                             "host_cpus": os.cpu_count()}
    if tc is not None and a.tc_variants == "default":
        r = run_tc(ctx, G22.L, a.steps, a.warmup, {})
        assert r["count"] == plain["count"], "the default options changed the triangle count"
        d = dominant(r["timing"])
        tc["heuristic_hub"] = {"value": G22.edges / r["wall_s"], "ms": r["wall_s"] * 1e3, "opts": {},
                               "speedup_vs_plain": plain["wall_s"] / r["wall_s"],
                               "roofline": roofline_obj(d, f"tc-rmat{a.tc_scale}-hub"),
                               "traffic_all_kernels": measured_traffic("tc-product-hub", f"tc-rmat{a.tc_scale}"),
                               "frac_all_kernels": (r["timing"]["bytes_alg"] / 1e9) / r["wall_s"] / HBM_PEAK_GBS,
                               "rows_per_acc": r["timing"]["rows_per_acc"]}
        log(f"tc heuristic_hub: {r['wall_s'] * 1e3:.2f} ms ({plain['wall_s'] / r['wall_s']:.3f}x plain)")
    if tc is not None and a.tc_variants == "opts":
        opts = json.loads(a.tc_opts)
        r = run_tc(ctx, G22.L, a.steps, a.warmup, opts)
        assert r["count"] == plain["count"], "the given options changed the triangle count"
        d = dominant(r["timing"])
        tc["given_opts"] = {"value": G22.edges / r["wall_s"], "ms": r["wall_s"] * 1e3, "opts": opts,
                            "speedup_vs_plain": plain["wall_s"] / r["wall_s"], "roofline": roofline_obj(d),
                            "frac_all_kernels": (r["timing"]["bytes_alg"] / 1e9) / r["wall_s"] / HBM_PEAK_GBS,
                            "rows_per_acc": r["timing"]["rows_per_acc"]}
        log(f"tc given opts {opts}: {r['wall_s'] * 1e3:.2f} ms ({plain['wall_s'] / r['wall_s']:.3f}x plain)")
    if tc is not None and a.tc_variants == "all":
        # learned selector: trained on timing samples the product generates
        # on other seeded graphs; then the global knobs autotuned on the
        # workload itself, as an autotuner serves it; both outside the timed
        # region
        from graphkokkos_amd import selector as SEL
        t0 = time.perf_counter()
        train_graphs = [HostGraph.rmat(18, 16, seed=a.seed + 1, **RMAT), HostGraph.grid(512),
                        HostGraph.chung_lu(1 << 18, 2.0, 16.0, seed=a.seed + 1)]
        rep = SEL.train_from_graphs(ctx, train_graphs, epochs=30, seed=a.seed)  # under the library defaults
        del train_graphs
        log(f"selector trained on {rep['samples']} rows, loss {rep['final_loss']:.3f}, "
            f"{time.perf_counter() - t0:.1f}s")
        dtune = ctx.upload(G22.L)
        tuned_h, _, tlog_h = SEL.autotune_tc(ctx, dtune, rounds=2, reps=3)
        tuned_l, _, tlog_l = SEL.autotune_tc(ctx, dtune, base={"selector": SEL_LEARNED}, rounds=2, reps=3,
                                             space=SEL.LEARNED_SPACE)
        dtune.free()
        log(f"autotuned: heuristic {tuned_h}, learned {tuned_l} ({len(tlog_h) + len(tlog_l)} trials)")
        res = {}
        for name, opts in tc_variants(tuned_h, tuned_l)[1:]:
            r = run_tc(ctx, G22.L, a.steps, a.warmup, opts)
            assert r["count"] == plain["count"], f"{name} changed the triangle count"
            d = dominant(r["timing"])
            res[name] = {"value": G22.edges / r["wall_s"], "ms": r["wall_s"] * 1e3, "opts": opts,
                         "speedup_vs_plain": plain["wall_s"] / r["wall_s"], "roofline": roofline_obj(d),
                         "frac_all_kernels": (r["timing"]["bytes_alg"] / 1e9) / (r["wall_s"]) / HBM_PEAK_GBS,
                         "rows_per_acc": r["timing"]["rows_per_acc"]}
            log(f"tc {name}: {r['wall_s'] * 1e3:.2f} ms ({plain['wall_s'] / r['wall_s']:.3f}x plain)")
        # the options whose HBM traffic profiles/tc_rmat22_tuned_counters.md
        # measured (what the autotuner settles on for this graph): measured
        # bytes over this run's time, beside the algorithmic fraction
        r = run_tc(ctx, G22.L, a.steps, a.warmup, PROFILED_TC_OPTS)
        assert r["count"] == plain["count"], "the profiled options changed the triangle count"
        tb = measured_traffic("tc-product-tuned", f"tc-rmat{a.tc_scale}")
        res["profiled_tuned"] = {
            "value": G22.edges / r["wall_s"], "ms": r["wall_s"] * 1e3, "opts": PROFILED_TC_OPTS,
            "speedup_vs_plain": plain["wall_s"] / r["wall_s"],
            "frac_all_kernels": (r["timing"]["bytes_alg"] / 1e9) / r["wall_s"] / HBM_PEAK_GBS,
            "traffic_all_kernels": tb,
            "hbm_frac_measured": (tb / 1e9) / r["wall_s"] / HBM_PEAK_GBS if tb else None,
            "rows_per_acc": r["timing"]["rows_per_acc"]}
        log(f"tc profiled_tuned: {r['wall_s'] * 1e3:.2f} ms ({plain['wall_s'] / r['wall_s']:.3f}x plain), "
            f"measured HBM {res['profiled_tuned']['hbm_frac_measured']}")
        # every variant against each fixed-selector baseline: plain (above),
        # the library default (hub bitmaps) and the autotuned heuristic
        for name in res:
            res[name]["speedup_vs_default"] = res["heuristic_hub"]["ms"] / res[name]["ms"]
            res[name]["speedup_vs_tuned_heuristic"] = res["tuned_heuristic"]["ms"] / res[name]["ms"]
        tc.update(res)
        tc["selector_training"] = rep
        if not a.no_learned_graphs:
            tc["other_graphs"] = run_learned_graphs(ctx)
    bc = None
    sources = None
    if "bc" in legs:
        rng = np.random.default_rng(a.seed)
        deg = np.diff(G20.A.Rst)
        cand = np.nonzero(deg > 0)[0]
        nsrc = 256 * a.bc_batches
        sources = np.sort(rng.choice(cand, size=nsrc, replace=False)).astype(np.int32)
        r = run_bc(ctx, G20.A, sources, 1)
        byts = bc_bytes(G20.n, G20.A.nz, r["levels"] // max(a.bc_batches, 1), 256) * a.bc_batches
        gbs = byts / 1e9 / r["wall_s"]
        traffic = measured_traffic("bc", f"bc-rmat{a.ktruss_scale}") if a.bc_batches == 4 else None
        bc = {"workload": f"batched Brandes BC, R-MAT scale {a.ktruss_scale}, {a.bc_batches} x 256 sources, FP64",
              "value": nsrc * E20 / r["wall_s"], "unit": "source-edges/s", "ms": r["wall_s"] * 1e3,
              "levels_total": r["levels"], "bytes_alg": byts, "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS,
              # the formula counts 24 B for every (edge, source) pair; the sweeps touch only the pairs
              # whose frontier bits match, so measured HBM bytes (rocprofv3) are reported beside it
              "traffic": traffic,
              "measured_gbs": traffic / 1e9 / r["wall_s"] if traffic else None,
              "forward_ms": r["timing"]["select_ms"], "backward_ms": r["timing"]["kernel_ms"] - r["timing"]["select_ms"]}
        log(f"bc: {r['wall_s'] * 1e3:.1f} ms for {nsrc} sources, {bc['value']:.3g} source-edges/s")

    batch = None
    if "batch" in legs:
        batch = run_batch(ctx, a)

    cmask = None
    if "cmask" in legs:
        cmask = run_cmask(ctx, a.cmask_scale, a.seed, a.steps, a.no_cpu, a.cpu_reps)

    cpu = None
    if not a.no_cpu:
        from oracle import oracle as O
        th = O.threads()
        reps = a.cpu_reps
        Gs = HostGraph.rmat(15, 16, seed=a.seed, **RMAT)
        # the reference loop: S<C>=C*C (parallel MSA) + prune, per k
        dt, (tr_cpu, kmax_cpu, _) = best_of(lambda: O.truss_decompose(Gs.A), reps)
        tr_gpu, kmax_gpu = ctx.truss_decompose(Gs.A)
        assert kmax_gpu == kmax_cpu and np.array_equal(tr_gpu, tr_cpu), "oracle and GPU truss numbers differ (s15)"
        cpu = {"value": Gs.edges / dt, "unit": "edges/s", "cores": th, "kind": "port",
               "sample": f"full truss decomposition of R-MAT scale 15 ef16 (same generator, |E|={Gs.edges}) "
                         f"by the reference loop (parallel MSA product + prune per k), best of {reps} after a "
                         f"warm-up: {dt:.2f}s; truss numbers equal the GPU's on the same graph; "
                         "the per-edge CPU cost grows with scale, so this overstates the CPU rate at scale 20"}
        log(f"cpu baseline ktruss: {cpu['value']:.3g} edges/s on {th} threads")
        if tc is not None:
            dt, cnt = best_of(lambda: O.tc_sandia(G22.L), reps)
            assert cnt == tc["triangles"], "oracle and GPU triangle counts differ"
            tc["cpu_baseline"] = {"value": G22.edges / dt, "unit": "edges/s", "cores": th, "kind": "port",
                                  "sample": f"the same triangle count (R-MAT scale {a.tc_scale}) by the parallel "
                                            f"MSA restatement, best of {reps} after a warm-up: {dt:.2f}s; "
                                            "count equal to the GPU's"}
            log(f"cpu baseline tc: {tc['cpu_baseline']['value']:.3g} edges/s")
        if batch is not None:
            batch["cpu_baseline"] = batch_cpu_baseline(a, th, *batch["_per_graph"], reps)
            log(f"cpu baseline batch: {batch['cpu_baseline']['value']:.3g} edges/s")
        if bc is not None:
            first = sources[:256]
            dt, bc_cpu = best_of(lambda: O.bc_batch(G20.A, first), reps)
            bc_gpu = ctx.bc_batch_f64(G20.A, first)
            err = np.abs(bc_gpu - bc_cpu) / np.maximum(np.abs(bc_cpu), 1.0)
            assert np.all(err <= 1e-9), f"oracle and GPU BC differ on the first batch: max rel err {err.max():.3e}"
            bc["first_batch_max_rel_err_vs_oracle"] = float(err.max())
            bc["cpu_baseline"] = {"value": 256 * E20 / dt, "unit": "source-edges/s", "cores": th, "kind": "port",
                                  "sample": f"the first batch of 256 sources on the same scale-{a.ktruss_scale} "
                                            f"graph, FP64 restatement, best of {reps} after a warm-up: {dt:.2f}s; "
                                            f"GPU scores for those sources within {err.max():.1e} relative"}
            log(f"cpu baseline bc: {bc['cpu_baseline']['value']:.3g} source-edges/s, GPU max rel err {err.max():.2e}")

    if a.calibrate_copy:
        calibrate_copy(ctx)

    if kt is not None:
        value = E20 / kt["wall_s"]
        result = {
            "metric": METRIC, "value": value, "unit": "edges/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": kt["wall_s"] * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "int32", "data": "synthetic (seeded R-MAT generated on the host)",
            "config": {"workload": f"k-truss full decomposition (k=3..max), R-MAT scale {a.ktruss_scale} ef16, "
                                   "a/b/c=.57/.19/.19, seed 1, degree-relabelled int32 CSR",
                       "vertices": G20.n, "edges": E20, "kmax": kt["kmax"], "rounds": kt["rounds"]},
            "roofline": roofline_obj(kt["dominant"], f"ktruss-rmat{a.ktruss_scale}"),
            "cpu_baseline": cpu,
            "ktruss": {"dev_ms": kt["dev_ms"], "kernel_ms": kt["timing"]["kernel_ms"],
                       "select_ms": kt["timing"]["select_ms"], "bytes_alg": kt["timing"]["bytes_alg"],
                       "triangles_counted_over_rounds": kt["timing"]["hits"]},
        }
    else:
        result = {"metric": METRIC, "value": None, "unit": "edges/s", "n_gpus": 1}
    if tc is not None:
        result["tc"] = tc
    if bc is not None:
        result["bc"] = bc
    if batch is not None:
        batch.pop("_per_graph", None)
        result["batch"] = batch
    if cmask is not None:
        result["cmask"] = cmask
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
