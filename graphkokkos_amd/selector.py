"""Learned accumulator selector and autotuner (host-side orchestration).

The network, its training and its inference run in libgraphhip.so
(selector.hip): a 16 -> 64 -> 64 -> 4 FP32 MLP over per-row features that
predicts each accumulator's cost of a row in log2 ns of device time.  This module
collects training samples from the product's own task timers, trains, saves
and loads the parameters, and runs the autotuner over the product's global
knobs (bin thresholds, tile widths, persistent grid sizes).
"""
from __future__ import annotations

import ctypes as C
import itertools
import time
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import check, default_opts, lib
from .algorithms import Context, DeviceCsr
from .graph import SpMtx

N_IN, N_HIDDEN, N_OUT = 16, 64, 4


def num_params() -> int:
    return int(lib().gk_model_num_params())


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def samples(ctx: Context, A: DeviceCsr, B: DeviceCsr | None = None, M: DeviceCsr | None = None,
            mode: int = 0, opts=None) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Features X (m x 16), labels T (m x 4, log2 ns) and validity
    (m x 4) of every row of C<M> = A*B; mode 0 = triangle counting.  The
    rows are timed under ``opts`` (a GkOpts or None for defaults) with each
    accumulator forced in turn."""
    B = B or A
    M = M or A
    m = A.m
    X = np.zeros((m, N_IN), np.float32)
    T = np.zeros((m, N_OUT), np.float32)
    V = np.zeros((m, N_OUT), np.uint8)
    check(lib().gk_selector_samples(ctx.handle, A.handle, B.handle, M.handle, mode,
                                    X.ctypes.data_as(C.POINTER(C.c_float)), T.ctypes.data_as(C.POINTER(C.c_float)),
                                    V.ctypes.data_as(C.POINTER(C.c_uint8)),
                                    C.byref(opts) if opts is not None else None), "gk_selector_samples")
    keep = V.any(axis=1)
    return X[keep], T[keep], V[keep]


def train(ctx: Context, X, T, V, epochs: int = 40, batch: int = 4096, lr: float = 2e-3, seed: int = 1) -> float:
    X, T = _f32(X), _f32(T)
    V = np.ascontiguousarray(V, np.uint8)
    loss = C.c_float(0)
    check(lib().gk_model_train(ctx.handle, X.ctypes.data_as(C.POINTER(C.c_float)),
                               T.ctypes.data_as(C.POINTER(C.c_float)), V.ctypes.data_as(C.POINTER(C.c_uint8)),
                               len(X), epochs, batch, lr, seed, C.byref(loss)), "gk_model_train")
    return float(loss.value)


def get_params(ctx: Context) -> np.ndarray:
    p = np.zeros(num_params(), np.float32)
    check(lib().gk_model_get(ctx.handle, p.ctypes.data_as(C.POINTER(C.c_float)), len(p)), "gk_model_get")
    return p


def set_params(ctx: Context, p) -> None:
    p = _f32(p)
    check(lib().gk_model_set(ctx.handle, p.ctypes.data_as(C.POINTER(C.c_float)), len(p)), "gk_model_set")


def forward(ctx: Context, X) -> np.ndarray:
    X = _f32(X)
    Y = np.zeros((len(X), N_OUT), np.float32)
    check(lib().gk_model_forward(ctx.handle, X.ctypes.data_as(C.POINTER(C.c_float)), len(X),
                                 Y.ctypes.data_as(C.POINTER(C.c_float))), "gk_model_forward")
    return Y


def grad(ctx: Context, X, T, V) -> tuple[np.ndarray, float]:
    X, T = _f32(X), _f32(T)
    V = np.ascontiguousarray(V, np.uint8)
    g = np.zeros(num_params() - 2 * N_IN, np.float32)
    loss = C.c_float(0)
    check(lib().gk_model_grad(ctx.handle, X.ctypes.data_as(C.POINTER(C.c_float)),
                              T.ctypes.data_as(C.POINTER(C.c_float)), V.ctypes.data_as(C.POINTER(C.c_uint8)),
                              len(X), g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)), "gk_model_grad")
    return g, float(loss.value)


def save(path, params) -> None:
    np.save(Path(path), _f32(params), allow_pickle=False)


def load(path) -> np.ndarray:
    return np.load(Path(path), allow_pickle=False)


def train_from_graphs(ctx: Context, graphs, epochs: int = 40, seed: int = 1, mode: int = 0, opts=None) -> dict:
    """Sample every graph's L (triangle counting) under ``opts``, train,
    return a report."""
    Xs, Ts, Vs = [], [], []
    t0 = time.perf_counter()
    for g in graphs:
        d = ctx.upload(g.L if mode == 0 else g.A)
        X, T, V = samples(ctx, d, mode=mode, opts=opts)
        Xs.append(X)
        Ts.append(T)
        Vs.append(V)
        d.free()
    X, T, V = np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Vs)
    t1 = time.perf_counter()
    loss = train(ctx, X, T, V, epochs=epochs, seed=seed)
    t2 = time.perf_counter()
    return {"samples": int(len(X)), "sample_s": t1 - t0, "train_s": t2 - t1, "final_loss": loss}


# ---- autotuner ----------------------------------------------------------------

SEARCH_SPACE = {
    "mca_max_mask": [0, 8, 16, 32, 48, 64],
    "small_max_mask": [32, 64, 96, 128],
    "blocks_small": [12, 16, 24, 32, 48],
    "blocks_long": [4, 8, 16, 24],
    "blocks_mca": [4, 8, 16, 32],
    "split_flop": [1 << 16, 1 << 18, 1 << 20],
    "hub_window": [0, 4096, 8192, 16384, 32768],
    "hub_probe": [0, 2, 4, 8],
}


# the learned selector adds the margin by which it must predict a gain
# before it leaves the fixed-threshold choice (milli-log2 units)
LEARNED_SPACE = {**SEARCH_SPACE, "learned_margin": [0, 125, 250, 500, 1000]}


def time_tc(ctx: Context, L: DeviceCsr, opts, reps: int = 3) -> float:
    ctx.tc_count(L, opts)  # warm-up
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.tc_count(L, opts)
    ctx.sync()
    return (time.perf_counter() - t0) / reps


def coordinate_descent(measure, space: dict, base: dict | None = None, rounds: int = 2,
                       progress=None) -> tuple[dict, float, list]:
    """Minimise measure(options dict) one knob at a time over `space`;
    returns (best options, best value, trial log).  progress(options, value)
    is called after every trial when given."""
    best = dict(base or {})
    best_t = measure(best)
    log = [(dict(best), best_t)]
    if progress:
        progress(best, best_t)
    for _ in range(rounds):
        improved = False
        for key, values in space.items():
            for v in values:
                if best.get(key, getattr(default_opts(), key)) == v:
                    continue
                cand = dict(best)
                cand[key] = v
                t = measure(cand)
                log.append((dict(cand), t))
                if progress:
                    progress(cand, t)
                if t < best_t * 0.98:
                    best, best_t, improved = cand, t, True
        if not improved:
            break
    return best, best_t, log


def autotune_tc(ctx: Context, L: DeviceCsr, base: dict | None = None, rounds: int = 2, reps: int = 3,
                space: dict | None = None) -> tuple[dict, float, list]:
    """Coordinate descent over the product's global knobs, timing triangle
    counting on L; returns (best options, best seconds, trial log)."""
    return coordinate_descent(lambda o: time_tc(ctx, L, default_opts(**o), reps), space or SEARCH_SPACE, base,
                              rounds)


BATCH_SPACE = {
    "batch_blocks": [2, 4, 8, 16],
    "batch_order": [0, 1],
    "small_max_mask": [32, 64, 128],
    "mca_max_mask": [0, 8, 16],
    "blocks_small": [12, 24, 32],
    "blocks_mca": [8, 16, 32, 64],
    "split_flop": [1 << 16, 1 << 18],
}


def autotune_batch(ctx: Context, A: DeviceCsr, voff, k: int = 4, base: dict | None = None, rounds: int = 2,
                   reps: int = 2, space: dict | None = None, progress=None) -> tuple[dict, float, list]:
    """Coordinate descent over the batching knobs (peeling workgroups per CU,
    graph order) and the support product's bin thresholds and tile/grid
    sizes, timing the captured batch step; returns (best options, best ms,
    trial log)."""
    def measure(o):
        return ctx.batch_tc_ktruss(A, voff, k=k, use_graph=True, reps=reps, opts=default_opts(**o))["ms"]
    return coordinate_descent(measure, space or BATCH_SPACE, base, rounds, progress)
