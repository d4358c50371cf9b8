"""The oracle spread over worker processes, for the GPU tests at full line
density whose oracle samples are tens of core-seconds.

Workers are spawned (the parent has initialised the GPU, so it is never
forked), keep one BLAS thread each and rebuild the grid from (nu_min,
nu_max, step) themselves.  A task is one (tile, layer range) of
``oracle.lbl_oracle.optical_depth`` or one layer range of
``oracle.adjoint_oracle.vjp_layers``.  Each task is sent only the lines that
the oracle's own band mask keeps for the task's grid points (nu0 within the
cutoff plus |shift| x the largest layer pressure of the points' edges,
forward.py:210-220): the per-tile selection (forward.py:298-301) then takes the same lines in the
same order as from the full list.
"""

from __future__ import annotations

import multiprocessing as mp
import os
from concurrent.futures import ProcessPoolExecutor
from typing import Dict, Sequence

import numpy as np


def n_workers() -> int:
    """One worker per core of the host share, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _init() -> None:
    from threadpoolctl import threadpool_limits
    threadpool_limits(1)


def executor(n: int | None = None) -> ProcessPoolExecutor:
    return ProcessPoolExecutor(max_workers=n or n_workers(), mp_context=mp.get_context("spawn"),
                               initializer=_init)


def lines_near(lines: Dict[str, np.ndarray], lo: float, hi: float, cutoff: float,
               p_max: float) -> Dict[str, np.ndarray]:
    """The lines the oracle's band mask keeps for grid points in [lo, hi]."""
    nu0 = np.asarray(lines["nu0"])
    b = np.abs(np.asarray(lines["shift"])) * p_max
    m = (nu0 >= lo - cutoff - b) & (nu0 <= hi + cutoff + b)
    return {k: np.asarray(v)[m] for k, v in lines.items()}


def _layer_chunks(n_layers: int, n_chunks: int):
    edges = np.linspace(0, n_layers, min(n_chunks, n_layers) + 1).round().astype(int)
    return [list(range(a, b)) for a, b in zip(edges[:-1], edges[1:]) if b > a]


def _sub(a, ks):
    return np.asarray(a)[ks]


def _od_task(task):
    from oracle import lbl_oracle as O
    (lo, hi, step), tile, ks, lines, tips, mm, p, t, vmr, col, cutoff, tw = task
    nu = O.grid(lo, hi, step)
    ts = O.tile_samples(nu, tw)
    a, b = tile * ts, min(tile * ts + ts, nu.size)
    od = O.optical_depth(nu, lines, tips, mm, _sub(p, ks), _sub(t, ks),
                         {g: _sub(v, ks) for g, v in vmr.items()},
                         {g: _sub(v, ks) for g, v in col.items()}, cutoff, tw, tiles=[tile])
    return tile, ks, od[:, a:b]


def optical_depth_tiles(pool, grid, tiles: Sequence[int], lines, tips, mm, p, t, vmr, col,
                        cutoff: float, tile_width: float, chunks_per_tile: int | None = None):
    """{tile: od[n_layers, tile points]} of the oracle on the grid
    (nu_min, nu_max, step) for the given tiles and every layer."""
    from oracle import lbl_oracle as O
    lo, hi, step = grid
    nu = O.grid(lo, hi, step)
    ts = O.tile_samples(nu, tile_width)
    n_layers = len(p)
    p_max = float(np.max(p))
    per = chunks_per_tile or max(1, -(-2 * n_workers() // len(tiles)))
    tasks = []
    for tl in tiles:
        a, b = tl * ts, min(tl * ts + ts, nu.size)
        near = lines_near(lines, float(nu[a]), float(nu[b - 1]), cutoff, p_max)
        for ks in _layer_chunks(n_layers, per):
            tasks.append(((lo, hi, step), tl, ks, near, tips, mm, p, t, vmr, col, cutoff, tile_width))
    out = {tl: np.zeros((n_layers, min(tl * ts + ts, nu.size) - tl * ts)) for tl in tiles}
    for tl, ks, od in pool.map(_od_task, tasks):
        out[tl][ks] = od
    return out


def _vjp_task(task):
    from oracle import adjoint_oracle as A
    from oracle import lbl_oracle as O
    (lo, hi, step), ks, lines, tips, mm, p, t, vmr, air, cutoff, tw, W = task
    nu = O.grid(lo, hi, step)
    return ks, A.vjp_layers(nu, lines, tips, mm, p, t, vmr, air, cutoff, tw, W, ks)


def adjoint_vjp(pool, grid, lines, tips, mm, p, t, vmr, air, cutoff: float, tile_width: float,
                mus, G):
    """oracle.adjoint_oracle.vjp on the grid (nu_min, nu_max, step), the
    forward's layers and the gradient's layers split over the pool."""
    from oracle import adjoint_oracle as A
    from oracle import lbl_oracle as O
    lo, hi, step = grid
    nu = O.grid(lo, hi, step)
    near = lines_near(lines, float(nu[0]), float(nu[-1]), cutoff, float(np.max(p)))
    col = {g: np.asarray(v) * air for g, v in vmr.items()}
    ts = O.tile_samples(nu, tile_width)
    tiles = list(range(-(-nu.size // ts)))
    od = optical_depth_tiles(pool, grid, tiles, near, tips, mm, p, t, vmr, col, cutoff,
                             tile_width)
    tau = np.concatenate([od[tl] for tl in tiles], axis=1).sum(0)
    Tm = np.stack([np.exp(-tau / m) for m in mus])
    W = A.weight(Tm, G, mus)
    n_layers = len(p)
    gT = np.zeros(n_layers)
    gV = {g: np.zeros(n_layers) for g in vmr}
    tasks = [((lo, hi, step), ks, near, tips, mm, p, t, vmr, air, cutoff, tile_width, W)
             for ks in _layer_chunks(n_layers, 2 * n_workers())]
    for ks, (gt, gv) in pool.map(_vjp_task, tasks):
        gT[ks] = gt[ks]
        for g in gv:
            gV[g][ks] = gv[g][ks]
    return gT, gV, Tm
