"""Parser of the canonical tree blob written by frep_tree_export.

Blob layout (u32 words): 'FRPT', version, n_levels, dims[n], counts[n], then
per level, per node in (x, y, z) order: x, y, z, k, k instruction words, the
leaf mask (dim^3/32 words) and, for non-last levels, the interior mask
(dim^3/32 words) and dim^3 child entries (canonical index in the next level,
0xFFFFFFFF where there is none).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MAGIC = 0x54505246

# Flop weights per executed tape op: point evaluation / interval evaluation.
# CONST and COPY move data and count zero; MUL is 4 products + 6 min/max
# comparisons, DIV adds a reciprocal pair, MIN/MAX add the two choice
# comparisons, SIN/COS count the range checks and two evaluations.
POINT_W = np.array([0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0], dtype=np.float64)
INTERVAL_W = np.array([0, 12, 12, 2, 2, 10, 12, 4, 4, 2, 2, 4, 0], dtype=np.float64)


@dataclass
class Level:
    dim: int
    coords: np.ndarray      # (n, 3)
    tapes: list             # list of np.ndarray [k, instrs]
    leaf: np.ndarray        # (n, W)
    interior: np.ndarray | None
    child: np.ndarray | None


def parse(blob: np.ndarray) -> list[Level]:
    b = np.asarray(blob, dtype=np.uint32)
    if int(b[0]) != MAGIC or int(b[1]) != 1:
        raise ValueError("not a canonical tree blob")
    nl = int(b[2])
    dims = [int(v) for v in b[3:3 + nl]]
    counts = [int(v) for v in b[3 + nl:3 + 2 * nl]]
    pos = 3 + 2 * nl
    levels = []
    for L in range(nl):
        dim = dims[L]
        cells = dim ** 3
        W = cells // 32
        last = L == nl - 1
        n = counts[L]
        coords = np.zeros((n, 3), dtype=np.uint32)
        leaf = np.zeros((n, W), dtype=np.uint32)
        interior = None if last else np.zeros((n, W), dtype=np.uint32)
        child = None if last else np.zeros((n, cells), dtype=np.uint32)
        tapes = []
        for j in range(n):
            coords[j] = b[pos:pos + 3]
            pos += 3
            k = int(b[pos])
            tapes.append(b[pos:pos + k + 1].copy())
            pos += k + 1
            leaf[j] = b[pos:pos + W]
            pos += W
            if not last:
                interior[j] = b[pos:pos + W]
                pos += W
                child[j] = b[pos:pos + cells]
                pos += cells
        levels.append(Level(dim, coords, tapes, leaf, interior, child))
    if pos != len(b):
        raise ValueError("trailing words in blob")
    return levels


def level_flops(levels: list[Level]) -> list[dict]:
    """Algorithmic work per level: cells evaluated and flops executed."""
    out = []
    for L, lv in enumerate(levels):
        last = L == len(levels) - 1
        w = POINT_W if last else INTERVAL_W
        cells = lv.dim ** 3
        flops = 0.0
        ops = 0
        for t in lv.tapes:
            op = t[1:] & 0xFF
            flops += float(w[op].sum())
            ops += len(op)
        out.append({"level": L, "nodes": len(lv.tapes), "cells": cells * len(lv.tapes),
                    "flops": flops * cells, "op_evals": float(ops) * cells,
                    "kind": "point" if last else "interval"})
    return out


def occupancy(levels: list[Level]) -> np.ndarray:
    """Flattened finest-grid occupancy (x, y, z), for small trees only."""
    dims = [lv.dim for lv in levels]
    F = int(np.prod(dims))
    occ = np.zeros((F, F, F), dtype=np.uint8)

    def rec(L, j):
        lv = levels[L]
        dim = lv.dim
        span = int(np.prod(dims[L + 1:])) if L + 1 < len(dims) else 1
        for idx in range(dim ** 3):
            x, y, z = idx // (dim * dim), (idx // dim) % dim, idx % dim
            gx, gy, gz = (int(lv.coords[j][0]) * dim + x, int(lv.coords[j][1]) * dim + y,
                          int(lv.coords[j][2]) * dim + z)
            if (int(lv.leaf[j][idx >> 5]) >> (idx & 31)) & 1:
                occ[gx * span:(gx + 1) * span, gy * span:(gy + 1) * span,
                    gz * span:(gz + 1) * span] = 1
            elif lv.interior is not None and (int(lv.interior[j][idx >> 5]) >> (idx & 31)) & 1:
                rec(L + 1, int(lv.child[j][idx]))

    rec(0, 0)
    return occ


FILE_MAGIC = 0x46505246   # 'FRPF'
FILE_HEADER = 18


def file_words(grid: dict, consts, slot_count: int, blob) -> np.ndarray:
    """A tree file (frep_tree_save's layout) from a grid, the tape's constant
    pool and slot count, and a canonical blob."""
    consts = np.asarray(consts, dtype=np.float32)
    h = np.zeros(FILE_HEADER, dtype=np.uint32)
    h[0], h[1], h[2] = FILE_MAGIC, 1, len(grid["dims"])
    h[3:3 + len(grid["dims"])] = grid["dims"]
    h[11:14] = np.asarray(grid["low"], dtype=np.float32).view(np.uint32)
    h[14] = np.float32(grid["world_size"]).view(np.uint32)
    h[15] = np.float32(grid.get("tape_time", 0.0)).view(np.uint32)
    h[16], h[17] = slot_count, len(consts)
    return np.concatenate([h, consts.view(np.uint32), np.asarray(blob, dtype=np.uint32)])


def parse_file(words):
    """(grid dict, constants, slot count, canonical blob) of a tree file."""
    w = np.asarray(words, dtype=np.uint32)
    if int(w[0]) != FILE_MAGIC or int(w[1]) != 1:
        raise ValueError("not a tree file")
    nl = int(w[2])
    grid = {"dims": [int(d) for d in w[3:3 + nl]],
            "low": [float(v) for v in w[11:14].view(np.float32)],
            "world_size": float(w[14:15].view(np.float32)[0]),
            "tape_time": float(w[15:16].view(np.float32)[0])}
    nc = int(w[17])
    consts = w[FILE_HEADER:FILE_HEADER + nc].view(np.float32).copy()
    return grid, consts, int(w[16]), w[FILE_HEADER + nc:].copy()
