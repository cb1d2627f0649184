"""Autotuner of the collision engines' launch shapes.

For a plan and a batch size it times the candidate settings of the engine the
plan uses and keeps the fastest:
  * fused forms with one workgroup per (cell, radial node) (FP32 3D 32^3, 3D 16^3 and
    2D 64^2; FP64 2D 32^2): cells per fused launch (3D 32^3 64 .. 512, the others
    256 .. 4096) and the tile shape, i.e. how many workgroups share a (cell, radial
    node)'s angles (1, 2, 4, 8; splits are tried while the launch has fewer than 2048
    workgroups, so that small batches fill the GPU);
  * the FP32 2D 16^2 / 32^2 form: cells per launch and rounds of quadrature pairs per
    workgroup (4 .. 32), which set its workgroups per cell;
  * staged form: the scratch budget (256 MiB .. 4 GiB) and the quadrature pairs per
    chunk (automatic, 16, 64), which set the items per pass launch.
A tile shape changes only the order of a cell's sums, so tuned and default results
agree to the precision's rounding; for a fixed setting a cell's result is the same
in any batch.
Results are cached in memory and in a JSON file (``INELASTIC_TUNE_CACHE`` or
``~/.cache/inelastic_mi355x_tune.json``) keyed by (dim, n, nr, m, precision,
batch bucket), so a process tunes each shape once.
"""
import json
import os
import pathlib

_MEM = {}


def _cache_path():
    p = os.environ.get("INELASTIC_TUNE_CACHE")
    return pathlib.Path(p) if p else pathlib.Path.home() / ".cache" / "inelastic_mi355x_tune.json"


def _load_disk():
    try:
        return json.loads(_cache_path().read_text())
    except Exception:
        return {}


def _store_disk(table):
    try:
        path = _cache_path()
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(table, indent=1, sort_keys=True))
    except OSError:
        pass


def bucket(batch):
    """Power-of-two batch bucket."""
    b = 1
    while b < batch:
        b *= 2
    return b


def key_of(coll, batch):
    r, _ = coll.sm.rquad()
    sigma, _ = coll.sm.circ_or_sphr_quad()
    return "d%d_n%d_nr%d_m%d_%s_b%d" % (coll.ndim, coll.sm.nv, len(r), sigma.shape[0],
                                        coll.precision, bucket(batch))


def _splits(coll, batch):
    """Angle splits worth timing: 1, and more while the launch stays under 2048 workgroups."""
    r, _ = coll.sm.rquad()
    return [s for s in (1, 2, 4, 8) if s == 1 or batch * len(r) * s // 2 < 2048]


def candidates(coll, batch):
    d, n, prec = coll.ndim, coll.sm.nv, coll.precision
    fp32 = prec in ("fp32", "fp32k")
    if fp32 and d == 3 and n == 32:
        return [{"fused_cells": c, "angle_split": s} for c in (64, 128, 256, 512) if c <= max(64, batch)
                for s in _splits(coll, batch)]
    if fp32 and d == 2 and n in (16, 32):
        return [{"fused_cells": c, "pair_rounds": r} for c in (256, 1024, 4096) if c <= max(256, batch)
                for r in (4, 8, 16, 32)]
    if (fp32 and ((d == 2 and n == 64) or (d == 3 and n == 16))) or (prec == "fp64" and d == 2 and n == 32):
        return [{"fused_cells": c, "angle_split": s} for c in (256, 1024, 4096) if c <= max(256, batch)
                for s in _splits(coll, batch)]
    return [{"scratch_budget": b << 20, "chunk_pairs": cp} for b in (256, 1024, 4096) for cp in (0, 16, 64)]


def apply(coll, setting):
    coll.set_tuning(chunk_pairs=setting.get("chunk_pairs", 0), fused_cells=setting.get("fused_cells", 0),
                    scratch_budget=setting.get("scratch_budget", 0))
    coll.set_tiles(angle_split=setting.get("angle_split", 1), pair_rounds=setting.get("pair_rounds", 0))


def tune(coll, f_dev, with_loss=True, reps=2, force=False):
    """Time every candidate on f_dev, apply and cache the fastest; returns it."""
    from . import device as dev
    t = dev.torch()
    d = coll.ndim
    batch = int(f_dev.numel() // (coll.sm.nv ** d))
    k = key_of(coll, batch)
    if not force:
        if k in _MEM:
            apply(coll, _MEM[k])
            return _MEM[k]
        disk = _load_disk()
        if k in disk:
            _MEM[k] = disk[k]["setting"]
            apply(coll, _MEM[k])
            return _MEM[k]
    results = []
    for cand in candidates(coll, batch):
        apply(coll, cand)
        coll.collide(f_dev, with_loss=with_loss)  # warm-up: sizes scratch
        t.cuda.synchronize()
        e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            coll.collide(f_dev, with_loss=with_loss)
        e1.record()
        t.cuda.synchronize()
        results.append((e0.elapsed_time(e1) / reps, cand))
    ms, best = min(results, key=lambda x: x[0])
    apply(coll, best)
    _MEM[k] = best
    disk = _load_disk()
    disk[k] = {"setting": best, "ms": ms, "all": [(m, c) for m, c in results]}
    _store_disk(disk)
    return best
