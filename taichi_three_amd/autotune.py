"""Autotuner for the batched renderer's bin sizes and launch shapes.

Each candidate Tuning (faces per binning workgroup, the giant-face tile
threshold, persistent grid sizes) is built into a render plan for the same
uploaded scenes, warmed up, and timed over several graph replays with device
events on the stream the graph runs on; the fastest wins.  Results are kept
per batch signature (resolutions and face counts, bucketed) for the process
and can be written to / read from a JSON file.
"""
import json
import math

from .batch import BatchRenderer, BatchScenes, Tuning

__all__ = ['candidates', 'tune', 'signature']

_CACHE = {}
LDS_BUDGET = 160 * 1024 - 1024


def candidates(max_tiles_screen=8160, cus=256):
    """The search space: bin sizes x giant threshold x persistent grids (grids in
    multiples of the CU count, 0 = one workgroup per work item)."""
    out = []
    for bf in (1024, 2048, 4096, 8192):
        # a binning workgroup keeps a counter per tile of the largest screen and
        # a span per face in LDS (the library's budget: 159 KB of the CU's 160)
        if (max_tiles_screen + bf) * 4 > LDS_BUDGET:
            continue
        for mt in (16, 64):
            for grids in ((0, 0, 0), (2 * cus, 4 * cus, 4 * cus), (cus, 8 * cus, 8 * cus)):
                out.append(Tuning(bf, mt, *grids))
    return out


def signature(scenes):
    """Batch signature: the sorted (resolution, log2 face bucket) pairs."""
    return tuple(sorted((W, H, int(math.log2(max(1, T)))) for (W, H), T in zip(scenes.res, scenes.nfaces)))


def time_plan(plan, reps=5, warmup=2):
    import torch
    plan.capture()
    for _ in range(warmup):
        plan.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        plan.replay()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(scenes, space=None, reps=5, cache_file=None, log=None):
    """Returns (best Tuning, [(Tuning, ms)]) for `scenes` (BatchScenes or workloads)."""
    scenes = scenes if isinstance(scenes, BatchScenes) else BatchScenes(scenes)
    sig = signature(scenes)
    if cache_file:
        try:
            with open(cache_file) as f:
                for k, v in json.load(f).items():
                    _CACHE.setdefault(k, (Tuning(*v['best']), [(Tuning(*t), ms) for t, ms in v['results']]))
        except (OSError, ValueError, KeyError):
            pass
    key = repr(sig)
    if key in _CACHE:
        return _CACHE[key]
    max_tiles = max(((W + 15) // 16) * ((H + 15) // 16) for W, H in scenes.res)
    results = []
    for t in (space or candidates(max_tiles)):
        plan = BatchRenderer(scenes, t)
        ms = time_plan(plan, reps)
        plan.close()
        results.append((t, ms))
        if log:
            log(f'{t.as_dict()}: {ms:.3f} ms')
    best = min(results, key=lambda r: r[1])[0]
    _CACHE[key] = (best, results)
    if cache_file:
        data = {k: {'best': list(b), 'results': [[list(t), ms] for t, ms in r]} for k, (b, r) in _CACHE.items()}
        with open(cache_file, 'w') as f:
            json.dump(data, f)
    return best, results
