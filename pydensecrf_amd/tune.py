"""Autotuning of the launch-shape knobs of a model, with a JSON cache.

Knobs
  splat chunk   contributors summed by one lane group before partial rows are
                recombined (dcrf_set_splat_chunk).  Small chunks balance the heavy
                vertices of bilateral lattices; large ones avoid recombination work.
  hash load     entries / slots of the lattice-build hash table (dcrf_set_hash_load):
                a lower load shortens probe sequences at the cost of memory.
The splat chunk is timed on the model itself (a few inference iterations, HIP events
on the current stream); the hash load by rebuilding the model's lattices through a
caller-supplied builder.  Winners are cached under a key of the problem shape (labels,
feature dimensions, log2 of the variable count), in
$PYDENSECRF_AMD_TUNE_CACHE or ~/.cache/pydensecrf_amd/tune.json.
"""
import json
import math
import os

from . import capi

CHUNKS = (16, 32, 64, 128, 256)
LOADS = (0.5, 0.35, 0.25)


def cache_path():
    return os.environ.get("PYDENSECRF_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "pydensecrf_amd", "tune.json"))


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db):
    path = cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def shape_key(model):
    dims = [model.term_info(k)[0] for k in range(model.nterms)]
    return "L%d_d%s_n%d" % (model.L, "-".join(map(str, dims)), int(round(math.log2(max(model.N, 1)))))


def _time_inference(model, iters, reps):
    import torch
    Q = torch.empty(model.N * model.L, device="cuda")
    s = torch.cuda.current_stream()
    model.inference_device(iters, Q.data_ptr(), None, s.cuda_stream)  # warm-up (and layout build)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(s)
    for _ in range(reps):
        model.inference_device(iters, Q.data_ptr(), None, s.cuda_stream)
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_splat_chunk(model, iters=3, reps=3, candidates=CHUNKS, use_cache=True):
    """Pick the fastest splat chunk for this model, apply it and return (chunk, timings)."""
    key = shape_key(model)
    db = _load() if use_cache else {}
    hit = db.get(key, {}).get("splat_chunk")
    if hit:
        model.set_splat_chunk(int(hit))
        return int(hit), None
    times = {}
    for c in candidates:
        model.set_splat_chunk(c)
        times[c] = _time_inference(model, iters, reps)
    best = min(times, key=times.get)
    model.set_splat_chunk(best)
    if use_cache:
        db.setdefault(key, {})["splat_chunk"] = best
        _store(db)
    return best, times


def tune_hash_load(build, candidates=LOADS, use_cache=True):
    """build(load) -> model builds a model with that hash load factor (applied before its
    terms are added).  Returns (best load, {load: build seconds})."""
    import time
    times, key = {}, None
    for load in candidates:
        t0 = time.perf_counter()
        m = build(load)
        times[load] = time.perf_counter() - t0
        key = shape_key(m)
        m.close()
    best = min(times, key=times.get)
    if use_cache and key:
        db = _load()
        db.setdefault(key, {})["hash_load"] = best
        _store(db)
    return best, times


def apply_cached(model):
    """Apply the cached knobs of this model's shape, if any.  Returns the entry."""
    ent = _load().get(shape_key(model), {})
    if "splat_chunk" in ent:
        model.set_splat_chunk(int(ent["splat_chunk"]))
    return ent


def cached_hash_load(L, dims, N):
    key = "L%d_d%s_n%d" % (L, "-".join(map(str, dims)), int(round(math.log2(max(N, 1)))))
    return _load().get(key, {}).get("hash_load")
