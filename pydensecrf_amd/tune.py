"""Autotuning of a model's scheduling knobs, with a JSON cache.

Knobs of a built model (timed on the model itself: a few inference iterations, HIP
events on the current stream; coordinate descent, one knob at a time with the others
at their current best):
  splat_chunk   contributors summed by one lane group before partial rows are
                recombined (dcrf_set_splat_chunk).  Small chunks balance the heavy
                vertices of bilateral lattices; large ones avoid recombination work.
  blur_pairs    two lattice axes per blur pass (the default) or one (dcrf_set_blur_pairs).
  blur_steps    most steps per multi-step pass, 2 or 3 (dcrf_set_blur_steps).
  meanfield     the mean-field kernel: 0 automatic (fused where available), 3 fused over
                LDS-staged pixel runs (dcrf_set_meanfield_kernel).
  q_blocks      Q between inference iterations in 8 x 8 pixel blocks (dcrf_set_q_blocks).
  blur_lanes    (labels per lane, below this many vertices) of the blur passes
                (dcrf_set_blur_lanes; (0, 0) = the default lane split).
Knobs of the build (they must be set before the first pairwise term; timed through a
caller-supplied builder):
  splat_tile    pixel tile side of the splat's segment order (dcrf_set_splat_tile).
  hash_load     entries / slots of the lattice-build hash table (dcrf_set_hash_load).

Every knob leaves the results as they are, up to the summation order of the splat's
partial rows (splat_chunk, splat_tile); the others are bit-identical.  A knob whose
values do not apply to a model (a lane split that does not divide the label count)
is skipped for it.

Winners of a model's knobs are cached under a key of the problem shape (labels, feature
dimensions, log2 of the variable count); build knobs under (labels, log2 of the
variable count), so that a new model can take them before its terms exist.  The cache
is $PYDENSECRF_AMD_TUNE_CACHE or ~/.cache/pydensecrf_amd/tune.json.  Models apply the
cached winners of their shape by themselves (capi.Model before its first inference,
the drop-in classes at creation and before their first inference); set
PYDENSECRF_AMD_TUNE=0 to switch that off.
"""
import json
import math
import os
import time

CHUNKS = (16, 32, 64, 128, 256)
LOADS = (0.5, 0.35, 0.25)
TILES = (128, 160, 192, 64, 256)
MODEL_KNOBS = {
    "splat_chunk": CHUNKS,
    "blur_pairs": (True, False),
    "blur_steps": (2, 3),
    "meanfield": (0, 3),
    "q_blocks": (False, True),
    "blur_lanes": ((0, 0), (1, 50000), (1, 0)),
}
DEFAULTS = {"splat_chunk": 32, "blur_pairs": True, "blur_steps": 2, "meanfield": 0, "q_blocks": False,
            "blur_lanes": (0, 0)}


def enabled():
    return os.environ.get("PYDENSECRF_AMD_TUNE", "1") != "0"


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


def _log2n(N):
    return int(round(math.log2(max(N, 1))))


def shape_key(model):
    dims = [model.term_info(k)[0] for k in range(model.nterms)]
    return "L%d_d%s_n%d" % (model.L, "-".join(map(str, dims)), _log2n(model.N))


def build_key(L, N):
    return "build_L%d_n%d" % (L, _log2n(N))


def _set(model, knob, value):
    if knob == "splat_chunk":
        model.set_splat_chunk(int(value))
    elif knob == "blur_pairs":
        model.set_blur_pairs(bool(value))
    elif knob == "blur_steps":
        model.set_blur_steps(int(value))
    elif knob == "meanfield":
        model.set_meanfield_kernel(int(value))
    elif knob == "q_blocks":
        model.set_q_blocks(bool(value))
    elif knob == "blur_lanes":
        model.set_blur_lanes(int(value[0]), int(value[1]))
    else:
        raise ValueError("unknown knob %r" % knob)


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


def tune_model(model, iters=3, reps=3, knobs=tuple(MODEL_KNOBS), use_cache=True):
    """Pick the fastest value of each knob (coordinate descent in the given order), apply
    the winners and cache them.  Returns ({knob: value}, {knob: {value: ms}}); with a
    cache hit the timings are None."""
    key = shape_key(model)
    db = _load() if use_cache else {}
    hit = db.get(key, {})
    if use_cache and all(k in hit for k in knobs):
        for k in knobs:
            _set(model, k, hit[k])
        return {k: hit[k] for k in knobs}, None
    best, times = {}, {}
    for k in knobs:
        times[k] = {}
        for v in MODEL_KNOBS[k]:
            try:
                _set(model, k, v)
            except ValueError:
                continue
            times[k][v] = _time_inference(model, iters, reps)
        best[k] = min(times[k], key=times[k].get)
        _set(model, k, best[k])
        if isinstance(best[k], tuple):
            best[k] = list(best[k])  # as JSON keeps it
    if use_cache:
        db.setdefault(key, {}).update(best)
        _store(db)
    return best, times


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


def _tune_build(build, knob, candidates, iters, reps, use_cache, time_inference):
    times, key = {}, None
    for v in candidates:
        t0 = time.perf_counter()
        m = build(v)
        t = time.perf_counter() - t0
        times[v] = _time_inference(m, iters, reps) if time_inference else t
        key = build_key(m.L, m.N)
        m.close()
    best = min(times, key=times.get)
    if use_cache and key:
        db = _load()
        db.setdefault(key, {})[knob] = best
        _store(db)
    return best, times


def tune_splat_tile(build, candidates=TILES, iters=3, reps=3, use_cache=True):
    """build(side) -> model builds a model with that splat tile side (applied before its
    terms are added).  Returns (best side, {side: ms per inference})."""
    return _tune_build(build, "splat_tile", candidates, iters, reps, use_cache, True)


def tune_hash_load(build, candidates=LOADS, use_cache=True):
    """build(load) -> model builds a model with that hash load factor (applied before its
    terms are added).  Returns (best load, {load: build seconds})."""
    return _tune_build(build, "hash_load", candidates, 0, 0, use_cache, False)


def apply_cached(model):
    """Apply the cached model knobs of this model's shape, if any.  Returns the entry."""
    ent = _load().get(shape_key(model), {})
    for k in MODEL_KNOBS:
        if k in ent:
            _set(model, k, ent[k])
    return ent


def cached_build_knobs(L, N):
    """{splat_tile, hash_load} cached for a model of L labels and N variables."""
    return _load().get(build_key(L, N), {})


def apply_build_knobs(model):
    """Apply the cached build knobs to a model that has no pairwise term yet."""
    ent = cached_build_knobs(model.L, model.N)
    if "splat_tile" in ent:
        model.set_splat_tile(int(ent["splat_tile"]))
    if "hash_load" in ent:
        model.set_hash_load(float(ent["hash_load"]))
    return ent


def cached_hash_load(L, dims, N):
    return cached_build_knobs(L, N).get("hash_load")
