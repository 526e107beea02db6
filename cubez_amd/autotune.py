"""The resolver's split between its one-wavefront and multi-wavefront
kernels, chosen by measurement (cbz_autotune_resolver) and cached on disk.

The split is a scheduling choice with no effect on results: of the worlds
of at least `min_contacts` contacts, the `max_worlds` largest resolve on a
kernel of four wavefronts per world (one per SIMD, with every contact's
data on chip), the others on the one-wavefront kernel.  Which is fastest
depends on the distribution of contact counts over the worlds (a few large
worlds set the step's critical path; many large worlds make the
multi-wavefront kernel, one world per CU, a throughput limit), so it is
measured on the scene itself, from its current (settled) state.

The cache key is the scene's shape and the step kind: precision, world and
body counts, a coarse histogram of per-world contact counts, and whether
the step is the surrogate's.  The cache file is $CUBEZ_AMD_TUNE_CACHE, or
~/.cache/cubez_amd/resolver_split.json."""
import json
import os

import numpy as np

CANDIDATES = ((0, 0), (257, 8), (257, 16), (257, 32), (257, 64), (257, 128))


def cache_path():
    p = os.environ.get("CUBEZ_AMD_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "cubez_amd", "resolver_split.json")


def scene_key(world, precision, surrogate):
    """Precision, step kind, worlds, bodies, and the per-world contact-count
    histogram (buckets of 256 contacts) of the current state's last
    narrowphase."""
    counts = world.contact_counts()
    hist = np.bincount(np.minimum(counts // 256, 8), minlength=9)
    return "|".join([precision, "sur" if surrogate else "exact", str(world.n_worlds), str(world.n_bodies),
                     ",".join(str(int(h)) for h in hist)])


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune(world, dt, iter_per_contact=4, precision="f64", surrogate=False, steps=5, candidates=CANDIDATES,
         tol=0.06, polish=4, use_cache=True):
    """Set the fastest split for this scene on `world` (from the cache when
    the scene's key is there, else by measurement from the current state,
    which is restored) and return {"split", "ms_per_step", "cached", "key"}."""
    world.narrowphase()  # contact counts of the current state for the key
    key = scene_key(world, precision, surrogate)
    path = cache_path()
    cache = _load(path) if use_cache else {}
    if key in cache:
        split = tuple(cache[key]["split"])
        world.set_resolver_split(*split)
        return {"split": split, "ms_per_step": cache[key]["ms_per_step"], "cached": True, "key": key}
    res = world.autotune_resolver(dt, iter_per_contact, steps, candidates, surrogate=surrogate, tol=tol,
                                  polish=polish)
    if use_cache:
        cache = _load(path)
        cache[key] = {"split": list(res["split"]), "ms_per_step": res["ms_per_step"]}
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(cache, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        except OSError:
            pass  # an unwritable cache only costs a re-measurement
    return {"split": res["split"], "ms_per_step": res["ms_per_step"], "cached": False, "key": key}
