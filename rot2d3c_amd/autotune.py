"""Per-size schedule autotuner with a JSON cache.

A schedule is (rows, blocks per CU, column radix form) for one grid size.
The first two are the launch shape of the row kernel: (2, 0) is the
non-persistent kernel (one block per 2 rows); (rows, blocks per CU) with
blocks per CU 1 or 2 the persistent kernel with 2 rows per block, or 1 when
n >= 512.  The third is the radix schedule of the column transforms from
n = 512: 0 the four-step form (512-point wave-local sub-transforms and one
radix n/512 exchange), 1 radix-8 Stockham passes.  Launch shapes of one
compiled kernel give bit-identical results; different kernels or column
forms agree within rounding.  The tuner only looks for the fastest, and its
cached choice keeps a machine's runs reproducible.  r2d3c_autotune times each candidate on the device; the choice is
cached per (GPU architecture, CU count, n) in a JSON file, by default
~/.cache/rot2d3c_amd/schedules.json (R2D3C_SCHEDULE_CACHE overrides it).
"""
import ctypes
import json
import os
import threading

import torch

from . import _lib
from ._lib import check

DEFAULT = (2, 0, 0)
# every schedule r2d3c_autotune may try, in its order (invalid ones for a
# size are rejected by the library)
CANDIDATES = tuple((rw, bpc, cs) for cs in (0, 1)
                   for rw, bpc in ((2, 0), (2, 1), (2, 2), (1, 1), (1, 2)))
_lock = threading.Lock()


def cache_path():
    return os.environ.get("R2D3C_SCHEDULE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "rot2d3c_amd", "schedules.json")


def device_key(device):
    pr = torch.cuda.get_device_properties(torch.device(device))
    return "%s:%d" % (getattr(pr, "gcnArchName", pr.name), pr.multi_processor_count)


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(table):
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp%d" % os.getpid()
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only costs a re-tune next time


def tune(plan_handle, reps=20, device=None):
    """Time every schedule on the plan (scratch state, best of three trials
    of `reps` steps each),
    keep the fastest in the plan and return ((rows, blocks_per_cu, column_form),
    ms/step)."""
    ms = ctypes.c_double()
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    check(_lib.lib().r2d3c_autotune(plan_handle, int(reps), ctypes.byref(ms), stream),
          "r2d3c_autotune")
    rw, bpc = ctypes.c_int(), ctypes.c_int()
    check(_lib.lib().r2d3c_plan_get_schedule(plan_handle, ctypes.byref(rw), ctypes.byref(bpc)),
          "r2d3c_plan_get_schedule")
    cs = ctypes.c_int()
    check(_lib.lib().r2d3c_plan_get_column_radix(plan_handle, ctypes.byref(cs)),
          "r2d3c_plan_get_column_radix")
    return (rw.value, bpc.value, cs.value), ms.value


def schedule(n, device, reps=20, retune=False):
    """The cached schedule of grid size n on this device, tuned on first use."""
    from .spectral import Plan
    device = torch.device(device)
    key = "%s:%d" % (device_key(device), int(n))
    with _lock:
        table = _load()
        # an entry without the column form predates its tuning: tune again
        if not retune and key in table and len(table[key]["schedule"]) == 3:
            return tuple(table[key]["schedule"])
        with torch.cuda.device(device):
            sched, ms = tune(Plan.get(n, device).handle, reps, device)
        table = _load()
        table[key] = {"schedule": list(sched), "ms_per_step": ms}
        _store(table)
        return sched


def apply(plan_handle, n, device, reps=20):
    """Set the cached (or freshly tuned) schedule of size n on a plan."""
    rw, bpc, cs = schedule(n, device, reps)
    check(_lib.lib().r2d3c_plan_set_schedule(plan_handle, rw, bpc), "r2d3c_plan_set_schedule")
    check(_lib.lib().r2d3c_plan_set_column_radix(plan_handle, cs),
          "r2d3c_plan_set_column_radix")
    return rw, bpc, cs
