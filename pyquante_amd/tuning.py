"""Stored autotuner results: per-class bra-run caps for named workloads.

`Engine.autotune(D)` times each class's J/K kernel for the candidate run
lengths and returns a 21 x 2 table (class, FP64/FP32 tiles).  A table stored
here under a workload name can be replayed with `Engine.set_group_caps`, so
that repeated runs of that workload launch the same kernels without re-tuning.
The tables live in tuning.json next to this file.
"""
import json
import os

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def load(name):
    """The stored 21 x 2 cap table for workload `name`, or None."""
    if not os.path.exists(_PATH):
        return None
    with open(_PATH) as f:
        table = json.load(f).get(name)
    if table is None:
        return None
    caps = table["caps"]
    if len(caps) != 21 or any(len(c) != 2 for c in caps):
        raise ValueError("malformed tuning entry %r" % name)
    return caps


def store(name, caps, **meta):
    """Store a cap table (as returned by Engine.autotune) under `name`."""
    data = {}
    if os.path.exists(_PATH):
        with open(_PATH) as f:
            data = json.load(f)
    data[name] = {"caps": [[int(a), int(b)] for a, b in caps], **meta}
    with open(_PATH, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
