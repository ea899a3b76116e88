"""log Z comparisons over independent NS runs.

The walk's move schedule (one move index per sub-step) is shared by every
chain and replica of an engine (the reference's ``_move_schedule``,
sampling/nested_sampling.py:136-148), so the replicas of one engine are not
independent runs: they share how many sub-steps of each move type they get.
The unit of independence is the engine (its own seed, hence its own
schedule).  A comparison therefore takes each engine's mean log Z over its
replicas as one sample and applies Welch's t to the engine means."""

from __future__ import annotations

import json

import numpy as np


def welch(means_a, means_b):
    """(difference of means b - a, its standard error, t)."""
    a, b = np.asarray(means_a, float), np.asarray(means_b, float)
    d = b.mean() - a.mean()
    se = float(np.hypot(a.std(ddof=1) / np.sqrt(len(a)), b.std(ddof=1) / np.sqrt(len(b))))
    return float(d), se, float(d / se)


def report(name, means_a, means_b, extra=None, labels=("plain", "surrogate")):
    """Print one ``LOGZ_RESULT`` JSON line and return t."""
    d, se, t = welch(means_a, means_b)
    rec = {"config": name, "delta_log_z": d, "sigma": se, "z": t,
           f"log_z_{labels[0]}": float(np.mean(means_a)),
           f"log_z_{labels[1]}": float(np.mean(means_b)),
           "engines": [len(means_a), len(means_b)],
           f"engine_means_{labels[0]}": [round(float(x), 4) for x in means_a],
           f"engine_means_{labels[1]}": [round(float(x), 4) for x in means_b],
           **(extra or {})}
    print("LOGZ_RESULT " + json.dumps(rec), flush=True)
    return t
