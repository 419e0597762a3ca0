"""Shared loaders for the committed golden fixtures and the problems they were
recorded on (the generator script is tests/golden/make_fixtures.py)."""
from __future__ import annotations

import json
import os

import numpy as np

from cfdr_amd import cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELD_NAMES = ["Ux", "Uy", "p", "k", "omega", "nuT", "F1", "F2", "Pk", "CDkw"]


def load(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


def small_channel(max_iterations=40000):
    """The 1x32 geometric channel (ratio 1.15) every small fixture uses."""
    return cases.channel_problem(ny=32, ratio=1.15, n_probes=30, tol=1e-6,
                                 max_iterations=max_iterations)


def small_hill(pressure_solver):
    return cases.hill_problem(nx=30, ny=16, max_iterations=400, pressure_solver=pressure_solver)


def max_rel(a, b):
    """Largest relative difference, measured against b (0 when both are empty)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def max_scaled(a, b):
    """Largest difference scaled by the field's own magnitude (for fields that
    cross zero, where a pointwise relative error is meaningless)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale
