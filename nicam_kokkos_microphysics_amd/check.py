"""The reference's check table (PROF_val_check, src/mod_debug.cpp:363-410) and the
pointwise tolerance used by the parity tests.

The sums are taken in one fixed order (numpy pairwise over the flattened
[l][k][ij] array) so that two runs on the same data print identical tables.
"""
from __future__ import annotations

import numpy as np

from .fields import CHECK_FIELDS, field_view


def prof_val_check_line(name: str, computed, reference) -> str:
    d = np.asarray(computed, dtype=np.float64) - np.asarray(reference, dtype=np.float64)
    return (f"Checking [{name}] Max = {d.max():.16e}; Min = {d.min():.16e}; Sum = {d.sum():.16e}; ")


def check_table(state, reference, fields=CHECK_FIELDS) -> list:
    out = []
    for key in fields:
        label = key.split(":")[1] if ":" in key else key
        out.append(prof_val_check_line(label, field_view(state, key), field_view(reference, key)))
    return out


def maxabs_error(a, b) -> tuple:
    """(max |a-b|, max |b|, ratio) for one field."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    err = float(np.max(np.abs(a - b))) if a.size else 0.0
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    return err, scale, (err / scale if scale > 0 else err)


def compare_states(state, reference, keys, rel_tol: float) -> dict:
    """Per-field max error relative to the field's max-abs; returns failures only."""
    bad = {}
    for key in keys:
        err, scale, ratio = maxabs_error(field_view(state, key), field_view(reference, key))
        if not (ratio <= rel_tol):
            bad[key] = (err, scale, ratio)
    return bad
