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


def field_errors(state, reference, keys) -> dict:
    """Per-field (max |a-b|, max |b|, ratio) for every key."""
    return {key: maxabs_error(field_view(state, key), field_view(reference, key)) for key in keys}


def column_budgets(state, dgz) -> dict:
    """Column integrals over the interior levels k = 1..94 of dry+moist mass (rhog),
    total water (sum of the 6 tracers of rhogq) and internal energy (rhoge), each
    weighted by the layer thickness dgz; shape [l][ij] per budget.  The surface
    precipitation of the step is the water that left the column."""
    dz = np.asarray(dgz, dtype=np.float64)[1:95]
    rhog = np.asarray(state["rhog"], dtype=np.float64)[:, 1:95, :]
    rhogq = np.asarray(state["rhogq"], dtype=np.float64)[:, :, 1:95, :]
    rhoge = np.asarray(state["rhoge"], dtype=np.float64)[:, 1:95, :]
    return {
        "mass": np.einsum("lki,k->li", rhog, dz),
        "water": np.einsum("lnki,k->li", rhogq, dz),
        "energy": np.einsum("lki,k->li", rhoge, dz),
    }


def budget_drift(state, reference, dgz) -> dict:
    """max over columns of |budget(state) - budget(reference)| / max |budget(reference)|."""
    a = column_budgets(state, dgz)
    b = column_budgets(reference, dgz)
    out = {}
    for k in a:
        scale = float(np.max(np.abs(b[k])))
        out[k] = float(np.max(np.abs(a[k] - b[k]))) / scale if scale > 0 else 0.0
    return out
