"""Diagnostics on downloaded fields (src/measures.jl:6-21)."""
from __future__ import annotations

import numpy as np


def weightednorm(dg, q, p=2, componentwise=False):
    """MJ-weighted p-norm over all components (src/measures.jl:6-21)."""
    a = q.numpy() if hasattr(q, "numpy") and not isinstance(q, np.ndarray) else np.asarray(q)
    if p == np.inf:
        n = np.abs(a).max(axis=1)
        return n if componentwise else float(n.max())
    n = np.array([np.sum(dg.MJ * np.abs(f) ** p) ** (1.0 / p) for f in a])
    return n if componentwise else float(np.sum(n ** p) ** (1.0 / p))
