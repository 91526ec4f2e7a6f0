"""Site-resolved inputs of the LLG kernels (subset of magex.structure on the hot path)."""
from __future__ import annotations

import numpy as np


def site_temperature(system, T):
    """Per-site temperature array, or the scalar T when no map is set (magex/structure.py:234-239)."""
    sT = getattr(system, "site_T", None)
    if sT is None:
        return T
    return np.where(np.isnan(sT), T, sT)
