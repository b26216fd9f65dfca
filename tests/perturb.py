"""One-ulp perturbations of a right-hand side, used to measure how far the
oracle's own iteration count moves under rounding-level changes (the size of
the change that a different summation order makes)."""
import numpy as np


def ulp_perturbed(b, seed):
    """b with one ulp added or taken away in a random eighth of its entries."""
    rng = np.random.default_rng(seed)
    sel = rng.random(b.size) < 0.125
    up = rng.random(b.size) < 0.5
    out = np.array(b, dtype=np.float64, copy=True)
    out[sel & up] = np.nextafter(out[sel & up], np.inf)
    out[sel & ~up] = np.nextafter(out[sel & ~up], -np.inf)
    return out
