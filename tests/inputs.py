"""Seeded synthetic inputs (re-exported from the package, shared with bench.py)."""
from mctdhb_amd.inputs import *  # noqa: F401,F403
from mctdhb_amd.inputs import coefficients, hermite_gauss, onebody, rng, twobody  # noqa: F401
