"""Drop-in ``src`` package for the reference's callers.

With ``triangle_counting_hll_amd/dropin`` ahead of the reference checkout on
``sys.path`` (for example ``PYTHONPATH=<this repo>/triangle_counting_hll_amd/dropin:<reference
checkout>``), the reference's own imports resolve unchanged:

* ``src.gpu_graph``, ``src.gpu_hll`` and ``src.gpu_hybrid`` (the GPU modules
  imported by experiment/exp_gpu.py:101-102) are this package's HIP-backed
  modules ``triangle_counting_hll_amd.gpu_graph`` / ``gpu_hll`` /
  ``gpu_hybrid`` -- the same module objects, private names included;
* every other ``src`` module (``src.exact``, ``src.utils``, ``src.graph`` ...)
  comes from the next ``src`` directory on ``sys.path``, the reference's own.
"""

import os
import pkgutil
import sys

# the repository root, so that triangle_counting_hll_amd imports from here
_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__)))))
if _ROOT not in sys.path:
    sys.path.append(_ROOT)

__path__ = pkgutil.extend_path(__path__, __name__)
