"""``src.gpu_graph`` of the reference, served by ``triangle_counting_hll_amd.gpu_graph``
(this name is bound to that module object, so every public and private name of
it is visible, and state such as warm-up flags is shared)."""

import sys

from triangle_counting_hll_amd import gpu_graph as _impl

sys.modules[__name__] = _impl
