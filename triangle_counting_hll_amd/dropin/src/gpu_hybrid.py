"""``src.gpu_hybrid`` of the reference, served by ``triangle_counting_hll_amd.gpu_hybrid``
(this name is bound to that module object, so every public and private name of
it is visible, and state such as warm-up flags is shared)."""

import sys

from triangle_counting_hll_amd import gpu_hybrid as _impl

sys.modules[__name__] = _impl
