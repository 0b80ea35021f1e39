"""The drop-in ``src`` package: a reference-style checkout's imports resolve
to this package's GPU modules, and its other ``src`` modules still load."""

import os
import subprocess
import sys
import textwrap

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "triangle_counting_hll_amd", "dropin")


def test_reference_imports_resolve(tmp_path):
    # a stand-in checkout laid out like the reference: its own src package
    # with CPU modules, and GPU modules that need numba (absent here)
    ref = tmp_path / "checkout"
    (ref / "src").mkdir(parents=True)
    (ref / "src" / "__init__.py").write_text("")
    (ref / "src" / "utils.py").write_text("MARK = 'checkout utils'\n")
    (ref / "src" / "gpu_hybrid.py").write_text("from numba import cuda\n")
    code = textwrap.dedent("""
        from src.utils import MARK
        from src.gpu_hybrid import gpu_exact_triangle_count, gpu_hybrid_triangle_count
        from src.gpu_hybrid import _THREADS, _require_supported_block_size
        from src.gpu_hll import build_sketches_kernel, alloc_sketch_arrays
        from src.gpu_graph import to_csr_numpy, to_csr_gpu
        import src.gpu_hybrid
        import triangle_counting_hll_amd.gpu_hybrid as impl
        assert src.gpu_hybrid is impl and _THREADS == 256
        regs, has = alloc_sketch_arrays(5, 4)
        assert regs.shape == (5 * 16,) and has.shape == (5,)
        print(MARK)
    """)
    env = dict(os.environ, PYTHONPATH=os.pathsep.join([DROPIN, str(ref)]))
    out = subprocess.run([sys.executable, "-c", code], cwd=str(tmp_path), env=env,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "checkout utils"
