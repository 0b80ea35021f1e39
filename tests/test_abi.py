"""CPU-side checks of the C ABI library: it loads, and exports what include/tchll.h declares."""

import os
import re

import pytest

from triangle_counting_hll_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "tchll.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("tc_exact_count", "tc_hll_build", "tc_union_estimate", "tc_hybrid_count",
                 "tc_csr_build_begin", "tc_csr_build_finish", "tc_degree_buckets",
                 "tc_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name


def test_binding_table_matches_header():
    assert sorted(_lib.EXPORTED) == _declared()


def test_no_error_before_any_call():
    lib = _lib.load()
    assert lib.tc_abi_version() == 1
    assert lib.tc_last_error() in (b"", None)


def test_status_codes_map_to_reference_exceptions():
    with pytest.raises(ValueError):
        _lib.check(_lib.TC_ERR_INVALID)
    with pytest.raises(MemoryError):
        _lib.check(_lib.TC_ERR_OOM)
    with pytest.raises(RuntimeError):
        _lib.check(_lib.TC_ERR_HIP)
    _lib.check(_lib.TC_OK)


def test_gpu_entry_points_fail_loudly_without_device():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from triangle_counting_hll_amd import gpu_graph, gpu_hll, gpu_hybrid

    with pytest.raises(RuntimeError, match="GPU"):
        gpu_hybrid.gpu_exact_triangle_count([0, 1, 1], [1], [0], [1])
    with pytest.raises(RuntimeError, match="GPU"):
        gpu_graph.build_csr_gpu([[0, 1]])
    with pytest.raises(RuntimeError, match="GPU"):
        gpu_hll.union_estimate_pairs([0] * 16, [0], [0], 4)
