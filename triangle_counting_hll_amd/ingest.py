"""Edge-list ingest on the device: SNAP-style text -> raw pairs -> oriented CSR.

Mirrors the GPU experiment's loader, ``experiment/exp_gpu.py``:

* ``parse_edge_list_gpu`` / ``load_edge_list_gpu`` replace the two Python passes
  over ``_iter_edges_raw`` (:131-143, :291-313): the text is parsed by HIP
  kernels (``tc_parse_edge_list``) with the same line rules -- breaks at
  ``\\n``/``\\r``, leading whitespace stripped, lines starting with a comment
  character (``COMMENT_CHARS = "#%"``, :93) or with fewer than two fields
  skipped, the first two fields read as ``int()``, self-loops dropped.
* ``load_graph_gpu`` is ``_load_graph`` (:264-389): the same 11-tuple, with the
  dedup / relabel / orientation / CSR steps done by ``gpu_graph.build_csr_gpu``
  and every array left on the device.

Fields that ``int()`` rejects raise ``ValueError`` naming the file and line, as
the reference does; negative ids raise ``ValueError`` and ids above
``INT32_MAX`` raise ``OverflowError`` (the reference stores ids in an int32
array, :300, where such ids go wrong).
"""

from __future__ import annotations

import ctypes
import gzip
import time
from typing import Optional

import numpy as np

from . import _lib, gpu_graph

COMMENT_CHARS = "#%"  # experiment/exp_gpu.py:93

_KIND_MSG = {2: "expected integer node ids", 3: "negative node id",
             4: "node id does not fit int32"}


class EdgeListError(ValueError):
    """A line the parser rejects; ``offset`` is its byte offset in the parsed text."""

    def __init__(self, msg: str, offset: int, kind: int):
        super().__init__(msg)
        self.offset = offset
        self.kind = kind


def _raise_parse_error(offset: int, kind: int, where: str):
    msg = f"{where}: {_KIND_MSG.get(kind, 'bad line')}"
    if kind == 4:
        raise OverflowError(msg)
    raise EdgeListError(msg, offset, kind)


def _device_text(text, t):
    if isinstance(text, t.Tensor):
        if text.dtype != t.uint8 or not text.is_cuda:
            raise ValueError("a tensor text must be a uint8 CUDA tensor")
        return text.contiguous()
    if isinstance(text, str):
        text = text.encode()
    arr = np.frombuffer(memoryview(text), dtype=np.uint8) if not isinstance(text, np.ndarray) \
        else np.ascontiguousarray(text).view(np.uint8).reshape(-1)
    return t.from_numpy(arr.copy()).cuda()


@_lib.on_stream
def _parse_device(buf, n_bytes: int, comment_chars: str, out, stream=None):
    """Run tc_parse_edge_list on ``buf[:n_bytes]`` writing into ``out`` (int32 [cap, 2]).

    Returns ``(n_pairs, bad_offset, bad_kind)``; bad_offset is -1 when every line parsed.
    """
    lib = _lib.load()
    n = ctypes.c_int64()
    bad = ctypes.c_int64()
    kind = ctypes.c_int()
    rc = lib.tc_parse_edge_list(_lib.ptr(buf), n_bytes, comment_chars.encode(),
                                _lib.ptr(out) if out is not None else None,
                                out.shape[0] if out is not None else 0, ctypes.byref(n),
                                ctypes.byref(bad), ctypes.byref(kind), _lib.stream_ptr(stream))
    if rc != _lib.TC_OK and bad.value >= 0:
        return n.value, bad.value, kind.value
    _lib.check(rc)
    return n.value, -1, 0


def max_pairs(n_bytes: int) -> int:
    """Most pairs any text of ``n_bytes`` can hold ("u v" plus a break per line)."""
    return (n_bytes + 1) // 4 + 1


@_lib.on_stream
def parse_edge_list_gpu(text, comment_chars: str = COMMENT_CHARS, stream=None):
    """Edge-list text (bytes / str / uint8 array on the host, or a uint8 CUDA tensor)
    -> raw pairs ``int32[m, 2]`` on the device, in file order, self-loops dropped."""
    t = _lib.require_device()
    if isinstance(text, str):
        text = text.encode()
    buf = _device_text(text, t)
    nb = buf.numel()
    out = t.empty((max_pairs(nb), 2), dtype=t.int32, device=buf.device)
    n, bad, kind = _parse_device(buf, nb, comment_chars, out, stream)
    if bad >= 0:
        prefix = bytes(buf[:bad].cpu().numpy()) if isinstance(text, t.Tensor) \
            else bytes(memoryview(text)[:bad])
        _raise_parse_error(bad, kind, f"line {_line_number(prefix)}")
    return out[:n]


@_lib.on_stream
def count_edge_list_gpu(text, comment_chars: str = COMMENT_CHARS, stream=None) -> int:
    """Number of pairs ``parse_edge_list_gpu`` would return (the first, counting pass only)."""
    t = _lib.require_device()
    buf = _device_text(text, t)
    n, bad, kind = _parse_device(buf, buf.numel(), comment_chars, None, stream)
    if bad >= 0:
        _raise_parse_error(bad, kind, f"byte {bad}")
    return n


@_lib.on_stream
def format_edge_list_gpu(pairs, sep: str = "\t", newline: str = "\n", stream=None):
    """Raw pairs (``int32[m, 2]``, ids >= 0, host or device) -> edge-list text as a
    uint8 CUDA tensor, one ``u<sep>v<newline>`` line per pair (``tc_format_edge_list``)."""
    t = _lib.require_device()
    lib = _lib.load()
    p = _lib.as_device(pairs, t.int32, "pairs")
    if p.ndim != 2 or p.shape[1] != 2:
        raise ValueError(f"pairs must have shape (m, 2), got {tuple(p.shape)}")
    nb = ctypes.c_int64()
    s = _lib.stream_ptr(stream)
    _lib.check(lib.tc_format_edge_list(_lib.ptr(p), p.shape[0], sep.encode(), newline.encode(),
                                       None, 0, ctypes.byref(nb), s))
    out = t.empty(nb.value, dtype=t.uint8, device=p.device)
    _lib.check(lib.tc_format_edge_list(_lib.ptr(p), p.shape[0], sep.encode(), newline.encode(),
                                       _lib.ptr(out), out.numel(), ctypes.byref(nb), s))
    return out


def _line_number(prefix: bytes) -> int:
    """1-based line of the byte after ``prefix`` (universal newlines)."""
    return prefix.count(b"\n") + prefix.count(b"\r") - prefix.count(b"\r\n") + 1


def _open(path: str):
    return gzip.open(path, "rb") if str(path).endswith(".gz") else open(path, "rb")


def load_edge_list_gpu(path: str, comment_chars: str = COMMENT_CHARS,
                       chunk_bytes: int = 1 << 28):
    """Stream an edge-list file through the device parser in whole-line chunks.

    Returns raw pairs ``int32[m, 2]`` on the device (file order, self-loops
    dropped) -- the edge buffer the reference fills at exp_gpu.py:299-313.
    ``.gz`` files are decompressed on the host.  Runs on the current stream.
    """
    t = _lib.require_device()
    chunk_bytes = max(int(chunk_bytes), 1 << 12)
    pinned = t.empty(chunk_bytes, dtype=t.uint8, pin_memory=True)
    pinned_np = pinned.numpy()
    dev = t.empty(chunk_bytes, dtype=t.uint8, device="cuda")
    parts = []
    file_off = 0  # file offset of the current chunk
    carry = b""
    with _open(path) as f:
        eof = False
        while not eof or carry:
            data = carry
            while not eof and len(data) < chunk_bytes:
                got = f.read(chunk_bytes - len(data))
                if not got:
                    eof = True
                    break
                data += got
            if not data:
                break
            cut = len(data)
            if not eof:
                cut = max(data.rfind(b"\n"), data.rfind(b"\r")) + 1
                if cut == 0:  # a line longer than the chunk: grow the buffers
                    chunk_bytes *= 2
                    pinned = t.empty(chunk_bytes, dtype=t.uint8, pin_memory=True)
                    pinned_np = pinned.numpy()
                    dev = t.empty(chunk_bytes, dtype=t.uint8, device="cuda")
                    carry = data
                    continue
            piece, carry = data[:cut], data[cut:]
            nb = len(piece)
            pinned_np[:nb] = np.frombuffer(piece, dtype=np.uint8)
            dev[:nb].copy_(pinned[:nb], non_blocking=True)
            out = t.empty((max_pairs(nb), 2), dtype=t.int32, device="cuda")
            n, bad, kind = _parse_device(dev, nb, comment_chars, out)
            if bad >= 0:
                _raise_parse_error(bad, kind,
                                   f"{path}:{_file_line(path, file_off + bad)}")
            parts.append(out[:n].clone())
            del out
            file_off += nb
            if eof and not carry:
                break
    if not parts:
        return t.empty((0, 2), dtype=t.int32, device="cuda")
    return parts[0] if len(parts) == 1 else t.cat(parts)


def _file_line(path: str, offset: int) -> int:
    with _open(path) as f:
        return _line_number(f.read(offset))


def load_graph_gpu(graph_path: str, need_forward_adj: bool = False, keep_edges_np: bool = True,
                   dedup_edges: bool = True, row_ptr_dtype: str = "int32",
                   comment_chars: str = COMMENT_CHARS, chunk_bytes: int = 1 << 28):
    """``_load_graph`` (experiment/exp_gpu.py:264-389) on the device.

    Returns ``(row_ptr, col_idx, forward_deg, edge_u, edge_v, forward_adj,
    edges_keep, node_ids, n_nodes, n_oriented, t_load)`` like the reference,
    with the arrays as CUDA tensors: ``edges_keep`` is the sorted unique
    canonical ``(min, max)`` edge list in original ids (int32, or None when
    ``keep_edges_np`` is False) and ``forward_adj`` a host dict (only with
    ``need_forward_adj``).  Duplicates are always removed; the reference's
    ``dedup_edges=False`` asserts unique input, for which the result is the
    same.
    """
    t = _lib.require_device()
    del dedup_edges  # the device build always dedups; unique input gives the same CSR
    t0 = time.perf_counter()
    pairs = load_edge_list_gpu(graph_path, comment_chars, chunk_bytes)
    csr = gpu_graph.build_csr_gpu(pairs, row_ptr_dtype=row_ptr_dtype)
    del pairs
    edges_keep = None
    if keep_edges_np:
        lo = t.minimum(csr.edge_u, csr.edge_v).long()
        hi = t.maximum(csr.edge_u, csr.edge_v).long()
        order = t.argsort(lo * max(csr.n_nodes, 1) + hi)
        edges_keep = t.stack([csr.node_ids[lo[order]], csr.node_ids[hi[order]]], 1).to(t.int32)
    forward_adj: Optional[dict] = None
    if need_forward_adj:
        rp = csr.row_ptr.cpu().numpy().astype(np.int64)
        col = csr.col_idx.cpu().numpy()
        forward_adj = {i: col[rp[i]:rp[i + 1]].astype(np.int64) for i in range(csr.n_nodes)}
    t.cuda.synchronize()
    t_load = time.perf_counter() - t0
    return (csr.row_ptr, csr.col_idx, csr.degrees, csr.edge_u, csr.edge_v, forward_adj,
            edges_keep, csr.node_ids, csr.n_nodes, csr.n_oriented, t_load)
