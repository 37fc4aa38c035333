"""The reference's dataset files and folders (preprocess/util.py:12-42).

A vector file is a little-endian uint64 element count followed by that many
int32 values; a scalar file is one little-endian int32.  The multi-GPU
drivers' 64-bit variant (multi-GPU/sss-multi-GPU/main.cu:25-57) stores int64
values and an int64 scalar; its readers narrow to int32 (here a value that
does not fit raises instead of wrapping).  The folder writers
lay out the files the reference command lines read (TC.cu:236-245,
sss.cu:234-244, ir.cu:254-267), so a dataset made here runs under both.
"""
from __future__ import annotations

import os

import numpy as np


def write_i32_vec(path, values):
    v = np.ascontiguousarray(values, dtype="<i4")
    with open(path, "wb") as f:
        f.write(np.uint64(v.size).astype("<u8").tobytes())
        f.write(v.tobytes())


def read_i32_vec(path):
    with open(path, "rb") as f:
        head = f.read(8)
        if len(head) != 8:
            raise ValueError(f"{path}: truncated header")
        n = int(np.frombuffer(head, "<u8")[0])
        data = np.frombuffer(f.read(4 * n), "<i4")
        if data.size != n:
            raise ValueError(f"{path}: {data.size} of {n} values present")
    return data.astype(np.int32)


def write_int(path, value):
    with open(path, "wb") as f:
        f.write(np.int32(value).astype("<i4").tobytes())


def read_int(path):
    with open(path, "rb") as f:
        b = f.read(4)
    if len(b) != 4:
        raise ValueError(f"{path}: truncated scalar")
    return int(np.frombuffer(b, "<i4")[0])


def write_i64_vec(path, values):
    v = np.ascontiguousarray(values, dtype="<i8")
    with open(path, "wb") as f:
        f.write(np.uint64(v.size).astype("<u8").tobytes())
        f.write(v.tobytes())


def read_i64_vec(path):
    with open(path, "rb") as f:
        head = f.read(8)
        if len(head) != 8:
            raise ValueError(f"{path}: truncated header")
        n = int(np.frombuffer(head, "<u8")[0])
        data = np.frombuffer(f.read(8 * n), "<i8")
        if data.size != n:
            raise ValueError(f"{path}: {data.size} of {n} values present")
    return data.astype(np.int64)


def _narrow(values, path):
    v = np.asarray(values, np.int64)
    if v.size and (v.min() < -2**31 or v.max() >= 2**31):
        raise ValueError(f"{path}: a value does not fit int32")
    return v.astype(np.int32)


def read_vec_64(path):
    """read_vec_64 (multi-GPU/sss-multi-GPU/main.cu:25-44): int64 file -> int32 values."""
    return _narrow(read_i64_vec(path), path)


def write_int64(path, value):
    with open(path, "wb") as f:
        f.write(np.int64(value).astype("<i8").tobytes())


def read_num_64(path):
    """read_num_64 (multi-GPU/sss-multi-GPU/main.cu:46-57): one int64 -> int32."""
    with open(path, "rb") as f:
        b = f.read(8)
    if len(b) != 8:
        raise ValueError(f"{path}: truncated scalar")
    return int(_narrow([np.frombuffer(b, "<i8")[0]], path)[0])


def write_tc_folder(folder, offsets, cols):
    """Oriented CSR for `tc`: csr_offsets_tri / csr_cols_tri / num_*_tri."""
    os.makedirs(folder, exist_ok=True)
    write_i32_vec(os.path.join(folder, "csr_offsets_tri.bin"), offsets)
    write_i32_vec(os.path.join(folder, "csr_cols_tri.bin"), cols)
    write_int(os.path.join(folder, "num_nodes_tri.bin"), len(offsets) - 1)
    write_int(os.path.join(folder, "num_edges_tri.bin"), len(cols))


def write_sss_folder(folder, offsets, cols):
    """Symmetric CSR for `sss`: candidate pair e is (vertexs[e], csr_cols[e])."""
    os.makedirs(folder, exist_ok=True)
    offsets = np.asarray(offsets, np.int64)
    vertexs = np.repeat(np.arange(len(offsets) - 1, dtype=np.int32), np.diff(offsets))
    write_i32_vec(os.path.join(folder, "vertexs.bin"), vertexs)
    write_i32_vec(os.path.join(folder, "csr_offsets.bin"), offsets)
    write_i32_vec(os.path.join(folder, "csr_cols.bin"), cols)
    write_int(os.path.join(folder, "num_nodes.bin"), len(offsets) - 1)
    write_int(os.path.join(folder, "num_edges.bin"), len(cols))


def write_ir_folder(folder, index_offsets, index, query, query_offsets):
    """Inverted index and queries for `ir` (terms of a query term0 first)."""
    os.makedirs(folder, exist_ok=True)
    write_i32_vec(os.path.join(folder, "inverted_index_offsets.bin"), index_offsets)
    write_i32_vec(os.path.join(folder, "inverted_index.bin"), index)
    write_int(os.path.join(folder, "inverted_index_num.bin"), len(index_offsets) - 1)
    write_i32_vec(os.path.join(folder, "query.bin"), query)
    write_i32_vec(os.path.join(folder, "query_offsets.bin"), query_offsets)
    write_int(os.path.join(folder, "query_num.bin"), len(query_offsets) - 1)
