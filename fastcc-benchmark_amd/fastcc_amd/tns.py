"""FROSTT-style text tensors, read and written with the reference's rules.

read_tns follows Tensor<double>::Tensor(filename, has_header)
(read.hpp:58-106):
  - a line whose first character is not a digit is skipped (comments, blank
    lines, a leading sign);
  - with has_header the first remaining line is skipped too;
  - fields are separated by single spaces; every field but the last is an
    integer coordinate kept exactly as written (FROSTT files are 1-based and
    the reference does not shift them), the last field is the value;
  - every line must hold the same number of fields.
Malformed lines (a doubled or trailing space, a non-numeric field) raise
ValueError, where the reference's std::stoi / std::stod throw.

write_tns follows Tensor<double>::write (read.hpp:108-117): it appends a line
with the shape (per mode, the largest coordinate, as _infer_shape leaves it,
contract.hpp:124-144), each value followed by a space, then one line per
nonzero with every coordinate followed by a space and the value printed by a
default C++ stream (6 significant digits); full_precision=True writes values
round-trip exact instead.
"""
from __future__ import annotations

import numpy as np

from .workloads import SparseTensor


def read_tns(path: str, has_header: bool = False) -> SparseTensor:
    coords, vals = [], []
    order = 0
    skip = has_header
    with open(path, "r") as f:
        for line in f:
            line = line.rstrip("\n")
            if not line or not line[0].isdigit():
                continue
            if skip:
                skip = False
                continue
            fields = line.split(" ")
            if len(fields) < 2:
                raise ValueError(f"{path}: a nonzero needs coordinates and a value: {line!r}")
            c = [int(x) for x in fields[:-1]]
            v = float(fields[-1])
            if order == 0:
                order = len(c)
            elif len(c) != order:
                raise ValueError(f"{path}: {len(c)} coordinates where {order} were expected: {line!r}")
            coords.append(c)
            vals.append(v)
    if not coords:
        return SparseTensor(np.zeros((max(order, 1), 0), dtype=np.uint32), np.zeros(0))
    cs = np.asarray(coords, dtype=np.int64)
    if cs.min() < 0 or cs.max() > 0xFFFFFFFF:
        raise ValueError(f"{path}: coordinates must fit in 32 unsigned bits")
    return SparseTensor(np.ascontiguousarray(cs.T.astype(np.uint32)), np.asarray(vals, dtype=np.float64))


def write_tns(t: SparseTensor, path: str, full_precision: bool = False) -> None:
    shape = t.coords.max(axis=1) if t.nnz else np.zeros(t.order, dtype=np.uint32)
    fmt = (lambda v: repr(float(v))) if full_precision else (lambda v: "%g" % v)
    with open(path, "a") as f:
        f.write("".join(f"{int(s)} " for s in shape) + "\n")
        for n in range(t.nnz):
            f.write("".join(f"{int(t.coords[m, n])} " for m in range(t.order)) + fmt(t.vals[n]) + "\n")
