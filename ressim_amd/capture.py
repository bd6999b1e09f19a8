"""The reference's linear-system capture format, fim-capture-v2 and fim-capture-v3
(rs/fim/linear/capture.rs:149-640): a reader and a writer.

A capture holds one FIM linear system J x = rhs with where it came from (Newton
iteration, failure reason, dominant residual family and item), its block layout,
optionally the equation scaling, and in v3 the companion data of the OPM Flow
lifecycle study (storage blocks, true-IMPES weights and the reservoir/well partition
of J). Values are written as Rust's `{:e}` prints an f64: the shortest digits that
round-trip, so a written system reads back bit for bit.

The library writes the systems its linear stack solves (`ReservoirSimulator.
set_capture_dir`) and replays any capture through that stack
(`ReservoirSimulator.solve_captured_system`), as the reference's offline solver lab
(rs/fim/linear/solver_lab.rs) replays captures through its CPU stack.
"""
from __future__ import annotations

import decimal
import math
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import scipy.sparse as sp

# rs/fim/linear/flow_lifecycle.rs:20-26, the source fingerprint a v3 capture must carry
FLOW_SOURCE_TAG = "release/2026.04/final"
FLOW_SOURCE_COMMIT = "b82f21dba405286c4c4446614dd3bf9cdebf7a2c"
DUNE_ISTL_VERSION = "2.11.0"
TRUE_IMPES_PRESSURE_SCALE_BAR = 50.0


class CaptureError(ValueError):
    """A capture that does not parse or does not validate."""


@dataclass
class CaptureMetadata:  # FimCaptureMetadata, capture.rs:77-83
    newton_iteration: int
    failure_reason: str
    dominant_family: str
    dominant_item_index: int


@dataclass
class BlockLayout:  # FimLinearBlockLayout, rs/fim/linear/mod.rs:149-154
    cell_block_count: int
    cell_block_size: int
    well_bhp_count: int
    perforation_tail_start: int


@dataclass
class EquationScaling:  # rs/fim/scaling.rs, per-item scales of each equation family
    water: np.ndarray
    oil_component: np.ndarray
    gas_component: np.ndarray
    well_constraint: np.ndarray
    perforation_flow: np.ndarray

    FAMILIES = ("water", "oil_component", "gas_component", "well_constraint", "perforation_flow")

    def __eq__(self, other):
        return isinstance(other, EquationScaling) and all(
            np.array_equal(getattr(self, f), getattr(other, f)) for f in self.FAMILIES)


@dataclass
class FlowLifecycle:  # FimFlowLifecycleCapture, capture.rs:89-101
    source_tag: str
    source_commit: str
    dune_istl_version: str
    pressure_scale_bar: float
    storage_blocks: np.ndarray  # (cells, 3, 3)
    true_impes_weights: np.ndarray  # (cells, 3)
    reservoir_unknown_count: int
    j_rr: sp.csr_matrix
    j_rw: sp.csr_matrix
    j_wr: sp.csr_matrix
    j_ww: sp.csr_matrix


@dataclass
class CapturedSystem:  # FimCapturedSystem, capture.rs:106-114
    metadata: CaptureMetadata
    layout: Optional[BlockLayout]
    jacobian: sp.csr_matrix
    rhs: np.ndarray
    equation_scaling: Optional[EquationScaling] = None
    flow_lifecycle: Optional[FlowLifecycle] = field(default=None)


def format_f64(v: float) -> str:
    """v as Rust's `{:e}` formats an f64 ("1.5e-8", "4e0", "1e12", "0e0", "-0e0", "inf", "NaN")."""
    v = float(v)
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    # repr gives the shortest decimal that round-trips; restate it as d.ddd e x
    sign, digits, exp = decimal.Decimal(repr(v)).as_tuple()
    ds = "".join(map(str, digits)).rstrip("0") or "0"
    e10 = exp + len(digits) - 1 if ds != "0" else 0
    mant = ds[0] + ("." + ds[1:] if len(ds) > 1 else "")
    return ("-" if sign else "") + mant + "e" + str(e10)


def _parse_f64(text: str) -> float:
    try:
        return float(text)
    except ValueError as exc:
        raise CaptureError(str(exc)) from None


def _parse_usize(text: str) -> int:
    if not text.isdigit():
        raise CaptureError(f"invalid digit in {text!r}")
    return int(text)


def true_impes_weight(storage_block) -> np.ndarray:
    """rs/fim/linear/flow_lifecycle.rs:32-60: solve S^T w = e0 with the pressure column
    of the storage block scaled by 50 bar, then normalise w by its largest magnitude."""
    s = np.array(storage_block, dtype=np.float64).reshape(3, 3).copy()
    s[:, 0] *= TRUE_IMPES_PRESSURE_SCALE_BAR
    try:
        w = np.linalg.solve(s.T, np.array([1.0, 0.0, 0.0]))
    except np.linalg.LinAlgError:
        raise CaptureError("singular true-IMPES storage block") from None
    if not np.all(np.isfinite(w)):
        raise CaptureError("non-finite true-IMPES storage solve")
    scale = np.max(np.abs(w))
    if not np.isfinite(scale) or scale <= np.finfo(np.float64).eps:
        raise CaptureError("degenerate true-IMPES normalization")
    return w / scale


def validate_flow_lifecycle(flow: FlowLifecycle, layout: BlockLayout, jacobian: sp.csr_matrix) -> None:
    """rs/fim/linear/flow_lifecycle.rs:153-215: fingerprint, cardinalities, the weights
    recomputed from the storage blocks (within 1e-12) and the four-way partition of J
    (bit for bit)."""
    if (flow.source_tag != FLOW_SOURCE_TAG or flow.source_commit != FLOW_SOURCE_COMMIT
            or flow.dune_istl_version != DUNE_ISTL_VERSION
            or abs(flow.pressure_scale_bar - TRUE_IMPES_PRESSURE_SCALE_BAR) > np.finfo(np.float64).eps):
        raise CaptureError("Y2d6 source/config fingerprint mismatch")
    n = layout.cell_block_count
    if (layout.cell_block_size != 3 or len(flow.storage_blocks) != n or len(flow.true_impes_weights) != n
            or flow.reservoir_unknown_count != n * layout.cell_block_size):
        raise CaptureError("Y2d6 storage/layout cardinality mismatch")
    for c, (block, w) in enumerate(zip(flow.storage_blocks, flow.true_impes_weights)):
        d = np.max(np.abs(true_impes_weight(block) - np.asarray(w)))
        if d > 1e-12:
            raise CaptureError(f"Y2d6 true-IMPES weight mismatch at cell {c}: {format_f64(d)}")
    r = flow.reservoir_unknown_count
    wn = max(jacobian.shape[0] - r, 0)
    dims = [(flow.j_rr.shape, (r, r)), (flow.j_rw.shape, (r, wn)), (flow.j_wr.shape, (wn, r)),
            (flow.j_ww.shape, (wn, wn))]
    if any(a != b for a, b in dims):
        raise CaptureError("Y2d6 reservoir/well partition dimensions mismatch")
    full = sp.bmat([[flow.j_rr, flow.j_rw], [flow.j_wr, flow.j_ww]], format="csr") if wn else flow.j_rr.tocsr()
    a = full.toarray().view(np.uint64)
    b = jacobian.toarray().view(np.uint64)
    if a.shape != b.shape or not np.array_equal(a, b):
        bad = np.argwhere(a != b) if a.shape == b.shape else [(0, 0)]
        raise CaptureError(f"Y2d6 partition mismatch at ({bad[0][0]},{bad[0][1]})")


def _sparse_lines(name: str, m: sp.csr_matrix) -> List[str]:
    m = sp.csr_matrix(m)
    m.sort_indices()
    out = [f"matrix_{name} {m.shape[0]} {m.shape[1]} {m.nnz}"]
    for r in range(m.shape[0]):
        for k in range(m.indptr[r], m.indptr[r + 1]):
            out.append(f"{r} {m.indices[k]} {format_f64(m.data[k])}")
    return out


def capture_text(system: CapturedSystem) -> str:
    """The text of one capture file (capture.rs:198-314)."""
    md, layout, flow = system.metadata, system.layout, system.flow_lifecycle
    lines = ["fim-capture-v3" if flow is not None else "fim-capture-v2",
             f"newton_iteration {md.newton_iteration}", f"failure_reason {md.failure_reason}",
             f"dominant_family {md.dominant_family}", f"dominant_item_index {md.dominant_item_index}"]
    if layout is None:
        lines.append("layout none")
    else:
        lines.append(f"layout {layout.cell_block_count} {layout.cell_block_size} {layout.well_bhp_count} "
                     f"{layout.perforation_tail_start}")
    sc = system.equation_scaling
    if sc is None:
        lines.append("equation_scaling 0")
    else:
        lines.append("equation_scaling 1")
        for fam in EquationScaling.FAMILIES:
            vals = np.asarray(getattr(sc, fam), dtype=np.float64)
            lines.append(" ".join([fam, str(len(vals))] + [format_f64(v) for v in vals]))
    if flow is not None:
        lines += ["flow_lifecycle 1", f"flow_source_tag {flow.source_tag}", f"flow_source_commit {flow.source_commit}",
                  f"dune_istl_version {flow.dune_istl_version}",
                  f"pressure_scale_bar {format_f64(flow.pressure_scale_bar)}"]
        sb = np.asarray(flow.storage_blocks, dtype=np.float64).reshape(-1, 9)
        lines.append(f"storage_blocks {len(sb)}")
        lines += [" ".join(format_f64(v) for v in row) for row in sb]
        tw = np.asarray(flow.true_impes_weights, dtype=np.float64).reshape(-1, 3)
        lines.append(f"true_impes_weights {len(tw)}")
        lines += [" ".join(format_f64(v) for v in row) for row in tw]
        lines.append(f"reservoir_unknown_count {flow.reservoir_unknown_count}")
        for name in ("j_rr", "j_rw", "j_wr", "j_ww"):
            lines += _sparse_lines(name, getattr(flow, name))
    J = sp.csr_matrix(system.jacobian)
    J.sort_indices()
    lines += [f"rows {J.shape[0]}", f"cols {J.shape[1]}", f"nnz {J.nnz}", "rhs"]
    lines += [format_f64(v) for v in np.asarray(system.rhs, dtype=np.float64)]
    lines.append("triplets")
    for r in range(J.shape[0]):
        for k in range(J.indptr[r], J.indptr[r + 1]):
            lines.append(f"{r} {J.indices[k]} {format_f64(J.data[k])}")
    return "\n".join(lines) + "\n"


def write_capture(directory, sequence: int, system: CapturedSystem) -> str:
    """Write `system` to <directory>/fim_capture_<sequence:05>.txt, creating the
    directory (capture.rs:198-210); returns the path."""
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, f"fim_capture_{sequence:05d}.txt")
    with open(path, "w") as f:
        f.write(capture_text(system))
    return path


def parse_capture(content: str) -> CapturedSystem:
    """Parse one capture (capture.rs:346-645); raises CaptureError."""
    lines = iter(content.splitlines())

    def nxt():
        return next(lines, None)

    def fld(line, key):
        if line is None:
            raise CaptureError(f"missing {key}")
        if not line.startswith(key):
            raise CaptureError(f"expected {key}, got {line!r}")
        return line[len(key):].strip()

    header = nxt()
    if header is None:
        raise CaptureError("empty file")
    if header not in ("fim-capture-v2", "fim-capture-v3"):
        raise CaptureError(f"unexpected header {header!r}")
    has_flow = header == "fim-capture-v3"
    md = CaptureMetadata(_parse_usize(fld(nxt(), "newton_iteration")), fld(nxt(), "failure_reason"),
                         fld(nxt(), "dominant_family"), _parse_usize(fld(nxt(), "dominant_item_index")))
    lay = fld(nxt(), "layout")
    if lay == "none":
        layout = None
    else:
        parts = [_parse_usize(p) for p in lay.split()]
        if len(parts) != 4:
            raise CaptureError(f"layout expects 4 fields, got {len(parts)}")
        layout = BlockLayout(*parts)

    def scale_row(key):
        rest = fld(nxt(), key).split()
        if not rest:
            raise CaptureError(f"{key} missing count")
        count = _parse_usize(rest[0])
        vals = np.array([_parse_f64(p) for p in rest[1:]], dtype=np.float64)
        if len(vals) != count:
            raise CaptureError(f"{key} expects {count} values, got {len(vals)}")
        return vals

    flag = fld(nxt(), "equation_scaling")
    if flag == "0":
        scaling = None
    elif flag == "1":
        scaling = EquationScaling(*[scale_row(f) for f in EquationScaling.FAMILIES])
    else:
        raise CaptureError(f"unexpected equation_scaling flag {flag!r}")

    def dense_rows(key, cols):
        count = _parse_usize(fld(nxt(), key))
        rows = []
        for r in range(count):
            line = nxt()
            if line is None:
                raise CaptureError(f"{key} truncated at row {r}")
            vals = [_parse_f64(p) for p in line.split()]
            if len(vals) != cols:
                raise CaptureError(f"{key} row {r} expects {cols} values, got {len(vals)}")
            rows.append(vals)
        return np.array(rows, dtype=np.float64).reshape(count, cols)

    def sparse_matrix(name):
        dims = fld(nxt(), f"matrix_{name}").split()
        if len(dims) != 3:
            raise CaptureError(f"matrix_{name} expects rows cols nnz")
        nr, nc, nz = (_parse_usize(d) for d in dims)
        rr, cc, vv = [], [], []
        for k in range(nz):
            line = nxt()
            if line is None:
                raise CaptureError(f"matrix_{name} truncated at entry {k}")
            parts = line.split()
            if len(parts) != 3:
                raise CaptureError(f"matrix_{name} invalid entry {k}")
            r, c, v = _parse_usize(parts[0]), _parse_usize(parts[1]), _parse_f64(parts[2])
            if r >= nr or c >= nc:
                raise CaptureError(f"matrix_{name} invalid entry {k}")
            rr.append(r), cc.append(c), vv.append(v)
        return sp.csr_matrix((vv, (rr, cc)), shape=(nr, nc))

    flow = None
    if has_flow:
        if fld(nxt(), "flow_lifecycle") != "1":
            raise CaptureError("fim-capture-v3 requires flow_lifecycle 1")
        tag, commit, dune = fld(nxt(), "flow_source_tag"), fld(nxt(), "flow_source_commit"), fld(
            nxt(), "dune_istl_version")
        psb = _parse_f64(fld(nxt(), "pressure_scale_bar"))
        storage = dense_rows("storage_blocks", 9).reshape(-1, 3, 3)
        weights = dense_rows("true_impes_weights", 3)
        runk = _parse_usize(fld(nxt(), "reservoir_unknown_count"))
        flow = FlowLifecycle(tag, commit, dune, psb, storage, weights, runk, sparse_matrix("j_rr"),
                             sparse_matrix("j_rw"), sparse_matrix("j_wr"), sparse_matrix("j_ww"))
    rows, cols, nnz = (_parse_usize(fld(nxt(), k)) for k in ("rows", "cols", "nnz"))
    if nxt() != "rhs":
        raise CaptureError("expected rhs section")
    rhs = np.empty(rows)
    for r in range(rows):
        line = nxt()
        if line is None:
            raise CaptureError("rhs truncated")
        rhs[r] = _parse_f64(line.strip())
    if nxt() != "triplets":
        raise CaptureError("expected triplets section")
    rr, cc, vv = [], [], []
    for line in lines:
        line = line.strip()
        if not line:
            continue
        parts = line.split()
        if len(parts) < 1:
            raise CaptureError("triplet missing row")
        if len(parts) < 2:
            raise CaptureError("triplet missing col")
        if len(parts) < 3:
            raise CaptureError("triplet missing value")
        rr.append(_parse_usize(parts[0])), cc.append(_parse_usize(parts[1])), vv.append(_parse_f64(parts[2]))
    if len(vv) != nnz:
        raise CaptureError(f"expected {nnz} triplets, found {len(vv)}")
    if any(r >= rows for r in rr) or any(c >= cols for c in cc):
        raise CaptureError("triplet outside the matrix")
    J = sp.csr_matrix((np.array(vv, dtype=np.float64), (rr, cc)), shape=(rows, cols))  # duplicates summed
    system = CapturedSystem(md, layout, J, rhs, scaling, flow)
    if flow is not None:
        if layout is None:
            raise CaptureError("fim-capture-v3 requires a block layout")
        validate_flow_lifecycle(flow, layout, J)
    return system


def load_capture(path) -> CapturedSystem:
    with open(path) as f:
        content = f.read()
    try:
        return parse_capture(content)
    except CaptureError as exc:
        raise CaptureError(f"{path}: {exc}") from None


def load_captures(directory) -> List[CapturedSystem]:
    """Every fim_capture_*.txt in `directory`, sorted by file name (capture.rs:318-333)."""
    names = sorted(n for n in os.listdir(directory) if n.startswith("fim_capture_") and n.endswith(".txt"))
    return [load_capture(os.path.join(directory, n)) for n in names]
