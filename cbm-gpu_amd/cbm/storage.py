"""On-disk CBM structure: a built matrix saved once and loaded instead of
rebuilt.  The reference rebuilds the structure from the edge list in every
constructor (cbm/cbm4mm.py:15-19, cbm_init_ in cbm/cbm_extensions.cpp:60-345);
a file here holds everything a ``cbm4*`` object needs, so loading skips the
candidate scan, the arborescence and the device layout.

Format: one uncompressed ``.npz`` (numpy's zip of ``.npy`` arrays, read with
``allow_pickle=False``), version 5 (version 1 lacked the layout's 4-range
piece set, versions 1 and 2 its tree tails, versions 1 to 3 the units of the
one-launch forward, versions 1 to 4 the gather order; such files still load,
without them):

  meta               uint8   JSON text: format, version, n, alpha, the class
                             that saved it, build and layout statistics, the
                             layout depth, and a CRC-32 per array
  delta_crow         int32   [n + 1]  Delta as CSR, rows ascending
  delta_col          int32   [nnz]    columns ascending within a row
  delta_sign         int8    [nnz]    +1 / -1 (the product's scales are not stored)
  mca_offset         int32   [nb + 1] branch offsets      (init's mca_branches)
  mca_src_idx        int32   [E]      tree edges, parent  (init's mca_row_idx)
  mca_dst_idx        int32   [E]      tree edges, child   (init's mca_col_idx)
  plan/<name>        int32            the device layout of cbm_plan_create, one
                                      array per CBM_P_* slot (include/cbm_host.h)
  gather_order       int32   [m]      the short rows' gather order (cbm_gather_order),
                                      present when the saved matrix had one

The structure does not depend on the product (A, A D, D A D), so a file saved
by any ``cbm4*`` class loads into any of them.
"""
import json
import zlib
from dataclasses import dataclass, field

import numpy as np
import torch

from . import structure

FORMAT = "cbm-structure"
VERSION = 5
# Layout arrays (and statistics) that older files do not have, by version.
_NEW_IN_2 = ("fpc4_b", "fpc4_e", "fpc4_slot", "fpc4_range", "frs4_slot", "frs4_cnt", "bpc4_b", "bpc4_e", "bpc4_mid",
             "bpc4_slot", "bpc4_range", "brs4_slot", "brs4_cnt", "bpc4_row")
_NEW_STATS_IN_2 = ("fwd_pieces4", "bwd_pieces4")
_NEW_IN_3 = ("tf_frag", "tf_seg", "tf_pos", "tr_frag", "tr_seg", "tr_pos")
_NEW_STATS_IN_3 = ("tf_level", "tr_level")  # absent: no tail (the depth)
_NEW_IN_4 = ("un_ptr", "un_row", "un_lpar", "un_pc", "un_pcb", "un_pce", "un_seg_ptr", "un_seg", "un_of",
             "un_waited")
_NEW_STATS_IN_4 = ("units", "unit_levels")  # absent: no units (0)
_CORE = ("delta_crow", "delta_col", "delta_sign", "mca_offset", "mca_src_idx", "mca_dst_idx")
_DTYPES = {"delta_crow": np.int32, "delta_col": np.int32, "delta_sign": np.int8, "mca_offset": np.int32,
           "mca_src_idx": np.int32, "mca_dst_idx": np.int32}


class StructureFileError(ValueError):
    """A structure file is not in this format, or its contents are damaged."""


@dataclass
class Stored:
    """The contents of a structure file (CPU tensors)."""
    n: int
    alpha: int
    saved_by: str
    delta_crow: torch.Tensor
    delta_col: torch.Tensor
    delta_sign: torch.Tensor
    mca_offset: torch.Tensor
    mca_src_idx: torch.Tensor
    mca_dst_idx: torch.Tensor
    plan: structure.Plan
    build_stats: dict = field(default_factory=dict)
    gather_order: torch.Tensor = None

    def delta_values(self):
        """The +-1 values of Delta as FP32."""
        return self.delta_sign.to(torch.float32)


def _crc(a):
    return zlib.crc32(np.ascontiguousarray(a).view(np.uint8)) & 0xFFFFFFFF


def _np(t, dtype):
    return np.ascontiguousarray(torch.as_tensor(t).detach().cpu().numpy()).astype(dtype, copy=False)


def save(path, *, n, alpha, delta_crow, delta_col, delta_sign, mca_offset, mca_src_idx, mca_dst_idx, plan,
         build_stats=None, saved_by="", gather_order=None):
    """Writes a structure file.  ``plan`` is a :class:`cbm.structure.Plan`
    of the same matrix; ``delta_sign`` holds the signs of Delta's entries;
    ``gather_order`` (optional) the short rows' gather order."""
    arrays = {"delta_crow": _np(delta_crow, np.int32), "delta_col": _np(delta_col, np.int32),
              "delta_sign": _np(delta_sign, np.int8), "mca_offset": _np(mca_offset, np.int32),
              "mca_src_idx": _np(mca_src_idx, np.int32), "mca_dst_idx": _np(mca_dst_idx, np.int32)}
    for name in structure.PLAN_NAMES:
        arrays["plan/" + name] = _np(plan.arrays[name], np.int32)
    if gather_order is not None:
        arrays["gather_order"] = _np(gather_order, np.int32)
    n = int(n)
    if arrays["delta_crow"].shape != (n + 1,) or arrays["delta_col"].shape != arrays["delta_sign"].shape:
        raise ValueError("Delta arrays do not describe an n x n CSR matrix")
    if not np.all(np.abs(arrays["delta_sign"]) == 1):
        raise ValueError("delta_sign must hold +1 / -1 only")
    meta = {"format": FORMAT, "version": VERSION, "n": n, "alpha": int(alpha), "saved_by": str(saved_by),
            "build_stats": {k: int(v) for k, v in (build_stats or {}).items()},
            "plan_stats": {k: int(v) for k, v in plan.stats.items()}, "plan_depth": int(plan.depth),
            "crc32": {k: _crc(v) for k, v in arrays.items()}}
    arrays["meta"] = np.frombuffer(json.dumps(meta, sort_keys=True).encode(), dtype=np.uint8)
    with open(path, "wb") as fh:
        np.savez(fh, **arrays)


def load(path):
    """Reads and checks a structure file; returns :class:`Stored`."""
    try:
        z = np.load(path, allow_pickle=False)
    except (OSError, ValueError) as exc:
        raise StructureFileError(f"{path}: not a structure file ({exc})") from exc
    with z:
        if "meta" not in z.files:
            raise StructureFileError(f"{path}: no meta record")
        try:
            meta = json.loads(bytes(z["meta"]).decode())
        except (ValueError, UnicodeDecodeError) as exc:
            raise StructureFileError(f"{path}: unreadable meta record") from exc
        if meta.get("format") != FORMAT:
            raise StructureFileError(f"{path}: format {meta.get('format')!r}, expected {FORMAT!r}")
        version = meta.get("version")
        if version not in (1, 2, 3, 4, VERSION):
            raise StructureFileError(f"{path}: version {version}, this library reads 1 to {VERSION}")
        absent = ((set(_NEW_IN_2) if version < 2 else set()) | (set(_NEW_IN_3) if version < 3 else set())
                  | (set(_NEW_IN_4) if version < 4 else set()))
        names = list(_CORE) + ["plan/" + k for k in structure.PLAN_NAMES if k not in absent]
        if version >= 5 and "gather_order" in meta["crc32"]:
            names.append("gather_order")
        arrays = {"plan/" + k: np.zeros(0, np.int32) for k in absent}
        for name in names:
            if name not in z.files:
                raise StructureFileError(f"{path}: array {name} is missing")
            a = z[name]
            want = _DTYPES.get(name, np.int32)
            if a.dtype != want or a.ndim != 1:
                raise StructureFileError(f"{path}: array {name} is {a.dtype}[{a.ndim}-d], expected 1-d {want.__name__}")
            if _crc(a) != meta["crc32"].get(name):
                raise StructureFileError(f"{path}: array {name} fails its checksum")
            arrays[name] = a
    n = int(meta["n"])
    nnz = arrays["delta_col"].shape[0]
    e = arrays["mca_src_idx"].shape[0]
    if (arrays["delta_crow"].shape[0] != n + 1 or arrays["delta_sign"].shape[0] != nnz
            or int(arrays["delta_crow"][-1]) != nnz or arrays["mca_dst_idx"].shape[0] != e
            or arrays["plan/lvl_node"].shape[0] != n or arrays["plan/fwd_col"].shape[0] != nnz
            or arrays["plan/bwd_row"].shape[0] != nnz or arrays["plan/ch_node"].shape[0] != e):
        raise StructureFileError(f"{path}: array sizes disagree with n = {n}")
    if "gather_order" in arrays:
        order = arrays["gather_order"]
        lens = np.diff(arrays["plan/fwd_ptr"])
        short = np.nonzero(lens <= structure.LONG_ROW)[0]
        if order.shape[0] != short.shape[0] or not np.array_equal(np.sort(order), short):
            raise StructureFileError(f"{path}: the gather order is not an order of the short rows")
    t = {k: torch.from_numpy(v.copy()) for k, v in arrays.items()}
    stats = dict(meta["plan_stats"])
    for k in _NEW_STATS_IN_2:
        stats.setdefault(k, 0)
    for k in _NEW_STATS_IN_3:
        stats.setdefault(k, int(meta["plan_depth"]))
    for k in _NEW_STATS_IN_4:
        stats.setdefault(k, 0)
    plan = structure.Plan(n=n, depth=int(meta["plan_depth"]),
                          arrays={k: t["plan/" + k] for k in structure.PLAN_NAMES}, stats=stats)
    return Stored(n=n, alpha=int(meta["alpha"]), saved_by=meta.get("saved_by", ""), delta_crow=t["delta_crow"],
                  delta_col=t["delta_col"], delta_sign=t["delta_sign"], mca_offset=t["mca_offset"],
                  mca_src_idx=t["mca_src_idx"], mca_dst_idx=t["mca_dst_idx"], plan=plan,
                  build_stats=dict(meta["build_stats"]), gather_order=t.get("gather_order"))
