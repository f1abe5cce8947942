"""Device side of the structured layout (``layout.StructuredLayout``).

``HipGrid`` uploads the renumbering, the regular-row flags and the composite
rows of one mesh, and exposes the elliptic operator and the PCG pieces of
``Driver.helm_solve`` (``ib_driver_np.py:311-410``) on the internal numbering.
The driver moves the right-hand side and the warm start into this numbering
once per solve and the solution back, so every PCG iteration runs on dense,
coalesced edge and cell blocks.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as N
from .core import dev_stream, ptr
from .layout import (MAX_LEVELS, build_layout, composite_rows, dir_mask_internal, edge_classes, g1v_flags,
                     row_volumes, to_sell)


G3_SIGMA = 64


def check_level_blocks(levels):
    """The tiled kernels address each level's cell block (3 values per cell) and edge blocks with
    32-bit offsets from the block's base: refuse a level whose float64 block would reach 4 GiB."""
    for i, lv in enumerate(levels):
        nx, ny, nz = int(lv.nx), int(lv.ny), int(lv.nz)
        largest = max(3 * nx * ny * nz, nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1), (nx + 1) * (ny + 1) * nz)
        if 8 * largest >= 2 ** 32:
            raise ValueError(f"level {i} ({nx}x{ny}x{nz} cells) is too large for one GPU level block: "
                             f"{8 * largest} bytes >= 4 GiB")


class HipGrid:
    def __init__(self, core, precisions=("float64",), margin=1, layout=None):
        self.core = core
        dev = self.device = core.device
        self.lay = lay = layout if layout is not None else build_layout(core.mesh, core.geo, core.tables,
                                                                        margin=margin)
        self.ncell, self.nedge = lay.ncell, lay.nedge
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
        self._cell_reg = t(lay.reg_cell, torch.uint8)
        self._cell_reg_g1 = t(lay.reg_g1, torch.uint8)
        self._cell_g1v = t(g1v_flags(lay), torch.uint8)
        self._edge_reg = t(lay.reg_edge, torch.uint8)
        self._dir = t(np.ones(lay.nedge, np.uint8), torch.uint8)
        # reference -> internal gather: idx over internal slots; internal -> reference: idx over ref edges
        self._to_int_idx = t(lay.edge_inv, torch.int64)
        self._to_int_sig = t(lay.sigma_int, torch.int8)
        self._to_ref_idx = t(lay.edge_map, torch.int64)
        self._to_ref_sig = t(lay.edge_sigma, torch.int8)
        self._t = t
        self._comp = {}
        for name in ("g1", "vis", "g3"):
            # the G3 composite's rows gather and update edge vectors at the row: slices of 64
            # consecutive rows (no length sort across slices) share their cache lines
            self._comp[name] = self._sell(getattr(lay, name), G3_SIGMA if name == "g3" else None)
            if name == "vis":
                self._comp[name]["rw"] = t(row_volumes(lay, core.geo, self._comp[name]["rows_host"]), torch.float64)
        self._coef = {}
        self._struct = {}
        self.work = {}
        for prec in precisions:
            self.add_precision(prec)

    def _sell(self, comp, sigma=None):
        t = self._t
        c = to_sell(comp) if sigma is None else to_sell(comp, sigma=sigma)
        return {"rows": t(c.rows, torch.int32), "width": t(c.width, torch.int32), "sptr": t(c.sptr, torch.int64),
                "col": t(c.col, torch.int32), "n": int(c.width.shape[0]), "coef64": c.coef, "rows_host": c.rows}

    def add_precision(self, prec):
        if prec in self._struct:
            return
        dt = torch.float64 if prec == "float64" else torch.float32
        self._coef[prec] = {n: torch.from_numpy(np.ascontiguousarray(c["coef64"])).to(self.device, dt)
                            for n, c in self._comp.items()}
        w = {"face": torch.empty(1, dtype=dt, device=self.device),
             "cell_a": torch.zeros(3 * self.ncell, dtype=dt, device=self.device),
             "cell_b": torch.zeros(3 * self.ncell, dtype=dt, device=self.device),
             "partials": torch.zeros(8 * N.MAX_PARTIALS, dtype=torch.float64, device=self.device)}
        w["struct"] = N.Work(ptr(w["face"]), ptr(w["cell_a"]), ptr(w["cell_b"]), ptr(w["partials"]))
        self.work[prec] = w
        self._struct[prec] = self._make(prec)

    def _make(self, prec):
        g = N.Grid()
        g.dtype = N.F64 if prec == "float64" else N.F32
        lay = self.lay
        g.nlevels = len(lay.levels)
        if g.nlevels > MAX_LEVELS:
            raise ValueError("too many levels")
        check_level_blocks(lay.levels)
        g.ncell, g.nedge = lay.ncell, lay.nedge
        for i, lv in enumerate(lay.levels):
            L = g.lev[i]
            L.nx, L.ny, L.nz = lv.nx, lv.ny, lv.nz
            L.cell_off = lv.cell_off
            L.e_off[0], L.e_off[1], L.e_off[2] = lv.e_off
            L.h = lv.h
            for a in range(3):
                L.hole_lo[a], L.hole_hi[a] = int(lv.hole_lo[a]), int(lv.hole_hi[a])
        g.cell_reg, g.cell_reg_g1 = ptr(self._cell_reg), ptr(self._cell_reg_g1)
        g.edge_reg, g.dir = ptr(self._edge_reg), ptr(self._dir)
        g.cell_g1v = ptr(self._cell_g1v)
        c, k = self._comp, self._coef[prec]
        for pre, name in (("g1", "g1"), ("v", "vis"), ("g3", "g3")):
            cc = c[name]
            setattr(g, pre + "_nslice", cc["n"])
            setattr(g, pre + "_row", ptr(cc["rows"]))
            setattr(g, pre + "_width", ptr(cc["width"]))
            setattr(g, pre + "_sptr", ptr(cc["sptr"]))
            setattr(g, pre + "_col", ptr(cc["col"]))
            setattr(g, pre + "_coef", ptr(k[name]))
        g.v_rw = ptr(c["vis"]["rw"])
        cls = getattr(self, "_cls", None)
        if cls is not None and prec in cls["prec"]:
            g.edge_cls, g.cls_prec = ptr(self._edge_cls), ptr(cls["prec"][prec])
            g.g3_rprec = ptr(cls["g3_rprec"][prec])
            for i in range(MAX_LEVELS):
                g.prec_level[i], g.prec_level_g3[i] = float(cls["plev"][i]), float(cls["pg3"][i])
        g.g3_mask = None if getattr(self, "_g3_dir_free", False) else ptr(self._dir)
        return g

    def set_dir_mask(self, dir_mask_ref):
        """Dirichlet rows (and empty slots) are never written by the G3 passes: their output is
        zero (w = 0, r unchanged at 0), so the G3 composite keeps only the other rows."""
        m = dir_mask_internal(self.lay, dir_mask_ref)
        self._dir.copy_(torch.from_numpy(m))
        self._dir_host = m
        self._cls = None            # the preconditioner classes follow the mask: set them again
        g3 = self.lay.g3
        self._comp["g3"] = self._sell(composite_rows(g3, np.flatnonzero(m[np.asarray(g3.rows)] == 0)), G3_SIGMA)
        self._g3_dir_free = True
        for prec in self._struct:
            dt = torch.float64 if prec == "float64" else torch.float32
            self._coef[prec]["g3"] = torch.from_numpy(np.ascontiguousarray(self._comp["g3"]["coef64"])).to(
                self.device, dt)
            self._struct[prec] = self._make(prec)

    def set_preconditioner(self, prec_int_host, prec_tensors):
        """Classes of the Jacobi preconditioner (``layout.edge_classes``) for the PCG iteration.
        ``prec_int_host``: the float64 inverse diagonal in internal numbering; ``prec_tensors``:
        {precision: the device copy a pyib_cg of that precision points at}.  The iteration uses the
        classes only with those arrays; with any other it reads its prec as is."""
        dir_host = getattr(self, "_dir_host", None)
        if dir_host is None:
            raise RuntimeError("set_dir_mask comes before set_preconditioner")
        cls, plev, pg3 = edge_classes(self.lay, dir_host, prec_int_host)
        self._edge_cls = self._t(cls, torch.uint8)
        # the G3 composite's rows read prec at their row from this per-lane copy (gathered from the
        # same device arrays, so the values are those of prec[row])
        rows = self._t(np.asarray(self._comp["g3"]["rows_host"], np.int64), torch.int64)
        live = rows >= 0
        rprec = {p: torch.where(live, t.index_select(0, rows.clamp(min=0)), torch.zeros((), dtype=t.dtype,
                                                                                          device=t.device))
                 for p, t in prec_tensors.items()}
        self._cls = {"plev": plev, "pg3": pg3, "prec": dict(prec_tensors), "g3_rprec": rprec}
        for prec in self._struct:
            self._struct[prec] = self._make(prec)

    def struct(self, prec="float64"):
        return C.byref(self._struct[prec])

    def work_struct(self, prec="float64"):
        return C.byref(self.work[prec]["struct"])

    # ------------------------------------------------------- numbering
    def to_internal(self, x_ref, out=None, dtype=torch.float64):
        out = torch.empty(self.nedge, dtype=dtype, device=self.device) if out is None else out
        N.call("pyib_permute", _code(x_ref.dtype), ptr(x_ref), _code(out.dtype), ptr(out), C.c_int64(self.nedge),
               ptr(self._to_int_idx), ptr(self._to_int_sig), dev_stream())
        return out

    def to_reference(self, x_int, out=None, dtype=torch.float64):
        n = self.core.nedge
        out = torch.empty(n, dtype=dtype, device=self.device) if out is None else out
        N.call("pyib_permute", _code(x_int.dtype), ptr(x_int), _code(out.dtype), ptr(out), C.c_int64(n),
               ptr(self._to_ref_idx), ptr(self._to_ref_sig), dev_stream())
        return out

    def internal_from_host(self, x_ref):
        v = np.zeros(self.nedge)
        v[self.lay.edge_map] = self.lay.edge_sigma * np.asarray(x_ref)
        return torch.from_numpy(v).to(self.device)

    def flow_mat(self, p_int, rou, visc, dt, out=None, with_dot=False):
        """w = A p on the internal numbering; with ``with_dot`` also p.A p (float64, device)
        as the PCG forms it from the cell passes."""
        prec = "float64" if p_int.dtype == torch.float64 else "float32"
        out = torch.zeros(self.nedge, dtype=p_int.dtype, device=self.device) if out is None else out
        dot = torch.zeros(1, dtype=torch.float64, device=self.device) if with_dot else None
        N.call("pyib_grid_flow_mat", self.struct(prec), self.work_struct(prec), float(rou), float(visc),
               0.5 * float(dt), ptr(p_int), ptr(out), ptr(dot) if with_dot else C.c_void_p(0), dev_stream())
        return (out, dot) if with_dot else out


def _code(dtype):
    return N.F64 if dtype == torch.float64 else N.F32
