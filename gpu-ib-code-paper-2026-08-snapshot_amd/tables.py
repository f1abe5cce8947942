"""Gather tables of the incidence operators (host setup; built in C++ by ``pyib_tables``).

Each device kernel is a gather over one of these tables, so every output
element is written by exactly one thread (no atomics) and visits its inputs in
the same order as the reference's ``bincount`` over the flat link list
(``ib_core_np.py:129-157``).

* curl  (face <- edge): the face's links in link order (4 per face, then its
  ``link_fe_insert`` links), signed edge index.
* rot   (edge <- face): the edge's links in link order, signed face index.
* cf    (cell <- face): first the faces whose first owner is the cell, then
  those whose second owner it is, each in face order (``bincount(c1) -
  bincount(c2)``); ``other`` is the opposite owner or, for boundary face k,
  the ghost index ``ncell + k`` (``ib_driver_np.py:112-118``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N


@dataclass
class Tables:
    ncell: int
    nface: int
    nedge: int
    nghost: int
    curl_ptr: np.ndarray
    curl_idx: np.ndarray
    rot_ptr: np.ndarray
    rot_idx: np.ndarray
    cf_ptr: np.ndarray
    cf_mid: np.ndarray
    cf_face: np.ndarray
    cf_other: np.ndarray
    f2c: np.ndarray
    bf: np.ndarray          # boundary faces in face order (ghost k <-> bf[k])


def build_tables(geo) -> Tables:
    """The tables by stable counting sorts in C++ (``pyib_tables``, host only)."""
    nface, nedge, ncell = geo.nface, geo.nedge, geo.ncell
    if max(nface, nedge, ncell) >= 2 ** 31 - 1:
        raise ValueError("mesh too large for int32 indices")
    lf_face = np.ascontiguousarray(geo.lf_face, np.int64)
    lf_edge = np.ascontiguousarray(geo.lf_edge, np.int64)
    sgn = np.ascontiguousarray(geo.fe_sign_link, np.float64)
    F2C = np.ascontiguousarray(geo.F2C, np.int64)
    nlink = lf_face.shape[0]
    ncf = nface + int(np.count_nonzero(geo.m1))
    if max(nlink, ncf) >= 2 ** 31 - 1:
        raise ValueError("table too large for int32 offsets")
    curl_ptr, rot_ptr = np.empty(nface + 1, np.int32), np.empty(nedge + 1, np.int32)
    curl_idx, rot_idx = np.empty(nlink, np.int32), np.empty(nlink, np.int32)
    cf_ptr, cf_mid = np.empty(ncell + 1, np.int32), np.empty(ncell, np.int32)
    cf_face, cf_other = np.empty(ncf, np.int32), np.empty(ncf, np.int32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    N.call("pyib_tables", C.c_int64(nface), C.c_int64(nedge), C.c_int64(ncell), C.c_int64(nlink), p(lf_face),
           p(lf_edge), p(sgn), p(F2C), p(curl_ptr), p(curl_idx), p(rot_ptr), p(rot_idx), p(cf_ptr), p(cf_mid),
           p(cf_face), p(cf_other))
    bf = np.flatnonzero(~np.asarray(geo.m1))
    return Tables(ncell, nface, nedge, int(bf.shape[0]), curl_ptr, curl_idx, rot_ptr, rot_idx, cf_ptr, cf_mid,
                  cf_face, cf_other, F2C.astype(np.int32), bf)


def ghost_kinds(geo, bc_face, bf, cavity, lid_mask=None, zf_mask=None):
    """Ghost rule per boundary face (``ib_driver_np.py:275-291``):
    0 inlet (family 1, zero tangential velocity), 1 zero gradient with the
    normal flux imposed, 2 no-slip wall, 3 moving lid, 4 spanwise copy."""
    if cavity:
        kind = np.full(bf.shape[0], 2, np.int32)
        if lid_mask is not None:
            kind[lid_mask] = 3
        if zf_mask is not None:
            kind[zf_mask] = 4
        return kind
    fam = np.asarray(bc_face)[bf]
    return np.where(fam == 1, 0, 1).astype(np.int32)
