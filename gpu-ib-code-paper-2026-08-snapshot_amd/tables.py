"""Gather tables of the incidence operators (host setup, NumPy).

Each device kernel is a gather over one of these tables, so every output
element is written by exactly one thread (no atomics) and visits its inputs in
the same order as the reference's ``bincount`` over the flat link list
(``ib_core_np.py:174-202``).

* curl  (face <- edge): the face's links in link order (4 per face, then its
  ``link_fe_insert`` links), signed edge index.
* rot   (edge <- face): the edge's links in link order, signed face index.
* cf    (cell <- face): first the faces whose first owner is the cell, then
  those whose second owner it is, each in face order (``bincount(c1) -
  bincount(c2)``); ``other`` is the opposite owner or, for boundary face k,
  the ghost index ``ncell + k`` (``ib_driver_np.py:112-118``).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def signed(idx, sign):
    """Encode a +-1 sign into an index: v >= 0 -> +x[v], v < 0 -> -x[-v-1]."""
    idx = np.asarray(idx, np.int64)
    return np.where(np.asarray(sign) > 0, idx, -idx - 1).astype(np.int32)


def _ptr(rows, n):
    counts = np.bincount(rows, minlength=n)
    ptr = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=ptr[1:])
    if ptr[-1] >= 2 ** 31:
        raise ValueError("table too large for int32 offsets")
    return ptr.astype(np.int32)


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
    nface, nedge, ncell = geo.nface, geo.nedge, geo.ncell
    if max(nface, nedge, ncell) >= 2 ** 31 - 1:
        raise ValueError("mesh too large for int32 indices")
    lf_face, lf_edge, sgn = geo.lf_face, geo.lf_edge, geo.fe_sign_link
    o = np.argsort(lf_face, kind="stable")
    curl_ptr = _ptr(lf_face, nface)
    curl_idx = signed(lf_edge[o], sgn[o])
    o = np.argsort(lf_edge, kind="stable")
    rot_ptr = _ptr(lf_edge, nedge)
    rot_idx = signed(lf_face[o], sgn[o])
    del o
    F2C = geo.F2C
    m1 = geo.m1
    bf = np.flatnonzero(~m1)
    ghost = np.full(nface, -1, np.int64)
    ghost[bf] = ncell + np.arange(bf.shape[0])
    f = np.arange(nface)
    fi = f[m1]
    cell = np.concatenate([F2C[:, 0], F2C[m1, 1]])
    part = np.concatenate([np.zeros(nface, np.int8), np.ones(fi.shape[0], np.int8)])
    face = np.concatenate([f, fi])
    other = np.concatenate([np.where(m1, F2C[:, 1], ghost), F2C[m1, 0]])
    order = np.lexsort((face, part, cell))
    cell, part, face, other = cell[order], part[order], face[order], other[order]
    cf_ptr = _ptr(cell, ncell)
    first = np.bincount(cell[part == 0], minlength=ncell)
    cf_mid = (cf_ptr[:-1].astype(np.int64) + first).astype(np.int32)
    return Tables(ncell, nface, nedge, int(bf.shape[0]), curl_ptr, curl_idx, rot_ptr, rot_idx, cf_ptr, cf_mid,
                  face.astype(np.int32), other.astype(np.int32), F2C.astype(np.int32), bf)


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
