"""On-disk cache of the host setup (mesh, geometry, gather tables, structured layout).

Building the 41M-edge nested mesh and its geometry takes minutes of host
time, the same order as the reference's recorded ``setup_seconds``.  When the
environment variable ``PYIB_HIP_SETUP_CACHE`` names a directory, the setup
objects are stored there as plain ``.npy`` arrays plus a JSON file (no
pickles) under a key made of the mesh parameters and the sha256 of the
setup sources, and later processes load them instead of rebuilding.
"""
from __future__ import annotations

import dataclasses
import hashlib
import json
import os
from pathlib import Path

import numpy as np

_SOURCES = ("mesh.py", "geometry.py", "tables.py", "layout.py", "setup_cache.py", "csrc/geometry.cpp",
            "csrc/mesh.cpp", "csrc/layout.cpp")


def cache_dir():
    d = os.environ.get("PYIB_HIP_SETUP_CACHE")
    return Path(d) if d else None


def key_for(params: dict, here: Path = None) -> str:
    """Key of a cached object: the mesh parameters and the bytes of every source that shapes it
    (``here`` is the package directory)."""
    here = Path(__file__).resolve().parent if here is None else Path(here)
    h = hashlib.sha256(json.dumps(params, sort_keys=True).encode())
    for name in _SOURCES:
        h.update((here / name).read_bytes())
    return h.hexdigest()[:24]


def _to_json(v):
    if isinstance(v, (tuple, list)):
        return [_to_json(x) for x in v]
    if isinstance(v, (np.integer,)):
        return int(v)
    if isinstance(v, (np.floating,)):
        return float(v)
    return v


def save(obj, path: Path):
    path.mkdir(parents=True, exist_ok=True)
    meta = {}
    items = obj.__dict__ if not dataclasses.is_dataclass(obj) else {f.name: getattr(obj, f.name)
                                                                     for f in dataclasses.fields(obj)}
    for name, v in items.items():
        if isinstance(v, np.ndarray):
            np.save(path / f"{name}.npy", v, allow_pickle=False)
            meta[name] = {"kind": "array"}
        elif dataclasses.is_dataclass(v) and not isinstance(v, type):
            save(v, path / name)
            meta[name] = {"kind": "object", "class": type(v).__name__}
        elif isinstance(v, list) and v and dataclasses.is_dataclass(v[0]):
            meta[name] = {"kind": "objlist", "class": type(v[0]).__name__,
                          "items": [{f.name: _to_json(getattr(x, f.name)) for f in dataclasses.fields(x)} for x in v]}
        else:
            meta[name] = {"kind": "value", "value": _to_json(v)}
    tmp = path / "meta.json.tmp"
    tmp.write_text(json.dumps(meta))
    os.replace(tmp, path / "meta.json")


def load(cls, path: Path, registry: dict):
    meta = json.loads((path / "meta.json").read_text())
    vals = {}
    for name, m in meta.items():
        if m["kind"] == "array":
            vals[name] = np.load(path / f"{name}.npy", allow_pickle=False)
        elif m["kind"] == "object":
            vals[name] = load(registry[m["class"]], path / name, registry)
        elif m["kind"] == "objlist":
            c = registry[m["class"]]
            vals[name] = [c(**{k: (tuple(v) if isinstance(v, list) else v) for k, v in it.items()}) for it in m["items"]]
        else:
            v = m["value"]
            vals[name] = tuple(tuple(x) if isinstance(x, list) else x for x in v) if isinstance(v, list) else v
    if dataclasses.is_dataclass(cls):
        return cls(**vals)
    obj = cls.__new__(cls)
    obj.__dict__.update(vals)
    return obj


def cached(kind, params, build, cls, registry=None):
    """Return build() or its cached copy under PYIB_HIP_SETUP_CACHE."""
    d = cache_dir()
    if d is None:
        return build()
    path = d / f"{kind}-{key_for(params)}"
    if (path / "meta.json").exists():
        return load(cls, path, registry or {})
    obj = build()
    try:
        save(obj, path)
    except OSError as exc:          # a full or read-only cache disk only costs the next rebuild
        import shutil
        import warnings
        shutil.rmtree(path, ignore_errors=True)
        warnings.warn(f"setup cache: {kind} not stored in {d} ({exc})", RuntimeWarning, stacklevel=2)
    return obj
