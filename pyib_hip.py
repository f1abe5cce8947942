"""Import shim: loads the package directory ``gpu-ib-code-paper-2026-08-snapshot_amd/``
(whose name is not a Python identifier) under the importable name ``pyib_hip``."""
import importlib.util as _ilu
import os as _os
import sys as _sys

_DIR = _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "gpu-ib-code-paper-2026-08-snapshot_amd")
_spec = _ilu.spec_from_file_location("pyib_hip", _os.path.join(_DIR, "__init__.py"),
                                     submodule_search_locations=[_DIR])
_mod = _ilu.module_from_spec(_spec)
_sys.modules["pyib_hip"] = _mod
_spec.loader.exec_module(_mod)
