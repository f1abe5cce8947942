"""Citations of the reference (``file.py:N`` / ``file.py:N-M``) in this repository's sources and
documents point inside the cited file, and the drop-in boundary's key citations start at the
function or class they stand for.  Reads the reference checkout (PYIB_REFERENCE, default
/root/reference) and is skipped where it is absent."""
import os
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path(os.environ.get("PYIB_REFERENCE", "/root/reference"))
pytestmark = pytest.mark.skipif(not (REF / "src" / "pyib").is_dir(), reason="needs the reference checkout")

# this repository's own sources and documents (other tracked files quote the reference on their own)
OURS_DIRS = ("gpu-ib-code-paper-2026-08-snapshot_amd/", "include/", "oracle/", "tests/", "profiles/")
OURS_FILES = {"bench.py", "__graft_entry__.py", "pyib_hip.py", "DESIGN.md", "INTEGRATION.md", "README.md"}
CITE = re.compile(r"\b([A-Za-z_][A-Za-z0-9_]*\.py):(\d+(?:-\d+)?(?:,\s*\d+(?:-\d+)?)*)")

# the boundary's citations: (file, first line) -> the definition that starts there
KEY = {
    ("ib_core_np.py", 23): "class IBCore", ("ib_core_np.py", 24): "def __init__",
    ("ib_core_np.py", 74): "Modify_geom", ("ib_core_np.py", 129): "def rot",
    ("ib_core_np.py", 133): "def curl", ("ib_core_np.py", 136): "def cdiv",
    ("ib_core_np.py", 144): "def integr_c2f", ("ib_core_np.py", 150): "def interp_f2c",
    ("ib_core_np.py", 168): "def flow_mat", ("ib_core_np.py", 177): "def cgrad_cellscalar",
    ("ib_driver_np.py", 78): "def diag_preconditioner", ("ib_driver_np.py", 94): "class Driver",
    ("ib_driver_np.py", 267): "def calc_vel", ("ib_driver_np.py", 293): "def conv_diff",
    ("ib_driver_np.py", 311): "def helm_solve", ("ib_driver_np.py", 412): "def step",
    ("ib_immersed.py", 12): "def delta_h3", ("ib_immersed.py", 22): "class ImmersedBoundary",
    ("ib_immersed.py", 87): "def _interp_stencil", ("ib_immersed.py", 103): "def _spread_stencil",
    ("ib_immersed.py", 113): "def direct_force", ("ib_immersed.py", 122): "def prepare_implicit",
    ("ib_immersed.py", 152): "def implicit_force",
    ("ib_sparse.py", 69): "def build_A", ("ib_kinematics.py", 31): "class Oscillation",
    ("ib_interior.py", 113): "class ClosedSphereLoadCorrection", ("ib_interior.py", 216): "def evaluate",
    ("ib_interior.py", 268): "def advance",
    ("ib_case.py", 262): "def validate_config", ("ib_case.py", 633): "def build_runtime",
    ("ib_case.py", 1129): "def diagnostics", ("ib_case.py", 1413): "def run_case",
    ("ib_case.py", 566): "def rotating_sphere_velocity", ("ib_case.py", 914): "def load_checkpoint",
    ("make_euler_mesh.py", 330): "def make_nested_mesh",
    ("make_euler_mesh.py", 787): "def _finalize_hanging_node_connectivity",
    ("validate_ib_result.py", 1168): "def _validate_history", ("validate_ib_result.py", 1269): "def validate_result",
    ("ib_postprocess.py", 49): "def recover_pressure",
}


def _reference_files():
    out = {}
    for p in REF.rglob("*.py"):
        out.setdefault(p.name, []).append(p)
    ours = {Path(f).name for f in subprocess.check_output(["git", "ls-files"], cwd=ROOT, text=True).split()}
    # a name that is also one of this repository's files is ambiguous in a citation
    return {k: v[0] for k, v in out.items() if len(v) == 1 and k not in ours}


def _citations():
    files = _reference_files()
    tracked = subprocess.check_output(["git", "ls-files"], cwd=ROOT, text=True).split()
    for f in tracked:
        if not (f.startswith(OURS_DIRS) or f in OURS_FILES) or f.endswith((".npz", ".so")):
            continue
        try:
            text = (ROOT / f).read_text(errors="replace")
        except (OSError, UnicodeDecodeError):
            continue
        for m in CITE.finditer(text):
            name = m.group(1)
            if name not in files:
                continue
            for part in re.split(r",\s*", m.group(2)):
                lo, _, hi = part.partition("-")
                yield f, name, files[name], int(lo), int(hi or lo)


def test_every_citation_lies_inside_its_file():
    bad, n = [], 0
    lengths = {}
    for f, name, path, lo, hi in _citations():
        n += 1
        if path not in lengths:
            lengths[path] = len(path.read_text(errors="replace").splitlines())
        if not 1 <= lo <= hi <= lengths[path]:
            bad.append(f"{f}: {name}:{lo}-{hi} (the file has {lengths[path]} lines)")
    assert n > 200
    assert not bad, "\n".join(bad)


def test_boundary_citations_start_at_their_definition():
    lines = {}
    seen, bad = set(), []
    for f, name, path, lo, hi in _citations():
        want = KEY.get((name, lo))
        if want is None:
            continue
        if path not in lines:
            lines[path] = path.read_text(errors="replace").splitlines()
        seen.add((name, lo))
        if want not in lines[path][lo - 1]:
            bad.append(f"{f}: {name}:{lo} is {lines[path][lo - 1].strip()!r}, not {want!r}")
    assert not bad, "\n".join(bad)
    assert len(seen) >= 20
    # the key table itself is right about the reference
    for (name, lo), want in KEY.items():
        p = _reference_files()[name]
        assert want in p.read_text(errors="replace").splitlines()[lo - 1], (name, lo, want)
