"""``pyib-run`` for the HIP path (mirrors ``src/pyib/solver/cfdagent_ib.py:25-78``).

    python -m pyib_hip.cli CASE.json --outdir DIR [--resume DIR/checkpoint.npz]
        [--device hip] [--precision float64|mixed] [--steps N] [--overwrite] [--quiet]

The reference selects NumPy or CuPy with ``--device cpu|cuda``; this runner
adds ``--device hip`` and executes only that device.  Configuration, solver
and invariant errors print ``CFDAgent-IB error: ...`` and exit with status 2.
"""
from __future__ import annotations

import argparse
import copy
import json
import sys
from pathlib import Path


def _read_raw_config(path: Path) -> dict:
    try:
        raw = json.loads(path.read_text(encoding="utf-8"))
    except OSError as exc:
        raise SystemExit(f"cannot read case file {path}: {exc}") from exc
    except json.JSONDecodeError as exc:
        raise SystemExit(f"invalid JSON in {path}: {exc}") from exc
    if not isinstance(raw, dict):
        raise SystemExit("case file must contain a JSON object")
    return raw


def parse_args(argv=None):
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("case", type=Path, help="versioned JSON case file")
    p.add_argument("--outdir", type=Path, required=True, help="run output directory")
    p.add_argument("--resume", type=Path, help="checkpoint.npz to continue")
    p.add_argument("--device", choices=("cpu", "cuda", "hip"), help="override runtime.device")
    p.add_argument("--precision", choices=("float64", "float32", "mixed"), help="override runtime.precision")
    p.add_argument("--steps", type=int, help="override the target total step count")
    p.add_argument("--overwrite", action="store_true", help="replace known run artifacts")
    p.add_argument("--quiet", action="store_true")
    return p.parse_args(argv)


def main(argv=None) -> int:
    args = parse_args(argv)
    raw = copy.deepcopy(_read_raw_config(args.case))
    rt = raw.setdefault("runtime", {})
    if args.device is not None:
        rt["device"] = args.device
    if args.precision is not None:
        rt["precision"] = args.precision
    if args.steps is not None:
        if args.steps <= 0:
            raise SystemExit("--steps must be positive")
        raw.setdefault("time", {})["steps"] = args.steps
    from .case import CaseConfigError, run_case, validate_config
    try:
        config = validate_config(raw)
        run_case(config, args.outdir, resume=args.resume, overwrite=args.overwrite, quiet=args.quiet)
    except (CaseConfigError, FileExistsError, RuntimeError, FloatingPointError) as exc:
        print(f"CFDAgent-IB error: {exc}", file=sys.stderr)
        return 2
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
