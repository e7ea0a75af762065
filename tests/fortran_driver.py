"""Helpers for the tests of the config.h-driven Fortran driver
(ls-sph-benchmarks_amd/bin/lssph_<case>, built from tests/fortran_cases/<case>.h)."""
from __future__ import annotations

import csv
import os
import subprocess
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[1]
BIN = REPO / "ls-sph-benchmarks_amd" / "bin"


def binary(case: str) -> Path:
    exe = BIN / f"lssph_{case}"
    if not exe.exists():
        raise FileNotFoundError(f"{exe} is not built (run __graft_entry__.build())")
    return exe


def run(case: str, cwd: Path, env: dict | None = None, timeout: int = 120) -> subprocess.CompletedProcess:
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([str(binary(case))], cwd=cwd, env=e, capture_output=True, text=True, timeout=timeout)


def read_state(path: Path, n: int, has_tem: bool) -> dict:
    """A snapshot in the write_data layout (write_data_mod.f90:43-52)."""
    with open(path, "rb") as f:
        out = {"ptype": np.fromfile(f, "<i4", n)}
        for k in ("x", "y", "u", "v", "rho", "pre"):
            out[k] = np.fromfile(f, "<f8", n)
        if has_tem:
            out["tem"] = np.fromfile(f, "<f8", n)
    return out


def write_state(path: Path, st, has_tem: bool) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "wb") as f:
        np.asarray(st.ptype, "<i4").tofile(f)
        for k in ("x", "y", "u", "v", "rho", "pre"):
            np.asarray(getattr(st, k), "<f8").tofile(f)
        if has_tem:
            np.asarray(st.tem, "<f8").tofile(f)


def read_parameters(path: Path) -> dict:
    """parameters.csv parsed the way the analysis package parses it
    (int, else float with d/D exponents, else string)."""
    out = {}
    with open(path, encoding="utf-8") as f:
        reader = csv.reader(f, delimiter=",")
        next(reader)
        for row in reader:
            key, val = row[0].strip(), row[1].strip()
            try:
                out[key] = int(val)
            except ValueError:
                try:
                    out[key] = float(val.replace("d", "e").replace("D", "E"))
                except ValueError:
                    out[key] = val
    return out
