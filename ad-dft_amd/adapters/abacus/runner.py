"""Reading ABACUS's running_scf.log (reference `adapters/abacus/runner.py:
34-55`): the final total energy and the force and stress tables, to compare
a case against a published ABACUS run.  Running ABACUS itself is not part of
this package."""

from __future__ import annotations

import re
from pathlib import Path

import numpy as np

from addft.runtime.units import HA_TO_EV


def _table(lines, header):
    """Rows of three trailing numbers after the line holding `header`, up to
    the first blank line that follows a row or the next TOTAL- header."""
    for i, line in enumerate(lines):
        if header not in line:
            continue
        rows = []
        for s in (x.strip() for x in lines[i + 1:]):
            if not s:
                if rows:
                    break
                continue
            if "TOTAL-" in s:
                break
            try:
                rows.append([float(x) for x in s.split()[-3:]])
            except ValueError:
                continue  # separators and column titles
        return np.array(rows, dtype=np.float64).reshape(-1, 3)
    raise ValueError(f"no {header!r} table in the log")


def parse_running_scf(path) -> dict:
    text = Path(path).read_text(errors="ignore")
    m = re.search(r"FINAL_ETOT_IS\s+([-+0-9.eE]+)", text)
    if m is None:
        raise ValueError(f"{path}: no FINAL_ETOT_IS line")
    lines = text.splitlines()
    e_ev = float(m.group(1))
    return dict(etot_ev=e_ev, etot_ha=e_ev / HA_TO_EV,
                forces=_table(lines, "TOTAL-FORCE (eV/Angstrom)"),
                stress=_table(lines, "TOTAL-STRESS (kbar)")[:3])
