"""Per-step metrics in the reference's CSV format.

RigidBodySystem.exportDataToFile (src/mergingBodies3D/RigidBodySystem.java:
495-542) writes, while "save CSV" is on, one line per step to
<sceneName>.csv (<sceneName>_merged.csv with merging on): a header on the
first call, then the body count, the contact count and the phase timings in
seconds, each line introduced by a newline and a space.  Numbers are printed
as java.lang.Double.toString prints them, so scripts written for the
reference's files (scenes3D/csv/) read these unchanged.
"""
from __future__ import annotations

import math
from decimal import Decimal

COLUMNS = ("#bodies", "#contacts", "detection", "warmstart", "LCPSolve", "updateCollections", "contactOrdering",
           "singleItPGS", "merging", "mergingBuild", "unmerging", "unmergingBuild", "computeTime")


def java_double_str(x: float) -> str:
    """Double.toString: the shortest decimal that reads back as x, in plain
    notation with at least one fractional digit for 1e-3 <= |x| < 1e7 and in
    computerized scientific notation (d.dddE-n) otherwise."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    t = Decimal(repr(abs(x))).as_tuple()
    digits = "".join(map(str, t.digits)).rstrip("0") or "0"
    e = t.exponent + len(t.digits) - 1          # power of ten of the leading digit
    if 1e-3 <= abs(x) < 1e7:
        if e >= 0:
            ip = digits[:e + 1].ljust(e + 1, "0")
            fp = digits[e + 1:] or "0"
        else:
            ip, fp = "0", "0" * (-e - 1) + digits
        return f"{sign}{ip}.{fp}"
    return f"{sign}{digits[0]}.{digits[1:] or '0'}E{e}"


def java_str(v) -> str:
    """PrintStream.print of an int or a double."""
    if isinstance(v, (int,)) and not isinstance(v, bool):
        return str(v)
    return java_double_str(v)


class CSVExporter:
    """The stream of exportDataToFile: open() writes the header, row() one
    step, close() ends the file."""

    def __init__(self, path: str):
        self.path = path
        self.stream = open(path, "w", newline="")
        self.stream.write(", ".join(COLUMNS) + "\n ")

    def row(self, values) -> None:
        if len(values) != len(COLUMNS):
            raise ValueError(f"expected {len(COLUMNS)} values, got {len(values)}")
        self.stream.write(", ".join(java_str(v) for v in values) + "\n ")

    def close(self) -> None:
        if self.stream is not None:
            self.stream.close()
            self.stream = None


def read_csv(path: str):
    """Reads a file in this format (the reference's or ours) into a list of
    dicts keyed by column name."""
    with open(path) as f:
        text = f.read()
    lines = [ln.strip() for ln in text.split("\n")]
    header = [h.strip() for h in lines[0].split(",")]
    rows = []
    for ln in lines[1:]:
        if not ln:
            continue
        vals = [v.strip() for v in ln.split(",")]
        rows.append({h: (int(v) if h.startswith("#") else float(v)) for h, v in zip(header, vals)})
    return rows
