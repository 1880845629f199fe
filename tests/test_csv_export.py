"""The CSV metrics file of RigidBodySystem.exportDataToFile
(RigidBodySystem.java:495-542): numbers printed as Java's Double.toString
prints them (checked against 2269 numbers the reference's own Java code
wrote into its scenes3D/csv files, tests/golden/java_double_strings.txt, and
against the documented notation boundaries), and the file layout (header,
then one line per step, each introduced by a newline and a space)."""
import glob
import math
import os

import pytest

from adaptivemerging_amd.csv_export import COLUMNS, CSVExporter, java_double_str, read_csv

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "java_double_strings.txt")
REF_CSV = "/root/reference/scenes3D/csv"


def test_reproduces_the_reference_printed_numbers():
    with open(GOLDEN) as f:
        tokens = [ln.strip() for ln in f if ln.strip()]
    assert len(tokens) > 2000
    bad = [(t, java_double_str(float(t))) for t in tokens if java_double_str(float(t)) != t]
    assert not bad, bad[:5]


@pytest.mark.parametrize("x,s", [
    (0.0, "0.0"), (-0.0, "-0.0"), (1.0, "1.0"), (123.0, "123.0"), (0.5, "0.5"),
    (1e-3, "0.001"), (9.999999999999998e-4, "9.999999999999998E-4"), (1e-4, "1.0E-4"),
    (9999999.0, "9999999.0"), (1e7, "1.0E7"), (1.25e8, "1.25E8"), (-2.5e-7, "-2.5E-7"),
    (0.1 + 0.2, "0.30000000000000004"), (math.inf, "Infinity"), (-math.inf, "-Infinity"), (math.nan, "NaN"),
])
def test_notation_boundaries(x, s):
    assert java_double_str(x) == s


def test_file_layout_and_reader(tmp_path):
    path = tmp_path / "scene_merged.csv"
    w = CSVExporter(str(path))
    w.row([328, 12, 0.0016224240000000001, 7.12e-7, 0.0, 0.0, 0.0, 0.0, 1.423e-6, 0.0, 1.6821e-5, 0.0, 0.001783946])
    w.row([329, 0] + [0.0] * 11)
    w.close()
    text = path.read_text()
    assert text.startswith(", ".join(COLUMNS) + "\n ")
    assert "\n 328, 12, 0.0016224240000000001, 7.12E-7, 0.0," in text
    rows = read_csv(str(path))
    assert [r["#bodies"] for r in rows] == [328, 329] and rows[0]["warmstart"] == 7.12e-7


@pytest.mark.skipif(not os.path.isdir(REF_CSV), reason="the reference's csv folder is not present")
def test_reads_the_reference_files():
    f = os.path.join(REF_CSV, "conditional_acceptance_revisions", "no_sleeping", "tower25platform_merged.csv")
    rows = read_csv(f)
    assert len(rows) > 1000 and rows[0]["#bodies"] == 328
    assert all(set(r) == set(COLUMNS) for r in rows)
    assert len(glob.glob(os.path.join(REF_CSV, "**", "*.csv"), recursive=True)) >= 10
