"""Writes java_double_strings.txt: numbers exactly as the reference's Java
code printed them (Double.toString through PrintStream.print), taken from
the metric files the reference ships (scenes3D/csv/**/*.csv).  Every
distinct token with a decimal point or an exponent is collected, sorted, and
every 200th kept, plus every token in scientific notation with an exponent
of -10 or below and every 100th of those with more than 17 significant
characters (the values printed with a non-terminating tail)."""
import glob
import os
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
tokens = set()
for f in glob.glob(os.path.join(ref, "scenes3D", "csv", "**", "*.csv"), recursive=True):
    with open(f) as fh:
        for line in fh:
            if line.startswith("#"):
                continue
            for tok in line.strip().split(", "):
                if tok and ("." in tok or "E" in tok):
                    tokens.add(tok)
ordered = sorted(tokens)
keep = set(ordered[::200])
keep |= {t for t in ordered if "E-1" in t and len(t.split("E")[1]) >= 3}
longs = [t for t in ordered if len(t.replace("-", "").replace(".", "").split("E")[0]) > 17]
keep |= set(longs[::100])
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "java_double_strings.txt")
with open(out, "w") as fh:
    for t in sorted(keep):
        fh.write(t + "\n")
print(len(tokens), "distinct tokens,", len(keep), "kept")
