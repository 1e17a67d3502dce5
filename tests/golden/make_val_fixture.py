"""Writes val_tgv_dissipation.json from the reference's own validation data:
the Taylor-Green dissipation series of val/*.txt (Re=1600, nu=6.25e-4,
dt=2e-4, one sample every 1000 steps, val/comparison_diss.m:5,10), every row
of every series, as (t, dissipation) pairs.  Run in a checkout that has the
reference at /root/reference; the JSON is the committed fixture."""
import glob
import json
import os

VAL = "/root/reference/val"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "val_tgv_dissipation.json")


def main():
    series = {}
    for path in sorted(glob.glob(os.path.join(VAL, "*_ab.txt")) + glob.glob(os.path.join(VAL, "*_eulero.txt"))):
        rows = []
        for line in open(path):
            parts = line.replace(",", " ").split()
            if len(parts) >= 2:
                rows.append([float(parts[0]), float(parts[1])])
        series[os.path.basename(path)[:-4]] = rows
    doc = {"source": "reference val/*.txt: TGV Re=1600 dissipation vs time, nu=6.25e-4, dt=2e-4, samples every "
                     "1000 steps (val/comparison_diss.m:5,10); every row of each file",
           "series": series}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=0)


if __name__ == "__main__":
    main()
