"""Writes dtm1_example.json: the reference's example inputs (CalculationExample: settings.cfg,
objects, z_sig_2d, xyzVectorB, frequencies) and, for the example's frequencies, the rows of its
published result table (CalculationResults/625_500_reg_R_Z.txt: |Z| in Ohm and apparent
resistivity in Ohm*m per receiver and period). Data only; run here, where the reference is
mounted:  python tests/golden/make_dtm1_fixture.py"""
import json
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "dtm1_example.json"
sys.path.insert(0, str(Path(__file__).resolve().parents[2]))

from mtcalc_with_dfp_cocr_amd import model as M  # noqa: E402


def main():
    ex = REF / "CalculationExample"
    s = M.read_settings(ex / "settings.cfg")
    tok = (ex / "frequencies").read_text().split()
    freqs = [float(v) for v in tok[1:1 + int(tok[0])]]
    rows = []
    lines = (REF / "CalculationResults" / "625_500_reg_R_Z.txt").read_text().splitlines()
    cols = ["northing", "easting", "period", "Zxx", "Zxy", "Zyx", "Zyy", "Rxx", "Rxy", "Ryx", "Ryy"]
    for ln in lines[1:]:
        v = [float(t) for t in ln.split()]
        if len(v) != len(cols):
            continue
        if any(abs(v[2] * f - 1.0) < 1e-4 for f in freqs):
            rows.append(v)
    doc = {
        "source": "reference CalculationExample/* and CalculationResults/625_500_reg_R_Z.txt",
        "settings": [s.step_xy, s.step_z, s.sparse_xy, s.sparse_z, s.gap_xy, s.gap_z, s.far_xy,
                     s.far_z],
        "objects": M.read_objects(ex / "objects").tolist(),
        "layers": M.read_layers(ex / "z_sig_2d").tolist(),
        "receivers": M.read_receivers(ex / "xyzVectorB").tolist(),
        "frequencies": freqs,
        "published_columns": cols,
        "published": rows,
    }
    OUT.write_text(json.dumps(doc, indent=1) + "\n")
    print(f"{OUT}: {len(rows)} published rows for {len(freqs)} frequencies")


if __name__ == "__main__":
    main()
