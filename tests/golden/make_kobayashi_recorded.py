"""Writes tests/golden/kobayashi1_recorded.json: the reference's recorded
Kobayashi problem 1 line values (results/line_plot_results.py:20-32, 124-135,
228-239), read as text -- the script is not imported or run.

Kept: the Monte Carlo benchmark rows and the "dd" rows (DOMA's diamond-difference
outputs; the step-characteristics rows come from a scheme the reference's solver
does not have).  Row names are kobayashi_<line>_dd_angle_<np = na>_space_<r + 1>,
r = the deck's mesh refinement (cases/3D_kobayashi_1.xml).

    python tests/golden/make_kobayashi_recorded.py [/root/reference]
"""
import json
import re
import sys
from pathlib import Path

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
text = (ref / "results" / "line_plot_results.py").read_text()
rows = {}
for name, body in re.findall(r"^(kobayashi_1[abc]_(?:bench|dd_angle_\d+_space_\d+))\s*=\s*"
                             r"np\.array\(\[([^\]]*)\]\)", text, flags=re.M):
    rows[name] = [float(v) for v in body.split(",")]
points = [5.0, 15.0, 25.0, 35.0, 45.0, 55.0, 65.0, 75.0, 85.0]
out = {
    "source": "results/line_plot_results.py (reference), recorded DOMA outputs",
    "points": {"1a": [[5.0, y, 5.0] for y in points],
               "1b": [[v, v, v] for v in points],
               "1c": [[x, 55.0, 5.0] for x in points]},
    "origin": "Kobayashi coordinates (x, y, z) of the source corner; the deck's cube spans "
              "[0, 200] cm with the source on [90, 110], so a point maps to 100 + x etc.",
    "rows": rows,
}
dst = Path(__file__).resolve().parent / "kobayashi1_recorded.json"
dst.write_text(json.dumps(out, indent=1) + "\n")
print(f"{len(rows)} rows -> {dst}")
