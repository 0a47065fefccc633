"""Writes lut_literals.json: the literal constant tables of the reference's
src/spatial/lut.rs (OOB_OCTANT, BITMAP_MASK_FOR_OCTANT_LUT, BITMAP_INDEX_LUT,
OCTANT_STEP_RESULT_LUT, RAY_TO_NODE_OCCUPANCY_BITMASK_LUT) as plain numbers.

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_lut_fixture.py
The JSON is the fixture; this script only re-derives it.
"""
import json
import re
import sys
from pathlib import Path

SRC = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/spatial/lut.rs")
OUT = Path(__file__).resolve().parent / "lut_literals.json"


def table(text: str, name: str) -> list:
    m = re.search(rf"pub\(crate\) const {name}: [^=]*=\s*(\[.*?\]);\s*\n", text, re.S)
    if not m:
        raise SystemExit(f"table {name} not found")
    body = re.sub(r"//[^\n]*", "", m.group(1))
    body = re.sub(r",(\s*\])", r"\1", body)  # trailing commas
    return json.loads(body)


def main() -> None:
    text = SRC.read_text()
    oob = int(re.search(r"pub\(crate\) const OOB_OCTANT: u8 = (\d+);", text).group(1))
    out = {
        "source": "src/spatial/lut.rs",
        "OOB_OCTANT": oob,
        "BITMAP_MASK_FOR_OCTANT_LUT": [int(v, 16) for v in re.findall(
            r"0x[0-9A-Fa-f]+", re.search(r"BITMAP_MASK_FOR_OCTANT_LUT: \[u64; 8\] = \[(.*?)\];",
                                         text, re.S).group(1))],
        "BITMAP_INDEX_LUT": table(text, "BITMAP_INDEX_LUT"),
        "OCTANT_STEP_RESULT_LUT": table(text, "OCTANT_STEP_RESULT_LUT"),
        "RAY_TO_NODE_OCCUPANCY_BITMASK_LUT": table(text, "RAY_TO_NODE_OCCUPANCY_BITMASK_LUT"),
    }
    OUT.write_text(json.dumps(out))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
