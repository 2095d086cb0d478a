"""Write the EIcore B-H curve samples of the reference's inductor example as
a JSON fixture.

Reads examples/onelab_inductor/BH.pro as text and extracts the number lists
of MatEIcore_b() and MatEIcore_h() — the samples the reference's ONELAB
interface feeds to InterpolatedNonlinearBHCurve<64> when the nonlinear
material is enabled (onelab_interface/magnetostatics_onelab.hpp:556-585,
4549-4558).  The fixture holds only the two sample lists.

Usage: python tests/golden/extract_bh_curve.py [reference_root]
"""
import json
import re
import sys
from pathlib import Path


def samples(text, name):
    m = re.search(re.escape(name) + r"\(\)\s*=\s*\{([^}]*)\}", text)
    if not m:
        raise SystemExit(f"{name}() not found")
    return [float(v) for v in re.split(r"[\s,]+", m.group(1).strip()) if v]


def main():
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    text = (root / "examples" / "onelab_inductor" / "BH.pro").read_text()
    b, h = samples(text, "MatEIcore_b"), samples(text, "MatEIcore_h")
    assert len(b) == len(h) >= 2
    out = Path(__file__).with_name("eicore_bh_curve.json")
    out.write_text(json.dumps({"source": "examples/onelab_inductor/BH.pro (MatEIcore_b, MatEIcore_h)",
                               "b": b, "h": h}, indent=1) + "\n")
    print(f"wrote {out} ({len(b)} samples)")


if __name__ == "__main__":
    main()
