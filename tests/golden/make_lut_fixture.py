"""Write tests/golden/lut_table.json: the 65-entry (sin, cos) table of the resampler.

The table is data from Einstein@Home erp_utilities.cpp (GPL-2.0-or-later) that the
reference keeps in include/brp/lut.hpp:4-5.  Its values are not re-derivable from
a rounding rule (for example entry 30 of the sine is 0.195091 while entry 2 is
0.195090), so bit-exact resampling indices need exactly these floats.  This script
parses the numbers out of the reference header (run in the build container, where
the reference is mounted) and stores them as a fixture, so that the tests can pin
both the product's and the oracle's copy of the table without reading the
reference at run time.
"""
import json
import pathlib
import re
import sys

SRC = pathlib.Path("/root/reference/include/brp/lut.hpp")
OUT = pathlib.Path(__file__).with_name("lut_table.json")


def main() -> int:
    text = SRC.read_text()
    table = {}
    for name in ("sinSamples", "cosSamples"):
        body = re.search(name + r"\[\]\s*=\s*\{([^}]*)\}", text).group(1)
        table[name] = [float(v.strip().rstrip("f")) for v in body.split(",")]
        assert len(table[name]) == 65, name
    OUT.write_text(json.dumps({"source": "include/brp/lut.hpp:4-5", **table}, indent=1) + "\n")
    print("wrote", OUT)
    return 0


if __name__ == "__main__":
    sys.exit(main())
