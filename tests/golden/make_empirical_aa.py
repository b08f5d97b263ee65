"""Write ``empirical_aa.json``: the LG and WAG tables of the reference, as data.

Reads ``torchtree/evolution/substitution_model/amino_acid.py`` under the reference checkout as
text (nothing is imported or run) and extracts the four numeric tables: LG and WAG stationary
frequencies (20 values, alphabetical one-letter order ACDEFGHIKLMNPQRSTVWY, the order of
``AminoAcidDataType``) and exchangeabilities (190 values, upper triangle row by row). The
published sources are Le & Gascuel (2008) for LG and Whelan & Goldman (2001) for WAG.

    python tests/golden/make_empirical_aa.py [/root/reference]
"""
import json
import re
import sys
from pathlib import Path


def tables(text):
    out = {}
    for cls in ("LG", "WAG"):
        body = text.split(f"class {cls}(")[1].split("\nclass ")[0]
        lists = re.findall(r"torch\.tensor\(\s*\[([^\]]*)\]", body)
        freqs, rates = ([float(x) for x in re.findall(r"[0-9]*\.[0-9]+", s)] for s in lists[:2])
        out[cls] = {"frequencies": freqs, "rates": rates}
    return out


if __name__ == "__main__":
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    src = ref / "torchtree" / "evolution" / "substitution_model" / "amino_acid.py"
    data = tables(src.read_text())
    for name, t in data.items():
        assert len(t["frequencies"]) == 20 and len(t["rates"]) == 190, name
    out = Path(__file__).with_name("empirical_aa.json")
    out.write_text(json.dumps(data, indent=1) + "\n")
    print(f"wrote {out}")
