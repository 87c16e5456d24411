"""Writes cie1931_1nm.csv: the CIE 1931 2-degree colour matching functions at 1 nm from 360 to
830 nm, as the reference holds them in Source/Falcor/Utils/Color/Spectra.inl (CIE_lambda, CIE_X,
CIE_Y, CIE_Z; 471 samples). The table is data the reference ships; the CSV is the fixture.
Run in a checkout that has the reference tree: python tests/golden/make_cie1931.py [/root/reference]."""
from __future__ import annotations

import re
import sys
from pathlib import Path


def table(text: str, name: str) -> list[float]:
    m = re.search(r"const float " + name + r"\[kCIESampleCount\] = \{(.*?)\};", text, re.S)
    body = re.sub(r"//[^\n]*", "", m.group(1))
    return [float(v) for v in body.replace("\n", " ").split(",") if v.strip()]


def main(ref: Path) -> Path:
    text = (ref / "Source/Falcor/Utils/Color/Spectra.inl").read_text()
    cols = [table(text, n) for n in ("CIE_lambda", "CIE_X", "CIE_Y", "CIE_Z")]
    assert all(len(c) == 471 for c in cols), [len(c) for c in cols]
    out = Path(__file__).with_name("cie1931_1nm.csv")
    rows = ["lambda_nm,x,y,z"] + [",".join(repr(v) for v in r) for r in zip(*cols)]
    out.write_text("\n".join(rows) + "\n")
    return out


if __name__ == "__main__":
    print(main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")))
