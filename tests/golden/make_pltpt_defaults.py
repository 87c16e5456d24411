"""Writes pltpt_defaults.json: the PLTPT pass's dictionary keys and default values, read from the
reference's own source text (run in a checkout that has the reference tree; the JSON is the fixture).

Keys: the `const std::string kX = "name";` table of Source/RenderPasses/PLTPT/PLTPT.cpp; the member
each key sets: its parseDictionary branch (PLTPT.cpp:147-175); the default: that member's initialiser
in PLTPT.h, with the symbolic initialisers resolved from the headers that define them
(SAMPLE_GENERATOR_DEFAULT in Utils/Sampling/SampleGeneratorType.slangh, DebugViewType::none in
RenderPasses/PLTPT/DebugViewType.slangh, EmissiveLightSamplerType::Power in
Rendering/Lights/EmissiveLightSamplerType.slangh).
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path


def main(ref: Path) -> dict:
    src = ref / "Source"
    cpp = (src / "RenderPasses/PLTPT/PLTPT.cpp").read_text()
    hdr = (src / "RenderPasses/PLTPT/PLTPT.h").read_text()
    names = dict(re.findall(r'const std::string (k\w+) = "(\w+)";', cpp))
    setters = dict(re.findall(r"key == (k\w+)\) (m\w+) = value;", cpp))
    inits = dict(re.findall(r"^\s*[\w:<>]+\s+(m\w+)\s*=\s*([^;]+);", hdr, re.M))
    sgt = (src / "Falcor/Utils/Sampling/SampleGeneratorType.slangh").read_text()
    sg = dict(re.findall(r"#define (SAMPLE_GENERATOR_\w+)\s+(\w+)", sgt))
    elst = (src / "Falcor/Rendering/Lights/EmissiveLightSamplerType.slangh").read_text()
    el = dict(re.findall(r"^\s*(\w+)\s*=\s*(\w+),", elst, re.M))
    dvt = (src / "RenderPasses/PLTPT/DebugViewType.slangh").read_text()
    dv = dict(re.findall(r"^\s*(\w+)\s*=\s*(\w+),", dvt, re.M))

    def value(expr: str):
        e = expr.strip().split("//")[0].strip()
        if e.startswith("(uint32_t)"):
            e = e[len("(uint32_t)"):]
        while e in sg:
            e = sg[e]
        if e.startswith("EmissiveLightSamplerType::"):
            e = el[e.split("::")[1]]
        if e.startswith("DebugViewType::"):
            e = dv[e.split("::")[1]]
        if e in ("true", "false"):
            return int(e == "true")
        return float(e.rstrip("f")) if any(c in e for c in ".e") else int(e, 0)

    out = {}
    for kname, member in setters.items():
        out[names[kname]] = value(inits[member])
    return out


if __name__ == "__main__":
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    d = main(ref)
    path = Path(__file__).with_name("pltpt_defaults.json")
    path.write_text(json.dumps(d, indent=1, sort_keys=True) + "\n")
    print(path, len(d), "keys")
