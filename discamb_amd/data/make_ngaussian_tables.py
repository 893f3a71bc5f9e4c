#!/usr/bin/env python3
"""Generate ``ngaussian_tables.json`` and ``../csrc/ngauss_tables.h``: N-Gaussian atomic form
factors f(h) = sum_i a_i exp(-b_i h^2 / 4) + c (h = |h|, 1/Angstrom).

The numbers are published tables: X-ray 5-Gaussian fits of Waasmaier & Kirfel (1995, Acta Cryst.
A51, 416), X-ray 4-Gaussian fits of International Tables vol. C (1992), electron 5-Gaussian fits
of International Tables vol. C and of cctbx.  DiSCaMB ships them as C arrays in
``src/Scattering/NGaussianFormFactorsTable.cpp`` (tables "Waasmeier-Kirfel", "IT92",
"electron-IT", "electron-cctbx"); this script reads those arrays as text (run it only where the
reference checkout exists) and writes them out unchanged.  The library never reads the reference.
"""
import json
import os
import re

REF = os.environ.get("DISCAMB_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
TABLES = [("Waasmeier-Kirfel", "Waasmeier_Kirfel", 5), ("IT92", "X_Ray_It95", 4),
          ("electron-IT", "Electron_IT", 5), ("electron-cctbx", "Electron_cctbx", 5)]
NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


def parse(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    out = {}
    for name, cname, n in TABLES:
        m = re.search(r"FormFactorData<%d>\s+%s\s*\[[^\]]*\]\s*=\s*\{" % (n, cname), text)
        if not m:
            raise SystemExit("table %s not found" % cname)
        start, depth, i = m.end(), 1, m.end()
        while depth:
            depth += {"{": 1, "}": -1}.get(text[i], 0)
            i += 1
        body = text[start:i - 1]
        entries = []
        for e in re.finditer(r'\{\s*"([^"]+)"\s*,\s*\{([^}]*)\}\s*,\s*\{([^}]*)\}\s*,\s*(%s)\s*\}' % NUM, body):
            a = [float(v) for v in re.findall(NUM, e.group(2))]
            b = [float(v) for v in re.findall(NUM, e.group(3))]
            if len(a) != n or len(b) != n:
                raise SystemExit("bad entry %s in %s" % (e.group(1), cname))
            entries.append({"label": e.group(1), "a": a, "b": b, "c": float(e.group(4))})
        out[name] = entries
    return out


def main():
    src = os.path.join(REF, "src", "Scattering", "NGaussianFormFactorsTable.cpp")
    tables = parse(open(src).read())
    with open(os.path.join(HERE, "ngaussian_tables.json"), "w") as f:
        json.dump({"source": "Waasmaier-Kirfel 1995; International Tables vol. C; cctbx (as tabulated "
                             "in DiSCaMB NGaussianFormFactorsTable.cpp)",
                   "formula": "f(h) = sum_i a_i exp(-b_i h^2 / 4) + c", "tables": tables}, f, indent=0)
    lines = ["// Generated by discamb_amd/data/make_ngaussian_tables.py - N-Gaussian atomic form factors",
             "// f(h) = sum_i a_i exp(-b_i h^2 / 4) + c  (published tables: Waasmaier & Kirfel 1995,",
             "// International Tables vol. C, cctbx).  Do not edit.",
             "#pragma once", "", "namespace discamb_amd {", "",
             "struct NGaussEntry {", "    const char* label;", "    double a[5], b[5], c;", "};", ""]
    for name, cname, n in TABLES:
        lines.append("// table \"%s\" (%d Gaussians)" % (name, n))
        lines.append("static const NGaussEntry kNGauss_%s[] = {" % cname)
        for e in tables[name]:
            a = e["a"] + [0.0] * (5 - n)
            b = e["b"] + [0.0] * (5 - n)
            lines.append('    {"%s", {%s}, {%s}, %r},' % (e["label"], ", ".join(repr(v) for v in a),
                                                          ", ".join(repr(v) for v in b), e["c"]))
        lines.append("};")
        lines.append("")
    lines.append("}  // namespace discamb_amd")
    with open(os.path.join(HERE, "..", "csrc", "ngauss_tables.h"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print({k: len(v) for k, v in tables.items()})


if __name__ == "__main__":
    main()
