#!/usr/bin/env python3
"""Generate ``wfn_tables.json``: Clementi-Roetti atomic Slater-orbital wavefunctions and
default deformation-valence parameters for the neutral atoms H..Kr and the ions of the table
(Li+ .. Zn2+, O-, F-, S-, Cl-, Se-, Br-).

The numbers are published tables (Clementi & Roetti 1974, At. Data Nucl. Data Tables 14, 177;
Clementi & Raimondi 1963, J. Chem. Phys. 38, 2686).  DiSCaMB ships them as C arrays in
``src/HC_Model/ClementiRoettiData.cpp`` and ``src/HC_Model/ClementiRaimondiData.cpp``; this
script reads those arrays as text (run it only where the reference checkout exists) and
applies the same post-processing the reference applies when it builds its wavefunction bank:

* orbital configuration parsing and core/valence split, including its rule that cations left
  with 28 or 46 electrons have no valence orbitals (``SlaterOrbitalWfnData.cpp:13-160``),
* exponents Bohr^-1 -> Angstrom^-1 with 1/0.52917721092, Slater normalisation
  ``N = (2a)^(k+1) sqrt(2a/(2k+2)!)`` and orbital renormalisation
  (``SlaterOrbitalWfnData.cpp:162-256``, ``HC_WfnData.h:33-37``),
* deformation-valence powers n_l and exponent zeta, UBDB variant for P and S
  (``DeformationValenceParameters.cpp:69-160, 303-345``; Clementi-Raimondi exponents scaled
  by 1.8897261245, ``ClementiRaimondiData.cpp:106-135``).

The output is plain data used by the library at run time; the library never reads the
reference.
"""
import json
import math
import os
import re
import sys

REF = os.environ.get("DISCAMB_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))

SYMBOLS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si",
           "P", "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu",
           "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"]


def period_of(z):
    return 1 if z <= 2 else 2 if z <= 10 else 3 if z <= 18 else 4


def block_of(z):
    if z in (1, 2, 3, 4, 11, 12, 19, 20):
        return "S"
    if 21 <= z <= 30:
        return "D"
    return "P"


def strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def parse_configurations(text):
    m = re.search(r"cr_configuration_data\[\d+\]\s*=\s*\{(.*?)\};", text, re.S)
    confs = {}
    for s in re.findall(r'"([^"]*)"', m.group(1)):
        label, conf = s.split()
        occ = {}
        for tok in conf.split(")"):
            if not tok:
                continue
            name, n_el = tok.split("(")
            n_el = int(n_el)
            if name == "K":
                occ[(1, 0)] = 2
            elif name == "L":
                occ[(2, 0)] = 2
                occ[(2, 1)] = 6
            elif name == "M":
                occ[(3, 0)] = 2
                occ[(3, 1)] = 6
                occ[(3, 2)] = 10
            else:
                if n_el > 0:
                    occ[(int(name[0]), "SPD".index(name[1]))] = n_el
        confs[label] = occ
    return confs


def parse_wfns(text):
    start = text.index("LocalDef_Wfn wfns[]")
    body = text[text.index("{", start):]
    tokens = re.findall(r'"[^"]*"|[-+]?\d*\.\d+(?:[eE][-+]?\d+)?|[-+]?\d+', body)
    pos = 0
    out = []
    while pos < len(tokens) and tokens[pos].startswith('"'):
        label = tokens[pos].strip('"')
        z, charge, n_orb = int(tokens[pos + 1]), int(tokens[pos + 2]), int(tokens[pos + 3])
        pos += 4
        orbs = []
        for _ in range(n_orb):
            n, l, n_terms = int(tokens[pos]), int(tokens[pos + 1]), int(tokens[pos + 2])
            pos += 3
            terms = []
            for _ in range(n_terms):
                terms.append((float(tokens[pos]), int(tokens[pos + 1]), float(tokens[pos + 2])))
                pos += 3
            orbs.append((n, l, terms))
        out.append((label, z, charge, orbs))
    return out


def parse_raimondi(text):
    m = re.search(r"cr_exponents\[\]\[11\]\s*=\s*\{(.*?)\};", text, re.S)
    rows = re.findall(r"\{([^{}]*)\}", m.group(1))
    return [[float(v) for v in r.split(",") if v.strip()] for r in rows]


# subshell column order of the Clementi-Raimondi table: 1S 2S 2P 3S 3P 4S 3D 4P 5S 4D 5P
RAIMONDI_COLUMNS = [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (4, 0), (3, 2), (4, 1)]


def valence_orbitals(z, charge, occ):
    n_el = z - charge
    if n_el in (0, 2, 10, 18, 36, 54, 86):
        return []
    if charge > 0 and n_el in (28, 46):
        return []
    blk, per = block_of(z), period_of(z)
    val = []
    if blk in "SP":
        val.append((per, 0))
    if blk == "P":
        val.append((per, 1))
    if blk == "D":
        if occ.get((3, 2), 0) != 10 and (3, 2) in occ:
            val.append((3, 2))
        elif occ.get((3, 2), 0) == 10 and any(k[0] == 4 for k in occ):
            val.append((4, 0))
    return val


def sto_norm(k, a):
    return (2 * a) ** (k + 1) * math.sqrt(2 * a / math.factorial(2 * (k + 1)))


def def_val_powers(z):
    per, blk = period_of(z), block_of(z)
    if per == 1:
        return [0, 1, 2, 3, 4]
    if per == 2:
        return [2, 2, 2, 3, 4]
    if per == 3 or (per == 4 and blk == "D"):
        return [4] * 5
    return [6] * 5


def main():
    cr = strip_comments(open(os.path.join(REF, "src/HC_Model/ClementiRoettiData.cpp")).read())
    rai = strip_comments(open(os.path.join(REF, "src/HC_Model/ClementiRaimondiData.cpp")).read())
    confs = parse_configurations(open(os.path.join(REF, "src/HC_Model/ClementiRoettiData.cpp")).read())
    wfns = parse_wfns(cr)
    rexp = parse_raimondi(rai)
    bohr_to_ang = 1.0 / 0.52917721092
    raimondi_scale = 1.8897261245
    table = {}
    for label, z, charge, orbs in wfns:
        if z > 36:
            continue
        occ = confs[label]
        val = valence_orbitals(z, charge, occ)
        entry = {"label": label, "Z": z, "charge": charge, "orbitals": [], "occupancy": [],
                 "core": [], "valence": []}
        for i, (n, l, terms) in enumerate(orbs):
            (entry["valence"] if (n, l) in val else entry["core"]).append(i)
            entry["occupancy"].append(occ.get((n, l), 0))
            ex = [t[2] * bohr_to_ang for t in terms]
            pw = [t[1] for t in terms]
            co = [sto_norm(pw[j], ex[j]) * terms[j][0] for j in range(len(terms))]
            ov = 0.0
            for a in range(len(co)):
                for b in range(len(co)):
                    nn = pw[a] + pw[b] + 2
                    ov += co[a] * co[b] * math.factorial(nn) / (ex[a] + ex[b]) ** (nn + 1)
            s = math.sqrt(1.0 / ov)
            entry["orbitals"].append({"n": n, "l": l, "coefficients": [c * s for c in co],
                                      "exponents": ex, "powers": pw})
        # deformation-valence parameters
        shells = {}
        for col, (n, l) in enumerate(RAIMONDI_COLUMNS):
            if col < len(rexp[z - 1]):
                shells[(n, l)] = raimondi_scale * rexp[z - 1][col]
        blk = block_of(z)
        if not val:
            top = max(n for (n, l) in shells)
            if blk == "S":
                zeta = 2 * shells[(top, 0)]
            elif blk == "P":
                zeta = (shells[(top, 0)] + 3 * shells[(top, 1)]) / 2.0
            else:
                zeta = 2 * shells[(3, 2)]
        else:
            tot = sum(occ[v] for v in val)
            zeta = 2 * sum(occ[v] * shells[v] for v in val) / tot
        powers = def_val_powers(z)
        ubdb = list(powers)
        symbol = "".join(ch for ch in label if ch.isalpha())     # UBDB rule keyed on the element
        if symbol == "P":
            ubdb = [4, 6, 6, 6, 6]
        if symbol == "S":
            ubdb = [1, 2, 4, 6, 8]
        entry["def_val_exponent"] = zeta
        entry["def_val_powers_standard"] = powers
        entry["def_val_powers_ubdb"] = ubdb
        table[label] = entry
    with open(os.path.join(HERE, "wfn_tables.json"), "w") as f:
        json.dump({"source": "Clementi-Roetti (1974) STO wavefunctions, Clementi-Raimondi (1963) "
                             "exponents; units Angstrom", "entries": table}, f, indent=1)
    print("wrote", len(table), "entries")


if __name__ == "__main__":
    sys.exit(main())
