#!/usr/bin/env python3
"""Extract the measured spectral data tables that the render path needs.

The render path and the CPU oracle must use the very same measured data:
  * CIE 1931 2-degree colour matching functions, 360..830 nm at 1 nm
    (reference: src/filianore/color/cie.h:8-761),
  * the 32-sample RGB -> spectrum basis functions (Smits 1999) used by
    from_reflectance_rgb / from_illuminance_rgb (cie.h:763-952),
  * complex refractive index (eta, k) tables of the nine forge metals
    (reference: src/filianore/color/metallurgy.h:11-1018).

These are numeric data, not code: this script parses the float literals out of
the reference headers, rounds each literal to float32 exactly as a C++
compiler would, and writes them as hexadecimal float literals into
``fl_tables.inc`` (a C/C++/HIP include consumed by both the product library
and the oracle).  Only this script reads the reference; the generated file is
committed so that nothing at run time needs the reference tree.

Usage: python gen_tables.py [/path/to/reference] > fl_tables.inc
"""
import os
import re
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
COLOR = os.path.join(REF, "src", "filianore", "color")

ARRAY_RE = re.compile(r"const\s+float\s+(\w+)\s*\[[^\]]*\]\s*=\s*\{(.*?)\};", re.S)
NUM_RE = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")


def parse_arrays(path):
    with open(path) as f:
        text = f.read()
    text = re.sub(r"//[^\n]*", "", text)
    out = {}
    for name, body in ARRAY_RE.findall(text):
        vals = [np.float32(float(tok)) for tok in NUM_RE.findall(body)]
        out[name] = vals
    return out


def emit(name, vals, lines):
    lines.append("static const float %s[%d] = {" % (name, len(vals)))
    row = []
    for v in vals:
        row.append(float(v).hex() + "f" if v != 0 else "0x0p+0f")
        if len(row) == 4:
            lines.append("    " + ", ".join(row) + ",")
            row = []
    if row:
        lines.append("    " + ", ".join(row) + ",")
    lines.append("};")


def main():
    cie = parse_arrays(os.path.join(COLOR, "cie.h"))
    met = parse_arrays(os.path.join(COLOR, "metallurgy.h"))
    lines = [
        "/* Generated by gen_tables.py: measured spectral data (float32, hex literals).",
        " * CIE 1931 CMFs 360..830 nm @1 nm; Smits RGB->spectrum basis (32 samples);",
        " * forge metal eta/k tables.  Do not edit by hand. */",
        "#pragma once",
        "#define FL_CIE_N 471",
        "#define FL_CIE_LAMBDA_MIN 360.0f",
        "#define FL_CIE_Y_INTEGRAL 106.856895f",
        "#define FL_RGB2SPEC_N 32",
    ]
    for src, dst in [("CIE_X", "FL_CIE_X"), ("CIE_Y", "FL_CIE_Y"), ("CIE_Z", "FL_CIE_Z"),
                     ("RGB2SpectLambda", "FL_RGB2SPEC_LAMBDA")]:
        assert len(cie[src]) == (471 if src.startswith("CIE") else 32), src
        emit(dst, cie[src], lines)
    basis = ["White", "Cyan", "Magenta", "Yellow", "Red", "Green", "Blue"]
    for kind in ["Refl", "Illum"]:
        for b in basis:
            src = "RGB%s2Spect%s" % (kind, b)
            assert len(cie[src]) == 32, src
            emit("FL_RGB%s_%s" % (kind.upper(), b.upper()), cie[src], lines)
    # forge order is GetMetalCoefficientsFromForge's switch (metallurgy.cpp:42-75)
    metals = ["Gold", "Copper", "Silver", "Aluminium", "Beryllium", "Chromium",
              "Mercury", "Iridium", "Lithium"]
    lines.append("#define FL_N_METALS %d" % len(metals))
    counts = []
    for m in metals:
        lam, eta, k = met[m + "_Lambda"], met[m + "_Eta"], met[m + "_K"]
        assert len(lam) == len(eta) == len(k)
        counts.append(len(lam))
        emit("FL_METAL_%s_LAMBDA" % m.upper(), lam, lines)
        emit("FL_METAL_%s_ETA" % m.upper(), eta, lines)
        emit("FL_METAL_%s_K" % m.upper(), k, lines)
    lines.append("static const int FL_METAL_COUNT[%d] = {%s};" % (len(metals), ", ".join(map(str, counts))))
    for part in ["LAMBDA", "ETA", "K"]:
        lines.append("static const float* const FL_METAL_%s[%d] = {%s};" % (
            part, len(metals), ", ".join("FL_METAL_%s_%s" % (m.upper(), part) for m in metals)))
    sys.stdout.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
