"""Writes fp_dft_tables.h: cos/sin of 2 pi m / R for the odd primes that the
FFT passes handle with register codelets.  Run: python gen_dft_tables.py"""
import math
import os

PRIMES = [7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47]
HERE = os.path.dirname(os.path.abspath(__file__))

lines = ["// Generated by gen_dft_tables.py: cos/sin(2 pi m / R), m < R, for the odd",
         "// prime radices with register codelets.  Do not edit.",
         "#pragma once", "", "namespace fp {", "", "template <int R>", "struct DftTab;", ""]
for R in PRIMES:
    c = ", ".join("%.17g" % math.cos(2 * math.pi * m / R) for m in range(R))
    s = ", ".join("%.17g" % math.sin(2 * math.pi * m / R) for m in range(R))
    lines += ["template <>", "struct DftTab<%d> {" % R,
              "    static constexpr double c[%d] = {%s};" % (R, c),
              "    static constexpr double s[%d] = {%s};" % (R, s), "};", ""]
lines += ["}  // namespace fp", ""]
open(os.path.join(HERE, "fp_dft_tables.h"), "w").write("\n".join(lines))
