"""Generates twiddles.hpp: exp(-2 pi i j / N), j < N, for the in-register
DFT sizes, as compile-time constants (cos/sin from Python's math module,
printed with 17 significant digits).  Run: python gen_twiddles.py > twiddles.hpp"""
import math

SIZES = (6, 9, 12, 16, 24, 32)

print("// twiddles.hpp -- generated by gen_twiddles.py: tw<N>::re/im[j] = cos/sin(-2 pi j/N)")
print("#pragma once")
print("namespace cfd {")
print("template <int N> struct tw;")
for n in SIZES:
    re = ", ".join("%.17g" % math.cos(-2 * math.pi * j / n) for j in range(n))
    im = ", ".join("%.17g" % math.sin(-2 * math.pi * j / n) for j in range(n))
    print("template <> struct tw<%d> {" % n)
    print("  static constexpr double re[%d] = {%s};" % (n, re))
    print("  static constexpr double im[%d] = {%s};" % (n, im))
    print("};")
print("}  // namespace cfd")
