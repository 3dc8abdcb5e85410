"""Writes pf_ziggurat.h: the 256-layer normal ziggurat tables (Marsaglia and Tsang) on
51-bit magnitudes, as Julia 1.5's randn uses them.  r = 3.6541528853610088 is the
outermost layer edge, v = r f(r) + sqrt(pi/2) erfc(r / sqrt 2) the area of every layer
(f(x) = exp(-x^2 / 2)); the layer edges x_i follow from f(x_i) = v / x_{i+1} + f(x_{i+1}).
Everything is computed in 60-digit arithmetic and rounded once:
  k[0] = floor(r / q 2^51) with q = v / f(r), k[i+1] = floor(x_i / x_{i+1} 2^51),
  w[0] = q / 2^51, w[i] = x_i / 2^51, f[i] = f(x_i), f[0] = 1.
Run from this directory: python3 gen_ziggurat.py > pf_ziggurat.h
"""
import mpmath


def tables():
    mpmath.mp.dps = 60
    r = mpmath.mpf("3.6541528853610088")
    f = lambda x: mpmath.exp(-x * x / 2)
    v = r * f(r) + mpmath.sqrt(mpmath.pi / 2) * mpmath.erfc(r / mpmath.sqrt(2))
    m1 = mpmath.mpf(2) ** 51
    xs = {255: r}
    x = r
    for i in range(254, 0, -1):
        x = mpmath.sqrt(-2 * mpmath.log(v / x + f(x)))
        xs[i] = x
    q = v / f(r)
    ki, wi, fi = [0] * 256, [0.0] * 256, [0.0] * 256
    ki[0] = int(mpmath.floor(r / q * m1))
    wi[0], wi[255] = float(q / m1), float(r / m1)
    fi[0], fi[255] = 1.0, float(f(r))
    for i in range(254, 0, -1):
        ki[i + 1] = int(mpmath.floor(xs[i] / xs[i + 1] * m1))
        fi[i] = float(f(xs[i]))
        wi[i] = float(xs[i] / m1)
    return ki, wi, fi


def main():
    ki, wi, fi = tables()
    out = ["// Generated by gen_ziggurat.py (see there for the construction); do not edit.",
           "#pragma once", "#include <cstdint>", "", "namespace pf {", ""]
    out.append("constexpr uint64_t kZigK[256] = {")
    out += ["    " + ", ".join(f"0x{k:016x}ULL" for k in ki[i:i + 4]) + "," for i in range(0, 256, 4)]
    out.append("};")
    for name, arr in (("kZigW", wi), ("kZigF", fi)):
        out.append(f"constexpr double {name}[256] = {{")
        out += ["    " + ", ".join(float.hex(x) for x in arr[i:i + 4]) + "," for i in range(0, 256, 4)]
        out.append("};")
    out += ["constexpr double kZigR = " + float.hex(3.6541528853610088) + ";", "",
            "}  // namespace pf", ""]
    print("\n".join(out))


if __name__ == "__main__":
    main()
