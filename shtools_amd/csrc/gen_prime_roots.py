"""Writes prime_roots.hpp: cos(2 pi q / R) and sin(2 pi q / R), q = 0 .. R-1,
for the prime radices R of the ring FFTs, rounded to double from 50-digit
values (mpmath).  Run from this directory: python gen_prime_roots.py"""
import mpmath

PRIMES = [11, 13, 17, 19, 23]  # SHTG_PRIME_RADICES of fftplan.hpp

mpmath.mp.dps = 50


def main():
    vals, offs, off = [], [], 0
    for R in PRIMES:
        offs.append((R, off))
        for q in range(R):
            a = 2 * mpmath.pi * q / R
            vals += [repr(float(mpmath.cos(a))), repr(float(mpmath.sin(a)))]
        off += 2 * R
    out = ["// Generated by gen_prime_roots.py: cos(2 pi q / R) and sin(2 pi q / R),",
           "// q = 0 .. R-1, for the prime radices R of the ring FFTs, rounded to",
           "// double from 50-digit values.",
           "#pragma once",
           "",
           "namespace shtg {",
           "",
           "// (cos, sin) pairs, one block of R pairs per prime radix R",
           "__constant__ const double kPrimeRoots[] = {"]
    for i in range(0, len(vals), 4):
        out.append("    " + ", ".join(vals[i:i + 4]) + ",")
    out += ["};",
            "",
            "// offset of radix R's block in kPrimeRoots (in doubles)",
            "__host__ __device__ constexpr int prime_roots_offset(int R)",
            "{",
            "    return " + " : ".join(f"R == {R} ? {o}" for R, o in offs) + " : -1;",
            "}",
            "",
            "}  // namespace shtg",
            ""]
    with open("prime_roots.hpp", "w") as fh:
        fh.write("\n".join(out))


if __name__ == "__main__":
    main()
