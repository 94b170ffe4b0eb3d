"""Writes cjx_pow5.h: for q in [-342, 308], the 128-bit mantissa T[q] of 5^q,
normalised to [2^127, 2^128) and truncated:

    T[q] = floor(5^q * 2^(127 - floor(log2(5^q))))

computed with Python's exact integers (5^q * 2^k for q >= 0, 2^k // 5^-q for
q < 0).  T[q] is exact for 0 <= q <= 55.  Run from the repository root:

    python3 cujson_amd/csrc/gen_pow5.py > cujson_amd/csrc/cjx_pow5.h
"""

QMIN, QMAX = -342, 308


def mantissa(q: int) -> int:
    if q >= 0:
        v = 5 ** q
        shift = 127 - (v.bit_length() - 1)  # floor(log2 5^q) = bit_length - 1
        return v << shift if shift >= 0 else v >> -shift
    d = 5 ** -q
    # floor(log2(1/d)) = -ceil(log2 d) = -(d.bit_length() - 1) - 1 unless d is a power of 2
    k = 127 + d.bit_length()  # 2^k / d lies in [2^127, 2^128)
    t = (1 << k) // d
    assert (1 << 127) <= t < (1 << 128)
    return t


def main():
    print("// cjx_pow5.h - generated by gen_pow5.py (do not edit): 128-bit truncated")
    print("// mantissas of 5^q, q in [-342, 308], normalised to [2^127, 2^128);")
    print("// {high 64 bits, low 64 bits}.")
    print("#pragma once")
    print("#include <stdint.h>")
    print("namespace cjx {")
    print(f"constexpr int kPow5Min = {QMIN}, kPow5Max = {QMAX};")
    print(f"constexpr uint64_t kPow5[{QMAX - QMIN + 1}][2] = {{")
    for q in range(QMIN, QMAX + 1):
        t = mantissa(q)
        assert t.bit_length() == 128
        print(f"    {{0x{t >> 64:016X}ull, 0x{t & ((1 << 64) - 1):016X}ull}},  // 5^{q}")
    print("};")
    print("}  // namespace cjx")


if __name__ == "__main__":
    main()
