"""Derives the double-precision constants of wam_math.h in exact rational
arithmetic and prints them as C hex-float literals:

    python wam_amd/csrc/gen_math_consts.py

pi comes from Machin's formula pi = 16 atan(1/5) - 4 atan(1/239) and ln 2 from
ln 2 = 2 atanh(1/3), each summed to 400 bits with integers; atan(j/8) from
the Euler series atan(x) = sum 2^(2n) (n!)^2 / (2n+1)! * x^(2n+1) / (1+x^2)^(n+1).
A split constant hi + mid + lo keeps hi (and mid) to a fixed number of
leading bits, so that a product with a small integer is exact.
"""
from fractions import Fraction

BITS = 400
ONE = 1 << BITS


def atan_inv(n):
    """atan(1/n) * 2^BITS (Taylor series in integers)."""
    s, t, k, sign = 0, ONE // n, 1, 1
    n2 = n * n
    while t:
        s += sign * (t // k)
        t //= n2
        k += 2
        sign = -sign
    return s


PI = Fraction(16 * atan_inv(5) - 4 * atan_inv(239), ONE)


def atanh_inv(n):
    s, t, k = 0, ONE // n, 1
    n2 = n * n
    while t:
        s += t // k
        t //= n2
        k += 2
    return s


LN2 = Fraction(2 * atanh_inv(3), ONE)
LN10 = 3 * LN2 + Fraction(2 * atanh_inv(9), ONE)     # ln 10 = ln 8 + ln(5/4), ln(5/4) = 2 atanh(1/9)


def atan_frac(x):
    """atan(x) for rational 0 <= x <= 1 (Euler's series), to ~400 bits."""
    x = Fraction(x)
    y = x * x / (1 + x * x)
    term = x / (1 + x * x)
    s, n = Fraction(0), 0
    while True:
        s += term
        n += 1
        term = term * y * Fraction(2 * n, 2 * n + 1)
        if term < Fraction(1, 1 << (BITS + 10)):
            break
    return s


def round_double(q):
    """Nearest double to the rational q (ties to even), as a Fraction."""
    if q == 0:
        return Fraction(0)
    sign = -1 if q < 0 else 1
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    scale = Fraction(2) ** (52 - e)
    m = a * scale
    fl = m.numerator // m.denominator
    rem = m - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return sign * Fraction(fl) / scale


def trunc_bits(q, bits):
    """q truncated toward zero to its leading `bits` significant bits."""
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    scale = Fraction(2) ** (bits - 1 - e)
    m = a * scale
    return (1 if q > 0 else -1) * Fraction(m.numerator // m.denominator) / scale


def hexf(q):
    return float(q).hex()


def fact(n):
    r = 1
    for i in range(2, n + 1):
        r *= i
    return r


def main():
    out = []
    pio2 = PI / 2
    p1 = trunc_bits(pio2, 36)
    p2 = trunc_bits(pio2 - p1, 36)
    p3 = round_double(pio2 - p1 - p2)
    out += [("WAM_PIO2_1", p1), ("WAM_PIO2_2", p2), ("WAM_PIO2_3", p3)]
    out += [("WAM_2_OVER_PI", round_double(2 / PI)), ("WAM_PI_D", round_double(PI)), ("WAM_PIO2_D", round_double(pio2))]
    l1 = trunc_bits(LN2, 44)
    out += [("WAM_LN2_HI", l1), ("WAM_LN2_LO", round_double(LN2 - l1)), ("WAM_INV_LN2", round_double(1 / LN2)),
            ("WAM_INV_LN10", round_double(1 / LN10))]
    for name, q in out:
        assert float(q) == q, name            # representable exactly
        print(f"#define {name} {hexf(q)}")
    print("/* 1/n! */")
    print("static const double WAM_INV_FACT[] = {" + ", ".join(hexf(round_double(Fraction(1, fact(n))))
                                                          for n in range(0, 19)) + "};")
    print("/* 1/(2n+1) */")
    print("static const double WAM_INV_ODD[] = {" + ", ".join(hexf(round_double(Fraction(1, 2 * n + 1)))
                                                         for n in range(0, 12)) + "};")
    print("/* atan(j/8), j = 0..8 */")
    print("static const double WAM_ATAN_J8[] = {" + ", ".join(hexf(round_double(atan_frac(Fraction(j, 8))))
                                                         for j in range(9)) + "};")


if __name__ == "__main__":
    main()
