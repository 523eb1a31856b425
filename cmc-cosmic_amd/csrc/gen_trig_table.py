"""Constants of the table-driven double-double sin/cos (dd_trig.hpp and its CPU copy
oracle/orc_trig.c): sin and cos of j*pi/128, j = 0..32, as (hi, lo) pairs with
hi = RN(v) and lo = RN(v - hi), and pi/128 split into H1 + H2 + H3 with H1 and H2
of 47 significant bits (so j*H1 and j*H2 are exact for |j| <= 32).  Computed with
mpmath at 200 bits; prints C initialisers.  Run: python gen_trig_table.py"""
import mpmath as mp

mp.mp.prec = 200


def split(v):
    hi = float(v)
    return hi, float(v - mp.mpf(hi))


def round_bits(v, bits):
    m, e = mp.frexp(v)
    return float(mp.ldexp(mp.nint(mp.ldexp(m, bits)), e - bits))


h = mp.pi / 128
H1 = round_bits(h, 47)
H2 = round_bits(h - H1, 47)
H3 = float(h - H1 - H2)
print(f"// pi/128 = H1 + H2 + H3\nconst double H1 = {H1.hex()}, H2 = {H2.hex()}, H3 = {H3.hex()};")
print("// {sin hi, sin lo, cos hi, cos lo} of j*pi/128, j = 0..32")
for j in range(33):
    s, c = mp.sin(j * h), mp.cos(j * h)
    sh, sl = split(s)
    ch, cl = split(c)
    print(f"    {{{sh.hex()}, {sl.hex()}, {ch.hex()}, {cl.hex()}}},")
