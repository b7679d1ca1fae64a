"""Generate the tables of de_pow.h (run once; output pasted into de_pow.h).

The tables belong to the table-driven pow of ARM's optimized-routines
(Szabolcs Nagy), the method glibc's pow has used since 2.28. They follow from
the published recipe and are computed here in 80-digit decimal arithmetic:

log table (128 subintervals of z in [0x1.69555p-1, 0x1.69555p0), split by the
top 7 fraction bits of bits(x) - 0x3fe6955500000000):
    1/c      = round(128/center)/128        if center < 1
             = round(256/center)/256        otherwise
    logc     = round(2^43 * log(c)) / 2^43  (so k*ln2hi + logc is exact)
    logctail = (double)(log(c) - logc)
exp table (2^(k/128) ~= H[k] * (1 + T[k])):
    H[k] = (double)2^(k/128), T[k] = (double)((2^(k/128) - H[k]) / H[k]),
    stored as bits(T[k]) and bits(H[k]) - (k << 45).
"""
import struct
from decimal import ROUND_HALF_EVEN, Decimal, getcontext

getcontext().prec = 80
N = 128
OFF = 0x3fe6955500000000


def _asd(u):
    return struct.unpack("<d", struct.pack("<Q", u))[0]


def _asu(d):
    return struct.unpack("<Q", struct.pack("<d", d))[0]


def log_table():
    rows = []
    two43 = Decimal(2) ** 43
    for i in range(N):
        lo = Decimal(_asd(OFF + i * (1 << 45)))
        hi = Decimal(_asd(OFF + (i + 1) * (1 << 45)))
        center = (lo + hi) / 2
        if center < 1:
            invc = (Decimal(N) / center).to_integral_value(ROUND_HALF_EVEN) / N
        else:
            invc = (Decimal(2 * N) / center).to_integral_value(ROUND_HALF_EVEN) / (2 * N)
        logc_exact = -invc.ln()
        logc = (logc_exact * two43).to_integral_value(ROUND_HALF_EVEN) / two43
        assert Decimal(float(invc)) == invc and Decimal(float(logc)) == logc
        rows.append((float(invc), float(logc), float(logc_exact - logc)))
    return rows


def exp_table():
    words = []
    for k in range(N):
        e = Decimal(2) ** (Decimal(k) / N)
        h = float(e)
        t = float((e - Decimal(h)) / Decimal(h))
        words.append((_asu(t), (_asu(h) - (k << 45)) & 0xFFFFFFFFFFFFFFFF))
    return words


def main():
    print("DE_POW_TABLE double kDePowLog[128 * 3] = {")
    for a, b, c in log_table():
        print("    %s, %s, %s," % (a.hex(), b.hex(), c.hex()))
    print("};")
    print("DE_POW_TABLE unsigned long long kDePowExp[256] = {")
    for t, h in exp_table():
        print("    0x%016xULL, 0x%016xULL," % (t, h))
    print("};")


if __name__ == "__main__":
    main()
