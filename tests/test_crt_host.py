"""The CRT form's integer arithmetic on the host (no GPU): the moduli are
pairwise coprime and odd, the operand width leaves the product range inside
half the moduli product (rows of 2-norm below 2^(beta-1), rounded entry by
entry: by Cauchy-Schwarz every product of two of them stays below
(2^(beta-1) + sqrt(K)/2)^2; past the 50-bit cap rows of maximum below
2^(beta-1): K 4^(beta-1)), and the reconstruction the kernels run (explicit
form for 11 and 12 moduli, Garner over pairs for 14) turns the residues of
exact integers back into those integers: within 2^12 + 2 ulp for 11 and 12
moduli (the operand rounding of a CRT product puts at least 2^44 into every
output), within 4 ulp for 14 (one rounding per Horner step).  Checked
against Python integers."""
import ctypes as C
import math
import random

import numpy as np
import pytest

from gain_engine import _lib


def moduli(nm):
    L = _lib.load()
    m = (C.c_int * nm)()
    _lib.check(L.gain_crt_moduli(nm, m))
    return list(m)


def width(nm, K):
    """(operand width, whether rows are scaled by their 2-norm)"""
    L = _lib.load()
    b, n = C.c_int(), C.c_int()
    _lib.check(L.gain_crt_beta(nm, K, C.byref(b), C.byref(n)))
    return b.value, bool(n.value)


def beta(nm, K):
    return width(nm, K)[0]


def reconstruct(nm, values):
    """Residues of every value (symmetric, as the kernels store them),
    packed four outputs per word, through gain_crt_reconstruct."""
    L = _lib.load()
    ms = moduli(nm)
    n4 = (len(values) + 3) // 4
    vals = list(values) + [0] * (4 * n4 - len(values))
    words = np.zeros((n4, nm), dtype=np.uint32)
    for g in range(n4):
        for i, m in enumerate(ms):
            w = 0
            for q in range(4):
                r = vals[4 * g + q] % m
                if r > (m - 1) // 2:
                    r -= m
                w |= (r & 0xFF) << (8 * q)
            words[g, i] = w
    out = np.zeros(4 * n4)
    _lib.check(L.gain_crt_reconstruct(
        nm, words.ctypes.data_as(C.POINTER(C.c_uint)), n4,
        out.ctypes.data_as(C.POINTER(C.c_double))))
    return out[:len(values)]


@pytest.mark.parametrize("nm", [11, 12, 14])
def test_moduli_and_width(nm):
    ms = moduli(nm)
    assert len(set(ms)) == nm and all(m % 2 == 1 and m <= 255 for m in ms)
    for i in range(nm):
        for j in range(i):
            assert math.gcd(ms[i], ms[j]) == 1
    M = math.prod(ms)
    for K in (64, 4032, 43008):
        b, norm = width(nm, K)
        if norm:
            # rows scaled by their 2-norm: the Cauchy-Schwarz range fits
            # with the reconstruction's margin (exact integers:
            # (2^b + sqrt(K))^2 < 2 M (1 - 7e-7), sqrt(K) bounded above by
            # isqrt(K) + 1 and below by isqrt(K))
            lim = 2 * M * (10 ** 7 - 7) // 10 ** 7
            assert b <= 50 and (2 ** b + math.isqrt(K) + 1) ** 2 < lim
            # and it is the widest that does
            assert (2 ** (b + 1) + math.isqrt(K)) ** 2 >= lim
            # about log2(K)/2 more bits than the bound by K times maxima
            assert 2 * K * 4 ** (b - 1) >= M
        else:
            # past the 50-bit cap: rows scaled by their maximum
            assert 2 * K * 4 ** (b - 1) < M
            assert b == 50 or 2 * K * 4 ** b >= M


def test_width_grows_with_norm_bound():
    """The Cauchy-Schwarz bound (row norms) against the bound by K times the
    row maxima (2 K 4^(beta-1) < M): at K = 43008 the width grows from 40
    to 47 bits with 12 moduli and from 36 to 43 with 11; 14 moduli would
    pass the 50-bit cap and keep the scaling by the maxima (47 and 49
    bits)."""
    assert width(12, 43008) == (47, True) and width(11, 43008) == (43, True)
    assert width(12, 4032) == (47, True) and width(11, 4032) == (43, True)
    assert width(14, 43008) == (47, False) and width(14, 4032) == (49, False)


@pytest.mark.parametrize("nm,K", [(11, 4032), (12, 4032), (12, 43008),
                                  (14, 4032), (14, 43008)])
def test_reconstruction_of_exact_integers(nm, K):
    rng = random.Random(nm * 100003 + K)
    b, norm = width(nm, K)
    # largest product of two integer rows (Cauchy-Schwarz, or K times the
    # product of the maxima)
    lim = ((2 ** (b - 1) + (math.isqrt(K) + 1) // 2) ** 2 if norm
           else K * 4 ** (b - 1))
    vals = [0, 1, -1, 2 ** 60, -(2 ** 60), lim, -lim, lim - 1]
    vals += [rng.randint(-lim, lim) for _ in range(3000)]
    vals += [rng.randint(-(2 ** 30), 2 ** 30) * rng.choice([1, 2 ** 20, 2 ** 40])
             for _ in range(3000)]
    got = reconstruct(nm, vals)
    for x, v in zip(got, vals):
        err = abs(x - float(v))
        if nm == 14:
            assert err <= 4 * math.ulp(float(v)), (nm, v, x)
        else:
            assert err <= 2 ** 12 + 2 * math.ulp(float(v)), (nm, v, x)


def test_reconstruction_rejects_bad_moduli_count():
    L = _lib.load()
    out = (C.c_double * 4)()
    words = (C.c_uint * 13)()
    assert L.gain_crt_reconstruct(13, words, 1, out) != 0
