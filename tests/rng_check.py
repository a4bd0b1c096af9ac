"""Checks of the production random stream (xoshiro128**, Blackman & Vigna,
ACM TOMS 47, 2021), independent of the device.

* `xoshiro128ss`: plain-Python restatement of the generator, with optional
  planted slips for sensitivity tests.
* `engine_period_is_maximal`: the state update is linear over GF(2); its
  period is 2^128 - 1 exactly when the minimal polynomial of one state bit's
  sequence has degree 128 and is primitive.  The polynomial comes from
  Berlekamp-Massey over 512 bits; primitivity is x^(2^128-1) = 1 and
  x^((2^128-1)/q) != 1 for every prime q dividing 2^128 - 1 (the Fermat
  numbers F0..F6: 3, 5, 17, 257, 65537, 641 * 6700417, 274177 * 67280421310721).
"""
M32 = 0xFFFFFFFF
PRIMES_2_128_MINUS_1 = (3, 5, 17, 257, 641, 65537, 274177, 6700417, 67280421310721)


def _rotl(x, k):
    return ((x << k) | (x >> (32 - k))) & M32


def step(s, slip=None):
    """One state update (s0, s1, s2, s3) -> new state; returns (output, state)."""
    s0, s1, s2, s3 = s
    out = (_rotl((s1 * 5) & M32, 7) * 9) & M32
    t = (s1 << (8 if slip == "shift" else 9)) & M32
    s2 ^= s0
    s3 ^= s1
    s1 ^= s2
    s0 ^= s3
    s2 ^= t
    s3 = _rotl(s3, 10 if slip == "rotate" else 11)
    return out, (s0, s1, s2, s3)


def step_order_slip(s):
    """`step` with s1 updated before s2 takes s0 (a plausible reordering slip)."""
    s0, s1, s2, s3 = s
    out = (_rotl((s1 * 5) & M32, 7) * 9) & M32
    t = (s1 << 9) & M32
    s3 ^= s1
    s1 ^= s2          # before s2 ^= s0
    s2 ^= s0
    s0 ^= s3
    s2 ^= t
    s3 = _rotl(s3, 11)
    return out, (s0, s1, s2, s3)


def xoshiro128ss(state, n, slip=None):
    out = []
    s = tuple(state)
    for _ in range(n):
        o, s = step_order_slip(s) if slip == "order" else step(s, slip)
        out.append(o)
    return out


def _berlekamp_massey(bits):
    """Connection polynomial (as an int, bit i = coefficient of x^i) and its degree."""
    c, b = 1, 1
    L, m = 0, 1
    for n in range(len(bits)):
        d = bits[n]
        for i in range(1, L + 1):
            d ^= ((c >> i) & 1) & bits[n - i]
        if d == 0:
            m += 1
        elif 2 * L <= n:
            t = c
            c ^= b << m
            L, b, m = n + 1 - L, t, 1
        else:
            c ^= b << m
            m += 1
    return c, L


def _mulmod(a, b, mod, deg):
    r = 0
    while b:
        if b & 1:
            r ^= a
        b >>= 1
        a <<= 1
        if (a >> deg) & 1:
            a ^= mod
    return r


def _powx(e, mod, deg):
    result, base = 1, 2  # the polynomials 1 and x
    while e:
        if e & 1:
            result = _mulmod(result, base, mod, deg)
        base = _mulmod(base, base, mod, deg)
        e >>= 1
    return result


def engine_period_is_maximal(slip=None, state=(0x12345678, 0x9ABCDEF0, 0x0F1E2D3C, 0x4B5A6978)):
    s = tuple(state)
    bits = []
    for _ in range(512):
        bits.append(s[0] & 1)
        _, s = step_order_slip(s) if slip == "order" else step(s, slip)
    conn, L = _berlekamp_massey(bits)
    if L != 128:
        return False
    # characteristic polynomial = reciprocal of the connection polynomial
    poly = int(bin(conn)[2:].zfill(L + 1)[::-1], 2)
    order = (1 << 128) - 1
    if _powx(order, poly, 128) != 1:
        return False
    return all(_powx(order // q, poly, 128) != 1 for q in PRIMES_2_128_MINUS_1)


# 2^256 - 1 = F0 F1 ... F7 (Fermat numbers); F5 = 641 * 6700417,
# F6 = 274177 * 67280421310721, F7 = 59649589127497217 * 5704689200685129054721
PRIMES_2_256_MINUS_1 = (3, 5, 17, 257, 641, 65537, 274177, 6700417, 67280421310721, 59649589127497217,
                        5704689200685129054721)


def is_probable_prime(n: int) -> bool:
    """Miller-Rabin with the first 13 prime bases (deterministic below 3.3e24)."""
    if n < 2:
        return False
    bases = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41)
    for p in bases:
        if n % p == 0:
            return n == p
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for a in bases:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def sequence_is_primitive(bits, degree: int, primes) -> bool:
    """True when the bit sequence obeys a linear recurrence whose minimal
    polynomial has the given degree and is primitive over GF(2), i.e. the
    generating linear state update has period 2^degree - 1."""
    conn, L = _berlekamp_massey(list(bits))
    if L != degree:
        return False
    poly = int(bin(conn)[2:].zfill(L + 1)[::-1], 2)
    order = (1 << degree) - 1
    if _powx(order, poly, degree) != 1:
        return False
    return all(_powx(order // q, poly, degree) != 1 for q in primes)
