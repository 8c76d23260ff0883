"""Pure-Python big-integer Fp2 arithmetic and brute-force definitions used by
the tests as closed forms (independent of the C oracle and the HIP library)."""
import numpy as np

P = (1 << 61) - 1


def fmul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def fadd(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def fsub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def fpow(a, e):
    r = (1, 0)
    while e:
        if e & 1:
            r = fmul(r, a)
        a = fmul(a, a)
        e >>= 1
    return r


def finv(a):
    n = (a[0] * a[0] + a[1] * a[1]) % P
    ni = pow(n, P - 2, P)
    return (a[0] * ni % P, -a[1] * ni % P)


def pairs(arr):
    """Interleaved uint64 array -> list of (re, im) ints."""
    a = [int(x) for x in np.asarray(arr).reshape(-1)]
    return [(a[2 * i], a[2 * i + 1]) for i in range(len(a) // 2)]


def soa_pairs(re, im):
    return [(int(a), int(b)) for a, b in zip(re, im)]


def chi(x, z):
    """eq(x, z) = prod_j (bit_j(x) ? z_j : 1 - z_j), LSB-first."""
    r = (1, 0)
    for j, zj in enumerate(z):
        r = fmul(r, zj if (x >> j) & 1 else fsub((1, 0), zj))
    return r


def mle(vals, z):
    """Brute-force MLE: sum_x chi(x, z) * vals[x]."""
    acc = (0, 0)
    for x, v in enumerate(vals):
        acc = fadd(acc, fmul(chi(x, z), v))
    return acc


def lagrange_eval(xs, ys, t):
    """Value at t of the polynomial through (xs[i], ys[i]) (xs small ints)."""
    acc = (0, 0)
    for i, (xi, yi) in enumerate(zip(xs, ys)):
        num, den = (1, 0), 1
        for j, xj in enumerate(xs):
            if j != i:
                num = fmul(num, fsub(t, (xj % P, 0)))
                den = den * (xi - xj) % P
        acc = fadd(acc, fmul(yi, fmul(num, (pow(den, P - 2, P), 0))))
    return acc


def rand_elems(rng, n):
    x = rng.integers(0, P, size=(n, 2), dtype=np.int64)
    return [(int(a), int(b)) for a, b in x]


def to_interleaved(elems):
    return np.array([c for e in elems for c in e], dtype=np.uint64)


def to_soa(elems):
    return (np.array([e[0] for e in elems], dtype=np.uint64),
            np.array([e[1] for e in elems], dtype=np.uint64))


# Boolean meaning of each gate type on inputs in {0, 1}
# (TP/libTruePix/gateprover.py gatefn_ of each class, read as truth tables).
GATE_TRUTH = {
    0: lambda x, y: x & y,          # mul
    1: lambda x, y: x + y,          # add (integer sum)
    2: lambda x, y: x - y,          # sub
    3: lambda x, y: x,              # mux left
    4: lambda x, y: y,              # mux right
    5: lambda x, y: x | y,          # or
    6: lambda x, y: x ^ y,          # xor
    7: lambda x, y: 1 - x,          # not
    8: lambda x, y: 1 - (x & y),    # nand
    9: lambda x, y: 1 - (x | y),    # nor
    10: lambda x, y: 1 - (x ^ y),   # nxor
    11: lambda x, y: (1 - x) & y,   # naab
    12: lambda x, y: x,             # pass
}
# Gates whose multilinear form has an xy term (degree 2 in the pair).
GATE_DEGREE2 = {0, 5, 6, 8, 9, 10, 11}
