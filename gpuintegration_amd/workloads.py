"""Seeded synthetic Genz workloads.

The generator is std::mt19937_64 (its output sequence is fixed by the C++
standard) with uniforms U = (x >> 11) * 2^-53.  Parameters are drawn per
integral in the order a_1..a_d, u_1..u_d (family and d first when they are
drawn too), so a C++ host can reproduce every workload bit for bit.
"""
import math

_MASK = (1 << 64) - 1


class MT19937_64:
    """std::mt19937_64 (Matsumoto-Nishimura 64-bit Mersenne twister)."""

    N, M = 312, 156
    MATRIX_A = 0xB5026F5AA96619E9
    UPPER, LOWER = 0xFFFFFFFF80000000, 0x7FFFFFFF

    def __init__(self, seed=5489):
        self.mt = [0] * self.N
        self.mt[0] = seed & _MASK
        for i in range(1, self.N):
            prev = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (prev ^ (prev >> 62)) + i) & _MASK
        self.idx = self.N

    def _twist(self):
        mt = self.mt
        for i in range(self.N):
            x = (mt[i] & self.UPPER) | (mt[(i + 1) % self.N] & self.LOWER)
            xa = x >> 1
            if x & 1:
                xa ^= self.MATRIX_A
            mt[i] = mt[(i + self.M) % self.N] ^ xa
        self.idx = 0

    def __call__(self):
        if self.idx >= self.N:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000  # b of std::mt19937_64 ([rand.predef])
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & _MASK

    def uniform(self):
        return (self() >> 11) * (1.0 / 9007199254740992.0)


FAMILIES = ["oscillatory", "product_peak", "corner_peak", "gaussian", "c0", "discontinuous"]


def gaussian_8d(seed=2024):
    """8D Genz Gaussian: a_i = 10 (1 + 0.05 (2U - 1)), u_i = 0.2 + 0.6 U."""
    g = MT19937_64(seed)
    d = 8
    a = [10.0 * (1.0 + 0.05 * (2.0 * g.uniform() - 1.0)) for _ in range(d)]
    u = [0.2 + 0.6 * g.uniform() for _ in range(d)]
    return "gaussian", a, u


def oscillatory_5d(seed=2024):
    """5D Genz oscillatory cos(2 pi u_1 + sum a_i x_i): a_i = 1 + 0.5 U, u_1 = U."""
    g = MT19937_64(seed)
    d = 5
    a = [1.0 + 0.5 * g.uniform() for _ in range(d)]
    u = [g.uniform()] + [0.0] * (d - 1)
    return "oscillatory", a, u


def mixed_6d(family, seed=2024):
    """6D corner peak / product peak: a_i = 5 + 10 U, u_i = U."""
    g = MT19937_64(seed)
    d = 6
    a = [5.0 + 10.0 * g.uniform() for _ in range(d)]
    u = [g.uniform() for _ in range(d)]
    return family, a, u


def batch(count=512, seed=2024):
    """Integrals of uniformly drawn family and d in {3..8}; epsrel log-uniform in
    [1e-6, 1e-3]."""
    g = MT19937_64(seed)
    out = []
    for _ in range(count):
        fam = FAMILIES[min(int(g.uniform() * 6), 5)]
        d = 3 + min(int(g.uniform() * 6), 5)
        if fam == "oscillatory":
            a = [1.0 + 0.5 * g.uniform() for _ in range(d)]
            u = [g.uniform() for _ in range(d)]
        elif fam == "gaussian":
            a = [10.0 * (1.0 + 0.05 * (2.0 * g.uniform() - 1.0)) for _ in range(d)]
            u = [0.2 + 0.6 * g.uniform() for _ in range(d)]
        elif fam == "discontinuous":
            a = [5.0 + 10.0 * g.uniform() for _ in range(d)]
            u = [0.3 + 0.4 * g.uniform() for _ in range(d)]
        else:
            a = [5.0 + 10.0 * g.uniform() for _ in range(d)]
            u = [g.uniform() for _ in range(d)]
        epsrel = 10.0 ** (-3.0 - 3.0 * g.uniform())
        out.append((fam, a, u, epsrel))
    return out


def small_batch(count=512, seed=7):
    """Many short integrals: the draws of batch() with d capped at 6 (the
    first six parameters kept) and epsrel raised to at least 1e-4."""
    return [(fam, a[:6], u[:6], max(eps, 1e-4)) for fam, a, u, eps in batch(count, seed)]


def gaussian_exact(a, u):
    v = 1.0
    for ai, ui in zip(a, u):
        v *= math.sqrt(math.pi) / (2.0 * ai) * (math.erf(ai * (1.0 - ui)) + math.erf(ai * ui))
    return v


def exact(family, a, u):
    """Closed-form integral over the unit cube of a Genz family member (the
    forms of genz_1abs_5d.cuh etc.; every one is a product over axes except
    the oscillatory phase and the corner peak's inclusion-exclusion)."""
    d = len(a)
    if family == "oscillatory":
        # Re[e^{i 2 pi u_1} prod_j (e^{i a_j} - 1) / (i a_j)]
        z = complex(math.cos(2 * math.pi * u[0]), math.sin(2 * math.pi * u[0]))
        for aj in a:
            z *= complex(math.sin(aj) / aj, (1.0 - math.cos(aj)) / aj)
        return z.real
    if family == "product_peak":
        v = 1.0
        for aj, uj in zip(a, u):
            v *= aj * (math.atan(aj * (1.0 - uj)) + math.atan(aj * uj))
        return v
    if family == "corner_peak":
        s = 0.0
        for c in range(1 << d):
            t = 1.0 + sum(a[j] for j in range(d) if (c >> j) & 1)
            s += (-1.0 if bin(c).count("1") & 1 else 1.0) / t
        return s / (math.factorial(d) * math.prod(a))
    if family == "gaussian":
        return gaussian_exact(a, u)
    if family == "c0":
        v = 1.0
        for aj, uj in zip(a, u):
            v *= (2.0 - math.exp(-aj * uj) - math.exp(-aj * (1.0 - uj))) / aj
        return v
    if family == "discontinuous":
        v = 1.0
        for aj, uj in zip(a, u):
            v *= (math.exp(aj * min(max(uj, 0.0), 1.0)) - 1.0) / aj
        return v
    raise ValueError(f"unknown family {family!r}")


def eval_flops_per_point(family, d):
    """Algorithmic FLOPs per rule point: point map (1 FMA per axis), the
    integrand, five weight FMAs.  exp is charged 24, cos 28, a division 4."""
    cf = {
        "gaussian": 4 * d + 24,
        "c0": 4 * d + 24,
        "oscillatory": 2 * d + 1 + 28,
        "product_peak": 4 * d + 1,
        "corner_peak": 2 * d + 5,
        "discontinuous": 3 * d + 24,
    }[family]
    return 2 * d + cf + 10


def mc_flops_per_sample(family, d):
    """Algorithmic FLOPs per m-Cubes sample: the point map (11 per axis:
    stratum offset, bin position, bin width and offset, two FMAs, the
    weight's two products), the integrand, the weight product, the square
    and the two running sums.  Random-number integer work is not counted."""
    return 11 * d + (eval_flops_per_point(family, d) - 2 * d - 10) + 4


def eval_flops_per_region(family, d):
    F = 1 + 8 * d + 2 * d * (d - 1) + 4 * d * (d - 1) + 4 * d * (d - 1) * (d - 2) // 3 + (1 << d)
    return F * eval_flops_per_point(family, d) + 160
