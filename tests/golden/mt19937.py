"""32-bit Mersenne Twister (Matsumoto & Nishimura 1998) and the libstdc++
mapping of std::uniform_real_distribution<double> onto it.

The reference's tests draw their points with std::mt19937(42) and
std::uniform_real_distribution<double>(a, b) (tests/interpolate/interpolate_01.cpp:160-177,
tests/IBTK/ghost_accumulation_01.cpp:131-135).  This module regenerates those
draws bit-for-bit so the golden outputs can be rebuilt from exact inputs
instead of the 16-digit printed coordinates.
"""


class MT19937:
    N, M = 624, 397

    def __init__(self, seed=5489):
        mt = [0] * self.N
        mt[0] = seed & 0xFFFFFFFF
        for i in range(1, self.N):
            mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.mt = mt
        self.idx = self.N

    def _twist(self):
        mt = self.mt
        for i in range(self.N):
            y = (mt[i] & 0x80000000) | (mt[(i + 1) % self.N] & 0x7FFFFFFF)
            v = mt[(i + self.M) % self.N] ^ (y >> 1)
            if y & 1:
                v ^= 0x9908B0DF
            mt[i] = v
        self.idx = 0

    def __call__(self):
        if self.idx >= self.N:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def generate_canonical(g):
    """std::generate_canonical<double, 53>(g) as libstdc++ implements it for a
    32-bit engine: k = 2 draws, sum = g1 + g2 * 2^32 in double, divided by 2^64,
    clamped below 1."""
    r = 4294967296.0
    s = 0.0
    t = 1.0
    for _ in range(2):
        s = s + float(g()) * t
        t = t * r
    ret = s / t
    if ret >= 1.0:
        import math
        ret = math.nextafter(1.0, 0.0)
    return ret


def uniform_real(g, a, b):
    """std::uniform_real_distribution<double>(a, b)(g) in libstdc++: canonical * (b - a) + a."""
    return generate_canonical(g) * (b - a) + a
