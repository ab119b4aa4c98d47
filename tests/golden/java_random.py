"""java.util.Random, restated from its documented algorithm (the 48-bit LCG of the Java SE API
specification: seed = (seed * 0x5DEECE66D + 0xB) mod 2^48), so fixtures can be generated from the
same seeds and draw sequences as the reference's kernel tests (e.g. `new Random(cardinality)` in
kernels/src/test/java/io/vecruntime/kernels/GroupedAggregationTest.java:76).

nextInt(origin, bound) follows the JDK 17+ default RandomGenerator method (bounded draw by
rejection on nextInt()); nextInt(bound), nextDouble() and nextBoolean() follow java.util.Random's
own specified algorithms.
"""

MULT = 0x5DEECE66D
ADD = 0xB
MASK = (1 << 48) - 1


def _i32(v):
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v & 0x80000000 else v


class JavaRandom:
    def __init__(self, seed):
        self.seed = (seed ^ MULT) & MASK

    def next(self, bits):
        self.seed = (self.seed * MULT + ADD) & MASK
        return _i32(self.seed >> (48 - bits))

    def next_int(self, bound=None):
        if bound is None:
            return self.next(32)
        if bound <= 0:
            raise ValueError("bound must be positive")
        if bound & (bound - 1) == 0:
            return _i32((bound * self.next(31)) >> 31)
        while True:
            bits = self.next(31)
            val = bits % bound
            if _i32(bits - val + (bound - 1)) >= 0:
                return val

    def next_int_range(self, origin, bound):
        r = self.next(32)
        n = _i32(bound - origin)
        m = _i32(n - 1)
        if n > 0 and (n & m) == 0:
            return _i32((r & m) + origin)
        if n > 0:
            u = (r & 0xFFFFFFFF) >> 1
            while True:
                r = u % n
                if _i32(u + m - r) >= 0:
                    break
                u = (self.next(32) & 0xFFFFFFFF) >> 1
            return _i32(r + origin)
        while r < origin or r >= bound:
            r = self.next(32)
        return r

    def next_double(self):
        return ((self.next(26) << 27) + self.next(27)) * (1.0 / (1 << 53))

    def next_boolean(self):
        return self.next(1) != 0

    def next_long(self):
        hi = self.next(32)
        lo = self.next(32)
        v = (hi << 32) + lo
        v &= (1 << 64) - 1
        return v - (1 << 64) if v >= (1 << 63) else v
