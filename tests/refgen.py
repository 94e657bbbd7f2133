"""Random generators of the reference's own unit tests, restated."""

MASK = (1 << 64) - 1


class XorShift64:
    """xorshift64 (13, 7, 17) of the reference's tests (src/dsp/lf.rs:311-317,
    src/dsp/cdef.rs:256-262): rnd(n) = state % n after one update."""

    def __init__(self, seed):
        self.s = seed & MASK

    def rnd(self, n):
        s = self.s
        s ^= (s << 13) & MASK
        s ^= s >> 7
        s ^= (s << 17) & MASK
        self.s = s
        return s % n
