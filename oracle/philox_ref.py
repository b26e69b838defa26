"""Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
easy as 1, 2, 3", SC 2011) and the device draw streams of the library's input
generation (mgkmc_cpp_amd/csrc/gen_kernels.hip), restated in Python.

TEST INFRASTRUCTURE ONLY: the checker of the device draws.  Pinned by the
published known-answer vectors (tests/test_device_inputs.py).
"""
import math

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    c = list(ctr)
    k0, k1 = key
    for r in range(10):
        if r:
            k0, k1 = (k0 + W0) & MASK, (k1 + W1) & MASK
        p0, p1 = M0 * c[0], M1 * c[2]
        c = [(p1 >> 32) ^ c[1] ^ k0, p1 & MASK, (p0 >> 32) ^ c[3] ^ k1, p0 & MASK]
    return c


class Stream:
    """Words for item i and purpose: counters (i lo, i hi, round, purpose)."""

    def __init__(self, item, purpose, seed):
        self.c0, self.c1 = item & MASK, (item >> 32) & MASK
        self.purpose = purpose
        self.key = (seed & MASK, (seed >> 32) & MASK)
        self.round = 0
        self.buf = []

    def word(self):
        if not self.buf:
            self.buf = philox4x32_10([self.c0, self.c1, self.round, self.purpose], self.key)
            self.round += 1
        return self.buf.pop(0)

    def uniform(self):
        return (self.word() + 0.5) / 4294967296.0

    def normal(self):
        u1, u2 = self.uniform(), self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def barrier_gaussian(i, seed, mean, std):
    return mean + std * Stream(i, 1, seed).normal()


def barrier_rayleigh(i, seed, loc, scale):
    return loc + scale * math.sqrt(-2.0 * math.log(1.0 - Stream(i, 1, seed).uniform()))


def planar_mode(i, seed, gamma0):
    s = Stream(i, 2, seed)
    return 0.5 * gamma0 * s.normal(), 0.5 * gamma0 * s.normal()
