"""Batched ADI configuration: block extents drawn with splitmix64.

Each of `nblocks` blocks gets three extents (x, y, z) uniform on {lo..hi}, drawn as
lo + next() % (hi - lo + 1) in that order, block after block. The initial field is drawn by
the application itself from a fresh splitmix64 with the same seed (adi_init, mode 1).
"""
from ops_amd.apps.tealeaf.decks import SplitMix64


def adi_sizes(seed=5, nblocks=48, lo=64, hi=256):
    g = SplitMix64(seed)
    out = []
    for _ in range(nblocks):
        out.append(tuple(lo + g.next() % (hi - lo + 1) for _ in range(3)))
    return out
