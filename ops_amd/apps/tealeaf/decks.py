"""TeaLeaf input decks for the benchmark configurations.

Rectangles are drawn from splitmix64 (Steele, Lea & Flood, "Fast splittable pseudorandom
number generators", OOPSLA 2014), not from a C++ distribution, so the same deck comes out
everywhere. For each of the `nrect` rectangles, six numbers are drawn in this order:
two x corners and two y corners uniform on [0, 10] (sorted), density ~ U[0.1, 10],
energy ~ U[1, 25]. States are painted in order (later rectangles win); the background is
density 100, energy 1e-4. Values are written with 17 significant digits, so atof
recovers them exactly.
"""

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))


def rectangles(seed=20261203, nrect=8):
    g = SplitMix64(seed)
    out = []
    for _ in range(nrect):
        xa, xb = g.uniform(0, 10), g.uniform(0, 10)
        ya, yb = g.uniform(0, 10), g.uniform(0, 10)
        rho, e = g.uniform(0.1, 10), g.uniform(1, 25)
        out.append((min(xa, xb), max(xa, xb), min(ya, yb), max(ya, yb), rho, e))
    return out


def tea_deck(cells=4096, steps=10, seed=20261203, nrect=8, eps=1e-10, max_iters=20000,
             solver="tl_use_cg", dt=0.004, extra=()):
    lines = ["*tea", " state 1 density=100.0 energy=0.0001"]
    for k, (x0, x1, y0, y1, rho, e) in enumerate(rectangles(seed, nrect)):
        lines.append(f" state {k + 2} density={rho!r} energy={e!r} geometry=rectangle "
                     f"xmin={x0!r} xmax={x1!r} ymin={y0!r} ymax={y1!r}")
    lines += [f" x_cells={cells}", f" y_cells={cells}", " xmin=0.0", " ymin=0.0",
              " xmax=10.0", " ymax=10.0", f" initial_timestep={dt!r}", f" end_step={steps}",
              f" tl_max_iters={max_iters}", f" {solver}", " tl_coefficient_inverse_density",
              f" tl_eps={eps!r}", " summary_frequency=0", *[f" {x}" for x in extra], "*endtea", ""]
    return "\n".join(lines)
