"""Restatement in numpy of the runtime's partitioned tridiagonal solve (the default for lines
of at most 256 rows: Thomas-style elimination within partitions, parallel cyclic reduction
of the reduced system), operation for operation, vectorised over lines. It is pinned against
dense solves in test_api.py; the GPU kernel is compared against it bit for bit, which checks
the kernel's indexing of partitions, reduced rows and line origins."""
import numpy as np

PP = 16  # partitions per line


def part_solve(a, b, c, d, reduced="pcr"):
    """Solves a_i x_{i-1} + b_i x_i + c_i x_{i+1} = d_i along the last axis of [lines, n]
    arrays (a_0 and c_{n-1} ignored) the way the partitioned kernel does: each of 16
    partitions of ceil(n/16) rows reduced to two rows in its first and last unknowns, the
    32-slot reduced system solved by parallel cyclic reduction (reduced="pcr", the y- and
    z-line kernel) or by a sequential sweep (reduced="thomas", the x-line kernel), the
    interior rows expanded."""
    a, b, c, d = (np.asarray(v, dtype=np.float64) for v in (a, b, c, d))
    L, n = a.shape
    m = (n + PP - 1) // PP
    x = np.zeros_like(d)
    S = 2 * PP
    sa, sb, sc, sd = np.zeros((S, L)), np.ones((S, L)), np.zeros((S, L)), np.zeros((S, L))
    parts = []  # per partition: (start, count, A, C, D)
    for p in range(PP):
        cnt = max(0, min(m, n - p * m))
        if cnt == 0:
            continue
        last = (p + 1) * m >= n
        s0 = p * m
        A, B, C, D = (v[:, s0:s0 + cnt].copy() for v in (a, b, c, d))
        cprev = np.zeros(L)
        dprev = np.zeros(L)
        aprev = -np.ones(L)
        for j in range(1, cnt):
            r = 1.0 / (B[:, j] - A[:, j] * cprev)
            D[:, j] = (D[:, j] - A[:, j] * dprev) * r
            A[:, j] = -(A[:, j] * aprev) * r
            C[:, j] = C[:, j] * r
            cprev, dprev, aprev = C[:, j], D[:, j], A[:, j]
        for j in range(cnt - 3, 0, -1):
            D[:, j] = D[:, j] - C[:, j] * D[:, j + 1]
            A[:, j] = A[:, j] - C[:, j] * A[:, j + 1]
            C[:, j] = -C[:, j] * C[:, j + 1]
        a0, b0, c0, d0 = a[:, s0], b[:, s0], c[:, s0], d[:, s0]
        if cnt == 1:
            r = 1.0 / b0
            fa, fc, fd = a0 * r, c0 * r, d0 * r
        else:
            if cnt == 2:
                d1, a1, e1 = np.zeros(L), np.zeros(L), -np.ones(L)
            else:
                d1, a1, e1 = D[:, 1], A[:, 1], C[:, 1]
            r = 1.0 / (b0 - c0 * a1)
            fa, fc, fd = a0 * r, -(c0 * e1) * r, (d0 - c0 * d1) * r
        f = p if m == 1 else 2 * p  # lines of at most 16 rows: slot i is row i
        sa[f] = 0.0 if p == 0 else fa
        sc[f] = 0.0 if (last and cnt == 1) else fc
        sd[f] = fd
        if cnt >= 2:
            sa[2 * p + 1] = A[:, cnt - 1]
            sc[2 * p + 1] = 0.0 if last else C[:, cnt - 1]
            sd[2 * p + 1] = D[:, cnt - 1]
        parts.append((p, s0, cnt, A, C, D))
    if reduced == "thomas":
        cpl, dpl = [], []
        for k in range(S):
            if k == 0:
                cpl.append(sc[k])
                dpl.append(sd[k])
            else:
                r = 1.0 / (1.0 - sa[k] * cpl[-1])
                cpl.append(sc[k] * r)
                dpl.append((sd[k] - sa[k] * dpl[-1]) * r)
        xs = np.zeros((S, L))
        for k in range(S - 1, -1, -1):
            xs[k] = dpl[k] if k == S - 1 else dpl[k] - cpl[k] * xs[k + 1]
        return expand(x, parts, xs, m)
    # parallel cyclic reduction (identity rows beyond the ends)
    s = 1
    while s < S:
        na, nb, nc, nd = sa.copy(), sb.copy(), sc.copy(), sd.copy()
        for i in range(S):
            if i - s >= 0:
                am, bm, cm, dm = sa[i - s], sb[i - s], sc[i - s], sd[i - s]
            else:
                am, bm, cm, dm = 0.0, 1.0, 0.0, 0.0
            if i + s < S:
                ap, bp, cp, dp = sa[i + s], sb[i + s], sc[i + s], sd[i + s]
            else:
                ap, bp, cp, dp = 0.0, 1.0, 0.0, 0.0
            k1 = sa[i] / bm
            k2 = sc[i] / bp
            na[i] = -(am * k1)
            nc[i] = -(cp * k2)
            nb[i] = sb[i] - cm * k1 - ap * k2
            nd[i] = sd[i] - dm * k1 - dp * k2
        sa, sb, sc, sd = na, nb, nc, nd
        s *= 2
    return expand(x, parts, sd / sb, m)


def expand(x, parts, xs, m):
    """Interior rows from the reduced solution xs (one value per slot)."""
    for p, s0, cnt, A, C, D in parts:
        xf = xs[p if m == 1 else 2 * p]
        xe = xs[2 * p + 1] if cnt >= 2 else xf
        x[:, s0] = xf
        for j in range(1, cnt):
            x[:, s0 + j] = xe if j == cnt - 1 else D[:, j] - A[:, j] * xf - C[:, j] * xe
    return x


def part_dims(inputs, ndim, sd, size, box, inc):
    """The partitioned solve applied to the solve box of flat arrays laid out as the
    apitest entry point lays them out (x fastest); returns the flat result (u + x with inc,
    else d with the box replaced by x)."""
    a, b, c, d, u = [v.reshape(tuple(size[::-1])) for v in inputs]
    res = (u if inc else d).copy()
    sl = tuple(slice(box[2 * k], box[2 * k + 1]) for k in range(ndim))[::-1]
    ax = ndim - 1 - sd
    A, B, C, D = (np.moveaxis(v[sl], ax, -1) for v in (a, b, c, d))
    shp = A.shape
    X = part_solve(*(v.reshape(-1, shp[-1]) for v in (A, B, C, D)),
                   reduced="thomas" if sd == 0 else "pcr").reshape(shp)
    R = np.moveaxis(res[sl], ax, -1)
    if inc:
        R[...] = R + X
    else:
        R[...] = X
    return res.ravel()
