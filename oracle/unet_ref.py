"""ORACLE — test infrastructure only. numpy restatement of the TeaLeaf warm-start U-Net
(ops_amd/apps/tealeaf/native/tealeaf_unet.hip): the same layers, parameter layout and
gradient routing, in float64. With bf16=True the values the device stores in BF16 (input,
weights, layer outputs, pooled maps, the masked output gradients fed to MFMA) are rounded
to BF16 (round to nearest even) at the same points, so the device results can be compared
within accumulation-order noise. With bf16=False the network is plain float64, which the
finite-difference check in the tests pins.
"""
import numpy as np

# (c1 upsampled, c2 same resolution, cout, level, relu)
SPEC = [(0, 16, 16, 0, 1), (0, 16, 32, 1, 1), (0, 32, 48, 2, 1), (0, 48, 64, 3, 1),
        (64, 48, 48, 2, 1), (48, 32, 32, 1, 1), (32, 16, 16, 0, 1), (0, 16, 16, 0, 0)]


def bf16(x):
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout():
    off, out = 0, []
    for c1, c2, co, lv, r in SPEC:
        cin = c1 + c2
        out.append((off, off + 9 * co * cin, co, cin))
        off += 9 * co * cin + co
    return out, off


def unpack(p):
    lay, _ = layout()
    Ws, bs = [], []
    for wo, bo, co, cin in lay:
        Ws.append(p[wo:wo + 9 * co * cin].reshape(co, 9, cin))
        bs.append(p[bo:bo + co])
    return Ws, bs


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def pool(x):
    return 0.25 * ((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + (x[:, 1::2, 0::2] + x[:, 1::2, 1::2]))


def shifted(x, dy, dx):
    """x[b, y + dy, x + dx, :] with zeros outside"""
    B, H, W, C = x.shape
    out = np.zeros_like(x)
    ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
    xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
    out[:, yd, xd] = x[:, ys, xs]
    return out


def conv(x, W, b):
    out = np.zeros(x.shape[:3] + (W.shape[0],))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        out += shifted(x, dy, dx) @ W[:, tap, :].T
    return out + b


def conv_t(g, W):
    """adjoint of conv w.r.t. its input"""
    out = np.zeros(g.shape[:3] + (W.shape[2],))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        out += shifted(g, -dy, -dx) @ W[:, tap, :]
    return out


def wgrad(g, x):
    """d loss / d W[co, tap, ci] = sum_p g[p, co] x[p + off(tap), ci]"""
    co, ci = g.shape[-1], x.shape[-1]
    dW = np.zeros((co, 9, ci))
    g2 = g.reshape(-1, co)
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        dW[:, tap, :] = g2.T @ shifted(x, dy, dx).reshape(-1, ci)
    return dW


def forward(p, inp, bf=True):
    """inp [B][H][W][16]; returns (prediction [B][H][W], cache)"""
    rnd = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    Ws, bs = unpack(np.asarray(p, dtype=np.float64))
    Ws = [rnd(w) for w in Ws]
    x = rnd(inp)
    act, ins = [None] * 8, [None] * 8
    for l, (c1, c2, co, lv, relu) in enumerate(SPEC):
        if l == 0:
            xi = x
        elif l <= 3:
            xi = rnd(pool(act[l - 1]))
        elif l <= 6:
            s1 = 3 if l == 4 else l - 1
            s2 = {4: 2, 5: 1, 6: 0}[l]
            xi = np.concatenate([up(act[s1]), act[s2]], axis=-1)
        else:
            xi = act[6]
        ins[l] = xi
        v = conv(xi, Ws[l], bs[l])
        if relu:
            v = np.maximum(v, 0.0)
        act[l] = v if l == 7 else rnd(v)
        if l == 7:
            pred = v[..., 0]
    return pred, dict(Ws=Ws, act=act, ins=ins, rnd=rnd)


def backward(p, inp, grad_out, bf=True):
    """parameter gradients (same layout as p) for d loss / d prediction = grad_out"""
    _, C = forward(p, inp, bf)
    Ws, act, ins, rnd = C["Ws"], C["act"], C["ins"], C["rnd"]
    lay, n = layout()
    gp = np.zeros(n)
    grad = [np.zeros_like(a) for a in act]
    grad[7][..., 0] = grad_out
    for l in range(7, -1, -1):
        c1, c2, co, lv, relu = SPEC[l]
        g = grad[l] * (act[l] > 0) if relu else grad[l]
        gm = rnd(g)
        wo, bo, _, cin = lay[l]
        gp[bo:bo + co] = g.reshape(-1, co).sum(axis=0)
        gp[wo:wo + 9 * co * cin] = wgrad(gm, ins[l]).ravel()
        if l == 0:
            break
        gin = conv_t(gm, Ws[l])
        if 1 <= l <= 3:
            grad[l - 1] += 0.25 * up(gin)
        elif 4 <= l <= 6:
            s1 = 3 if l == 4 else l - 1
            s2 = {4: 2, 5: 1, 6: 0}[l]
            grad[s2] += gin[..., c1:]
            gu = gin[..., :c1]
            grad[s1] += (gu[:, 0::2, 0::2] + gu[:, 0::2, 1::2]) + (gu[:, 1::2, 0::2] + gu[:, 1::2, 1::2])
        else:
            grad[6] += gin
    return gp
