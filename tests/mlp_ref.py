"""CPU FP64 restatement of the surrogate MLP (test infrastructure).

Same network and the same BF16 roundings as the GPU kernels (features,
weights, hidden activations and the dZ operands of the backward products are
rounded to BF16; everything else, and every accumulation, is FP64 here), so
the GPU gradients differ from these only by FP32 summation order.
"""
import numpy as np

D_IN, N_FREQ, D_FF, WIDTH, N_HID = 16, 32, 64, 256, 6


ROUND = True   # tests switch the BF16 roundings off to check the gradients by finite differences


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    if not ROUND:
        return np.asarray(a, np.float64)
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout(n_out=1):
    off, o = {}, 0
    for l in range(N_HID):
        k = D_FF if l == 0 else WIDTH
        off["W%d" % l] = (o, (k, WIDTH))
        o += k * WIDTH
        off["b%d" % l] = (o, (WIDTH,))
        o += WIDTH
    off["Wout"] = (o, (WIDTH, n_out) if n_out > 1 else (WIDTH,))
    o += WIDTH * n_out
    off["bout"] = (o, (n_out,))
    o += n_out
    return off, o


def unpack(p, n_out=1):
    off, n = layout(n_out)
    assert p.size == n
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in off.items()}


def _erf(x):
    from scipy.special import erf
    return erf(x)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def features(x, Bf, front=0):
    """Front 0: [sin, cos](2 pi B x) interleaved; front 1: the 16 inputs, then
    [sin, cos] of the first 24 projections."""
    # the projection is summed in FP32 in input order, as the feature kernel does
    x32, B32 = np.asarray(x, np.float32), np.asarray(Bf, np.float32)
    s = np.zeros((x32.shape[0], N_FREQ), np.float32)
    for d in range(D_IN):
        s = (s + B32[None, :, d] * x32[:, d:d + 1]).astype(np.float32)
    s = s.astype(np.float64)
    F = np.empty((x.shape[0], D_FF))
    if front == 0:
        F[:, 0::2] = np.sin(2 * np.pi * s)
        F[:, 1::2] = np.cos(2 * np.pi * s)
    else:
        nf = (D_FF - D_IN) // 2
        F[:, :D_IN] = x32
        F[:, D_IN::2] = np.sin(2 * np.pi * s[:, :nf])
        F[:, D_IN + 1::2] = np.cos(2 * np.pi * s[:, :nf])
    return bf16(F)


def forward_backward(p, Bf, x, y, n_out=1, front=0):
    """One member: returns (out, loss, grads as a flat array). y [n] for one
    output, [n][n_out] otherwise; the loss is the mean over all outputs."""
    P = unpack(p.astype(np.float64), n_out)
    Wb = {k: bf16(v) if k.startswith("W") else v for k, v in P.items()}
    H = [features(x, Bf, front)]
    Z = []
    for l in range(N_HID):
        z = H[-1] @ Wb["W%d" % l] + P["b%d" % l]
        Z.append(z)
        H.append(bf16(gelu(z)))
    if n_out == 1:
        out = H[-1] @ Wb["Wout"] + P["bout"][0]
        d = out - y
        loss = np.mean(d * d)
        g = 2.0 * d / d.size
        G = {"Wout": H[-1].T @ g, "bout": np.array([g.sum()])}
        dH = np.outer(g, Wb["Wout"])
    else:
        out = H[-1] @ Wb["Wout"] + P["bout"][None, :]
        d = out - y
        loss = np.mean(d * d)
        g = 2.0 * d / d.size
        gb = bf16(g)   # the output layer's backward products read dOut in BF16
        G = {"Wout": H[-1].T @ gb, "bout": g.sum(0)}
        dH = gb @ Wb["Wout"].T
    for l in range(N_HID - 1, -1, -1):
        dZ = dH * gelu_d(Z[l])
        G["b%d" % l] = dZ.sum(0)
        dZb = bf16(dZ)
        G["W%d" % l] = H[l].T @ dZb
        if l > 0:
            dH = dZb @ Wb["W%d" % l].T
    off, ntot = layout(n_out)
    flat = np.zeros(ntot)
    for k, (o, s) in off.items():
        flat[o:o + int(np.prod(s))] = np.asarray(G[k]).reshape(-1)
    return out, loss, flat
