"""CPU FP64 restatement of the step-proposal network (8 -> 64 -> 64 -> 64 -> 2,
ReLU) for checking the GPU kernels.  With ``bf16=True`` the weights,
normalised inputs and hidden activations are rounded to BF16 exactly where
the GPU rounds them (round to nearest even of the FP32 value), and the
backward pass treats the rounding as the identity, as the GPU does.
Normalised inputs are clamped to +-IN_CLAMP as on the GPU."""
import numpy as np

N_IN, H, N_OUT = 8, 64, 2
IN_CLAMP = 8.0  # normalised inputs are clamped to +-8 standard deviations
OFF = {}
_o = 0
for name, size in (("W1", H * N_IN), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * H), ("b3", H),
                   ("W4", N_OUT * H), ("b4", N_OUT)):
    OFF[name] = (_o, size)
    _o += size
N_PARAMS = _o


def bf16(x):
    """Round to the nearest BF16 (ties to even) via the FP32 bit pattern."""
    f = np.asarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    g = lambda n: p[OFF[n][0]:OFF[n][0] + OFF[n][1]]
    return (g("W1").reshape(H, N_IN), g("b1"), g("W2").reshape(H, H), g("b2"), g("W3").reshape(H, H), g("b3"),
            g("W4").reshape(N_OUT, H), g("b4"))


def loss_and_grad(params, x, t, in_mean, in_std, out_mean, out_std, bf16_round=True):
    """Mean over samples of 0.5 |y_n - t_n|^2 and its gradient (FP64)."""
    r = bf16 if bf16_round else (lambda v: np.asarray(v, dtype=np.float64))
    p = np.asarray(params, dtype=np.float32) if bf16_round else np.asarray(params, dtype=np.float64)
    # the GPU reads every parameter (biases included) from its BF16-rounded FP32 copy
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(r(p) if bf16_round else p)
    x32 = np.asarray(x, dtype=np.float32)
    if bf16_round:
        istd = (np.float32(1.0) / np.asarray(in_std, dtype=np.float32))
        z = (x32 - np.asarray(in_mean, dtype=np.float32)) * istd
        a0 = bf16(np.clip(z, -IN_CLAMP, IN_CLAMP))
        tn = ((np.asarray(t, dtype=np.float32) - np.asarray(out_mean, dtype=np.float32)) /
              np.asarray(out_std, dtype=np.float32)).astype(np.float64)
    else:
        a0 = np.clip((np.asarray(x, np.float64) - in_mean) / in_std, -IN_CLAMP, IN_CLAMP)
        tn = (np.asarray(t, np.float64) - out_mean) / out_std
    z1 = a0 @ W1.T + b1
    a1 = r(np.maximum(z1, 0))
    z2 = a1 @ W2.T + b2
    a2 = r(np.maximum(z2, 0))
    z3 = a2 @ W3.T + b3
    a3 = r(np.maximum(z3, 0))
    y = a3 @ W4.T + b4
    n = len(x)
    e = y - tn
    loss = 0.5 * np.sum(e * e) / n
    dy = e / n
    gW4, gb4 = dy.T @ a3, dy.sum(0)
    d3 = (dy @ W4) * (a3 > 0)
    gW3, gb3 = d3.T @ a2, d3.sum(0)
    d2 = (d3 @ W3) * (a2 > 0)
    gW2, gb2 = d2.T @ a1, d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)
    gW1, gb1 = d1.T @ a0, d1.sum(0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3, gW4.ravel(), gb4])
    return loss, grad, y * out_std + out_mean
