"""CPU restatement of the value network's arithmetic.

TEST INFRASTRUCTURE ONLY (the gradient checker of the GPU kernels; the
product never imports it). Float64 numpy with the GPU kernels' BF16 rounding
points made explicit, so the two differ only by float32 accumulation order:

  input rows, weights            -> bf16
  hidden layer l < last          -> bf16(relu(x W^T + b))
  output                         -> x W^T + b (float)
  loss                           masked Huber (delta 1) / sum(mask)
  d(output)                      -> bf16(mask * huber'(d) / sum(mask))
  dW_l = dZ_l^T A_l, db_l = colsum(dZ_l)
  dZ_{l-1}                       -> bf16((dZ_l W_l) * [A_l > 0])

With `quantize=False` it is the plain float64 MLP, whose gradients
tests/test_net_cpu.py pins against central finite differences.
"""
import numpy as np

IN, HID, OUT, LAYERS = 2706, 1024, 2652, 5


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def shapes(din=IN, hid=HID, dout=OUT, layers=LAYERS):
    dims = [din] + [hid] * (layers - 1) + [dout]
    return [(dims[l + 1], dims[l]) for l in range(layers)]


def unpack(params, shp):
    out, q = [], 0
    for (o, i) in shp:
        w = np.asarray(params[q:q + o * i], np.float64).reshape(o, i)
        q += o * i
        b = np.asarray(params[q:q + o], np.float64)
        q += o
        out.append((w, b))
    assert q == len(params)
    return out


def pack(grads):
    return np.concatenate([np.concatenate([w.ravel(), b]) for w, b in grads])


def forward(params, x, shp=None, quantize=True):
    shp = shp or shapes()
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params, shp)
    acts = [q(x)]
    for l, (w, b) in enumerate(layers):
        z = acts[-1] @ q(w).T + b
        if l + 1 < len(layers):
            acts.append(q(np.maximum(z, 0.0)))
        else:
            return acts, z


def huber(d):
    a = np.abs(d)
    return np.where(a <= 1.0, 0.5 * d * d, a - 0.5), np.where(a <= 1.0, d, np.sign(d))


def gradients(params, x, target, mask, shp=None, quantize=True):
    shp = shp or shapes()
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params, shp)
    acts, y = forward(params, x, shp, quantize)
    inv = 1.0 / np.sum(mask)
    h, dh = huber(y - target)
    loss = float(np.sum(mask * h) * inv)
    dz = q(mask * dh * inv)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        w, _ = layers[l]
        grads[l] = (dz.T @ acts[l], dz.sum(axis=0))
        if l > 0:
            dz = q((dz @ q(w)) * (acts[l] > 0))
    return loss, pack(grads)


def adam(params, grads, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1): moments, bias-corrected
    by 1 - b^t, and the update lr * m_hat / (sqrt(v_hat) + eps). Returns the
    new (params, m, v) in float64."""
    g = np.asarray(grads, np.float64)
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return np.asarray(params, np.float64) - lr * mh / (np.sqrt(vh) + eps), m, v
