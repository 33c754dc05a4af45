"""CPU restatement of the learned preconditioner's network (test infrastructure only).

The network has no counterpart in the reference (its Newton-GMRES is
unpreconditioned, NewtonKrylov.h:138-235); this module states the same
computation as the GPU kernels in numpy so that their forward pass and
parameter gradients can be checked:

    layer l:  out[p, co] = b[co] + sum_{t<5, ci} W[t, co, ci] * in[p + (t - 2) d_l, ci]
              (zero outside the profile of p; d_l the layer's dilation, 1 by
              default), ReLU between layers,
    loss:     mean over positions and real output channels of (out - target)^2.

Parameters are the flat FP32 vector of the library: per layer W[5][C[l+1]][C[l]]
then b[C[l+1]].  With bf16=True the tensors the GPU holds in BF16 (input,
weights, hidden activations, output gradient) are rounded to BF16
(round-to-nearest-even) and products accumulate in FP32-like double sums, as
on the matrix cores; with bf16=False everything is FP64, for finite
differences.
"""
import numpy as np

KT = 5


def to_bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    rounding = ((u >> 16) & 1) + 0x7FFF
    u = ((u + rounding) >> 16) << 16
    out = u.astype(np.uint32).view(np.float32)
    return np.where(np.isfinite(a), out, a).reshape(np.shape(x))


def unpack(params, channels):
    """List of (W[5][Cout][Cin], b[Cout]) per layer from the flat vector."""
    out, off = [], 0
    for l in range(len(channels) - 1):
        ci, co = channels[l], channels[l + 1]
        n = KT * co * ci
        W = np.asarray(params[off:off + n], dtype=np.float64).reshape(KT, co, ci)
        off += n
        b = np.asarray(params[off:off + co], dtype=np.float64)
        off += co
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def conv(x, W, b, d=1):
    """x [nprof][Z][Cin] -> [nprof][Z][Cout] (zero padding per profile), taps d apart."""
    nprof, Z, _ = x.shape
    xp = np.zeros((nprof, Z + 4 * d, x.shape[2]))
    xp[:, 2 * d:Z + 2 * d, :] = x
    out = np.zeros((nprof, Z, W.shape[1])) + (b if b is not None else 0.0)
    for t in range(KT):
        out += xp[:, t * d:t * d + Z, :] @ W[t].T
    return out


def conv_input_grad(g, W, d=1):
    """Gradient of sum(g * conv(x, W, d)) with respect to x."""
    nprof, Z, _ = g.shape
    gp = np.zeros((nprof, Z + 4 * d, g.shape[2]))
    gp[:, 2 * d:Z + 2 * d, :] = g
    out = np.zeros((nprof, Z, W.shape[2]))
    for t in range(KT):
        # x[p] feeds out[p - (t - 2) d] through W[t]
        out += gp[:, (4 - t) * d:(4 - t) * d + Z, :] @ W[t]
    return out


def conv_weight_grad(g, x, d=1):
    """Gradients (dW[5][Cout][Cin], db[Cout]) of sum(g * conv(x, W, d) + b)."""
    nprof, Z, Cin = x.shape
    xp = np.zeros((nprof, Z + 4 * d, Cin))
    xp[:, 2 * d:Z + 2 * d, :] = x
    dW = np.zeros((KT, g.shape[2], Cin))
    for t in range(KT):
        dW[t] = np.einsum("pzo,pzi->oi", g, xp[:, t * d:t * d + Z, :])
    return dW, g.sum(axis=(0, 1))


def forward(params, channels, x, bf16=True, dil=None):
    """Returns (output, activations, pre-activations)."""
    layers = unpack(params, channels)
    dil = dil or [1] * len(layers)
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    a = r(x).astype(np.float64)
    acts, pres = [a], []
    for l, (W, b) in enumerate(layers):
        pre = conv(a, r(W).astype(np.float64), b, dil[l])
        pres.append(pre)
        if l < len(layers) - 1:
            a = r(np.maximum(pre, 0.0)).astype(np.float64)
            acts.append(a)
    return pres[-1], acts, pres


def loss_and_grad(params, channels, x, target, creal, bf16=True, dil=None):
    """MSE over the first creal output channels and its gradient w.r.t. params."""
    layers = unpack(params, channels)
    dil = dil or [1] * len(layers)
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    out, acts, pres = forward(params, channels, x, bf16, dil)
    nprof, Z, C = out.shape
    d = np.zeros_like(out)
    d[..., :creal] = out[..., :creal] - target[..., :creal]
    denom = nprof * Z * creal
    loss = float((d ** 2).sum() / denom)
    g = r(2.0 * d / denom).astype(np.float64)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W = r(layers[l][0]).astype(np.float64)
        dW, db = conv_weight_grad(g, acts[l], dil[l])
        grads[l] = (dW, db)
        if l > 0:
            ga = conv_input_grad(g, W, dil[l])
            g = r(np.where(pres[l - 1] > 0, ga, 0.0)).astype(np.float64)
    return loss, pack(grads)
