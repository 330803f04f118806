"""CPU restatement (numpy, FP64) of the learned samplers' network -- test infrastructure:
only tests/ and __graft_entry__.smoke() use it, as the checker of the device kernels.

The network has no counterpart in the reference; what it must reproduce is its own
definition: inputs x[16] (7 used), three dense layers of 64 with GELU (tanh form), a dense
output layer of 2; parameters packed as W1[64][16] b1 W2[64][64] b2 W3[64][64] b3 W4[2][64] b4.
Pinned by: the gradient against central finite differences, GELU against the exact
x Phi(x) (the tanh form's known accuracy), bf16 rounding against exact bit patterns.
"""
import math

import numpy as np

IN, H, OUT = 16, 64, 2
P_W1 = 0
P_B1 = P_W1 + H * IN
P_W2 = P_B1 + H
P_B2 = P_W2 + H * H
P_W3 = P_B2 + H
P_B3 = P_W3 + H * H
P_W4 = P_B3 + H
P_B4 = P_W4 + OUT * H
NPARAM = P_B4 + OUT
GELU_C = math.sqrt(2.0 / math.pi)
GELU_A = 0.044715


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return [(p[P_W1:P_B1].reshape(H, IN), p[P_B1:P_W2]), (p[P_W2:P_B2].reshape(H, H), p[P_B2:P_W3]),
            (p[P_W3:P_B3].reshape(H, H), p[P_B3:P_W4]), (p[P_W4:P_B4].reshape(OUT, H), p[P_B4:NPARAM])]


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + GELU_A * z ** 3)))


def gelu_grad(z):
    th = np.tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * (1.0 + th) + 0.5 * z * (1.0 - th * th) * GELU_C * (1.0 + 3.0 * GELU_A * z * z)


def forward(p, X):
    h = np.asarray(X, dtype=np.float64)
    layers = unpack(p)
    for W, b in layers[:3]:
        h = gelu(h @ W.T + b)
    W, b = layers[3]
    return h @ W.T + b


def loss_grad(p, X, T, w0=1.0, w1=1.0):
    """mean over the batch of w0 (y0 - t0)^2 + w1 (y1 - t1)^2, and its gradient"""
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    B = X.shape[0]
    layers = unpack(p)
    hs, zs = [X], []
    h = X
    for W, b in layers[:3]:
        z = h @ W.T + b
        zs.append(z)
        h = gelu(z)
        hs.append(h)
    W4, b4 = layers[3]
    y = h @ W4.T + b4
    wv = np.array([w0, w1])
    d = y - T
    loss = float(np.sum(wv * d * d) / B)
    dz = 2.0 * wv * d / B
    grads = []
    dW = dz.T @ hs[3]
    grads.append((dW, dz.sum(axis=0)))
    dh = dz @ W4
    for k in (2, 1, 0):
        dz = dh * gelu_grad(zs[k])
        W, _ = layers[k]
        grads.append((dz.T @ hs[k], dz.sum(axis=0)))
        dh = dz @ W
    grads = grads[::-1]
    g = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, g


def bf16(x):
    """round to bfloat16 (nearest even), returned as float32 values"""
    a = np.asarray(x, dtype=np.float32).copy()
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def forward_bf16(p, X):
    """The device inference: bf16 weights, inputs and hidden activations, FP32-exact
    products (bf16 x bf16) accumulated here in FP64, biases and GELU in FP32."""
    layers = unpack(np.asarray(p, dtype=np.float32).astype(np.float64))
    h = bf16(X).astype(np.float64)
    for W, b in layers[:3]:
        z = h @ bf16(W).astype(np.float64).T + b.astype(np.float32)
        h = bf16(gelu(z.astype(np.float32).astype(np.float64))).astype(np.float64)
    W, b = layers[3]
    return h @ bf16(W).astype(np.float64).T + b.astype(np.float32)


def features(T_eV, Z, u1, u2, phase=0.0):
    """network inputs of samples at energies T_eV (eV) on targets Z with uniforms u1, u2 and
    the phase of T in the Seltzer-Berger energy grid"""
    T_eV = np.asarray(T_eV, dtype=np.float64)
    X = np.zeros((T_eV.size, IN), dtype=np.float32)
    X[:, 0] = (np.log10(T_eV) - 6.0) * 0.5
    Z = np.broadcast_to(np.asarray(Z), T_eV.shape)
    X[:, 1] = Z == 7
    X[:, 2] = Z == 8
    X[:, 3] = Z == 18
    X[:, 4] = 2.0 * np.asarray(u1) - 1.0
    X[:, 5] = 2.0 * np.asarray(u2) - 1.0
    X[:, 6] = 2.0 * np.asarray(phase) - 1.0
    return X
