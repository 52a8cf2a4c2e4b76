"""FP64 reference of the learned-bracket network (test infrastructure).

The same 5-64-64-64-1 SiLU network as pangu_mlp, in numpy FP64.  With
quantize=True every matrix-product operand is rounded to BF16 (round to
nearest even) exactly where the GPU rounds, so the GPU's FP32-accumulated
gradients can be compared tightly; with quantize=False it is the plain
FP64 network whose backward pass is checked against finite differences."""
import numpy as np

NH = 64
SIZES = [("W1", (5, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NH)),
         ("b3", (NH,)), ("W4", (NH,)), ("b4", (1,)), ("mean", (5,)), ("dev", (5,))]


def unpack(p):
    out, o = {}, 0
    for name, shp in SIZES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SIZES])


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFFFFFF
    return ((u >> 16) << 16).astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def loss_and_grad(p, feats, target, quantize=True):
    q = bf16 if quantize else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(p)
    n = target.size
    x = ((np.asarray(feats, np.float32) - P["mean"].astype(np.float32)) / P["dev"].astype(np.float32))
    X = q(x)
    W1, W2, W3, W4 = q(P["W1"]), q(P["W2"]), q(P["W3"]), q(P["W4"])
    Z1 = X @ W1 + P["b1"]
    A1 = q(silu(Z1))
    Z2 = A1 @ W2 + P["b2"]
    A2 = q(silu(Z2))
    Z3 = A2 @ W3 + P["b3"]
    A3 = q(silu(Z3))
    Z4 = A3 @ W4 + P["b4"][0]
    d = Z4 - target
    loss = np.mean(d * d)
    dZ4 = q(2.0 * d / n)
    g = {"mean": np.zeros(5), "dev": np.zeros(5)}
    g["W4"] = A3.T @ dZ4
    g["b4"] = np.array([dZ4.sum()])
    dA3 = np.outer(dZ4, W4)
    dZ3 = q(dA3 * silu_grad(Z3))
    g["W3"], g["b3"] = A2.T @ dZ3, dZ3.sum(0)
    dA2 = dZ3 @ W3.T
    dZ2 = q(dA2 * silu_grad(Z2))
    g["W2"], g["b2"] = A1.T @ dZ2, dZ2.sum(0)
    dA1 = dZ2 @ W2.T
    dZ1 = q(dA1 * silu_grad(Z1))
    g["W1"], g["b1"] = X.T @ dZ1, dZ1.sum(0)
    return loss, pack(g)


def forward(p, feats, quantize=True):
    q = bf16 if quantize else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(p)
    x = ((np.asarray(feats, np.float32) - P["mean"].astype(np.float32)) / P["dev"].astype(np.float32))
    A = q(x)
    for W, b in (("W1", "b1"), ("W2", "b2"), ("W3", "b3")):
        A = q(silu(A @ q(P[W]) + P[b]))
    z = A @ q(P["W4"]) + P["b4"][0]
    return 1.0 / (1.0 + np.exp(-z))
