"""numpy restatement of the learned node relaxation (features, forward pass,
gradient).  TEST INFRASTRUCTURE ONLY: the checker for bl_mlp.hip.

The network has no counterpart in the reference; the relaxation it imitates
is the mesh generator's equipotential stencil (setup_mesh.F90:797-863),
which oracle/bl_oracle.c restates.  What fixes the functions here:
  * features: closed forms on a uniform square grid (unit offsets, unit
    areas and Jacobians, zero mean) and invariance under translation and
    uniform scaling (tests/test_oracle_mlp.py);
  * forward: a hand-made network whose output is known in closed form;
  * gradient: central finite differences of the float64 loss without the
    BF16 rounding, parameter by parameter over every block of the vector.
"""
import numpy as np

NIN, NHID, NOUT = 26, 64, 2
OFF = {"W1": 0, "b1": 1664, "W2": 1728, "b2": 5824, "W3": 5888, "b3": 6016}
NPARAM = 6018
# stencil table order -> counter-clockwise from +k (P0..P7)
CCW = (4, 2, 1, 0, 3, 5, 6, 7)


def unpack(w):
    w = np.asarray(w)
    return (w[0:1664].reshape(64, 26), w[1664:1728], w[1728:5824].reshape(64, 64),
            w[5824:5888], w[5888:6016].reshape(2, 64), w[6016:6018])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([np.ravel(a) for a in (W1, b1, W2, b2, W3, b3)]).astype(np.float32)


def bf16(a):
    """round float32 values to the nearest BF16 (ties to even), as float32"""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def features(ring, x, y, nodes=None):
    """features (len(nodes), 26) float32 and h (float64) of the given nodes
    (default: every node with a stencil), same fp64 operation order as the
    device"""
    ring = np.asarray(ring).reshape(-1, 8)
    if nodes is None:
        nodes = np.nonzero(ring[:, 0] >= 0)[0]
    nb = ring[nodes]
    xn, yn = x[nodes], y[nodes]
    dx = x[nb] - xn[:, None]
    dy = y[nb] - yn[:, None]
    area, jac = [], []
    for q in range(4):
        a, b, c = CCW[2 * q], CCW[2 * q + 1], CCW[(2 * q + 2) & 7]
        area.append(0.5 * (dx[:, b] * (dy[:, c] - dy[:, a]) - (dx[:, c] - dx[:, a]) * dy[:, b]))
        jac.append(dx[:, a] * dy[:, c] - dx[:, c] * dy[:, a])
    h2 = (((area[0] + area[1]) + area[2]) + area[3]) / 4.0
    h = np.sqrt(np.maximum(h2, 1.0e-300))
    ih = 1.0 / h
    ih2 = 1.0 / (h * h)
    f = np.empty((len(nodes), NIN), np.float32)
    sx = np.zeros(len(nodes))
    sy = np.zeros(len(nodes))
    for i in range(8):
        f[:, 2 * i] = dx[:, i] * ih
        f[:, 2 * i + 1] = dy[:, i] * ih
        sx = sx + dx[:, i]
        sy = sy + dy[:, i]
    for q in range(4):
        f[:, 16 + q] = area[q] * ih2
        f[:, 20 + q] = jac[q] * ih2
    f[:, 24] = sx / 8.0 * ih
    f[:, 25] = sy / 8.0 * ih
    return f, h


def forward(w, x, rounding=True):
    """y (n, 2) and the intermediate (z1, a1, z2, a2) in float64; with
    rounding the weights, inputs and hidden activations are BF16 values as
    the matrix cores see them"""
    W1, b1, W2, b2, W3, b3 = (np.asarray(a, np.float64) for a in unpack(w))
    x = np.asarray(x, np.float32)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float32))
    if rounding:
        W1, W2, W3 = (bf16(a).astype(np.float64) for a in (W1, W2, W3))
        x = bf16(x)
    x = x.astype(np.float64)
    z1 = x @ W1.T + b1
    a1 = r(np.maximum(z1, 0.0)).astype(np.float64) if rounding else np.maximum(z1, 0.0)
    z2 = a1 @ W2.T + b2
    a2 = r(np.maximum(z2, 0.0)).astype(np.float64) if rounding else np.maximum(z2, 0.0)
    y = a2 @ W3.T + b3
    return y, (x, z1, a1, z2, a2, W1, W2, W3)


def loss_and_gradient(w, x, t, rounding=True):
    """mean((y - t)^2) over samples and outputs, and its gradient (float64)
    with respect to the parameters; the BF16 roundings pass gradients
    straight through"""
    y, (xr, z1, a1, z2, a2, W1, W2, W3) = forward(w, x, rounding)
    t = np.asarray(t, np.float64)
    n = len(x)
    r = y - t
    loss = float(np.sum(r * r) / (n * NOUT))
    d3 = 2.0 * r / (n * NOUT)
    d2 = (d3 @ W3) * (z2 > 0.0)
    d1 = (d2 @ W2) * (z1 > 0.0)
    g = np.concatenate([(d1.T @ xr).ravel(), d1.sum(0), (d2.T @ a1).ravel(), d2.sum(0),
                        (d3.T @ a2).ravel(), d3.sum(0)])
    return loss, g
