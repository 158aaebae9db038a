"""numpy restatement of the learned node relaxation (features, forward pass,
gradient).  TEST INFRASTRUCTURE ONLY: the checker for bl_mlp.hip.

The network has no counterpart in the reference; the relaxation it imitates
is the mesh generator's equipotential stencil (setup_mesh.F90:797-863),
which oracle/bl_oracle.c restates.  What fixes the functions here:
  * features: closed forms on a uniform square grid (unit offsets, unit
    areas and Jacobians, no one-sweep motion), invariance under translation
    and uniform scaling, and the one-sweep displacement [24..25] against the
    C oracle's relaxation sweep (tests/test_oracle_mlp.py);
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


def features(ring, x, y, nodes=None, rtype=3, om=1.0, tol=0.0, zerocut=1.0e-40):
    """features (len(nodes), 26) float32 and h (float32) of the given nodes
    (default: every node with a stencil): offsets in float64, the rest in
    float32 in the device's operation order; [24..25] one sweep of the
    relaxation rtype (1 LIN1, 2 LIN2, 3 EQUI) with om, tol, zerocut"""
    f32 = np.float32
    ring = np.asarray(ring).reshape(-1, 8)
    if nodes is None:
        nodes = np.nonzero(ring[:, 0] >= 0)[0]
    nb = ring[nodes]
    xn, yn = x[nodes], y[nodes]
    dx = (x[nb] - xn[:, None]).astype(f32)
    dy = (y[nb] - yn[:, None]).astype(f32)
    area = []
    for q in range(4):
        a, b, c = CCW[2 * q], CCW[2 * q + 1], CCW[(2 * q + 2) & 7]
        area.append(f32(0.5) * (dx[:, b] * (dy[:, c] - dy[:, a]) - (dx[:, c] - dx[:, a]) * dy[:, b]))
    h2 = (((area[0] + area[1]) + area[2]) + area[3]) / f32(4.0)
    h = np.sqrt(np.maximum(h2, f32(1e-30)))
    ih = f32(1.0) / h
    u = dx * ih[:, None]
    v = dy * ih[:, None]
    f = np.empty((len(nodes), NIN), f32)
    f[:, 0:16:2] = u
    f[:, 1:16:2] = v
    for q in range(4):
        a, b, c = CCW[2 * q], CCW[2 * q + 1], CCW[(2 * q + 2) & 7]
        f[:, 16 + q] = f32(0.5) * (u[:, b] * (v[:, c] - v[:, a]) - (u[:, c] - u[:, a]) * v[:, b])
        f[:, 20 + q] = u[:, a] * v[:, c] - u[:, c] * v[:, a]
    r = np.zeros((len(nodes), 8), f32)
    moves = np.ones(len(nodes), bool)
    if rtype == 3:
        r_psi, s_psi = f32(0.5) * (v[:, 4] - v[:, 3]), f32(0.5) * (u[:, 4] - u[:, 3])
        r_phi, s_phi = f32(0.5) * (v[:, 1] - v[:, 6]), f32(0.5) * (u[:, 1] - u[:, 6])
        zeta = s_psi * r_phi - s_phi * r_psi
        alph = s_psi * s_psi + r_psi * r_psi
        beta = s_phi * s_psi + r_phi * r_psi
        gamm = s_phi * s_phi + r_phi * r_phi
        moves = ~(np.abs(zeta).astype(np.float64) * (h.astype(np.float64) ** 2) < zerocut)
        fac = f32(2.0) * (alph + gamm)
        r[:, 0] = f32(0.5) * beta / fac
        r[:, 1] = alph / fac
        r[:, 2] = -r[:, 0]
        r[:, 3] = gamm / fac
        r[:, 4] = r[:, 3]
        r[:, 5] = r[:, 2]
        r[:, 6] = r[:, 1]
        r[:, 7] = r[:, 0]
    elif rtype == 1:
        r[:, 1] = r[:, 6] = 0.5
    else:
        r[:, 3] = r[:, 4] = 0.5
    su = np.zeros(len(nodes), f32)
    sv = np.zeros(len(nodes), f32)
    for i in range(8):
        su = su + r[:, i] * u[:, i]
        sv = sv + r[:, i] * v[:, i]
    hd = h.astype(np.float64)
    moves &= (np.abs(sv).astype(np.float64) * hd > tol) | (np.abs(su).astype(np.float64) * hd > tol)
    f[:, 24] = np.where(moves, f32(om) * su, f32(0))
    f[:, 25] = np.where(moves, f32(om) * sv, f32(0))
    return f, h


def standardise(x, norm):
    """the network's input (f - mu) * isig in float32"""
    x = np.asarray(x, np.float32)
    if norm is None:
        return x
    mu, isig = (np.asarray(a, np.float32) for a in norm[:2])
    return (x - mu) * isig


def residual(x, t):
    """what the network learns: the target's difference from one sweep
    x[:, 24:26], in float32"""
    return np.asarray(t, np.float32) - np.asarray(x, np.float32)[:, 24:26]


def stats(x, t):
    """the standardisation of pairs (x, t): per-column mean and (inverse)
    standard deviation of x and of residual(x, t); deviations ~0 keep scale
    1"""
    def ms(a):
        a = np.asarray(a, np.float64)
        m = a.mean(0)
        sd = np.sqrt(np.maximum((a * a).mean(0) - m * m, 0.0))
        return m, sd
    m, sd = ms(x)
    mt, sdt = ms(residual(x, t))
    isig = np.where(sd > 1e-8, 1.0 / np.where(sd > 0, sd, 1.0), 1.0)
    return (m.astype(np.float32), isig.astype(np.float32), mt.astype(np.float32),
            np.where(sdt > 1e-12, sdt, 1.0).astype(np.float32))


def forward(w, x, rounding=True, norm=None):
    """y (n, 2) and the intermediate (z1, a1, z2, a2) in float64; with
    rounding the weights, inputs and hidden activations are BF16 values as
    the matrix cores see them; with norm (mu, isig, mu_t, sig_t) the input is
    standardised first and y is the network's raw (standardised) output"""
    W1, b1, W2, b2, W3, b3 = (np.asarray(a, np.float64) for a in unpack(w))
    x = standardise(x, norm)
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


def predict(w, x, norm=None):
    """the displacement / h the device proposes: x[24..25] + mu_t + sig_t *
    N(x)"""
    y, _ = forward(w, x, True, norm)
    if norm is not None:
        y = norm[2].astype(np.float64) + norm[3].astype(np.float64) * y
    return np.asarray(x, np.float32)[:, 24:26].astype(np.float64) + y


def loss_and_gradient(w, x, t, rounding=True, norm=None):
    """mean((y - r)^2) over samples and outputs, r = residual(x, t), and its
    gradient (float64) with respect to the parameters; the BF16 roundings
    pass gradients straight through; with norm, y and r are standardised"""
    y, (xr, z1, a1, z2, a2, W1, W2, W3) = forward(w, x, rounding, norm)
    t = residual(x, t)
    if norm is not None:
        t = (t - norm[2]) / norm[3]
    t = t.astype(np.float64)
    n = len(x)
    r = y - t
    loss = float(np.sum(r * r) / (n * NOUT))
    d3 = 2.0 * r / (n * NOUT)
    d2 = (d3 @ W3) * (z2 > 0.0)
    d1 = (d2 @ W2) * (z1 > 0.0)
    g = np.concatenate([(d1.T @ xr).ravel(), d1.sum(0), (d2.T @ a1).ravel(), d2.sum(0),
                        (d3.T @ a2).ravel(), d3.sum(0)])
    return loss, g
