"""FP64 numpy restatement of the TreePM short-range force and of the learned
sub-cell correction (a 3-layer 3^3 convolution network on the PM meshes).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
kernels, never by the product path.

Conventions (shared with the HIP kernels):

* meshes are numpy arrays ``[iz, iy, ix, channel]`` (flat cell index
  iz*nm^2 + iy*nm + ix, channels last), periodic on every axis;
* a 3^3 convolution tap t = (dz+1)*9 + (dy+1)*3 + (dx+1) reads the input cell
  c + (dz, dy, dx); weights are ``w[t, c_in, c_out]`` (correlation form);
* the network is conv(8->16) SiLU, conv(16->16) SiLU, conv(16->3) linear;
  input channels: density contrast, potential minus its 7^3 box mean, the
  particles' mass-weighted offsets from the cell centre (3) and spread about
  their centre of mass, per mean cell mass, and two zero channels;
* CIC weights at particle positions follow crates/gadget-ng-pm/src/cic.rs:100-170
  (node-centred trilinear weights, periodic wrap of all 8 nodes).

Short range (crates/gadget-ng-treepm/src/short_range.rs):

* pair acceleration G m r_vec / r^3 * S(r) with r^2 = |r_vec|^2 + eps^2
  (short_range.rs:384-398), minimum image (short_range.rs:53-63), bare
  distance cut |r_vec| <= r_cut (short_range.rs:293-296);
* S = erfc(x) for the reference kernel, x = r / (sqrt(2) r_s); the split that
  exactly complements the long-range filter exp(-k^2 r_s^2 / 2)
  (crates/gadget-ng-pm/src/fft_poisson.rs:529-531) is
  S = erfc(x) + (2x/sqrt(pi)) exp(-x^2).  The reference omits the second term.

Pins live in tests/test_learned_oracle.py: convolution against
scipy.ndimage.correlate, gradients against central finite differences, the
short-range sum against brute-force Ewald forces (long range + exact short
range = Newton), the potential against a single-mode closed form.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erfc as _erfc

SQRT_PI = math.sqrt(math.pi)

TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
LAYERS = ((8, 16), (16, 16), (16, 3))


# ---------------------------------------------------------------------------
# bfloat16 rounding (round to nearest even on the float32 bit pattern)
# ---------------------------------------------------------------------------

def bf16_round(a) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float32."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    return out.reshape(f.shape)


def bf16_bits(a) -> np.ndarray:
    """bfloat16 bit patterns (uint16) of bf16_round(a)."""
    return (bf16_round(a).view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(b) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


# ---------------------------------------------------------------------------
# periodic 3^3 convolution and its gradients
# ---------------------------------------------------------------------------

def _shift(x, off):
    """y[c] = x[c + off] on the periodic mesh (axes 0..2 = z, y, x)."""
    dz, dy, dx = off
    return np.roll(x, shift=(-dz, -dy, -dx), axis=(0, 1, 2))


def conv3d(x, w, b=None):
    """z[c, o] = sum_t sum_i x[c + off_t, i] w[t, i, o] + b[o]   (FP64)."""
    x = np.asarray(x, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    cin, cout = w.shape[1], w.shape[2]
    flat = x.shape[0] * x.shape[1] * x.shape[2]
    z = np.zeros((flat, cout))
    for t, off in enumerate(TAPS):
        z += _shift(x, off).reshape(flat, cin) @ w[t]
    if b is not None:
        z += np.asarray(b, dtype=np.float64)
    return z.reshape(x.shape[:3] + (cout,))


def conv3d_backward_data(dz, w):
    """dx[c, i] = sum_t sum_o dz[c - off_t, o] w[t, i, o]   (FP64)."""
    dz = np.asarray(dz, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    cin, cout = w.shape[1], w.shape[2]
    flat = dz.shape[0] * dz.shape[1] * dz.shape[2]
    dx = np.zeros((flat, cin))
    for t, off in enumerate(TAPS):
        dx += _shift(dz, tuple(-o for o in off)).reshape(flat, cout) @ w[t].T
    return dx.reshape(dz.shape[:3] + (cin,))


def conv3d_backward_weights(x, dz):
    """dw[t, i, o] = sum_c x[c + off_t, i] dz[c, o];  db[o] = sum_c dz[c, o]."""
    x = np.asarray(x, dtype=np.float64)
    dz = np.asarray(dz, dtype=np.float64)
    cin, cout = x.shape[3], dz.shape[3]
    flat = x.shape[0] * x.shape[1] * x.shape[2]
    d2 = dz.reshape(flat, cout)
    dw = np.stack([_shift(x, off).reshape(flat, cin).T @ d2 for off in TAPS])
    return dw, d2.sum(axis=0)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


# ---------------------------------------------------------------------------
# the network
# ---------------------------------------------------------------------------

def init_params(seed: int = 1234, scale: float = 1.0):
    """He-style initialisation (normal, std = scale*sqrt(2/fan_in)); the last
    layer starts at a tenth of that so the initial correction is small."""
    g = np.random.default_rng(seed)
    p = {}
    for k, (ci, co) in enumerate(LAYERS, start=1):
        std = scale * math.sqrt(2.0 / (27 * ci)) * (0.1 if k == 3 else 1.0)
        p[f"w{k}"] = g.normal(0.0, std, (27, ci, co))
        p[f"b{k}"] = np.zeros(co)
    return p


def net_forward(x0, params, bf16: bool = False):
    """Forward pass.  With bf16=True the weights, the input and the hidden
    activations are rounded to bfloat16 where the HIP kernels store them, and
    the pre-activations stay at full precision.  Returns a dict with z1, x1,
    z2, x2, y (y = conv3 output, float64)."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    w = {k: r(v) if k.startswith("w") else np.asarray(v, dtype=np.float64) for k, v in params.items()}
    x0 = r(x0)
    z1 = conv3d(x0, w["w1"], w["b1"])
    x1 = r(silu(z1))
    z2 = conv3d(x1, w["w2"], w["b2"])
    x2 = r(silu(z2))
    y = conv3d(x2, w["w3"], w["b3"])
    return {"x0": np.asarray(x0, np.float64), "z1": z1, "x1": np.asarray(x1, np.float64), "z2": z2,
            "x2": np.asarray(x2, np.float64), "y": y, "w": w}


def net_backward(fw, dy, bf16: bool = False):
    """Gradients of a scalar loss with dL/dy = dy.  Rounding points mirror the
    HIP kernels: dy, dz2 and dz1 are stored in bfloat16."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    w = fw["w"]
    g = {}
    dz3 = np.asarray(r(dy), dtype=np.float64)
    g["w3"], g["b3"] = conv3d_backward_weights(fw["x2"], dz3)
    dz2 = np.asarray(r(conv3d_backward_data(dz3, w["w3"]) * silu_grad(fw["z2"])), dtype=np.float64)
    g["w2"], g["b2"] = conv3d_backward_weights(fw["x1"], dz2)
    dz1 = np.asarray(r(conv3d_backward_data(dz2, w["w2"]) * silu_grad(fw["z1"])), dtype=np.float64)
    g["w1"], g["b1"] = conv3d_backward_weights(fw["x0"], dz1)
    g["dz3"], g["dz2"], g["dz1"] = dz3, dz2, dz1
    return g


# ---------------------------------------------------------------------------
# CIC transfer between particles and multi-channel meshes
# ---------------------------------------------------------------------------

def _cic_nodes(pos, box: float, nm: int):
    """Yield (iz, iy, ix, weight) for the 8 nodes of every particle
    (cic.rs:100-170 weights, all indices wrapped)."""
    c = np.asarray(pos, dtype=np.float64) * (nm / box)
    i0 = np.floor(c)
    d = c - i0
    i0 = i0.astype(np.int64) % nm
    for a in (0, 1):
        wz = d[:, 2] if a else 1.0 - d[:, 2]
        iz = (i0[:, 2] + a) % nm
        for b in (0, 1):
            wy = d[:, 1] if b else 1.0 - d[:, 1]
            iy = (i0[:, 1] + b) % nm
            for e in (0, 1):
                wx = d[:, 0] if e else 1.0 - d[:, 0]
                ix = (i0[:, 0] + e) % nm
                yield iz, iy, ix, wz * wy * wx


def cic_gather(mesh, pos, box: float):
    """Per-particle trilinear gather of every channel of mesh[iz, iy, ix, c]."""
    nm = mesh.shape[0]
    out = np.zeros((len(pos), mesh.shape[3]))
    for iz, iy, ix, wt in _cic_nodes(pos, box, nm):
        out += wt[:, None] * mesh[iz, iy, ix]
    return out


def cic_scatter(values, pos, box: float, nm: int):
    """Transpose of cic_gather: mesh[node] += weight * values[p]."""
    values = np.asarray(values, dtype=np.float64)
    mesh = np.zeros((nm, nm, nm, values.shape[1]))
    for iz, iy, ix, wt in _cic_nodes(pos, box, nm):
        np.add.at(mesh, (iz, iy, ix), wt[:, None] * values)
    return mesh


def correction_loss(y, pos, box: float, target, f_ref, f0: float, out_scale: float):
    """Relative squared error of the predicted residual at the particles:

        pred_p = out_scale * CIC(y)(x_p)
        L = (1/P) sum_p |pred_p - target_p|^2 / (|f_ref_p|^2 + f0^2)

    Returns (L, dL/dy) with dL/dy on the mesh (the CIC transpose)."""
    pred = out_scale * cic_gather(y, pos, box)
    wgt = 1.0 / (np.sum(np.asarray(f_ref, np.float64) ** 2, axis=1) + f0 * f0)
    e = pred - target
    P = len(pos)
    loss = float(np.sum(wgt * np.sum(e * e, axis=1)) / P)
    gp = (2.0 / P) * wgt[:, None] * e * out_scale
    return loss, cic_scatter(gp, pos, box, y.shape[0])


def weighted_loss(y, pos, box: float, target, weight, out_scale: float, n_norm: int):
    """The device loss with explicit per-particle weights:
    L = (1/n_norm) sum_p w_p |out_scale CIC(y)(x_p) - target_p|^2; returns
    (L, dL/dy)."""
    pred = out_scale * cic_gather(y, pos, box)
    e = pred - target
    w = np.asarray(weight, np.float64)
    loss = float(np.sum(w * np.sum(e * e, axis=1)) / n_norm)
    gp = (2.0 / n_norm) * w[:, None] * e * out_scale
    return loss, cic_scatter(gp, pos, box, y.shape[0])


# ---------------------------------------------------------------------------
# mesh potential and network input
# ---------------------------------------------------------------------------

def potential_mesh(density, g: float, box: float, r_split: float = 0.0):
    """Phi(x) on the mesh with Phi_hat = -4 pi G rho_hat / k^2 (k = 0 dropped),
    times exp(-k^2 r_s^2 / 2) when r_split > 0; rho = density[iz, iy, ix]
    (mass per unit volume).  Follows fft_poisson.rs:509-549 without the
    gradient; frequencies by numpy's fftfreq (the sign at the Nyquist index
    does not enter k^2)."""
    nm = density.shape[0]
    f = np.fft.fftfreq(nm, d=1.0 / nm) * (2.0 * np.pi / box)
    kz, ky, kx = np.meshgrid(f, f, f, indexing="ij")
    k2 = kx * kx + ky * ky + kz * kz
    rh = np.fft.fftn(np.asarray(density, dtype=np.float64))
    with np.errstate(divide="ignore", invalid="ignore"):
        ph = np.where(k2 > 0, -4.0 * np.pi * g * rh / k2, 0.0)
    if r_split > 0:
        ph = ph * np.exp(-0.5 * k2 * r_split * r_split)
    return np.fft.ifftn(ph).real


def box_mean(a, half: int = 3):
    """Periodic (2 half + 1)^3 box mean of a mesh."""
    out = np.asarray(a, np.float64)
    for ax in range(3):
        out = sum(np.roll(out, s, axis=ax) for s in range(-half, half + 1)) / (2 * half + 1)
    return out


def particle_moments(pos, m, box: float, nm: int):
    """Per cell [iz, iy, ix]: D = sum m (f - 1/2) (3 components), Q = sum m
    |f - 1/2|^2 and S = sum m over the particles whose base cell floor(x nm /
    L) mod nm it is, f = in-cell fraction; returns (D, Q, S)."""
    c = np.asarray(pos, np.float64) * (nm / box)
    i0 = np.floor(c)
    f = c - i0 - 0.5
    i0 = i0.astype(np.int64) % nm
    flat = (i0[:, 2] * nm + i0[:, 1]) * nm + i0[:, 0]
    m = np.asarray(m, np.float64)
    M = nm ** 3
    D = np.stack([np.bincount(flat, m * f[:, a], M) for a in range(3)], axis=1)
    Q = np.bincount(flat, m * np.sum(f * f, axis=1), M)
    S = np.bincount(flat, m, M)
    shape = (nm, nm, nm)
    return D.reshape(shape + (3,)), Q.reshape(shape), S.reshape(shape)


def net_input(mass, potential, rho_mean: float, phi_scale: float, moments=None):
    """Channel 0 = mass/rho_mean - 1 (density contrast), channel 1 =
    (Phi - box7(Phi)) / phi_scale (the potential of the mass within a few
    cells), channels 2-4 = D / rho_mean and channel 5 = (Q - |D|^2/S) /
    rho_mean from particle_moments (zero without moments), channels 6-7 = 0;
    every channel is linear in the mass, as the PM force error is."""
    x = np.zeros(mass.shape + (LAYERS[0][0],))
    x[..., 0] = np.asarray(mass, np.float64) / rho_mean - 1.0
    x[..., 1] = (np.asarray(potential, np.float64) - box_mean(potential)) / phi_scale
    if moments is not None:
        D, Q, S = moments
        x[..., 2:5] = D / rho_mean
        with np.errstate(divide="ignore", invalid="ignore"):
            spread = np.where(S > 0, Q - np.sum(D * D, axis=-1) / np.where(S > 0, S, 1.0), 0.0)
        x[..., 5] = np.maximum(spread, 0.0) / rho_mean
    return x


# ---------------------------------------------------------------------------
# short-range force and Ewald sums
# ---------------------------------------------------------------------------

def split_factor(r, r_split: float, kernel: str = "exact"):
    x = np.asarray(r, dtype=np.float64) / (math.sqrt(2.0) * r_split)
    erfc = _erfc(x)
    if kernel == "reference":
        return erfc
    return erfc + (2.0 / SQRT_PI) * x * np.exp(-x * x)


def minimum_image(d, box: float):
    """short_range.rs:53-63."""
    d = d - box * np.round(d / box)
    d = np.where(d > 0.5 * box, d - box, d)
    return np.where(d < -0.5 * box, d + box, d)


def short_range_direct(pos, mass, box: float, r_split: float, r_cut: float, eps2: float = 0.0,
                       g: float = 1.0, kernel: str = "exact", idx=None):
    """Brute-force O(N K) periodic short-range accelerations of particles idx."""
    pos = np.asarray(pos, dtype=np.float64)
    mass = np.asarray(mass, dtype=np.float64)
    idx = np.arange(len(pos)) if idx is None else np.asarray(idx)
    out = np.zeros((len(idx), 3))
    for k, i in enumerate(idx):
        d = minimum_image(pos - pos[i], box)
        r2b = np.sum(d * d, axis=1)
        ok = (r2b <= r_cut * r_cut)
        ok[i] = False
        d, r2b, m = d[ok], r2b[ok], mass[ok]
        r2 = r2b + eps2
        r = np.sqrt(r2)
        s = split_factor(r, r_split, kernel)
        out[k] = np.sum((g * m * s / (r2 * r))[:, None] * d, axis=0)
    return out


def ewald_accelerations(pos, mass, box: float, g: float = 1.0, alpha=None, n_real: int = 3, k_max: int = 8):
    """Periodic Newtonian accelerations with a uniform neutralising background
    (the k = 0 mode dropped, as the PM solve does), by Ewald summation:
    real space over images |n| <= n_real boxes, Fourier space over
    |k_i| <= k_max (2 pi / L units).  For small N only."""
    pos = np.asarray(pos, dtype=np.float64)
    mass = np.asarray(mass, dtype=np.float64)
    L = float(box)
    a = (5.6 / L) if alpha is None else alpha
    n = len(pos)
    acc = np.zeros((n, 3))
    rng = range(-n_real, n_real + 1)
    images = np.array([(i, j, k) for i in rng for j in rng for k in rng], dtype=np.float64) * L
    kk = np.array([(i, j, k) for i in range(-k_max, k_max + 1) for j in range(-k_max, k_max + 1)
                   for k in range(-k_max, k_max + 1) if (i, j, k) != (0, 0, 0)], dtype=np.float64)
    kv = kk * (2.0 * np.pi / L)
    k2 = np.sum(kv * kv, axis=1)
    kf = (4.0 * np.pi / L ** 3) * np.exp(-k2 / (4.0 * a * a)) / k2
    erfc = _erfc
    for i in range(n):
        for j in range(n):
            rv = pos[j] - pos[i]
            if i != j:
                d = rv[None, :] + images
                r = np.sqrt(np.sum(d * d, axis=1))
                s = erfc(a * r) + (2.0 * a / SQRT_PI) * r * np.exp(-a * a * r * r)
                acc[i] += g * mass[j] * np.sum((s / r ** 3)[:, None] * d, axis=0)
            else:
                d = images[np.any(images != 0, axis=1)]
                r = np.sqrt(np.sum(d * d, axis=1))
                s = erfc(a * r) + (2.0 * a / SQRT_PI) * r * np.exp(-a * a * r * r)
                acc[i] += g * mass[j] * np.sum((s / r ** 3)[:, None] * d, axis=0)  # sums to 0 by symmetry
            acc[i] += g * mass[j] * np.sum((kf * np.sin(kv @ rv))[:, None] * kv, axis=0)
    return acc
