"""FP64 restatement of the reference's Zel'dovich initial conditions.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker, never by the
product path.

Mirrors crates/gadget-ng-core/src/ic_zeldovich.rs (paths relative to the
gadget-ng tree):

* :func:`hash_u64`          murmur3 finaliser (:69-76);
* :func:`mode_seed`         (:80-83);
* :func:`gaussian_std`      StdRng::seed_from_u64, two u53 uniforms,
                            Box-Muller cosine branch (:88-98);
* :func:`mode_int`          (:266-270);
* :func:`delta_kspace`      generate_delta_kspace (:287-369): a Gaussian per
                            mode, DC and every mode with a Nyquist component
                            zero, then delta(-k) = conj delta(k) with the
                            lexicographically smaller index as the source;
* :func:`power_law_sigma`   build_spectrum_fn, PowerLaw branch (:404-415);
* :func:`displacement`      delta_to_displacement (:541-603): Psi_a =
                            Re IFFT(i n_a delta / |n|^2) / n^3 * d;
* :func:`zeldovich`         the particle loop of zeldovich_ics (:733-766):
                            q = (i + 1/2) d, x = rem_euclid(q + s Psi, L),
                            p = s Psi * vel_factor;
* :func:`hubble_param`, :func:`growth_rate_f`, :func:`growth_factor_d`
                            crates/gadget-ng-core/src/cosmology.rs:509-520,
                            :598-610, :638-733;
* second order, crates/gadget-ng-core/src/ic_2lpt.rs:
  :func:`phi_second_derivatives` (:86-136), :func:`source_2lpt` (:146-159),
  :func:`source_to_psi2` (:181-244), :func:`factors_2lpt` (:414-443),
  :func:`zeldovich_2lpt` (the particle loop, :517-553).

The per-mode generator is rand's StdRng (rand 0.10.1: Cargo.lock:1940, the
workspace asks for rand = "0.10" at Cargo.toml:62), which is ChaCha12 (the
chacha20 crate), seeded by rand_core's seed_from_u64.  Neither crate is in the
reference tree; their published algorithms are restated here:

* seed_from_u64: eight PCG32 (XSH-RR 64/32) outputs, the LCG state advanced
  with multiplier 6364136223846793005 and increment 11634580027462260723
  before each output, stored as little-endian words of the 32-byte key;
* ChaCha12: the ChaCha block function (Bernstein 2008; RFC 8439 sec. 2.3) with
  6 double rounds, key words 4..11, a 64-bit block counter in words 12..13
  and a 64-bit stream id in words 14..15, both 0 for a fresh generator;
* next_u64: output words (0, 1), then (2, 3), low word first.

How these are pinned (tests/test_ic_oracle.py): the ChaCha block against the
RFC 8439 sec. 2.3.2 vector and the all-zero ChaCha20 keystream (RFC 8439
A.1 #1); ChaCha12 with next_u64 against the value-stability vector of rand's
StdRng (src/rngs/std.rs, test_stdrng_construction); the PCG32 step and
output function against the pcg32 reference demo (seed 42, sequence 54).
rand_core's increment constant and the order "advance, then output" are
restated from its source and not pinned by a published vector: parity
unpinned for that composition.  Everything downstream of the Gaussians is
pinned by closed forms (single planted modes) and the reference's own tests.

Index convention: the reference's flat index ix*n^2 + iy*n + iz (iz fastest);
arrays here are shaped (n, n, n) and indexed [ix, iy, iz].
"""
from __future__ import annotations

import math

import numpy as np

M32 = np.uint64(0xFFFFFFFF)
CHACHA_CONST = (0x61707865, 0x3320646E, 0x79622D32, 0x6B206574)
PCG_MUL = 6364136223846793005
PCG_INC = 11634580027462260723
CONJ_SALT = 0x9E3779B97F4A7C15


def _u64(x) -> np.ndarray:
    return np.asarray(x, dtype=np.uint64)


def hash_u64(x) -> np.ndarray:
    """murmur3 fmix64 (ic_zeldovich.rs:69-76), elementwise on uint64."""
    x = _u64(x).copy()
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(33)
        x *= np.uint64(0xFF51AFD7ED558CCD)
        x ^= x >> np.uint64(33)
        x *= np.uint64(0xC4CEB9FE1A85EC53)
        x ^= x >> np.uint64(33)
    return x


def mode_seed(seed: int, ix, iy, iz, n: int) -> np.ndarray:
    """hash(seed ^ hash(ix n^2 + iy n + iz)) (ic_zeldovich.rs:80-83)."""
    packed = _u64(ix) * np.uint64(n * n) + _u64(iy) * np.uint64(n) + _u64(iz)
    return hash_u64(np.uint64(seed) ^ hash_u64(packed))


def _rotl(x: np.ndarray, r: int) -> np.ndarray:
    return ((x << np.uint32(r)) | (x >> np.uint32(32 - r))).astype(np.uint32)


def chacha_block(words: np.ndarray, rounds: int) -> np.ndarray:
    """ChaCha block function on a (16, ...) uint32 input state: `rounds`
    rounds (column, then diagonal), then the input added word by word."""
    x = [np.array(w, dtype=np.uint32) for w in words]
    inp = [w.copy() for w in x]

    def qr(a, b, c, d):
        x[a] = (x[a] + x[b]).astype(np.uint32)
        x[d] = _rotl(x[d] ^ x[a], 16)
        x[c] = (x[c] + x[d]).astype(np.uint32)
        x[b] = _rotl(x[b] ^ x[c], 12)
        x[a] = (x[a] + x[b]).astype(np.uint32)
        x[d] = _rotl(x[d] ^ x[a], 8)
        x[c] = (x[c] + x[d]).astype(np.uint32)
        x[b] = _rotl(x[b] ^ x[c], 7)

    with np.errstate(over="ignore"):
        for _ in range(rounds // 2):
            qr(0, 4, 8, 12)
            qr(1, 5, 9, 13)
            qr(2, 6, 10, 14)
            qr(3, 7, 11, 15)
            qr(0, 5, 10, 15)
            qr(1, 6, 11, 12)
            qr(2, 7, 8, 13)
            qr(3, 4, 9, 14)
        return np.stack([(a + b).astype(np.uint32) for a, b in zip(x, inp)])


def pcg32_output(state) -> np.ndarray:
    """PCG XSH-RR output of a 64-bit state."""
    s = _u64(state)
    xs = (((s >> np.uint64(18)) ^ s) >> np.uint64(27)) & M32
    rot = s >> np.uint64(59)
    return ((xs >> rot) | (xs << ((np.uint64(32) - rot) & np.uint64(31)))) & M32


def pcg32_advance(state, mul: int = PCG_MUL, inc: int = PCG_INC) -> np.ndarray:
    with np.errstate(over="ignore"):
        return _u64(state) * np.uint64(mul) + np.uint64(inc)


def seed_from_u64_key(seed) -> np.ndarray:
    """rand_core's seed_from_u64 for a 32-byte seed: the 8 key words."""
    st = _u64(seed)
    words = []
    for _ in range(8):
        st = pcg32_advance(st)
        words.append(pcg32_output(st).astype(np.uint32))
    return np.stack(words)


def stdrng_first_u64s(key: np.ndarray):
    """The first two next_u64 of a fresh ChaCha12 generator with these key
    words (shape (8, ...))."""
    shape = key.shape[1:]
    state = [np.full(shape, c, dtype=np.uint32) for c in CHACHA_CONST]
    state += list(key)
    state += [np.zeros(shape, dtype=np.uint32)] * 4
    out = chacha_block(np.stack(state), 12).astype(np.uint64)
    u0 = out[0] | (out[1] << np.uint64(32))
    u1 = out[2] | (out[3] << np.uint64(32))
    return u0, u1


def gaussian_std(seed) -> np.ndarray:
    """ic_zeldovich.rs:93-98: u1 = max(u53, 1e-300), u2 = u53,
    sqrt(-2 ln u1) cos(2 pi u2)."""
    u0, u1 = stdrng_first_u64s(seed_from_u64_key(seed))
    a = (u0 >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    b = (u1 >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    a = np.maximum(a, 1e-300)
    return np.sqrt(-2.0 * np.log(a)) * np.cos((2.0 * math.pi) * b)


def mode_int(j, n: int) -> np.ndarray:
    """j for j <= n/2, else j - n (ic_zeldovich.rs:266-270)."""
    j = np.asarray(j, dtype=np.int64)
    return np.where(j <= n // 2, j, j - n)


def delta_kspace(n: int, seed: int, sigma_of_n2) -> np.ndarray:
    """generate_delta_kspace: complex128 array (n, n, n) [ix, iy, iz].

    sigma_of_n2(n2) gives spectrum_fn(sqrt(n2)) for integer |n|^2 arrays."""
    idx = np.arange(n)
    ix, iy, iz = np.meshgrid(idx, idx, idx, indexing="ij")
    nx, ny, nz = mode_int(ix, n), mode_int(iy, n), mode_int(iz, n)
    n2 = nx * nx + ny * ny + nz * nz
    half = n // 2
    live = (n2 != 0) & (np.abs(nx) != half) & (np.abs(ny) != half) & (np.abs(nz) != half)
    ms = mode_seed(seed, ix[live], iy[live], iz[live], n)
    g_r = gaussian_std(ms)
    g_i = gaussian_std(hash_u64(ms ^ np.uint64(CONJ_SALT)))
    sig = np.asarray(sigma_of_n2(n2[live]), dtype=np.float64)
    delta = np.zeros((n, n, n), dtype=np.complex128)
    delta[live] = sig * g_r + 1j * (sig * g_i)
    # step 2: the smaller flat index of each pair {k, -k} is the source
    lin = (ix * n + iy) * n + iz
    neg = (((-nx) % n) * n + ((-ny) % n)) * n + ((-nz) % n)
    src = live & (lin < neg)
    flat = delta.reshape(-1)
    flat[neg[src]] = np.conj(flat[lin[src]])
    return delta


def power_law_sigma(n: int, amplitude: float, spectral_index: float):
    """build_spectrum_fn, PowerLaw: A |n|^(n_s/4) / sqrt(n^3), 0 at |n| = 0."""
    inv = 1.0 / math.sqrt(float(n * n * n))

    def f(n2):
        n2 = np.asarray(n2)
        out = np.zeros(n2.shape)
        flat = out.reshape(-1)
        for i, v in enumerate(n2.reshape(-1)):
            na = math.sqrt(float(v))
            flat[i] = 0.0 if na <= 0.0 else amplitude * math.pow(na, spectral_index / 4.0) * inv
        return out
    return f


def sigma_table(n: int, sigma_of_n2) -> np.ndarray:
    """sigma[n2] for every integer |n|^2 of a mode without a Nyquist
    component: n2 = 0 .. 3 (n/2 - 1)^2."""
    top = 3 * (n // 2 - 1) ** 2
    return np.asarray(sigma_of_n2(np.arange(top + 1)), dtype=np.float64)


def displacement(delta: np.ndarray, box: float):
    """delta_to_displacement: [Psi_x, Psi_y, Psi_z], each (n, n, n)."""
    n = delta.shape[0]
    f = mode_int(np.arange(n), n).astype(np.float64)
    nx, ny, nz = np.meshgrid(f, f, f, indexing="ij")
    n2 = nx * nx + ny * ny + nz * nz
    n2s = np.where(n2 == 0.0, 1.0, n2)
    d = box / n
    out = []
    for na in (nx, ny, nz):
        ph = np.where(n2 == 0.0, 0.0, (-na * delta.imag / n2s) + 1j * (na * delta.real / n2s))
        # numpy's inverse includes the 1/n^3 the reference applies by hand
        out.append(np.fft.ifftn(ph).real * d)
    return out


def rem_euclid(x, l):
    r = np.fmod(x, l)
    return np.where(r < 0.0, r + abs(l), r)


def zeldovich(n: int, box: float, psi, vel_factor: float, scale: float = 1.0, lo: int = 0, hi=None):
    """Particles lo .. hi of the lattice: (pos, vel), each (hi - lo, 3)."""
    hi = n ** 3 if hi is None else min(hi, n ** 3)
    gid = np.arange(lo, hi)
    ix, rem = gid // (n * n), gid % (n * n)
    iy, iz = rem // n, rem % n
    d = box / n
    p = np.stack([psi[0][ix, iy, iz], psi[1][ix, iy, iz], psi[2][ix, iy, iz]], axis=1)
    if scale != 1.0:
        p = p * scale
    q = np.stack([(ix + 0.5) * d, (iy + 0.5) * d, (iz + 0.5) * d], axis=1)
    return rem_euclid(q + p, box), p * vel_factor


# -- second order (crates/gadget-ng-core/src/ic_2lpt.rs)

PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def phi_second_derivatives(delta: np.ndarray):
    """phi_second_derivatives (ic_2lpt.rs:86-136): [xx, yy, zz, xy, xz, yz]
    = Re IFFT(-n_a n_b / |n|^2 delta-hat)."""
    n = delta.shape[0]
    f = mode_int(np.arange(n), n).astype(np.float64)
    ns = np.meshgrid(f, f, f, indexing="ij")
    n2 = ns[0] ** 2 + ns[1] ** 2 + ns[2] ** 2
    n2s = np.where(n2 == 0.0, 1.0, n2)
    out = []
    for a, b in PAIRS:
        c = np.where(n2 == 0.0, 0.0, -ns[a] * ns[b] / n2s)
        out.append(np.fft.ifftn(c * delta.real + 1j * (c * delta.imag)).real)
    return out


def source_2lpt(phi):
    """build_2lpt_source (ic_2lpt.rs:146-159)."""
    xx, yy, zz, xy, xz, yz = phi
    return (xx * yy - xy * xy) + (yy * zz - yz * yz) + (zz * xx - xz * xz)


def source_to_psi2(src: np.ndarray, box: float):
    """source_to_psi2 (ic_2lpt.rs:181-244): Psi2_a = Re IFFT(-i n_a S-hat /
    |n|^2) d, every mode with DC or a Nyquist component zero."""
    n = src.shape[0]
    sk = np.fft.fftn(src)
    fi = mode_int(np.arange(n), n)
    f = fi.astype(np.float64)
    ns = np.meshgrid(f, f, f, indexing="ij")
    n2 = ns[0] ** 2 + ns[1] ** 2 + ns[2] ** 2
    nyq = np.abs(fi) == n // 2
    bad = (n2 == 0.0) | nyq[:, None, None] | nyq[None, :, None] | nyq[None, None, :]
    n2s = np.where(bad, 1.0, n2)
    d = box / n
    out = []
    for a in range(3):
        ph = np.where(bad, 0.0, (ns[a] * sk.imag / n2s) + 1j * (-ns[a] * sk.real / n2s))
        out.append(np.fft.ifftn(ph).real * d)
    return out


def factors_2lpt(c, a_init: float, enabled: bool = True):
    """(vel1, vel2, d2_over_d1sq) of zeldovich_2lpt_ics (ic_2lpt.rs:414-443):
    the cosmology is (omega_m, omega_lambda, h0) only; Omega_m(a) from H(a);
    f2 = 2 Omega_m(a)^(6/11); D2/D1^2 = -3/7 Omega_m(a)^(-1/143)."""
    cp = (c[0], c[1], c[2], -1.0, 0.0, 0.0) if enabled else (1.0, 0.0, c[2], -1.0, 0.0, 0.0)
    a = a_init if enabled else 1.0
    h_a = hubble_param(cp, a)
    f1 = growth_rate_f(cp, a)
    if enabled and cp[2] > 0.0:
        hr = (h_a / cp[2]) * (h_a / cp[2])
        om_a = max(cp[0] / (a * a * a) / hr, 0.0)
    else:
        om_a = 1.0
    f2 = 2.0 * math.pow(om_a, 6.0 / 11.0)
    d2 = -3.0 / 7.0 * math.pow(om_a, -1.0 / 143.0)
    return a * a * f1 * h_a, a * a * f2 * h_a * d2, d2


def zeldovich_2lpt(n: int, box: float, psi1, psi2, vel1: float, vel2: float, d2: float, scale: float = 1.0,
                   lo: int = 0, hi=None):
    """The particle loop of zeldovich_2lpt_ics (ic_2lpt.rs:517-553):
    x = rem_euclid(q + s Psi1 + D2/D1^2 s^2 Psi2, L), p = s Psi1 v1 + s^2 Psi2 v2."""
    hi = n ** 3 if hi is None else min(hi, n ** 3)
    gid = np.arange(lo, hi)
    ix, rem = gid // (n * n), gid % (n * n)
    iy, iz = rem // n, rem % n
    d = box / n
    p1 = np.stack([psi1[0][ix, iy, iz], psi1[1][ix, iy, iz], psi1[2][ix, iy, iz]], axis=1)
    p2 = np.stack([psi2[0][ix, iy, iz], psi2[1][ix, iy, iz], psi2[2][ix, iy, iz]], axis=1)
    if scale != 1.0:
        p1 = p1 * scale
        p2 = p2 * (scale * scale)
    q = np.stack([(ix + 0.5) * d, (iy + 0.5) * d, (iz + 0.5) * d], axis=1)
    return rem_euclid(q + p1 + d2 * p2, box), p1 * vel1 + p2 * vel2


# -- background cosmology (cosmology.rs), c = (omega_m, omega_lambda, h0, w0, wa, omega_nu)

def hubble_param(c, a: float) -> float:
    om, ol, h0, w0, wa, onu = c
    if abs(w0 + 1.0) < 1e-10 and abs(wa) < 1e-10:
        ode = ol
    else:
        ode = ol * math.pow(a, -3.0 * (1.0 + w0 + wa)) * math.exp(3.0 * wa * (a - 1.0))
    h_sq = (om + onu) / (a * a * a) + ode
    return h0 * math.sqrt(max(h_sq, 0.0))


def growth_rate_f(c, a: float) -> float:
    """Linder's Omega_m(a)^0.55 (cosmology.rs:598-610)."""
    if a <= 0.0:
        return 1.0
    h = hubble_param(c, a)
    if h <= 0.0 or c[2] <= 0.0:
        return 1.0
    hr = (h / c[2]) * (h / c[2])
    om_a = c[0] / (a * a * a) / hr
    return math.pow(max(om_a, 0.0), 0.55)


def _growth_cpt92(c, a: float) -> float:
    om, ol = c[0], c[1]
    a3 = a * a * a
    den = om + ol * a3
    if den <= 0.0:
        return a
    oma, ola = om / den, ol * a3 / den
    br = math.pow(max(oma, 0.0), 4.0 / 7.0) - ola + (1.0 + oma / 2.0) * (1.0 + ola / 70.0)
    if br <= 0.0:
        return a
    return a * (2.5 * oma / br)


def _omega_m_frac(c, a: float) -> float:
    h = hubble_param(c, a)
    if h <= 0.0 or c[2] <= 0.0:
        return 0.0
    hr = (h / c[2]) * (h / c[2])
    return (c[0] + c[5]) / (a * a * a) / max(hr, 1e-300)


def _dlnh_dlna(c, a: float) -> float:
    eps = max(1e-7, a * 1e-9)
    hp = hubble_param(c, min(a + eps, 1e9))
    hm = hubble_param(c, max(a - eps, 1e-18))
    h = hubble_param(c, a)
    if h <= 0.0:
        return 0.0
    return (a / h) * (hp - hm) / (2.0 * eps)


def _growth_ode(c, a: float) -> float:
    a0 = 1e-5
    if a <= a0:
        return max(a, 0.0)
    steps = 768
    h = (math.log(a) - math.log(a0)) / steps
    la, d, dp = math.log(a0), a0, a0

    def deriv(l, y, yp):
        aa = math.exp(l)
        return yp, -(2.0 + _dlnh_dlna(c, aa)) * yp + 1.5 * _omega_m_frac(c, aa) * y
    for _ in range(steps):
        k1 = deriv(la, d, dp)
        k2 = deriv(la + 0.5 * h, d + 0.5 * h * k1[0], dp + 0.5 * h * k1[1])
        k3 = deriv(la + 0.5 * h, d + 0.5 * h * k2[0], dp + 0.5 * h * k2[1])
        k4 = deriv(la + h, d + h * k3[0], dp + h * k3[1])
        d += h / 6.0 * (k1[0] + 2.0 * k2[0] + 2.0 * k3[0] + k4[0])
        dp += h / 6.0 * (k1[1] + 2.0 * k2[1] + 2.0 * k3[1] + k4[1])
        la += h
    return d


def growth_factor_d(c, a: float) -> float:
    """CPT92 for w0 = -1, wa = 0; otherwise the RK4 growth ODE in ln a."""
    if a <= 0.0:
        return 0.0
    if abs(c[3] + 1.0) < 1e-10 and abs(c[4]) < 1e-10:
        return _growth_cpt92(c, a)
    return _growth_ode(c, a)
