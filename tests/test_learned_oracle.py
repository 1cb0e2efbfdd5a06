"""Pins of oracle/learned_oracle.py (CPU only): each check is independent of
the formula it tests (another library, finite differences, Ewald sums,
closed forms, adjoint identities)."""
import math

import numpy as np
import pytest
from scipy import ndimage

import learned_oracle as lo


def _rand_mesh(n, c, seed):
    return np.random.default_rng(seed).normal(size=(n, n, n, c))


def test_bf16_round_known_values():
    one = 1.0
    assert lo.bf16_round(np.float32(one)) == one
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert lo.bf16_round(np.float32(1 + 2 ** -8)) == 1.0
    # 1 + 3*2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even -> 1 + 2^-6
    assert lo.bf16_round(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6
    assert lo.bf16_round(np.float32(1 + 2 ** -8 + 2 ** -20)) == 1 + 2 ** -7
    assert lo.bf16_round(np.float32(-2.5)) == -2.5
    b = lo.bf16_bits(np.array([1.0, -2.0, 0.0], np.float32))
    assert list(b) == [0x3F80, 0xC000, 0x0000]
    assert np.array_equal(lo.from_bf16_bits(b), [1.0, -2.0, 0.0])


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    x = np.random.default_rng(3).normal(size=4096).astype(np.float32) * 100
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(lo.bf16_round(x), ref)


@pytest.mark.parametrize("cin,cout", [(2, 16), (16, 3), (3, 5)])
def test_conv3d_matches_scipy_correlate(cin, cout):
    n = 6
    x = _rand_mesh(n, cin, 1)
    g = np.random.default_rng(2)
    w = g.normal(size=(27, cin, cout))
    b = g.normal(size=cout)
    got = lo.conv3d(x, w, b)
    ref = np.zeros_like(got)
    for i in range(cin):
        for o in range(cout):
            k = w[:, i, o].reshape(3, 3, 3)  # index (dz+1, dy+1, dx+1)
            ref[..., o] += ndimage.correlate(x[..., i], k, mode="wrap")
    ref += b
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_conv3d_tap_convention():
    """A single unit weight at tap (dz, dy, dx) = (1, 0, -1) copies x[c + (1, 0, -1)]."""
    n = 5
    x = _rand_mesh(n, 1, 4)
    w = np.zeros((27, 1, 1))
    w[(1 + 1) * 9 + (0 + 1) * 3 + (-1 + 1)] = 1.0
    z = lo.conv3d(x, w)
    for c in [(0, 0, 0), (4, 2, 0), (3, 4, 1)]:
        src = ((c[0] + 1) % n, c[1], (c[2] - 1) % n)
        assert z[c][0] == x[src][0]


def test_backward_data_is_the_adjoint():
    n, cin, cout = 5, 4, 3
    x = _rand_mesh(n, cin, 5)
    dz = _rand_mesh(n, cout, 6)
    w = np.random.default_rng(7).normal(size=(27, cin, cout))
    lhs = np.sum(lo.conv3d(x, w) * dz)
    rhs = np.sum(x * lo.conv3d_backward_data(dz, w))
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_backward_weights_is_the_adjoint():
    n, cin, cout = 5, 4, 3
    x = _rand_mesh(n, cin, 8)
    dz = _rand_mesh(n, cout, 9)
    g = np.random.default_rng(10)
    w = g.normal(size=(27, cin, cout))
    b = g.normal(size=cout)
    dw, db = lo.conv3d_backward_weights(x, dz)
    lhs = np.sum(lo.conv3d(x, w, b) * dz)
    assert lhs == pytest.approx(np.sum(dw * w) + np.sum(db * b), rel=1e-12)


def test_cic_gather_scatter_adjoint_and_mass():
    import oracle as O
    g = np.random.default_rng(11)
    nm, box = 8, 3.0
    pos = g.uniform(0, box, (200, 3))
    pos[:3] = [[0, 0, 0], [box * (1 - 1e-16), 0.5, 2.9999], [1.5, 1.5, 1.5]]
    v = g.normal(size=(200, 2))
    mesh = g.normal(size=(nm, nm, nm, 2))
    lhs = np.sum(lo.cic_gather(mesh, pos, box) * v)
    rhs = np.sum(mesh * lo.cic_scatter(v, pos, box, nm))
    assert lhs == pytest.approx(rhs, rel=1e-12)
    m = g.uniform(0.5, 2, 200)
    dep = lo.cic_scatter(m[:, None], pos, box, nm)[..., 0]
    assert np.allclose(dep, O.cic_assign(pos, m, box, nm), rtol=1e-12, atol=1e-13)


def _loss_of(params, x0, pos, box, target, fref, f0, osc):
    fw = lo.net_forward(x0, params)
    return lo.correction_loss(fw["y"], pos, box, target, fref, f0, osc)[0]


def test_network_gradients_against_finite_differences():
    g = np.random.default_rng(12)
    n, box = 4, 4.0
    x0 = g.normal(size=(n, n, n, lo.LAYERS[0][0]))
    params = lo.init_params(seed=3)
    params["w3"] *= 10.0
    for k in ("b1", "b2", "b3"):
        params[k] = g.normal(scale=0.1, size=params[k].shape)
    pos = g.uniform(0, box, (40, 3))
    target = g.normal(size=(40, 3))
    fref = g.normal(size=(40, 3))
    f0, osc = 0.7, 1.3
    fw = lo.net_forward(x0, params)
    loss, dy = lo.correction_loss(fw["y"], pos, box, target, fref, f0, osc)
    grads = lo.net_backward(fw, dy)
    h = 1e-6
    for name in ("w1", "b1", "w2", "b2", "w3", "b3"):
        flat = params[name].reshape(-1)
        for j in g.choice(flat.size, size=min(6, flat.size), replace=False):
            keep = flat[j]
            flat[j] = keep + h
            lp = _loss_of(params, x0, pos, box, target, fref, f0, osc)
            flat[j] = keep - h
            lm = _loss_of(params, x0, pos, box, target, fref, f0, osc)
            flat[j] = keep
            fd = (lp - lm) / (2 * h)
            an = grads[name].reshape(-1)[j]
            assert an == pytest.approx(fd, rel=2e-5, abs=1e-9), (name, j)


def test_silu_grad_against_finite_differences():
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    fd = (lo.silu(z + h) - lo.silu(z - h)) / (2 * h)
    assert np.allclose(lo.silu_grad(z), fd, rtol=1e-7, atol=1e-9)


def test_exact_split_factor_is_the_derivative_of_the_erfc_potential():
    """S(r) = -r^2 d/dr [erfc(r / sqrt2 r_s) / r], by central differences."""
    from scipy.special import erfc
    rs = 0.7
    r = np.linspace(0.05, 5.0, 200)
    h = 1e-6
    pot = lambda q: erfc(q / (math.sqrt(2) * rs)) / q
    fd = -r * r * (pot(r + h) - pot(r - h)) / (2 * h)
    assert np.allclose(lo.split_factor(r, rs, "exact"), fd, rtol=1e-7, atol=1e-10)
    assert not np.allclose(lo.split_factor(r, rs, "reference"), fd, rtol=1e-2)


def test_ewald_is_alpha_independent_and_conserves_momentum():
    g = np.random.default_rng(13)
    pos = g.uniform(0, 1, (8, 3))
    m = g.uniform(0.5, 1.5, 8)
    a1 = lo.ewald_accelerations(pos, m, 1.0)
    a2 = lo.ewald_accelerations(pos, m, 1.0, alpha=4.0, n_real=4, k_max=6)
    assert np.abs(a1 - a2).max() <= 1e-10 * np.abs(a1).max()
    assert np.abs((m[:, None] * a1).sum(0)).max() <= 1e-12 * np.abs(m[:, None] * a1).max()


def test_ewald_close_pair_is_newtonian_and_lattice_is_force_free():
    pos = np.array([[0.5, 0.5, 0.5], [0.51, 0.5, 0.5]])
    a = lo.ewald_accelerations(pos, np.ones(2), 1.0)
    assert a[0, 0] == pytest.approx(1.0 / 0.01 ** 2, rel=1e-4)
    # a simple cubic lattice of 8 equal masses: every particle sits at a symmetry centre
    lat = np.array([(i, j, k) for i in (0, 0.5) for j in (0, 0.5) for k in (0, 0.5)]) + 0.1
    assert np.abs(lo.ewald_accelerations(lat, np.ones(8), 1.0)).max() < 1e-9


def test_long_range_plus_exact_short_range_is_newton():
    """PM long range (Gaussian filter, CIC deconvolved) + exact short range
    reproduces the Ewald forces; the erfc-only kernel of short_range.rs:384-398
    misses about a third of the force for pairs near r_s."""
    import oracle as O
    g = np.random.default_rng(5)
    L, n, nm = 1.0, 12, 64
    pos = g.uniform(0, L, (n, 3))
    m = g.uniform(0.5, 1.5, n)
    ref = lo.ewald_accelerations(pos, m, L)
    rs = 4 * L / nm
    lr = O.treepm_long_range(pos, m, L, nm, r_split=rs, deconvolve=True, mode=O.MODE_FAITHFUL)
    rms = lambda a: math.sqrt(np.sum((a - ref) ** 2) / np.sum(ref ** 2))
    exact = lr + lo.short_range_direct(pos, m, L, rs, 5 * rs, kernel="exact")
    refk = lr + lo.short_range_direct(pos, m, L, rs, 5 * rs, kernel="reference")
    assert rms(exact) < 3e-3
    assert rms(refk) > 0.1


def test_short_range_cut_and_softening():
    pos = np.array([[0.0, 0.0, 0.0], [0.3, 0.0, 0.0], [0.0, 0.9, 0.0]])
    m = np.array([1.0, 2.0, 3.0])
    box, rs, rc, eps2 = 2.0, 0.2, 0.5, 0.01
    a = lo.short_range_direct(pos, m, box, rs, rc, eps2=eps2, idx=[0])
    r2 = 0.09 + eps2
    r = math.sqrt(r2)
    s = lo.split_factor(r, rs)
    # particle 2 is 0.9 away (beyond r_cut, and its image is 1.1 away): no contribution
    assert a[0] == pytest.approx([2.0 * s * 0.3 / (r2 * r), 0.0, 0.0], rel=1e-12)
    # minimum image: a particle at x = 1.9 acts from -0.1
    a2 = lo.short_range_direct(np.array([[0.0, 0, 0], [1.9, 0, 0]]), np.ones(2), box, rs, rc, idx=[0])
    assert a2[0, 0] < 0


def test_box_mean_and_net_input():
    g = np.random.default_rng(15)
    a = g.normal(size=(9, 9, 9))
    # brute-force 7^3 periodic mean at a few cells
    m = lo.box_mean(a)
    for c in [(0, 0, 0), (8, 4, 1), (3, 3, 7)]:
        idx = [(np.arange(-3, 4) + c[k]) % 9 for k in range(3)]
        assert m[c] == pytest.approx(a[np.ix_(*idx)].mean(), rel=1e-12)
    # a constant potential offset does not change the input
    mass = g.uniform(0, 2, (9, 9, 9))
    x1 = lo.net_input(mass, a, 1.0, 2.0)
    x2 = lo.net_input(mass, a + 123.0, 1.0, 2.0)
    assert np.allclose(x1, x2, atol=1e-12)
    assert np.allclose(x1[..., 0], mass - 1.0)


def test_potential_single_mode_closed_form():
    nm, box, g_, amp = 16, 3.0, 1.7, 0.4
    x = (np.arange(nm) + 0.0) * box / nm
    k = 2 * np.pi * 2 / box
    rho = np.broadcast_to(amp * np.cos(k * x)[None, None, :], (nm, nm, nm))
    phi = lo.potential_mesh(rho, g_, box)
    ref = -4 * np.pi * g_ * amp / k ** 2 * np.cos(k * x)
    assert np.allclose(phi[3, 5], ref, rtol=1e-12, atol=1e-12)
    rs = 0.3
    phil = lo.potential_mesh(rho, g_, box, r_split=rs)
    assert np.allclose(phil[0, 0], ref * math.exp(-0.5 * k * k * rs * rs), rtol=1e-12, atol=1e-12)


def test_potential_gradient_matches_the_oracle_force_meshes():
    """-grad Phi taken spectrally (Nyquist derivative zeroed) equals the C
    oracle's force meshes from the same density."""
    import oracle as O
    nm, box = 16, 5.0
    g = np.random.default_rng(14)
    mass = g.uniform(0, 1, (nm, nm, nm))
    dens = mass / (box / nm) ** 3
    phi = lo.potential_mesh(dens, 1.0, box)
    f = np.fft.fftfreq(nm, d=1.0 / nm)
    f[nm // 2] = 0.0
    kk = f * 2 * np.pi / box
    ph = np.fft.fftn(phi)
    fx = np.fft.ifftn(-1j * kk[None, None, :] * ph).real
    fz = np.fft.ifftn(-1j * kk[:, None, None] * ph).real
    ox, oy, oz = O.solve_forces(mass, 1.0, nm, box, mode=O.MODE_FAITHFUL)
    assert np.allclose(fx, ox, rtol=1e-10, atol=1e-10 * np.abs(ox).max())
    assert np.allclose(fz, oz, rtol=1e-10, atol=1e-10 * np.abs(oz).max())


def test_particle_moments_against_per_particle_sums():
    """particle_moments against a loop over particles, and the spread channel
    against the direct sum m |f - f_com|^2 (parallel-axis theorem)."""
    g = np.random.default_rng(8)
    nm, box = 4, 2.0
    pos = np.concatenate([g.uniform(-box, 2 * box, (60, 3)), g.normal(1.1, 0.05, (30, 3))])
    m = g.uniform(0.2, 2.0, len(pos))
    D, Q, S = lo.particle_moments(pos, m, box, nm)
    D2, Q2, S2 = np.zeros((nm, nm, nm, 3)), np.zeros((nm, nm, nm)), np.zeros((nm, nm, nm))
    groups = {}
    for p, mp in zip(pos, m):
        c = [p[a] * nm / box for a in range(3)]
        i = [int(math.floor(v)) % nm for v in c]
        f = np.array([v - math.floor(v) for v in c])
        key = (i[2], i[1], i[0])
        D2[key] += mp * (f - 0.5)
        Q2[key] += mp * np.sum((f - 0.5) ** 2)
        S2[key] += mp
        groups.setdefault(key, []).append((mp, f))
    assert np.allclose(D, D2, atol=1e-12) and np.allclose(Q, Q2, atol=1e-12) and np.allclose(S, S2, atol=1e-12)
    x = lo.net_input(np.ones((nm, nm, nm)), np.zeros((nm, nm, nm)), 0.5, 1.0, moments=(D, Q, S))
    for key, items in groups.items():
        ms = np.array([a for a, _ in items])
        fs = np.array([b for _, b in items])
        com = (ms[:, None] * fs).sum(0) / ms.sum()
        spread = np.sum(ms * np.sum((fs - com) ** 2, axis=1))
        assert x[key][5] == pytest.approx(spread / 0.5, rel=1e-9, abs=1e-12)
        assert np.allclose(x[key][2:5], (ms[:, None] * (fs - 0.5)).sum(0) / 0.5)
    # a lone particle has no spread; particles at cell centres give no moments
    centre = (np.array([[0.5, 1.5, 2.5]]) * box / nm)
    D, Q, S = lo.particle_moments(centre, [3.0], box, nm)
    assert np.abs(D).max() < 1e-12 and np.abs(Q).max() < 1e-12 and S.sum() == 3.0
