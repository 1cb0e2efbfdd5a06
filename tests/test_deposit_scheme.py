"""CPU emulation of the GPU deposit's tile/class scheme (no GPU needed).

k_deposit_tiles (gadget-ng_amd/csrc/deposit.hpp) gives each tile of T^3
cells to one workgroup.  A particle's base cell fixes its tile and a 3-bit
class (bit a set when the base cell is on the tile's last layer along axis
a); tile u reads its own particles and, from each lower neighbour u - d, the
classes containing every bit of d, as 14 (d, first class, end class) ranges;
a particle of range d adds only the targets whose offset a satisfies
(a & class) == d, at local index 0 along each bit of d.  This test reads the
three range tables from the kernel source and replays the scheme in numpy on
the kernel's own cell records (the FP64 cell index and 43-bit fraction of
cell_fix), for 1, 2 and 3 tiles per axis and inputs crowded onto tile
boundaries: every CIC target must be added exactly once, so the mesh equals
the oracle's deposit (cic.rs:37-93).  It also restates k_tile_ranges (the
per-tile layout of the 14 ranges that the deposit prologue reads) and the
record lookup built on it, and checks that they visit exactly the sorted
slots of each tile's ranges."""
import os
import re

import numpy as np
import pytest

import oracle

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gadget-ng_amd", "csrc",
                   "deposit.hpp")


def _table(name):
    text = open(SRC).read()
    m = re.search(name + r"\[14\]\s*=\s*\{([^}]*)\}", text)
    assert m, name
    return [int(v) for v in m.group(1).split(",")]


D_LIST, LO, HI = _table("kRangeDList"), _table("kRangeLo"), _table("kRangeHi")
FRAC_BITS = 43


def cell_fix(x, box, nm):
    """cell_fix of pm_kernels.hpp in numpy: c = (x nm / L) rem_euclid nm,
    i = floor(c), fraction rounded to 43 bits with the carry to the next cell."""
    c = x * (nm / box)
    out = ~((c >= 0.0) & (c < nm))
    c = np.where(out, np.fmod(c, float(nm)), c)
    c = np.where(c < 0.0, c + nm, c)
    f = np.floor(c)
    i = f.astype(np.int64)
    q = np.rint((c - f) * 2.0 ** FRAC_BITS).astype(np.int64)
    carry = q >= (1 << FRAC_BITS)
    q = np.where(carry, 0, q)
    i = i + carry
    i = np.where(i >= nm, 0, i)
    return i, q


def emulate(pos, mass, box, nm, T):
    nt = nm // T
    cells, fr = zip(*(cell_fix(pos[:, a], box, nm) for a in range(3)))
    ix, iy, iz = cells
    f = [q.astype(np.float64) * 2.0 ** -FRAC_BITS for q in fr]
    tile = ((iz // T) * nt + (iy // T)) * nt + ix // T
    cls = (ix % T == T - 1).astype(int) | ((iy % T == T - 1).astype(int) << 1) | ((iz % T == T - 1).astype(int) << 2)
    rho = np.zeros(nm ** 3)
    hits = np.zeros((pos.shape[0], 8), dtype=np.int64)
    for uz in range(nt):
        for uy in range(nt):
            for ux in range(nt):
                for r in range(14):
                    d = D_LIST[r]
                    dx, dy, dz = d & 1, (d >> 1) & 1, (d >> 2) & 1
                    src = (((uz - dz) % nt) * nt + (uy - dy) % nt) * nt + (ux - dx) % nt
                    sel = np.nonzero((tile == src) & (cls >= LO[r]) & (cls < HI[r]))[0]
                    for p in sel:
                        lx0, ly0, lz0 = ix[p] % T, iy[p] % T, iz[p] % T
                        cx, cy, cz = int(lx0 == T - 1), int(ly0 == T - 1), int(lz0 == T - 1)
                        wx = (1.0 - f[0][p], f[0][p])
                        wy = (1.0 - f[1][p], f[1][p])
                        wz = (1.0 - f[2][p], f[2][p])
                        for a in (0, 1):
                            for b in (0, 1):
                                for c in (0, 1):
                                    if (a & cz) != dz or (b & cy) != dy or (c & cx) != dx:
                                        continue
                                    lz = 0 if dz else lz0 + a
                                    ly = 0 if dy else ly0 + b
                                    lx = 0 if dx else lx0 + c
                                    assert lx < T and ly < T and lz < T
                                    g = ((uz * T + lz) * nm + (uy * T + ly)) * nm + ux * T + lx
                                    rho[g] += mass[p] * wx[c] * wy[b] * wz[a]
                                    hits[p, a * 4 + b * 2 + c] += 1
    return rho.reshape(nm, nm, nm), hits


def test_range_tables_cover_every_class_exactly_once_per_offset():
    """Each (d, class) with d a subset of the class bits is read by exactly one range."""
    for d in range(8):
        for cls in range(8):
            n = sum(1 for r in range(14) if D_LIST[r] == d and LO[r] <= cls < HI[r])
            assert n == (1 if (cls & d) == d else 0), (d, cls)


@pytest.mark.parametrize("nm,T", [(8, 8), (8, 4), (12, 4)])
def test_every_target_added_once_and_mesh_matches_the_oracle(nm, T):
    rng = np.random.default_rng(nm * 10 + T)
    box = float(nm) * 1.5
    h = box / nm
    n = 400
    pos = rng.uniform(-box, 2.0 * box, (n, 3))
    # crowd a third of the particles onto tile faces, edges and corners
    k = n // 3
    edge = rng.integers(1, nm // T + 1, (k, 3)) * T * h - rng.uniform(0.0, 0.999, (k, 3)) * h
    pos[:k] = np.where(rng.random((k, 3)) < 0.7, edge, pos[:k])
    pos[0] = (-1e-17, box, 2.0 * box)
    mass = rng.uniform(0.5, 2.0, n)
    rho, hits = emulate(pos, mass, box, nm, T)
    assert (hits == 1).all()
    ref = oracle.cic_assign(pos, mass, box, nm)
    assert np.abs(rho - ref).max() <= 1e-12 * mass.sum()


def test_cell_fix_matches_the_oracle_cell_coordinate():
    rng = np.random.default_rng(5)
    box, nm = 7.0, 24
    x = np.concatenate([rng.uniform(-3 * box, 3 * box, 5000), [-1e-17, 0.0, box, -box, 1e-300, box * (1 - 1e-16)]])
    i, q = cell_fix(x, box, nm)
    c = np.mod(x * (nm / box), nm)
    ref_i = np.floor(c).astype(np.int64) % nm
    ref_f = c - np.floor(c)
    # same cell unless the fraction rounded up to the next one
    up = (i != ref_i)
    assert np.all(((ref_i[up] + 1) % nm == i[up]) & (q[up] == 0) & (ref_f[up] > 1.0 - 2.0 ** -FRAC_BITS))
    assert np.abs(q[~up] * 2.0 ** -FRAC_BITS - ref_f[~up]).max() <= 2.0 ** -(FRAC_BITS + 1)


def tile_ranges(offset, src, r):
    """k_tile_ranges for tile rank r in numpy: the 14 ranges laid end to end;
    words 0-12 the running ends, 13 the total, 14 the slot delta of range 0,
    15 + k the change of the delta from range k to k + 1."""
    w = np.zeros(32, dtype=np.int64)
    pre, prev = 0, 0
    for k in range(14):
        b = src[8 * r + D_LIST[k]]
        beg = offset[b + LO[k]]
        length = offset[b + HI[k]] - beg
        delta = beg - pre
        pre += length
        if k < 13:
            w[k] = pre
        if k == 0:
            w[14] = delta
        else:
            w[14 + k] = delta - prev
        prev = delta
    w[13] = pre
    return w


@pytest.mark.parametrize("nt,seed", [(1, 0), (2, 1), (3, 2), (4, 3)])
def test_range_table_addresses_every_record_of_the_14_ranges(nt, seed):
    """The deposit prologue's range table and the record lookup built on it
    (rank r of flattened item q = the range ends at or below q; sorted slot =
    q + delta, summed from the steps; a wavefront of 64 items all below the
    first end skips the search) visit, for every tile, exactly the sorted
    slots of its 14 bucket ranges, each with its own range index.  Bucket
    counts are random (many empty) and tile ranks a random permutation."""
    rng = np.random.default_rng(seed)
    ntiles = nt ** 3
    rank = rng.permutation(ntiles)
    counts = rng.integers(0, 40, ntiles * 8) * (rng.random(ntiles * 8) < 0.7)
    offset = np.concatenate([[0], np.cumsum(counts)])
    src = np.zeros(8 * ntiles, dtype=np.int64)
    for t in range(ntiles):
        ux, uy, uz = t % nt, (t // nt) % nt, t // (nt * nt)
        for d in range(8):
            s = ((((uz - (d >> 2 & 1)) % nt) * nt + (uy - (d >> 1 & 1)) % nt) * nt + (ux - (d & 1)) % nt)
            src[8 * rank[t] + d] = rank[s] * 8
    for r in range(ntiles):
        w = tile_ranges(offset, src, r)
        want = sorted((int(q), k) for k in range(14)
                      for q in range(offset[src[8 * r + D_LIST[k]] + LO[k]], offset[src[8 * r + D_LIST[k]] + HI[k]]))
        got = []
        total = int(w[13])
        for q0 in range(0, total, 64):  # one wavefront: items q0 .. q0 + 63
            for q in range(q0, min(q0 + 64, total)):
                if q0 + 63 < w[0]:
                    rr, dl = 0, int(w[14])
                else:
                    ge = [q >= w[k] for k in range(13)]
                    rr = sum(ge)
                    dl = int(w[14]) + sum(int(w[15 + k]) for k in range(13) if ge[k])
                got.append((q + dl, rr))
        assert sorted(got) == want, r
