/*
 * pm_oracle.c -- FP64 CPU restatement of gadget-ng's periodic PM force solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only to check or to time the CPU baseline.  The product
 * library never links or calls it.
 *
 * What it restates (reference paths are relative to the gadget-ng tree):
 *   - CIC mass assignment            crates/gadget-ng-pm/src/cic.rs:37-93, 260-292
 *     (per-thread private grids + reduction, the Rayon form: cic.rs:355-401)
 *   - CIC force interpolation        crates/gadget-ng-pm/src/cic.rs:100-170, 298-346
 *   - Poisson solve                  crates/gadget-ng-pm/src/fft_poisson.rs:306-476
 *       rho = mass_cell / V_cell (:319-329), full complex 3-D DFT with passes
 *       x, y, z (:733-771), spectral kernel (:509-549), inverse passes z, y, x
 *       (:775-813), Re() * 1/nm^3 (:466-473)
 *   - wavenumber convention          fft_poisson.rs:724-730 (index nm/2 -> +nm/2)
 *   - filters                        fft_poisson.rs:528-536: exp(-k^2 r_s^2 / 2),
 *                                    exp(-k^2 eps^2)
 *   - solver orchestration           crates/gadget-ng-pm/src/solver.rs:74-140
 *   - TreePM long-range half         crates/gadget-ng-treepm/src/solver.rs:56-62, 83-90
 *
 * Two solve modes:
 *   GNGO_MODE_FAITHFUL  full complex Z2Z + Re(), as the reference does it; the
 *                       FFT passes run on one thread like the reference's
 *                       (the reference parallelises only deposit, spectral loop
 *                       and interpolation).
 *   GNGO_MODE_LEAN      real-to-complex half spectrum with the per-component
 *                       Nyquist rule: F_alpha is zeroed on the n_alpha = nm/2
 *                       plane.  Taking Re() of the full complex inverse equals
 *                       the Hermitian projection of F-hat, which is exactly
 *                       that zeroing, so both modes agree to rounding.  FFT
 *                       passes are threaded; memory is ~32 B per cell.
 *
 * Defined behaviours where the reference is silent or defective:
 *   - all 8 CIC indices wrap modulo nm (the serial reference does not wrap
 *     ix0, cic.rs:54-56,88; the Rayon form does, cic.rs:380-384);
 *   - optional CIC deconvolution (not in the reference): the Green's function
 *     is divided by W_CIC(k)^2 with W_CIC = prod_alpha sinc^2(pi n_alpha / nm),
 *     sinc(x) = sin(x)/x (Hockney & Eastwood: one window for assignment, one
 *     for interpolation).
 *
 * The FFT is this file's own mixed-radix Stockham transform (radices 4, 2, 3,
 * 5 and a generic odd radix), twiddles from long-double sin/cos.  The
 * reference uses rustfft 6.4.1 (crates/gadget-ng-pm/Cargo.toml:21), whose
 * plan-specific last-ulp rounding is not reproduced.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define GNGO_MODE_FAITHFUL 0
#define GNGO_MODE_LEAN 1

static int g_threads = 0; /* 0 = OpenMP default */

void gngo_set_threads(int n) { g_threads = n; }

static int nthreads(void) { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

int gngo_get_threads(void) { return nthreads(); }

/* ------------------------------------------------------------------------ */
/* 1-D mixed-radix Stockham FFT (double complex, interleaved re/im)          */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n;
    int nf;
    int fac[64];
    double *tw; /* tw[2m], tw[2m+1] = cos, -sin of 2 pi m / n */
} fft_plan;

static fft_plan *plan_create(int n) {
    fft_plan *p = (fft_plan *)calloc(1, sizeof(fft_plan));
    p->n = n;
    int r = n;
    while (r % 4 == 0) { p->fac[p->nf++] = 4; r /= 4; }
    while (r % 2 == 0) { p->fac[p->nf++] = 2; r /= 2; }
    for (int f = 3; r > 1; f += 2) {
        while (r % f == 0) { p->fac[p->nf++] = f; r /= f; }
    }
    p->tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int m = 0; m < n; ++m) {
        long double a = two_pi * (long double)m / (long double)n;
        p->tw[2 * m] = (double)cosl(a);
        p->tw[2 * m + 1] = (double)(-sinl(a));
    }
    return p;
}

static void plan_destroy(fft_plan *p) {
    if (!p) return;
    free(p->tw);
    free(p);
}

/* One line, in place; scratch holds 2n doubles.  inverse: e^{+2 pi i}. */
static void fft_line(const fft_plan *p, double *data, double *scratch, int inverse) {
    const int n = p->n;
    if (n == 1) return;
    double *in = data, *out = scratch;
    const double sg = inverse ? -1.0 : 1.0; /* sign applied to sin part */
    int ns = 1;
    double vr[64], vi[64], wr[64], wi[64];
    for (int s = 0; s < p->nf; ++s) {
        const int R = p->fac[s];
        const int nr = n / R;
        const int tstep = n / (ns * R);
        for (int j = 0; j < nr; ++j) {
            const int k = j % ns;
            for (int r = 0; r < R; ++r) {
                const int idx = j + r * nr;
                const int ti = r * k * tstep;
                const double c = p->tw[2 * ti], sn = sg * p->tw[2 * ti + 1];
                const double xr = in[2 * idx], xi = in[2 * idx + 1];
                vr[r] = xr * c - xi * sn;
                vi[r] = xr * sn + xi * c;
            }
            if (R == 4) {
                const double ar = vr[0] + vr[2], ai = vi[0] + vi[2];
                const double br = vr[0] - vr[2], bi = vi[0] - vi[2];
                const double cr = vr[1] + vr[3], ci = vi[1] + vi[3];
                /* d = (v1 - v3) * (-i) forward, * (+i) inverse */
                const double er = vr[1] - vr[3], ei = vi[1] - vi[3];
                const double dr = sg * ei, di = -sg * er;
                wr[0] = ar + cr; wi[0] = ai + ci;
                wr[1] = br + dr; wi[1] = bi + di;
                wr[2] = ar - cr; wi[2] = ai - ci;
                wr[3] = br - dr; wi[3] = bi - di;
            } else if (R == 2) {
                wr[0] = vr[0] + vr[1]; wi[0] = vi[0] + vi[1];
                wr[1] = vr[0] - vr[1]; wi[1] = vi[0] - vi[1];
            } else {
                /* generic radix: X_q = sum_r v_r exp(-+ 2 pi i r q / R) */
                const int step = n / R;
                for (int q = 0; q < R; ++q) {
                    double sr = 0.0, si = 0.0;
                    for (int r = 0; r < R; ++r) {
                        const int ti = ((r * q) % R) * step;
                        const double c = p->tw[2 * ti], sn = sg * p->tw[2 * ti + 1];
                        sr += vr[r] * c - vi[r] * sn;
                        si += vr[r] * sn + vi[r] * c;
                    }
                    wr[q] = sr;
                    wi[q] = si;
                }
            }
            const int base = (j / ns) * ns * R + k;
            for (int r = 0; r < R; ++r) {
                out[2 * (base + r * ns)] = wr[r];
                out[2 * (base + r * ns) + 1] = wi[r];
            }
        }
        double *t = in; in = out; out = t;
        ns *= R;
    }
    if (in != data) memcpy(data, in, sizeof(double) * 2 * (size_t)n);
}

/*
 * Transform every line of a 3-D complex array [nz][ny][nx] (x fastest) along
 * one axis.  Lines along y and z are gathered in blocks of adjacent x so the
 * strided reads touch whole cache lines.
 */
#define GNGO_BLK 8
static void fft_axis(double *data, int nx, int ny, int nz, int axis, int inverse, int threads) {
    const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    fft_plan *p = plan_create(n);
    if (axis == 0) {
        const int64_t lines = (int64_t)ny * nz;
#pragma omp parallel num_threads(threads)
        {
            double *scr = (double *)malloc(sizeof(double) * 2 * (size_t)n);
#pragma omp for schedule(static)
            for (int64_t l = 0; l < lines; ++l) fft_line(p, data + 2 * l * nx, scr, inverse);
            free(scr);
        }
    } else {
        const int64_t estride = axis == 1 ? (int64_t)nx : (int64_t)nx * ny;
        const int outer = axis == 1 ? nz : ny;
        const int64_t ostride = axis == 1 ? (int64_t)nx * ny : (int64_t)nx;
        const int nblk = (nx + GNGO_BLK - 1) / GNGO_BLK;
        const int64_t work = (int64_t)outer * nblk;
#pragma omp parallel num_threads(threads)
        {
            double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)n * GNGO_BLK);
            double *scr = (double *)malloc(sizeof(double) * 2 * (size_t)n);
#pragma omp for schedule(static)
            for (int64_t w = 0; w < work; ++w) {
                const int o = (int)(w / nblk);
                const int x0 = (int)(w % nblk) * GNGO_BLK;
                const int bw = (x0 + GNGO_BLK <= nx) ? GNGO_BLK : nx - x0;
                double *base = data + 2 * (o * ostride + x0);
                for (int i = 0; i < n; ++i)
                    for (int b = 0; b < bw; ++b) {
                        buf[2 * ((int64_t)b * n + i)] = base[2 * (i * estride + b)];
                        buf[2 * ((int64_t)b * n + i) + 1] = base[2 * (i * estride + b) + 1];
                    }
                for (int b = 0; b < bw; ++b) fft_line(p, buf + 2 * (int64_t)b * n, scr, inverse);
                for (int i = 0; i < n; ++i)
                    for (int b = 0; b < bw; ++b) {
                        base[2 * (i * estride + b)] = buf[2 * ((int64_t)b * n + i)];
                        base[2 * (i * estride + b) + 1] = buf[2 * ((int64_t)b * n + i) + 1];
                    }
            }
            free(buf);
            free(scr);
        }
    }
    plan_destroy(p);
}

/* 1-D transform of `count` contiguous lines of length n (exported for tests). */
int gngo_fft1d(double *data, int n, int64_t count, int inverse) {
    if (n <= 0 || count < 0) return 1;
    fft_plan *p = plan_create(n);
    double *scr = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int64_t l = 0; l < count; ++l) fft_line(p, data + 2 * l * n, scr, inverse);
    free(scr);
    plan_destroy(p);
    return 0;
}

/*
 * Full complex 3-D transform of an nm^3 array in the reference pass order:
 * forward x, y, z (fft_poisson.rs:733-771), inverse z, y, x (:775-813).
 * Unnormalised.
 */
int gngo_fft3d(double *data, int nm, int inverse, int threads) {
    if (nm <= 0) return 1;
    if (threads <= 0) threads = nthreads();
    if (!inverse) {
        fft_axis(data, nm, nm, nm, 0, 0, threads);
        fft_axis(data, nm, nm, nm, 1, 0, threads);
        fft_axis(data, nm, nm, nm, 2, 0, threads);
    } else {
        fft_axis(data, nm, nm, nm, 2, 1, threads);
        fft_axis(data, nm, nm, nm, 1, 1, threads);
        fft_axis(data, nm, nm, nm, 0, 1, threads);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* CIC deposit and interpolation                                            */
/* ------------------------------------------------------------------------ */

/*
 * Cell coordinate as in cic.rs:50-60: c = (x * inv_cell).rem_euclid(nm),
 * i0 = floor(c), d = c - i0.  rem_euclid can round up to exactly nm for tiny
 * negative inputs; the index is wrapped so that case lands on node 0.
 */
static inline void cell_coord(double x, double inv_cell, int nm, int *i0, double *d) {
    const double nmf = (double)nm;
    double c = fmod(x * inv_cell, nmf);
    if (c < 0.0) c += nmf;
    double f = floor(c);
    int i = (int)f;
    *d = c - f;
    if (i >= nm) i -= nm;
    if (i < 0) i += nm;
    *i0 = i;
}

/*
 * deposit_mode 0: serial, particle order (cic.rs:37-93).
 * deposit_mode 1: per-thread private nm^3 grids summed afterwards, the
 *                 structure of assign_rayon (cic.rs:355-401).
 * deposit_mode 2: particles bucketed by z plane, plane ranges owned by one
 *                 thread, even ranges then odd ranges (no private grids).
 * Modes differ only in floating-point summation order.
 */
int gngo_cic_assign(const double *pos, const double *mass, int64_t n, double box, int nm,
                    int deposit_mode, double *rho) {
    if (nm <= 0 || n < 0 || box <= 0.0) return 1;
    const int64_t nm2 = (int64_t)nm * nm, nm3 = nm2 * nm;
    const double inv_cell = (double)nm / box;
    const int threads = nthreads();
    memset(rho, 0, sizeof(double) * (size_t)nm3);
    if (deposit_mode == 0 || threads == 1 || n == 0) {
        for (int64_t i = 0; i < n; ++i) {
            int ix, iy, iz;
            double dx, dy, dz;
            cell_coord(pos[3 * i], inv_cell, nm, &ix, &dx);
            cell_coord(pos[3 * i + 1], inv_cell, nm, &iy, &dy);
            cell_coord(pos[3 * i + 2], inv_cell, nm, &iz, &dz);
            const double wx[2] = {1.0 - dx, dx}, wy[2] = {1.0 - dy, dy}, wz[2] = {1.0 - dz, dz};
            const int xs[2] = {ix, (ix + 1) % nm}, ys[2] = {iy, (iy + 1) % nm}, zs[2] = {iz, (iz + 1) % nm};
            const double m = mass[i];
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b)
                    for (int c = 0; c < 2; ++c)
                        rho[zs[a] * nm2 + ys[b] * nm + xs[c]] += m * wx[c] * wy[b] * wz[a];
        }
        return 0;
    }
    if (deposit_mode == 1) {
        double **grids = (double **)calloc((size_t)threads, sizeof(double *));
#pragma omp parallel num_threads(threads)
        {
            const int t = omp_get_thread_num();
            double *g = (double *)calloc((size_t)nm3, sizeof(double));
            grids[t] = g;
#pragma omp for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                int ix, iy, iz;
                double dx, dy, dz;
                cell_coord(pos[3 * i], inv_cell, nm, &ix, &dx);
                cell_coord(pos[3 * i + 1], inv_cell, nm, &iy, &dy);
                cell_coord(pos[3 * i + 2], inv_cell, nm, &iz, &dz);
                const double wx[2] = {1.0 - dx, dx}, wy[2] = {1.0 - dy, dy}, wz[2] = {1.0 - dz, dz};
                const double m = mass[i];
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b)
                        for (int c = 0; c < 2; ++c)
                            g[((iz + a) % nm) * nm2 + ((iy + b) % nm) * nm + (ix + c) % nm] +=
                                m * wx[c] * wy[b] * wz[a];
            }
        }
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t j = 0; j < nm3; ++j) {
            double s = 0.0;
            for (int t = 0; t < threads; ++t)
                if (grids[t]) s += grids[t][j];
            rho[j] = s;
        }
        for (int t = 0; t < threads; ++t) free(grids[t]);
        free(grids);
        return 0;
    }
    /* deposit_mode 2: bucket by base z plane (stable), then plane ranges */
    int64_t *count = (int64_t *)calloc((size_t)nm + 1, sizeof(int64_t));
    int *pz = (int *)malloc(sizeof(int) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        int iz;
        double dz;
        cell_coord(pos[3 * i + 2], inv_cell, nm, &iz, &dz);
        pz[i] = iz;
        count[iz + 1]++;
    }
    for (int z = 0; z < nm; ++z) count[z + 1] += count[z];
    int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)nm);
    memcpy(fill, count, sizeof(int64_t) * (size_t)nm);
    for (int64_t i = 0; i < n; ++i) order[fill[pz[i]]++] = i;
    free(fill);
    free(pz);
    int nchunk = 2 * threads;
    if (nchunk > nm) nchunk = nm - (nm % 2);
    if (nchunk < 2) nchunk = 1;
    for (int phase = 0; phase < (nchunk > 1 ? 2 : 1); ++phase) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
        for (int ch = phase; ch < nchunk; ch += (nchunk > 1 ? 2 : 1)) {
            const int z0 = (int)((int64_t)nm * ch / nchunk), z1 = (int)((int64_t)nm * (ch + 1) / nchunk);
            for (int64_t o = count[z0]; o < count[z1]; ++o) {
                const int64_t i = order[o];
                int ix, iy, iz;
                double dx, dy, dz;
                cell_coord(pos[3 * i], inv_cell, nm, &ix, &dx);
                cell_coord(pos[3 * i + 1], inv_cell, nm, &iy, &dy);
                cell_coord(pos[3 * i + 2], inv_cell, nm, &iz, &dz);
                const double wx[2] = {1.0 - dx, dx}, wy[2] = {1.0 - dy, dy}, wz[2] = {1.0 - dz, dz};
                const double m = mass[i];
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b)
                        for (int c = 0; c < 2; ++c)
                            rho[((iz + a) % nm) * nm2 + ((iy + b) % nm) * nm + (ix + c) % nm] +=
                                m * wx[c] * wy[b] * wz[a];
            }
        }
    }
    free(order);
    free(count);
    return 0;
}

/* cic.rs:100-170 / 408-454: 8-node trilinear gather per particle. */
int gngo_cic_interpolate(const double *fx, const double *fy, const double *fz, const double *pos,
                         int64_t n, double box, int nm, double *acc) {
    if (nm <= 0 || n < 0 || box <= 0.0) return 1;
    const int64_t nm2 = (int64_t)nm * nm;
    const double inv_cell = (double)nm / box;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        int ix, iy, iz;
        double dx, dy, dz;
        cell_coord(pos[3 * i], inv_cell, nm, &ix, &dx);
        cell_coord(pos[3 * i + 1], inv_cell, nm, &iy, &dy);
        cell_coord(pos[3 * i + 2], inv_cell, nm, &iz, &dz);
        const double wx[2] = {1.0 - dx, dx}, wy[2] = {1.0 - dy, dy}, wz[2] = {1.0 - dz, dz};
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int c = 0; c < 2; ++c) {
                    const double w = wx[c] * wy[b] * wz[a];
                    const int64_t idx = ((iz + a) % nm) * nm2 + ((iy + b) % nm) * nm + (ix + c) % nm;
                    ax += fx[idx] * w;
                    ay += fy[idx] * w;
                    az += fz[idx] * w;
                }
        acc[3 * i] = ax;
        acc[3 * i + 1] = ay;
        acc[3 * i + 2] = az;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Spectral kernel                                                          */
/* ------------------------------------------------------------------------ */

/* fft_poisson.rs:724-730 */
static inline int freq_index(int j, int nm) { return j <= nm / 2 ? j : j - nm; }

static inline double sinc(double x) { return x == 0.0 ? 1.0 : sin(x) / x; }

/* 1/W_CIC^2 factor per axis: 1/sinc^4(pi n / nm). */
static double *deconv_table(int nm) {
    const double pi = 3.14159265358979323846;
    double *t = (double *)malloc(sizeof(double) * (size_t)nm);
    for (int j = 0; j < nm; ++j) {
        const double s = sinc(pi * (double)freq_index(j, nm) / (double)nm);
        t[j] = 1.0 / (s * s * s * s);
    }
    return t;
}

/*
 * Green's function factor of fft_poisson.rs:516-541 for one mode:
 * returns -4 pi G * filter / k^2 (0 at DC).
 */
static inline double green(double kx, double ky, double kz, double four_pi_g, double rs2, int use_rs,
                           double e2, double deconv) {
    const double k2 = kx * kx + ky * ky + kz * kz;
    if (k2 < 1e-30) return 0.0;
    double filter = 1.0;
    if (use_rs) filter *= exp(-0.5 * k2 * rs2);
    if (e2 > 0.0) filter *= exp(-k2 * e2);
    filter *= deconv;
    return -four_pi_g * filter / k2;
}

/* ------------------------------------------------------------------------ */
/* Solve                                                                     */
/* ------------------------------------------------------------------------ */

/*
 * Faithful mode: the structure of solve_forces_impl (fft_poisson.rs:306-476).
 * fft_threads = 1 reproduces the reference's single-threaded FFT passes.
 */
static int solve_faithful(const double *density, double g, int nm, double box, double r_split,
                          double plummer_eps, int deconvolve, int fft_threads, double *fx, double *fy,
                          double *fz) {
    const int64_t nm2 = (int64_t)nm * nm, nm3 = nm2 * nm;
    const double cell = box / nm;
    const double rho_scale = 1.0 / (cell * cell * cell);
    const int threads = nthreads();
    double *rho = (double *)malloc(sizeof(double) * 2 * (size_t)nm3);
    if (!rho) return 2;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t i = 0; i < nm3; ++i) {
        rho[2 * i] = density[i] * rho_scale;
        rho[2 * i + 1] = 0.0;
    }
    gngo_fft3d(rho, nm, 0, fft_threads);

    const double pi = 3.14159265358979323846;
    const double dk = 2.0 * pi / box;
    const double four_pi_g = 4.0 * pi * g;
    const int use_rs = r_split > 0.0;
    const double rs2 = r_split * r_split;
    const double e2 = plummer_eps > 0.0 ? plummer_eps * plummer_eps : 0.0;
    double *dct = deconvolve ? deconv_table(nm) : NULL;
    double *f[3];
    for (int c = 0; c < 3; ++c) {
        f[c] = (double *)malloc(sizeof(double) * 2 * (size_t)nm3);
        if (!f[c]) return 2;
    }
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t flat = 0; flat < nm3; ++flat) {
        const int iz = (int)(flat / nm2), iy = (int)((flat / nm) % nm), ix = (int)(flat % nm);
        const double kx = dk * freq_index(ix, nm), ky = dk * freq_index(iy, nm), kz = dk * freq_index(iz, nm);
        const double dc = dct ? dct[ix] * dct[iy] * dct[iz] : 1.0;
        const double gf = green(kx, ky, kz, four_pi_g, rs2, use_rs, e2, dc);
        const double phr = rho[2 * flat] * gf, phi = rho[2 * flat + 1] * gf;
        /* F_alpha = -i k_alpha phi = (k_alpha phi_im, -k_alpha phi_re), fft_poisson.rs:543-548 */
        f[0][2 * flat] = kx * phi; f[0][2 * flat + 1] = -kx * phr;
        f[1][2 * flat] = ky * phi; f[1][2 * flat + 1] = -ky * phr;
        f[2][2 * flat] = kz * phi; f[2][2 * flat + 1] = -kz * phr;
    }
    free(rho);
    free(dct);
    const double norm = 1.0 / (double)nm3;
    double *outs[3] = {fx, fy, fz};
    for (int c = 0; c < 3; ++c) {
        gngo_fft3d(f[c], nm, 1, fft_threads);
        double *o = outs[c];
        const double *src = f[c];
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t i = 0; i < nm3; ++i) o[i] = src[2 * i] * norm;
        free(f[c]);
    }
    return 0;
}

/*
 * Lean mode: R2C half spectrum [nz][ny][nm/2+1].  The x pass transforms each
 * real row as a complex line of length nm and keeps indices 0..nm/2; the
 * inverse x pass rebuilds the full row by Hermitian symmetry and keeps Re().
 * Components are produced one at a time.  When `pos` is given, the force mesh
 * of each component is interpolated straight away (acc gets component c) and
 * no force mesh is kept; otherwise the meshes go to fx, fy, fz.
 */
static int solve_lean(const double *density, double g, int nm, double box, double r_split,
                      double plummer_eps, int deconvolve, double *fx, double *fy, double *fz,
                      const double *pos, int64_t npos, double *acc) {
    const int h = nm / 2 + 1;
    const int64_t nm2 = (int64_t)nm * nm, nm3 = nm2 * nm, mh = nm2 * h;
    const double cell = box / nm;
    const double rho_scale = 1.0 / (cell * cell * cell);
    const int threads = nthreads();
    double *spec = (double *)malloc(sizeof(double) * 2 * (size_t)mh);
    double *work = (double *)malloc(sizeof(double) * 2 * (size_t)mh);
    if (!spec || !work) return 2;
    /* x pass: real rows -> half rows */
#pragma omp parallel num_threads(threads)
    {
        fft_plan *p = plan_create(nm);
        double *line = (double *)malloc(sizeof(double) * 2 * (size_t)nm);
        double *scr = (double *)malloc(sizeof(double) * 2 * (size_t)nm);
#pragma omp for schedule(static)
        for (int64_t row = 0; row < nm2; ++row) {
            const double *src = density + row * nm;
            for (int x = 0; x < nm; ++x) {
                line[2 * x] = src[x] * rho_scale;
                line[2 * x + 1] = 0.0;
            }
            fft_line(p, line, scr, 0);
            memcpy(spec + 2 * row * h, line, sizeof(double) * 2 * (size_t)h);
        }
        free(line);
        free(scr);
        plan_destroy(p);
    }
    fft_axis(spec, h, nm, nm, 1, 0, threads);
    fft_axis(spec, h, nm, nm, 2, 0, threads);

    const double pi = 3.14159265358979323846;
    const double dk = 2.0 * pi / box;
    const double four_pi_g = 4.0 * pi * g;
    const int use_rs = r_split > 0.0;
    const double rs2 = r_split * r_split;
    const double e2 = plummer_eps > 0.0 ? plummer_eps * plummer_eps : 0.0;
    double *dct = deconvolve ? deconv_table(nm) : NULL;
    /* spec <- phi-hat (Green's function applied once) */
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t flat = 0; flat < mh; ++flat) {
        const int iz = (int)(flat / ((int64_t)nm * h)), iy = (int)((flat / h) % nm), ix = (int)(flat % h);
        const double kx = dk * freq_index(ix, nm), ky = dk * freq_index(iy, nm), kz = dk * freq_index(iz, nm);
        const double dc = dct ? dct[ix] * dct[iy] * dct[iz] : 1.0;
        const double gf = green(kx, ky, kz, four_pi_g, rs2, use_rs, e2, dc);
        spec[2 * flat] *= gf;
        spec[2 * flat + 1] *= gf;
    }
    free(dct);
    double *mesh = NULL;
    if (pos) {
        mesh = (double *)malloc(sizeof(double) * (size_t)nm3);
        if (!mesh) return 2;
    }
    double *outs[3] = {fx, fy, fz};
    const double norm = 1.0 / (double)nm3;
    for (int c = 0; c < 3; ++c) {
        /* work = -i k_alpha,eff phi, k_alpha,eff = 0 on the n_alpha = nm/2 plane */
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t flat = 0; flat < mh; ++flat) {
            const int iz = (int)(flat / ((int64_t)nm * h)), iy = (int)((flat / h) % nm), ix = (int)(flat % h);
            const int j = c == 0 ? ix : (c == 1 ? iy : iz);
            const double ka = (nm % 2 == 0 && j == nm / 2) ? 0.0 : dk * freq_index(j, nm);
            work[2 * flat] = ka * spec[2 * flat + 1];
            work[2 * flat + 1] = -ka * spec[2 * flat];
        }
        fft_axis(work, h, nm, nm, 2, 1, threads);
        fft_axis(work, h, nm, nm, 1, 1, threads);
        double *o = pos ? mesh : outs[c];
#pragma omp parallel num_threads(threads)
        {
            fft_plan *p = plan_create(nm);
            double *line = (double *)malloc(sizeof(double) * 2 * (size_t)nm);
            double *scr = (double *)malloc(sizeof(double) * 2 * (size_t)nm);
#pragma omp for schedule(static)
            for (int64_t row = 0; row < nm2; ++row) {
                const double *src = work + 2 * row * h;
                for (int x = 0; x < h; ++x) {
                    line[2 * x] = src[2 * x];
                    line[2 * x + 1] = src[2 * x + 1];
                }
                for (int x = h; x < nm; ++x) {
                    line[2 * x] = src[2 * (nm - x)];
                    line[2 * x + 1] = -src[2 * (nm - x) + 1];
                }
                fft_line(p, line, scr, 1);
                double *dst = o + row * nm;
                for (int x = 0; x < nm; ++x) dst[x] = line[2 * x] * norm;
            }
            free(line);
            free(scr);
            plan_destroy(p);
        }
        if (pos) {
            const double inv_cell = (double)nm / box;
#pragma omp parallel for num_threads(threads) schedule(static)
            for (int64_t i = 0; i < npos; ++i) {
                int ix, iy, iz;
                double dx, dy, dz;
                cell_coord(pos[3 * i], inv_cell, nm, &ix, &dx);
                cell_coord(pos[3 * i + 1], inv_cell, nm, &iy, &dy);
                cell_coord(pos[3 * i + 2], inv_cell, nm, &iz, &dz);
                const double wx[2] = {1.0 - dx, dx}, wy[2] = {1.0 - dy, dy}, wz[2] = {1.0 - dz, dz};
                double a = 0.0;
                for (int u = 0; u < 2; ++u)
                    for (int v = 0; v < 2; ++v)
                        for (int w = 0; w < 2; ++w)
                            a += mesh[((iz + u) % nm) * nm2 + ((iy + v) % nm) * nm + (ix + w) % nm] *
                                 (wx[w] * wy[v] * wz[u]);
                acc[3 * i + c] = a;
            }
        }
    }
    free(mesh);
    free(spec);
    free(work);
    return 0;
}

/*
 * Force meshes from a mass-per-cell mesh (fft_poisson.rs:272-290 front ends:
 * r_split <= 0 means no long/short filter, plummer_eps <= 0 means none).
 * fft_threads only affects GNGO_MODE_FAITHFUL (1 = the reference's serial FFT).
 */
int gngo_solve_forces(const double *density, double g, int nm, double box, double r_split,
                      double plummer_eps, int deconvolve, int mode, int fft_threads, double *fx,
                      double *fy, double *fz) {
    if (nm <= 0 || box <= 0.0) return 1;
    if (mode == GNGO_MODE_FAITHFUL)
        return solve_faithful(density, g, nm, box, r_split, plummer_eps, deconvolve,
                              fft_threads > 0 ? fft_threads : nthreads(), fx, fy, fz);
    return solve_lean(density, g, nm, box, r_split, plummer_eps, deconvolve, fx, fy, fz, NULL, 0, NULL);
}

/*
 * PmSolver::accelerations_for_indices (solver.rs:74-140) and the long-range
 * half of TreePmSolver (treepm/src/solver.rs:83-90, r_split > 0):
 * deposit all N particles, solve, interpolate at pos[idx[j]] for j < k
 * (idx == NULL means all N).  out is k x 3.  density_out (optional) receives
 * the mass-per-cell mesh.
 */
int gngo_pm_accelerations(const double *pos, const double *mass, int64_t n, const int64_t *idx,
                          int64_t k, double box, int nm, double g, double r_split, double plummer_eps,
                          int deconvolve, int mode, int deposit_mode, int fft_threads, double *out,
                          double *density_out) {
    if (nm <= 0 || box <= 0.0 || n < 0 || k < 0) return 1;
    if (!idx && k != n) return 1;
    const int64_t nm3 = (int64_t)nm * nm * nm;
    double *rho = density_out ? density_out : (double *)malloc(sizeof(double) * (size_t)nm3);
    if (!rho) return 2;
    int rc = gngo_cic_assign(pos, mass, n, box, nm, deposit_mode, rho);
    if (rc) return rc;
    double *apos = (double *)pos;
    if (idx) {
        apos = (double *)malloc(sizeof(double) * 3 * (size_t)(k > 0 ? k : 1));
        for (int64_t j = 0; j < k; ++j) {
            if (idx[j] < 0 || idx[j] >= n) return 1;
            apos[3 * j] = pos[3 * idx[j]];
            apos[3 * j + 1] = pos[3 * idx[j] + 1];
            apos[3 * j + 2] = pos[3 * idx[j] + 2];
        }
    }
    if (mode == GNGO_MODE_LEAN) {
        rc = solve_lean(rho, g, nm, box, r_split, plummer_eps, deconvolve, NULL, NULL, NULL, apos, k, out);
    } else {
        double *f = (double *)malloc(sizeof(double) * 3 * (size_t)nm3);
        if (!f) return 2;
        rc = gngo_solve_forces(rho, g, nm, box, r_split, plummer_eps, deconvolve, mode, fft_threads, f,
                               f + nm3, f + 2 * nm3);
        if (!rc) rc = gngo_cic_interpolate(f, f + nm3, f + 2 * nm3, apos, k, box, nm, out);
        free(f);
    }
    if (idx) free(apos);
    if (!density_out) free(rho);
    return rc;
}
