/*
 * fft_oracle.c - CPU restatement, in FP64 / long double, of the reference's FFT algorithms for the
 * aclfft execution path.  TEST INFRASTRUCTURE ONLY: the checker for tests/, the smoke check and
 * the CPU-baseline leg of the benchmark.  Nothing in the product library calls it.
 *
 * Conventions (reference, all pinned by tests/test_oracle.py):
 *   forward  X[k] = sum_n x[n] e^{-2 pi i nk/N}   (src/fft1_d/arch35/fft/fft1_d_c2c.cpp:49)
 *   backward x[n] = sum_k X[k] e^{+2 pi i nk/N}, unnormalised
 *            (golden ifft(x)*n: src/fft1_d/tests/fft1_d_data/gen_data.py:86)
 *   complex data interleaved re/im (include/cann_ops_fft.h:117-120)
 *
 * Functions and the reference code each one restates:
 *   oracle_dft_naive   brute-force DFT in long double with exact (n*k mod N) index reduction;
 *                      the mathematical definition, used to pin oracle_c2c.
 *   oracle_c2c         mixed-radix Stockham autosort, one stage per prime factor in ascending order
 *                      (radix list: src/fft1_d/arch35/mix/fft1_d_c2c_mix.cpp:49-68; stage formula
 *                      y[blk*len + q*prev + j] = sum_p DFT_r[q][p] tw[p*prev+j] x[blk*prev + j + p*N/r]:
 *                      src/fft1_d/arch35/mix/fft1_d_c2c_mix_kernel.h:1108-1176; twiddle
 *                      angle = sign*2*pi*p*j/len: src/fft1_d/arch35/mix/fft1_d_c2c_mix.cpp:99-125).
 *                      Any prime factor is accepted (a direct radix-p DFT stage).
 *   oracle_r2c         even N: pack z[m] = x[2m] + i x[2m+1], N/2-point FFT, Hermitian post-process
 *                      X[k] = (Z[k] + conj Z[N/2-k])/2 + W^k (Z[k] - conj Z[N/2-k])/(2i), with DC and
 *                      Nyquist imaginary parts 0 (src/rfft1_d/arch35/fft/rfft1_d_fft_kernel.h:39-228);
 *                      odd N: full complex FFT of (x, 0) truncated to N/2+1 bins (same file, odd path).
 *   oracle_c2r         Hermitian expansion of N/2+1 bins to N, N-point backward FFT, real part, no
 *                      scaling (src/irfft1_d/arch35/fft/irfft1_d_fft_kernel.h:40-180).
 *   oracle_c2c_2d      row transforms then column transforms of [batch][nx][ny]
 *                      (golden torch.fft.fft2: src/fft2_d/tests/fft2_d_data/gen_data.py:59-61).
 *   oracle_c2c_vertical column transforms of an [n][batch] matrix (vertical layout; golden
 *                      src/fft1_d/tests/fft1_d_data/gen_data.py:79-81).
 *   oracle_c2c_3d      3D transform of [batch][nx][ny][nz]: a 2D transform of every x-plane, then
 *                      the transforms along x (the reference has no 3D execution, its
 *                      aclfftMakePlan3d is a stub, lib/fft_plan_init_api.cpp:250-254; the
 *                      convention is that of its 2D golden, torch.fft.fftn with an unnormalised
 *                      inverse).
 *
 * Build: make -C oracle  (gcc, no dependencies) -> oracle/liboracle.so
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI_L 3.141592653589793238462643383279502884L

/* (cos, sin) of 2*pi*e/L in long double after exact integer reduction of e mod L */
static void cs2pi(long long e, long long L, long double* c, long double* s) {
    long long m = e % L;
    if (m < 0) m += L;
    long double a = 2.0L * OR_PI_L * (long double)m / (long double)L;
    *c = cosl(a);
    *s = sinl(a);
}

/* ---- brute force --------------------------------------------------------------------------- */
void oracle_dft_naive(const double* in, double* out, long long n, long long batch, int dir) {
    for (long long b = 0; b < batch; ++b) {
        const double* x = in + 2 * b * n;
        double* y = out + 2 * b * n;
        for (long long k = 0; k < n; ++k) {
            long double re = 0.0L, im = 0.0L;
            for (long long j = 0; j < n; ++j) {
                long double c, s;
                cs2pi((j * k) % n, n, &c, &s);
                long double wr = c, wi = (dir < 0) ? -s : s; /* e^{dir * 2 pi i jk/n} */
                long double xr = x[2 * j], xi = x[2 * j + 1];
                re += xr * wr - xi * wi;
                im += xr * wi + xi * wr;
            }
            y[2 * k] = (double)re;
            y[2 * k + 1] = (double)im;
        }
    }
}

/* ---- mixed-radix Stockham ------------------------------------------------------------------ */
static int factor_primes(long long n, long long* f) {
    int cnt = 0;
    for (long long p = 2; p * p <= n; ++p)
        while (n % p == 0) {
            f[cnt++] = p;
            n /= p;
        }
    if (n > 1) f[cnt++] = n;
    return cnt;
}

/* Per-length plan: prime radices in ascending order and, per stage, the inter-stage twiddles
 * e^{-2 pi i p j/len} (p < r, j < prev) and the DFT_r matrix e^{-2 pi i pq/r}, computed in long
 * double and stored in double (reference layouts: fft1_d_c2c_mix.cpp:74-125). */
typedef struct {
    long long n;
    int nf;
    long long f[64];
    double* tw[64];  /* r*prev complex, [p*prev + j] */
    double* dft[64]; /* r*r complex, [q*r + p] */
} or_plan;

static void plan_make(or_plan* pl, long long n) {
    pl->n = n;
    pl->nf = (n > 1) ? factor_primes(n, pl->f) : 0;
    long long len = 1;
    for (int s = 0; s < pl->nf; ++s) {
        const long long r = pl->f[s], prev = len;
        len *= r;
        pl->tw[s] = (double*)malloc(sizeof(double) * 2 * (size_t)(r * prev));
        pl->dft[s] = (double*)malloc(sizeof(double) * 2 * (size_t)(r * r));
        for (long long p = 0; p < r; ++p)
            for (long long j = 0; j < prev; ++j) {
                long double c, sn;
                cs2pi(p * j, len, &c, &sn);
                pl->tw[s][2 * (p * prev + j)] = (double)c;
#ifdef OR_PLANT_TWIDDLE_SIGN /* deliberately wrong build, used only to show the pins catch it */
                pl->tw[s][2 * (p * prev + j) + 1] = (double)sn;
#else
                pl->tw[s][2 * (p * prev + j) + 1] = (double)-sn;
#endif
            }
        for (long long q = 0; q < r; ++q)
            for (long long p = 0; p < r; ++p) {
                long double c, sn;
                cs2pi(p * q, r, &c, &sn);
                pl->dft[s][2 * (q * r + p)] = (double)c;
                pl->dft[s][2 * (q * r + p) + 1] = (double)-sn;
            }
    }
}
static void plan_free(or_plan* pl) {
    for (int s = 0; s < pl->nf; ++s) {
        free(pl->tw[s]);
        free(pl->dft[s]);
    }
}

/* One transform, in -> out (n complex each, no aliasing).  dir = -1 forward, +1 backward (the
 * backward tables are the conjugates of the forward ones). */
static void stockham_one(const or_plan* pl, const double* in, double* out, int dir, double* wa, double* wb) {
    const long long n = pl->n;
    const double* src = in;
    double* bufs[2] = {wa, wb};
    const double sg = (dir < 0) ? 1.0 : -1.0; /* multiplies the imaginary part of each table entry */
    if (pl->nf == 0) {
        memcpy(out, in, sizeof(double) * 2 * (size_t)n);
        return;
    }
    long long len = 1;
    for (int s = 0; s < pl->nf; ++s) {
        const long long r = pl->f[s];
        const long long prev = len;
        len *= r;
        const long long part = n / r;
        double* dst = (s == pl->nf - 1) ? out : bufs[s & 1];
        const long long blocks = n / len;
        const double* tw = pl->tw[s];
        const double* dft = pl->dft[s];
        double br[64], bi[64];
        double* brv = br;
        double* biv = bi;
        if (r > 64) {
            brv = (double*)malloc(sizeof(double) * (size_t)r);
            biv = (double*)malloc(sizeof(double) * (size_t)r);
        }
        for (long long blk = 0; blk < blocks; ++blk) {
            for (long long j = 0; j < prev; ++j) {
                for (long long p = 0; p < r; ++p) {
                    const long long ii = blk * prev + j + p * part;
                    const double xr = src[2 * ii], xi = src[2 * ii + 1];
                    const double tr = tw[2 * (p * prev + j)], ti = sg * tw[2 * (p * prev + j) + 1];
                    brv[p] = xr * tr - xi * ti;
                    biv[p] = xr * ti + xi * tr;
                }
                for (long long q = 0; q < r; ++q) {
                    double accr = 0.0, acci = 0.0;
                    const double* d = dft + 2 * q * r;
                    for (long long p = 0; p < r; ++p) {
                        const double dr = d[2 * p], di = sg * d[2 * p + 1];
                        accr += brv[p] * dr - biv[p] * di;
                        acci += brv[p] * di + biv[p] * dr;
                    }
                    const long long y = blk * len + q * prev + j;
                    dst[2 * y] = accr;
                    dst[2 * y + 1] = acci;
                }
            }
        }
        if (r > 64) {
            free(brv);
            free(biv);
        }
        src = dst;
    }
}

void oracle_c2c(const double* in, double* out, long long n, long long batch, int dir) {
    or_plan pl;
    plan_make(&pl, n);
    double* wa = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double* wb = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double* tmp = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    for (long long b = 0; b < batch; ++b) {
        stockham_one(&pl, in + 2 * b * n, tmp, dir, wa, wb);
        memcpy(out + 2 * b * n, tmp, sizeof(double) * 2 * (size_t)n);
    }
    free(wa);
    free(wb);
    free(tmp);
    plan_free(&pl);
}

/* ---- real transforms ----------------------------------------------------------------------- */
void oracle_r2c(const double* in, double* out, long long n, long long batch) {
    const long long h = n / 2 + 1;
    double* z = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double* Z = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    for (long long b = 0; b < batch; ++b) {
        const double* x = in + b * n;
        double* X = out + 2 * b * h;
        if (n % 2 == 0 && n >= 2) {
            const long long M = n / 2;
            for (long long m = 0; m < M; ++m) {
                z[2 * m] = x[2 * m];
                z[2 * m + 1] = x[2 * m + 1];
            }
            oracle_c2c(z, Z, M, 1, -1);
            for (long long k = 0; k <= M; ++k) {
                const long long k1 = k % M, k2 = (M - k) % M;
                long double zr = Z[2 * k1], zi = Z[2 * k1 + 1];
                long double cr = Z[2 * k2], ci = -Z[2 * k2 + 1]; /* conj Z[M-k] */
                long double er = (zr + cr) / 2, ei = (zi + ci) / 2;
                long double dr = (zr - cr) / 2, di = (zi - ci) / 2; /* (Z - conj)/2 */
                /* o = d / i = (di, -dr) */
                long double orr = di, oi = -dr;
                long double c, s;
                cs2pi(k, n, &c, &s);
                long double wr = c, wi = -s; /* W_n^k */
                long double xr = er + (wr * orr - wi * oi);
                long double xi = ei + (wr * oi + wi * orr);
                if (k == 0 || k == M) xi = 0.0L;
                X[2 * k] = (double)xr;
                X[2 * k + 1] = (double)xi;
            }
        } else {
            for (long long m = 0; m < n; ++m) {
                z[2 * m] = x[m];
                z[2 * m + 1] = 0.0;
            }
            oracle_c2c(z, Z, n, 1, -1);
            for (long long k = 0; k < h; ++k) {
                X[2 * k] = Z[2 * k];
                X[2 * k + 1] = (k == 0) ? 0.0 : Z[2 * k + 1];
            }
        }
    }
    free(z);
    free(Z);
}

void oracle_c2r(const double* in, double* out, long long n, long long batch) {
    const long long h = n / 2 + 1;
    double* Y = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double* y = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    for (long long b = 0; b < batch; ++b) {
        const double* X = in + 2 * b * h;
        for (long long k = 0; k < n; ++k) {
            if (k < h) {
                Y[2 * k] = X[2 * k];
                Y[2 * k + 1] = X[2 * k + 1];
            } else {
                Y[2 * k] = X[2 * (n - k)];
                Y[2 * k + 1] = -X[2 * (n - k) + 1];
            }
        }
        oracle_c2c(Y, y, n, 1, +1);
        for (long long m = 0; m < n; ++m) out[b * n + m] = y[2 * m];
    }
    free(Y);
    free(y);
}

/* ---- 2D and vertical ----------------------------------------------------------------------- */
void oracle_c2c_2d(const double* in, double* out, long long nx, long long ny, long long batch, int dir) {
    double* col = (double*)malloc(sizeof(double) * 2 * (size_t)nx);
    double* colo = (double*)malloc(sizeof(double) * 2 * (size_t)nx);
    oracle_c2c(in, out, ny, nx * batch, dir);
    for (long long b = 0; b < batch; ++b) {
        double* img = out + 2 * b * nx * ny;
        for (long long c = 0; c < ny; ++c) {
            for (long long r = 0; r < nx; ++r) {
                col[2 * r] = img[2 * (r * ny + c)];
                col[2 * r + 1] = img[2 * (r * ny + c) + 1];
            }
            oracle_c2c(col, colo, nx, 1, dir);
            for (long long r = 0; r < nx; ++r) {
                img[2 * (r * ny + c)] = colo[2 * r];
                img[2 * (r * ny + c) + 1] = colo[2 * r + 1];
            }
        }
    }
    free(col);
    free(colo);
}

void oracle_c2c_vertical(const double* in, double* out, long long n, long long cols, int dir) {
    double* col = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double* colo = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    for (long long c = 0; c < cols; ++c) {
        for (long long r = 0; r < n; ++r) {
            col[2 * r] = in[2 * (r * cols + c)];
            col[2 * r + 1] = in[2 * (r * cols + c) + 1];
        }
        oracle_c2c(col, colo, n, 1, dir);
        for (long long r = 0; r < n; ++r) {
            out[2 * (r * cols + c)] = colo[2 * r];
            out[2 * (r * cols + c) + 1] = colo[2 * r + 1];
        }
    }
    free(col);
    free(colo);
}

void oracle_c2c_3d(const double* in, double* out, long long nx, long long ny, long long nz, long long batch,
                   int dir) {
    const long long plane = ny * nz;
    double* tmp = (double*)malloc(sizeof(double) * 2 * (size_t)(nx * plane));
    oracle_c2c_2d(in, out, ny, nz, nx * batch, dir);
    for (long long b = 0; b < batch; ++b) {
        double* img = out + 2 * b * nx * plane;
        oracle_c2c_vertical(img, tmp, nx, plane, dir);
        memcpy(img, tmp, sizeof(double) * 2 * (size_t)(nx * plane));
    }
    free(tmp);
}
