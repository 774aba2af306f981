/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle.  Nothing in the fastddm_amd
 * product path links, loads or calls this file; only tests/, the smoke check
 * and bench.py's cpu_baseline leg do, as a checker and as the timed CPU
 * baseline.
 *
 * FP64 C restatement of the reference C++ core (src/cpp/ddm.cc), with an own
 * mixed-radix Stockham FFT in place of FFTW (the vendored FFTW gets its
 * config.h only from its own CMake build, lib/fftw-3.3.10/CMakeLists.txt:304,
 * which is not run here, so the reference core is not built).
 *
 *   oracle_ddm_fft   <- ddm_fft<T>   src/cpp/ddm.cc:250-453
 *   oracle_ddm_diff  <- ddm_diff<T>  src/cpp/ddm.cc:37-192
 *   rfft2 stage      <- fft2_create_plan src/cpp/helper_fftw.cc:25-56 and the
 *                       copy/window/normalise steps ddm.cc:288-323
 *   time-axis FFT    <- fft_create_plan  src/cpp/helper_fftw.cc:61-92 (forward,
 *                       executed twice, ddm.cc:364,386)
 *   output shift     <- make_shifted_isf src/cpp/helper_ddm.cc:41-103
 *
 * Pinning (tests/test_oracle.py): the FFT against numpy.fft and a brute-force
 * DFT; the full outputs against the numpy oracle (oracle/oracle_np.py, an
 * independent FFT), brute-force direct sums on tiny stacks, fft == diff mode,
 * nt independence, and the closed-form special cases.
 *
 * Build: oracle/Makefile -> oracle/liboracle.so (gcc -O2 -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double re, im;
} cpx;

/* ---------------------------------------------------------------------------
 * Mixed-radix Stockham FFT (forward, unnormalised), any length n >= 1.
 * Factors 4, 2, 3, 5, 7 are peeled first, any other prime is handled by a
 * direct O(p^2) DFT stage.  Twiddles: w[m] = exp(-2 pi i m / n).
 * ------------------------------------------------------------------------- */
typedef struct {
    int n;
    int nf;
    int fac[64];
    cpx *w;
} fftplan;

static void plan_init(fftplan *p, int n) {
    p->n = n;
    p->nf = 0;
    int r = n;
    static const int pref[] = {4, 2, 3, 5, 7};
    for (int i = 0; i < 5; ++i)
        while (r % pref[i] == 0) {
            p->fac[p->nf++] = pref[i];
            r /= pref[i];
        }
    for (int f = 11; r > 1; f += 2)
        while (r % f == 0) {
            p->fac[p->nf++] = f;
            r /= f;
        }
    p->w = (cpx *)malloc(sizeof(cpx) * (size_t)(n > 0 ? n : 1));
    for (int m = 0; m < n; ++m) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
        p->w[m].re = (double)cosl(a);
        p->w[m].im = (double)sinl(a);
    }
}

static void plan_free(fftplan *p) { free(p->w); }

/* One Stockham stage of radix R; src and dst hold n points. */
static void stage(const fftplan *p, int R, int ns, const cpx *src, cpx *dst) {
    const int n = p->n, m = n / R, step = n / (ns * R), wr = n / R;
    cpx v[64], o[64];
    cpx *vv = R <= 64 ? v : (cpx *)malloc(sizeof(cpx) * R);
    cpx *oo = R <= 64 ? o : (cpx *)malloc(sizeof(cpx) * R);
    for (int j = 0; j < m; ++j) {
        const int k = j % ns;
        for (int r = 0; r < R; ++r) {
            cpx x = src[j + r * m];
            const cpx t = p->w[(size_t)r * k * step % n];
            vv[r].re = x.re * t.re - x.im * t.im;
            vv[r].im = x.re * t.im + x.im * t.re;
        }
        if (R == 2) {
            oo[0].re = vv[0].re + vv[1].re;
            oo[0].im = vv[0].im + vv[1].im;
            oo[1].re = vv[0].re - vv[1].re;
            oo[1].im = vv[0].im - vv[1].im;
        } else if (R == 4) {
            cpx a = {vv[0].re + vv[2].re, vv[0].im + vv[2].im};
            cpx b = {vv[0].re - vv[2].re, vv[0].im - vv[2].im};
            cpx c = {vv[1].re + vv[3].re, vv[1].im + vv[3].im};
            cpx d = {vv[1].re - vv[3].re, vv[1].im - vv[3].im};
            /* forward: -i * d */
            oo[0].re = a.re + c.re;
            oo[0].im = a.im + c.im;
            oo[2].re = a.re - c.re;
            oo[2].im = a.im - c.im;
            oo[1].re = b.re + d.im;
            oo[1].im = b.im - d.re;
            oo[3].re = b.re - d.im;
            oo[3].im = b.im + d.re;
        } else {
            for (int s = 0; s < R; ++s) {
                double ar = 0.0, ai = 0.0;
                for (int r = 0; r < R; ++r) {
                    const cpx t = p->w[(size_t)((r * s) % R) * wr];
                    ar += vv[r].re * t.re - vv[r].im * t.im;
                    ai += vv[r].re * t.im + vv[r].im * t.re;
                }
                oo[s].re = ar;
                oo[s].im = ai;
            }
        }
        const int base = (j / ns) * ns * R + k;
        for (int r = 0; r < R; ++r) dst[base + r * ns] = oo[r];
    }
    if (vv != v) free(vv);
    if (oo != o) free(oo);
}

/* In-place forward FFT of x (n points); work holds n points of scratch. */
static void fft_exec(const fftplan *p, cpx *x, cpx *work) {
    cpx *src = x, *dst = work;
    int ns = 1;
    for (int f = 0; f < p->nf; ++f) {
        stage(p, p->fac[f], ns, src, dst);
        ns *= p->fac[f];
        cpx *t = src;
        src = dst;
        dst = t;
    }
    if (src != x) memcpy(x, src, sizeof(cpx) * (size_t)p->n);
}

/* ---------------------------------------------------------------------------
 * Input conversion.  dtype codes: 0 u8, 1 i16, 2 u16, 3 i32, 4 u32, 5 i64,
 * 6 u64, 7 f32, 8 f64 (the nine overloads of ddm_bindings.h:34-43).
 * ------------------------------------------------------------------------- */
static double load_px(const void *img, int dtype, size_t i) {
    switch (dtype) {
    case 0: return (double)((const uint8_t *)img)[i];
    case 1: return (double)((const int16_t *)img)[i];
    case 2: return (double)((const uint16_t *)img)[i];
    case 3: return (double)((const int32_t *)img)[i];
    case 4: return (double)((const uint32_t *)img)[i];
    case 5: return (double)((const int64_t *)img)[i];
    case 6: return (double)((const uint64_t *)img)[i];
    case 7: return (double)((const float *)img)[i];
    default: return ((const double *)img)[i];
    }
}

/* Normalised half-plane spectra of every frame: spec[t][y][kx], kx < nx/2+1,
 * scaled by 1/sqrt(nx*ny) (ddm.cc:288-323).  Rows are transformed two at a
 * time as one complex sequence (row a + i row b) and split afterwards. */
static void rfft2_frames(const void *img, int dtype, size_t L, size_t H, size_t W, size_t nx,
                         size_t ny, const double *window, cpx *spec, int nthreads) {
    const size_t nxh = nx / 2 + 1, plane = ny * nxh;
    const double norm = 1.0 / sqrt((double)(nx * ny));
    fftplan px, py;
    plan_init(&px, (int)nx);
    plan_init(&py, (int)ny);
#pragma omp parallel num_threads(nthreads)
    {
        cpx *z = (cpx *)malloc(sizeof(cpx) * (nx > ny ? nx : ny));
        cpx *wk = (cpx *)malloc(sizeof(cpx) * (nx > ny ? nx : ny));
#pragma omp for schedule(dynamic, 1)
        for (long long t = 0; t < (long long)L; ++t) {
            cpx *out = spec + (size_t)t * plane;
            memset(out, 0, sizeof(cpx) * plane);
            for (size_t y = 0; y < H; y += 2) {
                const int two = (y + 1 < H);
                for (size_t x = 0; x < nx; ++x) {
                    double a = 0.0, b = 0.0;
                    if (x < W) {
                        a = load_px(img, dtype, ((size_t)t * H + y) * W + x);
                        if (window) a *= window[y * W + x];
                        if (two) {
                            b = load_px(img, dtype, ((size_t)t * H + y + 1) * W + x);
                            if (window) b *= window[(y + 1) * W + x];
                        }
                    }
                    z[x].re = a;
                    z[x].im = b;
                }
                fft_exec(&px, z, wk);
                for (size_t k = 0; k < nxh; ++k) {
                    const cpx p = z[k], q = z[(nx - k) % nx];
                    /* X_a = (Z[k] + conj Z[-k]) / 2, X_b = (Z[k] - conj Z[-k]) / 2i */
                    out[y * nxh + k].re = 0.5 * (p.re + q.re);
                    out[y * nxh + k].im = 0.5 * (p.im - q.im);
                    if (two) {
                        out[(y + 1) * nxh + k].re = 0.5 * (p.im + q.im);
                        out[(y + 1) * nxh + k].im = -0.5 * (p.re - q.re);
                    }
                }
            }
            for (size_t k = 0; k < nxh; ++k) {
                for (size_t y = 0; y < ny; ++y) z[y] = out[y * nxh + k];
                fft_exec(&py, z, wk);
                for (size_t y = 0; y < ny; ++y) {
                    out[y * nxh + k].re = z[y].re * norm;
                    out[y * nxh + k].im = z[y].im * norm;
                }
            }
        }
        free(z);
        free(wk);
    }
    plan_free(&px);
    plan_free(&py);
}

/* fftshift along y of a (rows, ny, nxh) real array, written to out
 * (helper_ddm.cc:41-103): out[s][i][x] = raw[s][(i - ny/2) mod ny][x]. */
static void shift_store(const double *raw, double *out, size_t nslices, size_t ny, size_t nxh) {
    const size_t c = ny / 2;
    for (size_t s = 0; s < nslices; ++s)
        for (size_t i = 0; i < ny; ++i) {
            const size_t src = (i + ny - c) % ny;
            memcpy(out + (s * ny + i) * nxh, raw + (s * ny + src) * nxh, sizeof(double) * nxh);
        }
}

/* diff mode: ddm.cc:112-171. */
int oracle_ddm_diff(const void *img, int dtype, uint64_t L, uint64_t H, uint64_t W,
                    const uint32_t *lags, uint64_t nl, uint64_t nx, uint64_t ny,
                    const double *window, double *out, int nthreads) {
    const size_t nxh = nx / 2 + 1, nq = ny * nxh;
    cpx *spec = (cpx *)malloc(sizeof(cpx) * L * nq);
    double *raw = (double *)malloc(sizeof(double) * (nl + 2) * nq);
    if (!spec || !raw) {
        free(spec);
        free(raw);
        return 1;
    }
    rfft2_frames(img, dtype, L, H, W, nx, ny, window, spec, nthreads);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long long q = 0; q < (long long)nq; ++q) {
        for (size_t i = 0; i < nl; ++i) {
            const size_t dt = lags[i];
            double acc = 0.0;
            for (size_t t = 0; t + dt < L; ++t) {
                const cpx a = spec[(t + dt) * nq + q], b = spec[t * nq + q];
                acc += (a.re - b.re) * (a.re - b.re) + (a.im - b.im) * (a.im - b.im);
            }
            raw[i * nq + q] = acc / (double)(L - dt);
        }
        double ps = 0.0, mr = 0.0, mi = 0.0;
        for (size_t t = 0; t < L; ++t) {
            const cpx a = spec[t * nq + q];
            ps += a.re * a.re + a.im * a.im;
            mr += a.re;
            mi += a.im;
        }
        ps /= (double)L;
        mr /= (double)L;
        mi /= (double)L;
        raw[nl * nq + q] = ps;
        raw[(nl + 1) * nq + q] = ps - mr * mr - mi * mi;
    }
    shift_store(raw, out, nl + 2, ny, nxh);
    free(spec);
    free(raw);
    return 0;
}

/* fft mode: ddm.cc:329-426, chunks of `chunk` q series at a time. */
int oracle_ddm_fft(const void *img, int dtype, uint64_t L, uint64_t H, uint64_t W,
                   const uint32_t *lags, uint64_t nl, uint64_t nx, uint64_t ny, uint64_t nt,
                   uint64_t chunk, const double *window, double *out, int nthreads) {
    const size_t nxh = nx / 2 + 1, nq = ny * nxh;
    if (nt < L || chunk == 0) return 2;
    cpx *spec = (cpx *)malloc(sizeof(cpx) * L * nq);
    double *raw = (double *)malloc(sizeof(double) * (nl + 2) * nq);
    if (!spec || !raw) {
        free(spec);
        free(raw);
        return 1;
    }
    rfft2_frames(img, dtype, L, H, W, nx, ny, window, spec, nthreads);
    fftplan pt;
    plan_init(&pt, (int)nt);
    const double pre = 1.0 / sqrt((double)nt);
    const long long nchunks = (long long)((nq + chunk - 1) / chunk);
#pragma omp parallel num_threads(nthreads)
    {
        cpx *ws = (cpx *)malloc(sizeof(cpx) * chunk * nt);
        cpx *wk = (cpx *)malloc(sizeof(cpx) * nt);
        double *tmp = (double *)malloc(sizeof(double) * chunk);
        double *avg = (double *)malloc(sizeof(double) * chunk);
#pragma omp for schedule(dynamic, 1)
        for (long long ci = 0; ci < nchunks; ++ci) {
            const size_t q0 = (size_t)ci * chunk;
            const size_t nc = (q0 + chunk <= nq) ? chunk : nq - q0;
            /* gather F[t][q] / sqrt(nt), zero padded to nt (ddm.cc:348-361) */
            for (size_t c = 0; c < nc; ++c) {
                cpx *s = ws + c * nt;
                for (size_t t = 0; t < L; ++t) {
                    s[t].re = spec[t * nq + q0 + c].re * pre;
                    s[t].im = spec[t * nq + q0 + c].im * pre;
                }
                for (size_t t = L; t < nt; ++t) s[t].re = s[t].im = 0.0;
                fft_exec(&pt, s, wk); /* ddm.cc:364 */
                for (size_t t = 0; t < nt; ++t) { /* |.|^2, ddm.cc:367-373 */
                    const double a = s[t].re, b = s[t].im;
                    s[t].re = a * a + b * b;
                    s[t].im = 0.0;
                }
                /* |sum F|^2 / L^2 from bin 0 (ddm.cc:376-383) */
                avg[c] = s[0].re * (double)nt / ((double)L * (double)L);
                tmp[c] = 0.0;
                fft_exec(&pt, s, wk); /* forward again: real part = C(dt) (ddm.cc:386) */
            }
            /* two-ended running sum over t, lags visited in descending order
             * (ddm.cc:389-414) */
            size_t idx = 0;
            for (size_t t = 0; t < L; ++t) {
                for (size_t c = 0; c < nc; ++c) {
                    const cpx a = spec[t * nq + q0 + c], b = spec[(L - 1 - t) * nq + q0 + c];
                    tmp[c] += a.re * a.re + a.im * a.im + b.re * b.re + b.im * b.im;
                }
                if (idx < nl && L - 1 - t == lags[nl - 1 - idx]) {
                    const size_t dt = lags[nl - 1 - idx];
                    for (size_t c = 0; c < nc; ++c) {
                        cpx *s = ws + c * nt;
                        s[dt].re = (tmp[c] - 2.0 * s[dt].re) / (double)(L - dt);
                    }
                    ++idx;
                }
            }
            /* scatter D, PS, VAR (ddm.cc:417-425) */
            for (size_t c = 0; c < nc; ++c) {
                const cpx *s = ws + c * nt;
                for (size_t i = 0; i < nl; ++i) raw[i * nq + q0 + c] = s[lags[i]].re;
                const double ps = 0.5 * tmp[c] / (double)L;
                raw[nl * nq + q0 + c] = ps;
                raw[(nl + 1) * nq + q0 + c] = ps - avg[c];
            }
        }
        free(ws);
        free(wk);
        free(tmp);
        free(avg);
    }
    plan_free(&pt);
    shift_store(raw, out, nl + 2, ny, nxh);
    free(spec);
    free(raw);
    return 0;
}

/* Half-plane normalised spectra, unshifted: out (L, ny, nx/2+1) complex. */
int oracle_rfft2(const void *img, int dtype, uint64_t L, uint64_t H, uint64_t W, uint64_t nx,
                 uint64_t ny, const double *window, double *out_complex, int nthreads) {
    rfft2_frames(img, dtype, L, H, W, nx, ny, window, (cpx *)out_complex, nthreads);
    return 0;
}

/* Batched forward FFT of interleaved complex data, in place. */
int oracle_fft(double *data, uint64_t n, uint64_t batch) {
    fftplan p;
    plan_init(&p, (int)n);
    cpx *wk = (cpx *)malloc(sizeof(cpx) * n);
    for (uint64_t b = 0; b < batch; ++b) fft_exec(&p, (cpx *)data + b * n, wk);
    free(wk);
    plan_free(&p);
    return 0;
}
