/*
 * oracle.c — float64 CPU restatement of gpu-rf-pipeline's hot-path definitions.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle the GPU path is compared with (tests, smoke, bench's
 * CPU baseline). Nothing in the shipped library calls it.
 *
 * Every sum runs in the order of its definition, in float64. The FFT is an iterative radix-2
 * Cooley-Tukey transform with every twiddle computed directly by cos/sin of an exactly
 * reduced angle; other lengths use a direct DFT with the same twiddles.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double kTwoPi = 6.283185307179586476925286766559;

/* tools/reference.py:103-111 (np.convolve(x, taps)[:len(x)]); src/fir.cu:1-5. */
void orc_fir(const double* x, long n, const double* taps, long num_taps, double* y)
{
    for (long i = 0; i < n; ++i) {
        const long last = i < num_taps - 1 ? i : num_taps - 1;
        double re = 0.0;
        double im = 0.0;
        for (long k = 0; k <= last; ++k) {
            re += taps[k] * x[2 * (i - k)];
            im += taps[k] * x[2 * (i - k) + 1];
        }
        y[2 * i] = re;
        y[2 * i + 1] = im;
    }
}

/* tools/reference.py:114-125 (scipy.signal.correlate(b, a, 'full')); src/xcorr.cu:22-48. */
void orc_xcorr(const double* a, long na, const double* b, long nb, double* r)
{
    for (long i = 0; i < na + nb - 1; ++i) {
        const long lag = i - (na - 1);
        const long first = lag < 0 ? -lag : 0;
        const long end = (nb - lag < na) ? nb - lag : na;
        double re = 0.0;
        double im = 0.0;
        for (long m = first; m < end; ++m) {
            const double ar = a[2 * m], ai = a[2 * m + 1];
            const double br = b[2 * (m + lag)], bi = b[2 * (m + lag) + 1];
            /* b * conj(a) */
            re += br * ar + bi * ai;
            im += bi * ar - br * ai;
        }
        r[2 * i] = re;
        r[2 * i + 1] = im;
    }
}

static int is_pow2(long n)
{
    return n > 0 && (n & (n - 1)) == 0;
}

/* exp(sign 2 pi i e / n) for an integer e already reduced modulo n. */
static void unit(long e, long n, int sign, double* c, double* s)
{
    const double a = (double)sign * kTwoPi * (double)e / (double)n;
    *c = cos(a);
    *s = sin(a);
}

int orc_fft(double* data, long n, int sign)
{
    if (n < 1 || (sign != -1 && sign != 1)) {
        return -1;
    }
    if (n == 1) {
        return 0;
    }
    if (!is_pow2(n)) {
        double* out = (double*)malloc(sizeof(double) * 2 * (size_t)n);
        double* tw = (double*)malloc(sizeof(double) * 2 * (size_t)n);
        if (!out || !tw) {
            free(out);
            free(tw);
            return -2;
        }
        for (long m = 0; m < n; ++m) {
            unit(m, n, sign, &tw[2 * m], &tw[2 * m + 1]);
        }
        for (long k = 0; k < n; ++k) {
            double re = 0.0, im = 0.0;
            long e = 0;
            for (long j = 0; j < n; ++j) {
                const double xr = data[2 * j], xi = data[2 * j + 1];
                const double wr = tw[2 * e], wi = tw[2 * e + 1];
                re += xr * wr - xi * wi;
                im += xr * wi + xi * wr;
                e += k;
                if (e >= n) {
                    e -= n;
                }
            }
            out[2 * k] = re;
            out[2 * k + 1] = im;
        }
        memcpy(data, out, sizeof(double) * 2 * (size_t)n);
        free(out);
        free(tw);
        return 0;
    }
    /* bit reversal */
    for (long i = 1, j = 0; i < n; ++i) {
        long bit = n >> 1;
        for (; j & bit; bit >>= 1) {
            j ^= bit;
        }
        j ^= bit;
        if (i < j) {
            double t = data[2 * i];
            data[2 * i] = data[2 * j];
            data[2 * j] = t;
            t = data[2 * i + 1];
            data[2 * i + 1] = data[2 * j + 1];
            data[2 * j + 1] = t;
        }
    }
    double* tw = (double*)malloc(sizeof(double) * (size_t)n); /* n/2 complex */
    if (!tw) {
        return -2;
    }
    for (long m = 0; m < n / 2; ++m) {
        unit(m, n, sign, &tw[2 * m], &tw[2 * m + 1]);
    }
    for (long len = 2; len <= n; len <<= 1) {
        const long half = len >> 1;
        const long step = n / len;
        for (long i = 0; i < n; i += len) {
            for (long j = 0; j < half; ++j) {
                const double wr = tw[2 * (j * step)], wi = tw[2 * (j * step) + 1];
                double* u = data + 2 * (i + j);
                double* v = data + 2 * (i + j + half);
                const double vr = v[0] * wr - v[1] * wi;
                const double vi = v[0] * wi + v[1] * wr;
                v[0] = u[0] - vr;
                v[1] = u[1] - vi;
                u[0] += vr;
                u[1] += vi;
            }
        }
    }
    free(tw);
    return 0;
}

/* tools/reference.py:88-95. */
int orc_spectrum(const double* x, long n, long nfft, double* out)
{
    if (nfft < 1) {
        return -1;
    }
    const long frames = (n + nfft - 1) / nfft;
    memset(out, 0, sizeof(double) * 2 * (size_t)(frames * nfft));
    memcpy(out, x, sizeof(double) * 2 * (size_t)n);
    for (long f = 0; f < frames; ++f) {
        const int rc = orc_fft(out + 2 * f * nfft, nfft, -1);
        if (rc != 0) {
            return rc;
        }
    }
    return 0;
}

void orc_magnitude(const double* x, long n, double* mag)
{
    for (long i = 0; i < n; ++i) {
        mag[i] = hypot(x[2 * i], x[2 * i + 1]);
    }
}

/* tools/reference.py:128-146: u[t][m] = sum_p h[pM+m] x[tM-pM-m], then M * ifft = the
 * unnormalised inverse DFT (src/channelize.cu:3-11). */
int orc_channelize(const double* x, long n, const double* taps, long num_taps, long channels,
                   double* out)
{
    if (channels < 1 || num_taps < 1) {
        return -1;
    }
    const long m_count = channels;
    const long frames = (n + m_count - 1) / m_count;
    const long branches = (num_taps + m_count - 1) / m_count;
    for (long t = 0; t < frames; ++t) {
        double* u = out + 2 * t * m_count;
        for (long m = 0; m < m_count; ++m) {
            double re = 0.0, im = 0.0;
            for (long p = 0; p < branches; ++p) {
                const long l = p * m_count + m;
                const long s = t * m_count - l;
                if (l >= num_taps || s < 0) {
                    continue;
                }
                re += taps[l] * x[2 * s];
                im += taps[l] * x[2 * s + 1];
            }
            u[2 * m] = re;
            u[2 * m + 1] = im;
        }
        const int rc = orc_fft(u, m_count, 1);
        if (rc != 0) {
            return rc;
        }
    }
    return 0;
}

/* tools/reference.py:149-162. */
int orc_gcc_phat(const double* a, long na, const double* b, long nb, double* r)
{
    if (na < 1 || nb < 1) {
        return -1;
    }
    const long lags = na + nb - 1;
    long L = 1;
    while (L < lags) {
        L *= 2;
    }
    double* fa = (double*)calloc(2 * (size_t)L, sizeof(double));
    double* fb = (double*)calloc(2 * (size_t)L, sizeof(double));
    if (!fa || !fb) {
        free(fa);
        free(fb);
        return -2;
    }
    memcpy(fa, a, sizeof(double) * 2 * (size_t)na);
    memcpy(fb, b, sizeof(double) * 2 * (size_t)nb);
    orc_fft(fa, L, -1);
    orc_fft(fb, L, -1);
    for (long k = 0; k < L; ++k) {
        /* cross = B conj(A); weighted = cross / |cross|, 0 where |cross| == 0 */
        const double ar = fa[2 * k], ai = fa[2 * k + 1];
        const double br = fb[2 * k], bi = fb[2 * k + 1];
        const double cr = br * ar + bi * ai;
        const double ci = bi * ar - br * ai;
        const double mag = hypot(cr, ci);
        fb[2 * k] = mag > 0.0 ? cr / mag : 0.0;
        fb[2 * k + 1] = mag > 0.0 ? ci / mag : 0.0;
    }
    orc_fft(fb, L, 1);
    /* lags -(na-1) .. -1 wrap to the end of the circular result */
    for (long i = 0; i < lags; ++i) {
        const long src = i < na - 1 ? L - (na - 1) + i : i - (na - 1);
        r[2 * i] = fb[2 * src] / (double)L;
        r[2 * i + 1] = fb[2 * src + 1] / (double)L;
    }
    free(fa);
    free(fb);
    return 0;
}

long orc_peak(const double* r, long n, double* magnitude)
{
    long best = 0;
    double best_mag = -1.0;
    for (long k = 0; k < n; ++k) {
        const double m = hypot(r[2 * k], r[2 * k + 1]);
        if (m > best_mag) {
            best_mag = m;
            best = k;
        }
    }
    if (magnitude) {
        *magnitude = best_mag;
    }
    return best;
}

double orc_doa(const long* lags, long count, double spacing, double speed, double rate)
{
    double sum_it = 0.0, sum_ii = 0.0;
    for (long i = 1; i <= count; ++i) {
        sum_it += (double)i * (double)lags[i - 1] / rate;
        sum_ii += (double)(i * i);
    }
    double s = speed * (sum_it / sum_ii) / spacing;
    s = s < -1.0 ? -1.0 : (s > 1.0 ? 1.0 : s);
    return asin(s) * 180.0 / 3.14159265358979323846;
}
