// Multithreaded FP32 CPU path: the CPU baseline the GPU numbers are reported beside. It is test
// and measurement infrastructure (bench.py's cpu_baseline leg and tests/ only), not part of the
// library. Same definitions as the float64 oracle (oracle.c), in float32 with float32
// accumulation, split across std::threads:
//   cpu_fir_f32         causal direct FIR (tools/reference.py:103-111, src/fir.cu:1-12), output
//                       ranges per thread with a T-1 sample halo;
//   cpu_channelize_f32  polyphase branch sums then the M-point inverse DFT per frame
//                       (tools/reference.py:128-146, src/channelize.cu:3-11), frame ranges per
//                       thread;
//   cpu_gcc_phat_peaks  GCC-PHAT (tools/reference.py:149-162, src/xcorr.cu:219-233) and the
//                       first peak of |r| per pair (src/main.cpp:390-401), pairs per thread.
// The FFT is an iterative radix-2 transform over a float table of exp(-2 pi i k / 2^22) made in
// double precision once.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <functional>
#include <thread>
#include <vector>

namespace {

using cf = std::complex<float>;

constexpr int kTableLog2 = 22;  // longest transform: 2^22 points

const std::vector<cf>& twiddles()
{
    static const std::vector<cf> table = [] {
        const long n = 1L << kTableLog2;
        std::vector<cf> t(static_cast<std::size_t>(n / 2));
        for (long k = 0; k < n / 2; ++k) {
            const double a = -2.0 * M_PI * static_cast<double>(k) / static_cast<double>(n);
            t[static_cast<std::size_t>(k)] = cf(static_cast<float>(std::cos(a)),
                                                static_cast<float>(std::sin(a)));
        }
        return t;
    }();
    return table;
}

// In-place FFT of 2^log2n points; sign -1 forward, +1 inverse (unnormalised).
void fft(cf* x, int log2n, int sign)
{
    const long n = 1L << log2n;
    for (long i = 1, j = 0; i < n; ++i) {
        long bit = n >> 1;
        for (; j & bit; bit >>= 1) {
            j ^= bit;
        }
        j ^= bit;
        if (i < j) {
            std::swap(x[i], x[j]);
        }
    }
    const std::vector<cf>& w = twiddles();
    for (int s = 1; s <= log2n; ++s) {
        const long half = 1L << (s - 1);
        const long stride = 1L << (kTableLog2 - s);
        for (long k = 0; k < n; k += 2 * half) {
            for (long j = 0; j < half; ++j) {
                cf t = w[static_cast<std::size_t>(j * stride)];
                if (sign > 0) {
                    t = std::conj(t);
                }
                const cf u = x[k + j];
                const cf v = x[k + j + half] * t;
                x[k + j] = u + v;
                x[k + j + half] = u - v;
            }
        }
    }
}

// Runs body(begin, end) over [0, count) split into `threads` contiguous ranges.
void parallel_ranges(long count, int threads, const std::function<void(long, long)>& body)
{
    threads = std::max(1, std::min<int>(threads, static_cast<int>(std::max(1L, count))));
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        const long b = count * t / threads;
        const long e = count * (t + 1) / threads;
        pool.emplace_back(body, b, e);
    }
    for (std::thread& th : pool) {
        th.join();
    }
}

}  // namespace

extern "C" {

int cpu_threads(void)
{
    const unsigned n = std::thread::hardware_concurrency();
    return n == 0 ? 1 : static_cast<int>(n);
}

// y[i] = sum_{k <= i} taps[k] x[i - k], x and y interleaved I/Q float32.
int cpu_fir_f32(const float* x, long n, const float* taps, long num_taps, float* y, int threads)
{
    if (n < 0 || num_taps < 1) {
        return -1;
    }
    std::vector<float> rev(static_cast<std::size_t>(num_taps));
    for (long k = 0; k < num_taps; ++k) {
        rev[static_cast<std::size_t>(k)] = taps[num_taps - 1 - k];
    }
    const long halo = num_taps - 1;
    parallel_ranges(n, threads, [&](long s, long e) {
        // Deinterleaved input from s - halo (zeros before the capture), so output i reads
        // xr[i - s .. i - s + halo] against the reversed taps.
        const long len = e - s + halo;
        std::vector<float> xr(static_cast<std::size_t>(len) + 32, 0.0f);
        std::vector<float> xi(static_cast<std::size_t>(len) + 32, 0.0f);
        for (long j = 0; j < len; ++j) {
            const long src = s - halo + j;
            if (src >= 0) {
                xr[static_cast<std::size_t>(j)] = x[2 * src];
                xi[static_cast<std::size_t>(j)] = x[2 * src + 1];
            }
        }
        constexpr int B = 16;
        const float* h = rev.data();
        long i = s;
        for (; i + B <= e; i += B) {
            float ar[B] = {};
            float ai[B] = {};
            const float* pr = xr.data() + (i - s);
            const float* pi = xi.data() + (i - s);
            for (long k = 0; k < num_taps; ++k) {
                const float hk = h[k];
                for (int b = 0; b < B; ++b) {
                    ar[b] += hk * pr[k + b];
                    ai[b] += hk * pi[k + b];
                }
            }
            for (int b = 0; b < B; ++b) {
                y[2 * (i + b)] = ar[b];
                y[2 * (i + b) + 1] = ai[b];
            }
        }
        for (; i < e; ++i) {
            float ar = 0.0f, ai = 0.0f;
            for (long k = 0; k < num_taps; ++k) {
                ar += h[k] * xr[static_cast<std::size_t>(i - s + k)];
                ai += h[k] * xi[static_cast<std::size_t>(i - s + k)];
            }
            y[2 * i] = ar;
            y[2 * i + 1] = ai;
        }
    });
    return 0;
}

// out[t][k] = sum_m u[t][m] exp(+2 pi i k m / M), u[t][m] = sum_p taps[pM + m] x[tM - pM - m].
int cpu_channelize_f32(const float* x, long n, const float* taps, long num_taps, long channels,
                       float* out, int threads)
{
    if (channels < 1 || num_taps < 1 || n < 0) {
        return -1;
    }
    const long m_count = channels;
    const long frames = (n + m_count - 1) / m_count;
    const long branches = (num_taps + m_count - 1) / m_count;
    int log2m = 0;
    while ((1L << log2m) < m_count) {
        ++log2m;
    }
    const bool pow2 = (1L << log2m) == m_count && log2m <= kTableLog2;
    // Exact DFT factors for other M.
    std::vector<cf> dft;
    if (!pow2) {
        dft.resize(static_cast<std::size_t>(m_count));
        for (long k = 0; k < m_count; ++k) {
            const double a = 2.0 * M_PI * static_cast<double>(k) / static_cast<double>(m_count);
            dft[static_cast<std::size_t>(k)] = cf(static_cast<float>(std::cos(a)),
                                                  static_cast<float>(std::sin(a)));
        }
    }
    const cf* xs = reinterpret_cast<const cf*>(x);
    cf* zs = reinterpret_cast<cf*>(out);
    parallel_ranges(frames, threads, [&](long f0, long f1) {
        std::vector<cf> u(static_cast<std::size_t>(m_count));
        for (long t = f0; t < f1; ++t) {
            for (long m = 0; m < m_count; ++m) {
                cf acc(0.0f, 0.0f);
                for (long p = 0; p < branches; ++p) {
                    const long l = p * m_count + m;
                    const long s = t * m_count - l;
                    if (l < num_taps && s >= 0) {
                        acc += taps[l] * xs[s];
                    }
                }
                u[static_cast<std::size_t>(m)] = acc;
            }
            cf* z = zs + t * m_count;
            if (pow2) {
                fft(u.data(), log2m, 1);
                std::copy(u.begin(), u.end(), z);
            } else {
                for (long k = 0; k < m_count; ++k) {
                    cf acc(0.0f, 0.0f);
                    for (long m = 0; m < m_count; ++m) {
                        acc += u[static_cast<std::size_t>(m)]
                               * dft[static_cast<std::size_t>((k * m) % m_count)];
                    }
                    z[k] = acc;
                }
            }
        }
    });
    return 0;
}

// For each pair: r = GCC-PHAT(a, b) over lags -(na-1)..(nb-1); peak[i] = the first index of the
// largest |r| (lag = peak - (na - 1)). Pairs are handed to threads through a shared counter.
int cpu_gcc_phat_peaks(int count, const float* const* a, const long* na, const float* const* b,
                       const long* nb, long* peak, int threads)
{
    for (int i = 0; i < count; ++i) {
        const long lags = na[i] + nb[i] - 1;
        if (na[i] < 1 || nb[i] < 1 || lags > (1L << kTableLog2)) {
            return -1;
        }
    }
    (void)twiddles();
    std::atomic<int> next(0);
    const auto worker = [&](long, long) {
        std::vector<cf> fa, fb;
        for (int i = next++; i < count; i = next++) {
            const long lags = na[i] + nb[i] - 1;
            int log2l = 0;
            while ((1L << log2l) < lags) {
                ++log2l;
            }
            const long L = 1L << log2l;
            fa.assign(static_cast<std::size_t>(L), cf(0.0f, 0.0f));
            fb.assign(static_cast<std::size_t>(L), cf(0.0f, 0.0f));
            const cf* ai = reinterpret_cast<const cf*>(a[i]);
            const cf* bi = reinterpret_cast<const cf*>(b[i]);
            std::copy(ai, ai + na[i], fa.begin());
            std::copy(bi, bi + nb[i], fb.begin());
            fft(fa.data(), log2l, -1);
            fft(fb.data(), log2l, -1);
            for (long k = 0; k < L; ++k) {
                const cf c = fb[static_cast<std::size_t>(k)] * std::conj(fa[static_cast<std::size_t>(k)]);
                const float mag = std::abs(c);
                fb[static_cast<std::size_t>(k)] = mag > 0.0f ? c / mag : cf(0.0f, 0.0f);
            }
            fft(fb.data(), log2l, 1);
            long best = 0;
            float best_mag = -1.0f;
            for (long j = 0; j < lags; ++j) {
                const long src = j < na[i] - 1 ? L - (na[i] - 1) + j : j - (na[i] - 1);
                const float m = std::abs(fb[static_cast<std::size_t>(src)]);
                if (m > best_mag) {
                    best_mag = m;
                    best = j;
                }
            }
            peak[i] = best;
        }
    };
    parallel_ranges(threads, threads, worker);
    return 0;
}

}  // extern "C"
