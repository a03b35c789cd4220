/*
 * oracle.h — float64 CPU restatement of gpu-rf-pipeline's hot-path definitions.
 *
 * TEST INFRASTRUCTURE ONLY. This is the oracle the GPU results are compared with. Only the
 * tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the library
 * (librfhip.so) never links or calls it.
 *
 * Complex arrays are interleaved doubles [re0, im0, re1, im1, ...] (numpy complex128 layout).
 * Each function cites the reference definition it restates (tools/reference.py, src/...).
 */
#ifndef RF_ORACLE_H
#define RF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* y[n] = sum_{k<=min(n,T-1)} taps[k] x[n-k]; tools/reference.py:103-111, src/fir.cu:1-5. */
void orc_fir(const double* x, long n, const double* taps, long num_taps, double* y);

/* r[i] = sum_n b[n+k] conj(a[n]), k = i-(na-1); tools/reference.py:114-125, src/xcorr.cu:3-7. */
void orc_xcorr(const double* a, long na, const double* b, long nb, double* r);

/* In-place unnormalised DFT of n points: sign -1 forward (exp(-2 pi i nk/N)), +1 inverse.
 * Radix-2 for powers of two, direct summation otherwise. Returns 0 on success. */
int orc_fft(double* data, long n, int sign);

/* Framed forward FFT, last frame zero-padded; tools/reference.py:88-95. out: frames*nfft. */
int orc_spectrum(const double* x, long n, long nfft, double* out);

/* |x|; tools/reference.py:98-100. */
void orc_magnitude(const double* x, long n, double* mag);

/* Polyphase branch sums and an M-point unnormalised inverse DFT per frame;
 * tools/reference.py:128-146, src/channelize.cu:3-11. out: ceil(n/M)*M values. */
int orc_channelize(const double* x, long n, const double* taps, long num_taps, long channels,
                   double* out);

/* GCC-PHAT; tools/reference.py:149-162. r: na+nb-1 values in xcorr's lag order. */
int orc_gcc_phat(const double* a, long na, const double* b, long nb, double* r);

/* First index of the largest |r[k]| (np.argmax order, src/main.cpp:390-401). */
long orc_peak(const double* r, long n, double* magnitude);

/* Least-squares angle in degrees from per-channel lags (channel i >= 1 behind channel 0);
 * tools/reference.py:171-178, src/main.cpp:402-406. */
double orc_doa(const long* lags, long count, double spacing, double speed, double rate);

#ifdef __cplusplus
}
#endif

#endif /* RF_ORACLE_H */
