/*
 * sfft_oracle.h -- CPU oracle for the Signalsmith FFT transform path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (fft_amd/, include/)
 * links, loads or calls this library.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py use it, and only
 * as the checker.
 *
 * It restates, in plain C and double precision, the algorithm of the
 * reference header signalsmith-fft.h (file:line citations on each function
 * in sfft_oracle.c):
 *   - the size helpers (validSize / sizeMinimum / sizeMaximum and the
 *     RealFFT variants with their quirks),
 *   - the complex FFT: trial-division factoriser, recursive step planner
 *     with radix-4 merging and twiddle tables, mixed-radix digit-reversal
 *     permutation, and the radix-2/3/4/generic DIT passes,
 *   - RealFFT / ModifiedRealFFT pack, half-size complex FFT and the
 *     split post-/pre-process with DC and Nyquist packed into bin 0.
 * Plus a long-double O(N^2) DFT used to pin the restatement by brute force.
 *
 * Pinning (each function is checked by tests/test_oracle.py):
 *   size helpers        -> bit-exact against the reference size tables
 *                          (tests/golden, n in [0, 70000]) and the
 *                          reference test "Sizes" (tests/00-fft.cpp:168-187)
 *   complex fft / ifft  -> reference FFT<double> golden vectors (rel L2
 *                          <= 1e-14), brute-force long-double DFT, single-bin
 *                          closed form, linearity, ifft(fft(x)) = N x,
 *                          Parseval
 *   real fft / ifft     -> reference RealFFT<double> / ModifiedRealFFT<double>
 *                          golden vectors, complex FFT of the real input
 *                          (tests/01-real.cpp:17-100), half-bin shift identity
 *   dft_ld              -> closed forms (single bins, delta, constant)
 */
#ifndef SFFT_ORACLE_H
#define SFFT_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Size helpers (integer exact). */
size_t oracle_size_minimum(size_t n);
size_t oracle_size_maximum(size_t n);
size_t oracle_real_size_minimum(size_t n);
size_t oracle_real_size_maximum(size_t n);

/* Complex FFT, double precision, interleaved (re, im).
 * batch transforms of length n, contiguous with stride n.
 * inverse != 0 gives the unnormalised inverse (round trip = n * x).
 * Out of place: in and out must not overlap.  Returns 0 on success. */
int oracle_fft_c2c(size_t n, size_t batch, const double *in, double *out, int inverse);

/* Real FFT, double precision.  n even.
 * forward: in = batch x n reals, out = batch x n/2 complex (interleaved).
 *          out[0] = (DC, Nyquist) unless half_shift.
 * inverse: in = batch x n/2 complex, out = batch x n reals.
 * half_shift selects ModifiedRealFFT (bins at k + 1/2). */
int oracle_fft_real(size_t n, size_t batch, const double *in, double *out,
                    int inverse, int half_shift);

/* Brute-force O(n^2) DFT in long double, output rounded to double.
 * sign = -1 forward (e^{-2 pi i nk/N}), +1 inverse (unnormalised). */
int oracle_dft_ld(size_t n, const double *in, double *out, int sign);

#ifdef __cplusplus
}
#endif

#endif
