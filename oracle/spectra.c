/*
 * spectra.c -- oracle restatement of the absorption spectrum.  TEST INFRASTRUCTURE.
 *
 * The reference forms C(t) = <psi(0)|psi(t)> with cublasZdotc after every
 * stored step (spectraClass.cu:509-571), multiplies it by a Gaussian
 * window exp(-0.5((n+1)dt/sigma)^2) for n < T (gpuSpectraHelper.cu:16-42,
 * spectraClass.cu:585-603), runs an unnormalised inverse-direction cuFFT of
 * length T+1 (spectraClass.cu:617-626) and scales by |mu|^2/T
 * (spectraClass.cu:641-686).  The oracle zero-pads to nfft and puts the
 * frequency axis on k/(nfft dt).
 */
#include "dq_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

int orc_fft(int n, double *x, int sign)
{
    if (n <= 0 || (n & (n - 1))) return -1;
    /* bit reversal */
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t0 = x[2 * i], t1 = x[2 * i + 1];
            x[2 * i] = x[2 * j]; x[2 * i + 1] = x[2 * j + 1];
            x[2 * j] = t0; x[2 * j + 1] = t1;
        }
    }
    const double pi = atan(1.0) * 4.0;
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; ++k) {
            double ang = sign * 2.0 * pi * (double)k / (double)len;
            double wr = cos(ang), wi = sin(ang);
            for (int i = 0; i < n; i += len) {
                double *a = x + 2 * (i + k), *b = x + 2 * (i + k + half);
                double tr = b[0] * wr - b[1] * wi, ti = b[0] * wi + b[1] * wr;
                b[0] = a[0] - tr; b[1] = a[1] - ti;
                a[0] = a[0] + tr; a[1] = a[1] + ti;
            }
        }
    }
    return 0;
}

void orc_dft(int n, const double *x, double *y, int sign)
{
    const double pi = atan(1.0) * 4.0;
    for (int k = 0; k < n; ++k) {
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < n; ++j) {
            long long kj = ((long long)k * j) % n;
            double ang = sign * 2.0 * pi * (double)kj / (double)n;
            double c = cos(ang), s = sin(ang);
            sr += x[2 * j] * c - x[2 * j + 1] * s;
            si += x[2 * j] * s + x[2 * j + 1] * c;
        }
        y[2 * k] = sr;
        y[2 * k + 1] = si;
    }
}

int orc_spectrum(int T, const double *corr, double dt_fs, double sigma_fs,
                 double mu2, int nfft, double *omega, double *spec)
{
    if (T + 1 > nfft) return -1;
    memset(spec, 0, sizeof(double) * 2 * nfft);
    for (int n = 0; n <= T; ++n) {
        double w = 1.0;
        if (sigma_fs > 0.0 && n < T) {
            double t = dt_fs * (double)(n + 1);
            w = exp(-0.5 * (t / sigma_fs) * (t / sigma_fs));
        }
        spec[2 * n] = corr[2 * n] * w;
        spec[2 * n + 1] = corr[2 * n + 1] * w;
    }
    if (orc_fft(nfft, spec, +1) != 0) return -1;
    /* spectraClass.cu:147-151 */
    double inv_fs_to_inv_cm = 1. / (299792458. * 100. * 1.e-15);
    double scale = mu2 / (double)T;
    for (int k = 0; k < nfft; ++k) {
        int kk = k > nfft / 2 ? k - nfft : k;
        if (omega) omega[k] = inv_fs_to_inv_cm * (double)kk / (dt_fs * (double)nfft);
        spec[2 * k] *= scale;
        spec[2 * k + 1] *= scale;
    }
    return 0;
}
