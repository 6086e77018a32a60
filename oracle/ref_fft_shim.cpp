// ref_fft_shim.cpp -- C entry points over the reference's own CPU FFT.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) together
// with Source/FastFourierTransformCPU/FastFourierTransform.cpp read in place
// from the reference checkout; the result lands in oracle/_ref/ and is used
// only to pin the oracle's FFT restatement (ctr_oracle.c) against the
// reference's compiled code on the same inputs.
#include <vector>

#include "FastFourierTransformCPU/FastFourierTransform.h"

extern "C" {

// FFT::fft_radix2 (FastFourierTransform.cpp:49-88) on split re/im arrays.
int ref_fft_radix2(double *re, double *im, int n)
{
    std::vector<FFT::complex> v(n);
    for (int i = 0; i < n; i++) v[i] = FFT::complex(re[i], im[i]);
    try {
        FFT::fft_radix2(v);
    } catch (...) {
        return -1;
    }
    for (int i = 0; i < n; i++) { re[i] = v[i].r; im[i] = v[i].i; }
    return 0;
}

// FFT::inverse_fft_radix2 (FastFourierTransform.cpp:90-106).
int ref_inverse_fft_radix2(double *re, double *im, int n)
{
    std::vector<FFT::complex> v(n);
    for (int i = 0; i < n; i++) v[i] = FFT::complex(re[i], im[i]);
    try {
        FFT::inverse_fft_radix2(v);
    } catch (...) {
        return -1;
    }
    for (int i = 0; i < n; i++) { re[i] = v[i].r; im[i] = v[i].i; }
    return 0;
}

// FFT::fft_shift (FastFourierTransform.cpp:113-120).
void ref_fft_shift(double *re, double *im, int n, int shift)
{
    std::vector<FFT::complex> v(n);
    for (int i = 0; i < n; i++) v[i] = FFT::complex(re[i], im[i]);
    FFT::fft_shift(v, shift);
    for (int i = 0; i < n; i++) { re[i] = v[i].r; im[i] = v[i].i; }
}

double ref_pi(void) { return FFT::PI; }
}
