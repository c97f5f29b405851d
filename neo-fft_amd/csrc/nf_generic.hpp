// nf_generic.hpp -- transforms of block lengths outside the 8/16/32-point
// codelets: a direct DFT per output (binary64 twiddles rounded to binary32,
// index-order sums), and per-lane codelet-order helpers for the axes of a
// generic block whose length is 8, 16 or 32.
#pragma once
#include "nf_device.hpp"
#include "nf_fft.hpp"

namespace nf {


__device__ __forceinline__ bool radix_length(int n) { return n == 8 || n == 16 || n == 32; }

// Output k of a direct DFT of n values x[j * st] (real when xi is null):
// binary64 twiddles rounded to binary32, index-order sums.
__device__ __forceinline__ void direct_term(int n, const float* xr, const float* xi, int st, int k, const float* tw,
                                            bool fwd, float& out_r, float& out_i) {
  float sr = 0.0f, si = 0.0f;
  for (int j = 0; j < n; ++j) {
    const int e = (j * k) % n;
    const float c = tw[2 * e];
    const float s = fwd ? -tw[2 * e + 1] : tw[2 * e + 1];
    const float r = xr[j * st], i = xi ? xi[j * st] : 0.0f;
    const float pr = r * c - i * s;
    const float pi = r * s + i * c;
    sr = sr + pr;
    si = si + pi;
  }
  out_r = sr;
  out_i = si;
}

template <int N>
__device__ __forceinline__ void row_forward_radix(const float* x, float* orr, float* oi) {
  float re[N], im[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    re[j] = x[j];
    im[j] = 0.0f;
  }
  cfft<N, true>(re, im);
#pragma unroll
  for (int k = 0; k < N / 2 + 1; ++k) {
    orr[k] = re[k];
    oi[k] = (k == N / 2) ? 0.0f : im[k];
  }
}

template <int N, bool FWD>
__device__ __forceinline__ void line_radix(float* lr, float* li, int st) {
  float re[N], im[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    re[j] = lr[j * st];
    im[j] = li[j * st];
  }
  cfft<N, FWD>(re, im);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    lr[j * st] = re[j];
    li[j * st] = im[j];
  }
}

template <int N>
__device__ __forceinline__ void row_inverse_radix(const float* zr, const float* zi, float fct, const float* wxs,
                                                  float* out, bool& bad) {
  constexpr int K = N / 2 + 1;
  float re[N], im[N];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    re[k] = zr[k];
    im[k] = zi[k];
  }
#pragma unroll
  for (int k = K; k < N; ++k) {
    re[k] = re[N - k];
    im[k] = 0.0f - im[N - k];
  }
  cfft<N, false>(re, im);
#pragma unroll
  for (int x = 0; x < N; ++x) {
    const float v = re[x] * fct;
    bad |= !isfinite(v);
    out[x] = v * wxs[x];
  }
}

// In-place transform of every column (n = bh values, stride K) of one block.
__device__ __forceinline__ void columns(int BH, int K, float* re, float* im, float* tr, float* ti, const float* twy,
                                        bool fwd, int tid, int nthreads) {
  if (radix_length(BH)) {
    for (int k = tid; k < K; k += nthreads) {
      if (BH == 8) fwd ? line_radix<8, true>(re + k, im + k, K) : line_radix<8, false>(re + k, im + k, K);
      else if (BH == 16) fwd ? line_radix<16, true>(re + k, im + k, K) : line_radix<16, false>(re + k, im + k, K);
      else fwd ? line_radix<32, true>(re + k, im + k, K) : line_radix<32, false>(re + k, im + k, K);
    }
    __syncthreads();
  } else {
    for (int t = tid; t < BH * K; t += nthreads) {
      const int ky = t / K, k = t - ky * K;
      direct_term(BH, re + k, im + k, K, ky, twy, fwd, tr[t], ti[t]);
    }
    __syncthreads();
    for (int t = tid; t < BH * K; t += nthreads) {
      re[t] = tr[t];
      im[t] = ti[t];
    }
    __syncthreads();
  }
}


}  // namespace nf
