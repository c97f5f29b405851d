// nf_fft.hpp -- per-lane radix-2 complex transforms of length 8, 16 and 32,
// shared by the gfx950 kernels and the host plan builder.
//
// Operation order follows the reference's AVX2 codelets so that results are
// bit-identical to them up to compiler contraction: bit-reversed radix-2 DIT
// (src/spectral/codelets/fft16x16_avx2.cpp:27-201), the 8-point variant with
// its sqrt(1/2) rotations (fft8x8_avx2.cpp:49-140), and the 32-point
// even/odd split over two 16-point transforms (fft32x32_avx2.cpp:251-275).
// Twiddles are binary32 cos/sin of the binary32 angle k*PI/8 (k*PI/16), with
// PI the binary32 constant, as the codelets initialise them
// (fft16x16_avx2.cpp:11-25); the complex twiddle product is
// (fma(ir,wr,-(ii*wi)), fma(ir,wi,ii*wr)) (fft16x16_avx2.cpp:64-69).
// Everything else is compiled with -ffp-contract=off.
#pragma once
#include <cmath>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define NF_HD __host__ __device__ __forceinline__
#else
#define NF_HD inline
#endif

namespace nf {

struct Twiddle {
  static constexpr float w16r[8] = {0x1p+0f, 0x1.d906bcp-1f, 0x1.6a09e6p-1f, 0x1.87de2ap-2f,
                                    -0x1.777a5cp-25f, -0x1.87de3p-2f, -0x1.6a09e6p-1f, -0x1.d906cp-1f};
  static constexpr float w16i[8] = {0x0p+0f, 0x1.87de2cp-2f, 0x1.6a09e6p-1f, 0x1.d906bcp-1f,
                                    0x1p+0f, 0x1.d906bcp-1f, 0x1.6a09e6p-1f, 0x1.87de2p-2f};
  static constexpr float w32r[16] = {0x1p+0f, 0x1.f6297cp-1f, 0x1.d906bcp-1f, 0x1.a9b662p-1f,
                                     0x1.6a09e6p-1f, 0x1.1c73b2p-1f, 0x1.87de2ap-2f, 0x1.8f8b78p-3f,
                                     -0x1.777a5cp-25f, -0x1.8f8b84p-3f, -0x1.87de3p-2f, -0x1.1c73b8p-1f,
                                     -0x1.6a09e6p-1f, -0x1.a9b664p-1f, -0x1.d906cp-1f, -0x1.f6297ep-1f};
  static constexpr float w32i[16] = {0x0p+0f, 0x1.8f8b84p-3f, 0x1.87de2cp-2f, 0x1.1c73b4p-1f,
                                     0x1.6a09e6p-1f, 0x1.a9b664p-1f, 0x1.d906bcp-1f, 0x1.f6297ep-1f,
                                     0x1p+0f, 0x1.f6297cp-1f, 0x1.d906bcp-1f, 0x1.a9b66p-1f,
                                     0x1.6a09e6p-1f, 0x1.1c73b2p-1f, 0x1.87de2p-2f, 0x1.8f8b82p-3f};
};

NF_HD void bf2(float& xr, float& xi, float& yr, float& yi) {
  const float ur = xr, ui = xi, vr = yr, vi = yi;
  xr = ur + vr;
  xi = ui + vi;
  yr = ur - vr;
  yi = ui - vi;
}

template <bool FWD>
NF_HD void cmul(float& orr, float& oi, float ir, float ii, float wr, float wi_val) {
  const float wi = FWD ? -wi_val : wi_val;
  const float p = ii * wi, q = ii * wr;
  orr = fmaf(ir, wr, -p);
  oi = fmaf(ir, wi, q);
}

template <bool FWD>
NF_HD void mulj(float& orr, float& oi, float ir, float ii) {
  if (FWD) {
    orr = ii;
    oi = 0.0f - ir;
  } else {
    orr = 0.0f - ii;
    oi = ir;
  }
}

// Combine butterfly u +/- v with v already multiplied.
NF_HD void bfv(float& ur, float& ui, float& lr, float& li, float vr, float vi) {
  const float a = ur, b = ui;
  ur = a + vr;
  ui = b + vi;
  lr = a - vr;
  li = b - vi;
}

template <bool FWD>
NF_HD void fft8(float* r, float* i) {
  const float s = 0.70710678118654752440f;
  float ar[8] = {r[0], r[4], r[2], r[6], r[1], r[5], r[3], r[7]};
  float ai[8] = {i[0], i[4], i[2], i[6], i[1], i[5], i[3], i[7]};
  bf2(ar[0], ai[0], ar[1], ai[1]);
  bf2(ar[2], ai[2], ar[3], ai[3]);
  bf2(ar[4], ai[4], ar[5], ai[5]);
  bf2(ar[6], ai[6], ar[7], ai[7]);
  bf2(ar[0], ai[0], ar[2], ai[2]);
  bf2(ar[4], ai[4], ar[6], ai[6]);
  {
    float tr, ti;
    mulj<FWD>(tr, ti, ar[3], ai[3]);
    ar[3] = tr;
    ai[3] = ti;
    mulj<FWD>(tr, ti, ar[7], ai[7]);
    ar[7] = tr;
    ai[7] = ti;
  }
  bf2(ar[1], ai[1], ar[3], ai[3]);
  bf2(ar[5], ai[5], ar[7], ai[7]);
  bf2(ar[0], ai[0], ar[4], ai[4]);
  {
    const float ur = ar[5], ui = ai[5];
    if (FWD) {
      ar[5] = s * (ur + ui);
      ai[5] = s * (ui - ur);
    } else {
      ar[5] = s * (ur - ui);
      ai[5] = s * (ui + ur);
    }
  }
  bf2(ar[1], ai[1], ar[5], ai[5]);
  {
    float tr, ti;
    mulj<FWD>(tr, ti, ar[6], ai[6]);
    ar[6] = tr;
    ai[6] = ti;
  }
  bf2(ar[2], ai[2], ar[6], ai[6]);
  {
    const float ur = ar[7], ui = ai[7];
    if (FWD) {
      ar[7] = s * (ui - ur);
      ai[7] = s * ((0.0f - ur) - ui);
    } else {
      ar[7] = s * ((0.0f - ur) - ui);
      ai[7] = s * (ur - ui);
    }
  }
  bf2(ar[3], ai[3], ar[7], ai[7]);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    r[k] = ar[k];
    i[k] = ai[k];
  }
}

template <bool FWD>
NF_HD void fft16(float* r, float* i) {
  constexpr int br[16] = {0, 8, 4, 12, 2, 10, 6, 14, 1, 9, 5, 13, 3, 11, 7, 15};
  float ar[16], ai[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    ar[k] = r[br[k]];
    ai[k] = i[br[k]];
  }
#pragma unroll
  for (int k = 0; k < 16; k += 2) bf2(ar[k], ai[k], ar[k + 1], ai[k + 1]);
#pragma unroll
  for (int g = 0; g < 16; g += 4) {
    bf2(ar[g], ai[g], ar[g + 2], ai[g + 2]);
    float vr, vi;
    mulj<FWD>(vr, vi, ar[g + 3], ai[g + 3]);
    bfv(ar[g + 1], ai[g + 1], ar[g + 3], ai[g + 3], vr, vi);
  }
#pragma unroll
  for (int g = 0; g < 16; g += 8) {
    bf2(ar[g], ai[g], ar[g + 4], ai[g + 4]);
    float vr, vi;
    cmul<FWD>(vr, vi, ar[g + 5], ai[g + 5], Twiddle::w16r[2], Twiddle::w16i[2]);
    bfv(ar[g + 1], ai[g + 1], ar[g + 5], ai[g + 5], vr, vi);
    mulj<FWD>(vr, vi, ar[g + 6], ai[g + 6]);
    bfv(ar[g + 2], ai[g + 2], ar[g + 6], ai[g + 6], vr, vi);
    cmul<FWD>(vr, vi, ar[g + 7], ai[g + 7], Twiddle::w16r[6], Twiddle::w16i[6]);
    bfv(ar[g + 3], ai[g + 3], ar[g + 7], ai[g + 7], vr, vi);
  }
  bf2(ar[0], ai[0], ar[8], ai[8]);
#pragma unroll
  for (int k = 1; k < 8; ++k) {
    float vr, vi;
    if (k == 4)
      mulj<FWD>(vr, vi, ar[12], ai[12]);
    else
      cmul<FWD>(vr, vi, ar[8 + k], ai[8 + k], Twiddle::w16r[k], Twiddle::w16i[k]);
    bfv(ar[k], ai[k], ar[8 + k], ai[8 + k], vr, vi);
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    r[k] = ar[k];
    i[k] = ai[k];
  }
}

template <bool FWD>
NF_HD void fft32(float* r, float* i) {
  float er[16], ei[16], orr[16], oi[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    er[m] = r[2 * m];
    ei[m] = i[2 * m];
    orr[m] = r[2 * m + 1];
    oi[m] = i[2 * m + 1];
  }
  fft16<FWD>(er, ei);
  fft16<FWD>(orr, oi);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    float vr, vi;
    cmul<FWD>(vr, vi, orr[k], oi[k], Twiddle::w32r[k], Twiddle::w32i[k]);
    r[k] = er[k] + vr;
    i[k] = ei[k] + vi;
    r[k + 16] = er[k] - vr;
    i[k + 16] = ei[k] - vi;
  }
}

template <int N, bool FWD>
NF_HD void cfft(float* r, float* i) {
  static_assert(N == 8 || N == 16 || N == 32, "transform length");
  if constexpr (N == 8) fft8<FWD>(r, i);
  else if constexpr (N == 16) fft16<FWD>(r, i);
  else fft32<FWD>(r, i);
}

// Temporal-axis transforms of DFTTest's 3D spectrum, same operation order as
// the reference's dense small-volume butterflies (src/spectral/
// fft_backend.cpp:397-420) for depths 3, 4, 5 and its centre-slice inverse
// (:526-540). cscale multiplies both parts by a float constant, crot
// multiplies by -i (forward) or +i (inverse).
struct Cpx {
  float r, i;
};
NF_HD Cpx cadd(Cpx a, Cpx b) { return {a.r + b.r, a.i + b.i}; }
NF_HD Cpx csub(Cpx a, Cpx b) { return {a.r - b.r, a.i - b.i}; }
NF_HD Cpx cscale(Cpx a, float s) { return {a.r * s, a.i * s}; }
NF_HD Cpx crot(Cpx z, bool fwd) { return fwd ? Cpx{z.i, -z.r} : Cpx{-z.i, z.r}; }

struct TemporalK {
  static constexpr float k3 = 0.86602540378443864676f;
  static constexpr float k5a = 0.30901699437494742410f, k5b = 0.80901699437494742410f;
  static constexpr float k5c = 0.95105651629515357212f, k5d = 0.58778525229247312917f;
};

NF_HD void temporal_butterfly(int n, Cpx* v, bool fwd) {
  if (n == 3) {
    const Cpx sum = cadd(v[1], v[2]), diff = csub(v[1], v[2]);
    const Cpx center = csub(v[0], cscale(sum, 0.5f));
    const Cpx side = crot(cscale(diff, TemporalK::k3), fwd);
    v[0] = cadd(v[0], sum);
    v[1] = cadd(center, side);
    v[2] = csub(center, side);
  } else if (n == 4) {
    const Cpx es = cadd(v[0], v[2]), ed = csub(v[0], v[2]);
    const Cpx os = cadd(v[1], v[3]), od = crot(csub(v[1], v[3]), fwd);
    v[0] = cadd(es, os);
    v[2] = csub(es, os);
    v[1] = cadd(ed, od);
    v[3] = csub(ed, od);
  } else {
    const Cpx a = cadd(v[1], v[4]), b = cadd(v[2], v[3]), c = csub(v[1], v[4]), d = csub(v[2], v[3]);
    const Cpx base1 = csub(cadd(v[0], cscale(a, TemporalK::k5a)), cscale(b, TemporalK::k5b));
    const Cpx base2 = cadd(csub(v[0], cscale(a, TemporalK::k5b)), cscale(b, TemporalK::k5a));
    const Cpx side1 = crot(cadd(cscale(c, TemporalK::k5c), cscale(d, TemporalK::k5d)), fwd);
    const Cpx side2 = crot(csub(cscale(c, TemporalK::k5d), cscale(d, TemporalK::k5c)), fwd);
    v[0] = cadd(cadd(v[0], a), b);
    v[1] = cadd(base1, side1);
    v[4] = csub(base1, side1);
    v[2] = cadd(base2, side2);
    v[3] = csub(base2, side2);
  }
}

// Centre slice of the depth-3/5 inverse with the reference's operation order.
NF_HD Cpx temporal_inverse_center(int n, const Cpx* v) {
  if (n == 3) {
    const Cpx sum = cadd(v[1], v[2]), diff = csub(v[1], v[2]);
    const Cpx center = csub(v[0], cscale(sum, 0.5f));
    const Cpx side = cscale(diff, TemporalK::k3);
    return cadd(center, Cpx{-side.i, side.r});
  }
  const Cpx a = cadd(v[1], v[4]), b = cadd(v[2], v[3]), c = csub(v[1], v[4]), d = csub(v[2], v[3]);
  const Cpx base = cadd(csub(v[0], cscale(a, TemporalK::k5b)), cscale(b, TemporalK::k5a));
  const Cpx side = csub(cscale(c, TemporalK::k5d), cscale(d, TemporalK::k5c));
  return cadd(base, Cpx{-side.i, side.r});
}

// Direct DFT term k of length n: sum_j x[j] * (c_e, +-s_e), e = j*k mod n,
// with tw[2e] = float(cos(2 pi e / n)), tw[2e+1] = float(sin(2 pi e / n)).
NF_HD Cpx dft_term(int n, const Cpx* x, int k, const float* tw, bool fwd) {
  float sr = 0.0f, si = 0.0f;
  for (int j = 0; j < n; ++j) {
    const int e = (j * k) % n;
    const float c = tw[2 * e];
    const float s = fwd ? -tw[2 * e + 1] : tw[2 * e + 1];
    const float pr = x[j].r * c - x[j].i * s;
    const float pi = x[j].r * s + x[j].i * c;
    sr = sr + pr;
    si = si + pi;
  }
  return {sr, si};
}

// Host-side dispatch by runtime length (plan creation only): the radix-2
// codelet order for 8/16/32 points, else a direct DFT with binary64 twiddles
// rounded to binary32 and index-order sums (the order of the device path for
// those lengths).
inline bool cfft_supported(int n) { return n == 8 || n == 16 || n == 32; }
inline void cfft_host(int n, float* r, float* i, bool fwd) {
  if (n == 8) {
    fwd ? fft8<true>(r, i) : fft8<false>(r, i);
  } else if (n == 16) {
    fwd ? fft16<true>(r, i) : fft16<false>(r, i);
  } else if (n == 32) {
    fwd ? fft32<true>(r, i) : fft32<false>(r, i);
  } else if (n > 1) {
    const double pi = 3.1415926535897932384626433832795;
    float tw[2 * 256], xr[256], xi[256];
    for (int e = 0; e < n; ++e) {
      tw[2 * e] = (float)std::cos(2.0 * pi * (double)e / (double)n);
      tw[2 * e + 1] = (float)std::sin(2.0 * pi * (double)e / (double)n);
      xr[e] = r[e];
      xi[e] = i[e];
    }
    Cpx x[256];
    for (int e = 0; e < n; ++e) x[e] = Cpx{xr[e], xi[e]};
    for (int k = 0; k < n; ++k) {
      const Cpx v = dft_term(n, x, k, tw, fwd);
      r[k] = v.r;
      i[k] = v.i;
    }
  }
}

}  // namespace nf
