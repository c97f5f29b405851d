// nf_dfttest.hip -- DFTTest spatial (tbsize=1) path on gfx950.
//
// One workgroup owns an output tile [X0,X1)x[Y0,Y1) (cover coordinates) and
// walks the block rows that touch it from top to bottom. For each block row
// it stages the reflected, decoded source rows in LDS, runs every covering
// block through window -> 2D real FFT -> gain -> inverse -> synthesis window
// with one lane per block row/column transform, and adds the block outputs
// into an LDS ring of B accumulator rows in the reference's canonical order
// (block row, then block column; reference plan.cpp:591-602 and
// docs/specs/phase-5/dfttest/kernel-temporal-ola.md). Rows that no later
// block row touches are converted and stored. No global atomics, no global
// intermediate.
#include <type_traits>

#include "nf_device.hpp"
#include "nf_fft.hpp"
#include "nf_gain.hpp"
#include "nf_launch.hpp"
#include "nf_worklist.hpp"

namespace nf {

namespace {

// One output tile of job J; ORs NF_FLAG_NONFINITE into *flags on a non-finite
// sample or intermediate.
template <int B>
__device__ __forceinline__ void t1_tile(const DftJob& J, int tile, unsigned* flags) {
  extern __shared__ float smem[];
  constexpr int K = B / 2 + 1;
  const int tx = tile % J.tiles_x, ty = tile / J.tiles_x;
  const int S = J.S;
  const int tid = threadIdx.x, nthreads = blockDim.x;

  const int vx0 = J.gx.offset, vx1 = J.gx.offset + J.gx.length;
  const int vy0 = J.gy.offset, vy1 = J.gy.offset + J.gy.length;
  const int X0 = max(vx0, (J.tile_x0 + tx) * J.tile_w), X1 = min(vx1, (J.tile_x0 + tx + 1) * J.tile_w);
  const int Y0 = max(vy0, (J.tile_y0 + ty) * J.tile_h), Y1 = min(vy1, (J.tile_y0 + ty + 1) * J.tile_h);
  const int bx_lo = ceil_div_int(X0 - B + 1, S), bx_hi = (X1 - 1) / S;
  const int by_lo = ceil_div_int(Y0 - B + 1, S), by_hi = (Y1 - 1) / S;
  const int nbx = bx_hi - bx_lo + 1;
  const int WW = (nbx - 1) * S + B;
  const int cx0 = bx_lo * S;
  const int TW = X1 - X0;

  // LDS carve-up (floats). ww_max is an odd row stride; spectra are real and
  // imaginary planes with row stride KS and block stride BS chosen so that
  // the row and the column transforms both touch 32 distinct banks; windows
  // use a row stride of B + 1. A block's inverse rows overwrite its own
  // spectrum rows: output x of row y sits at re[y*KS + x] for x < KS, else
  // at im[y*KS + x - KS].
  constexpr int KS = t1_row_stride(B);
  constexpr int BS = t1_block_stride(B);
  constexpr int OS = B + 1;
  float* sre = smem;                                   // [nbx_max][BS]
  float* sim = sre + J.nbx_max * BS;                   // [nbx_max][BS]
  float* win = sim + J.nbx_max * BS;                   // [B][ww_max]
  float* ring = win + B * J.ww_max;                    // [B][tile_w]
  float* dc = ring + B * J.tile_w;                     // [nbx_max][2]
  float* hw = dc + 2 * J.nbx_max;                      // [B][OS]
  float* hs = hw + B * OS;                             // [B][OS]

  bool bad = false;
  for (int i = tid; i < B * B; i += nthreads) {
    const int y = i / B, x = i - y * B;
    hw[y * OS + x] = J.h[i];
    hs[y * OS + x] = J.hsyn[i];
  }
  for (int i = tid; i < B * J.tile_w; i += nthreads) ring[i] = 0.0f;
  const float fct = 1.0f / ((float)B * (float)B);
  const float volume = 1.0f * (float)B * (float)B;
  __syncthreads();

  for (int by = by_lo; by <= by_hi; ++by) {
    const int oy = by * S;
    // A: decoded, reflected source window [B][WW]; lanes own columns
    {
      const int R = nthreads >= WW ? nthreads / WW : 1;
      const int rl = R > 1 ? tid / WW : 0;
      for (int c = (R > 1 ? tid - rl * WW : tid); c < WW && rl < R; c += nthreads) {
        const int sx = reflect(cx0 + c, J.gx);
        for (int r = rl; r < B; r += R) {
          const int sy = reflect(oy + r, J.gy);
          const float v = load_sample(J.src, J.src_stride, J.storage, sy, sx) * J.dscale;
          bad |= !isfinite(v);
          win[r * J.ww_max + c] = v;
        }
      }
    }
    __syncthreads();
    // B: windowed rows -> forward row transforms, bins 0..B/2
    for (int t = tid; t < nbx * B; t += nthreads) {
      const int bi = t / B, y = t - bi * B;
      float re[B], im[B];
      const float* src = win + y * J.ww_max + bi * S;
#pragma unroll
      for (int x = 0; x < B; ++x) {
        re[x] = src[x] * hw[y * OS + x];
        im[x] = 0.0f;
      }
      cfft<B, true>(re, im);
      float* dr = sre + bi * BS + y * KS;
      float* di = sim + bi * BS + y * KS;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        dr[k] = re[k];
        di[k] = (k == B / 2) ? 0.0f : im[k];
      }
    }
    __syncthreads();
    // C: forward column transforms
    for (int t = tid; t < nbx * K; t += nthreads) {
      const int bi = t / K, k = t - bi * K;
      float re[B], im[B];
      float* cr = sre + bi * BS + k;
      float* ci = sim + bi * BS + k;
#pragma unroll
      for (int y = 0; y < B; ++y) {
        re[y] = cr[y * KS];
        im[y] = ci[y * KS];
      }
      cfft<B, true>(re, im);
#pragma unroll
      for (int y = 0; y < B; ++y) {
        cr[y * KS] = re[y];
        ci[y * KS] = im[y];
      }
      if (k == 0) {
        dc[2 * bi] = re[0];
        dc[2 * bi + 1] = im[0];
      }
    }
    __syncthreads();
    // D: gain (zmean residual) and inverse column transforms
    auto stage_d = [&](auto wiener) {
      for (int t = tid; t < nbx * K; t += nthreads) {
        const int bi = t / K, k = t - bi * K;
        const float gs = J.grid ? (J.mean_scale * dc[2 * bi]) / J.grid[0].x : 0.0f;
        float re[B], im[B];
        float* cr = sre + bi * BS + k;
        float* ci = sim + bi * BS + k;
#pragma unroll
        for (int y = 0; y < B; ++y) {
          re[y] = cr[y * KS];
          im[y] = ci[y * KS];
          const int bin = y * K + k;
          float mr = 0.0f, mi = 0.0f;
          if (J.grid) {
            const float2 gv = J.grid[bin];
            mr = gs * gv.x;
            mi = gs * gv.y;
          }
          apply_gain<decltype(wiener)::value>(J.g, bin, re[y], im[y], mr, mi, bad);
        }
        cfft<B, false>(re, im);
#pragma unroll
        for (int y = 0; y < B; ++y) {
          cr[y * KS] = re[y];
          ci[y * KS] = im[y];
        }
      }
    };
    if (gain_is_wiener(J.g)) stage_d(std::true_type{});
    else stage_d(std::false_type{});
    __syncthreads();
    // E: inverse row transforms with Hermitian extension, synthesis window
    for (int t = tid; t < nbx * B; t += nthreads) {
      const int bi = t / B, y = t - bi * B;
      float re[B], im[B];
      float* rr = sre + bi * BS + y * KS;
      float* ri = sim + bi * BS + y * KS;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        re[k] = rr[k];
        im[k] = ri[k];
      }
#pragma unroll
      for (int k = K; k < B; ++k) {
        re[k] = re[B - k];
        im[k] = 0.0f - im[B - k];
      }
      cfft<B, false>(re, im);
#pragma unroll
      for (int x = 0; x < B; ++x) {
        float v = re[x] * fct;
        bad |= !isfinite(v);
        if (J.tmode1) v = volume * v;
        if (x < KS) rr[x] = v * hs[y * OS + x];
        else ri[x - KS] = v * hs[y * OS + x];
      }
    }
    __syncthreads();
    // F: overlap-add into the ring, block columns in ascending order; lanes
    // own output columns (their covering block range is computed once)
    const int RT = nthreads >= TW ? nthreads / TW : 1;
    const int rlt = RT > 1 ? tid / TW : 0;
    const int c_first = RT > 1 ? tid - rlt * TW : tid;
    {
      const int r0 = max(oy, Y0), r1 = min(oy + B, Y1);
      const int nr = r1 - r0;
      for (int c = c_first; c < TW && rlt < RT; c += nthreads) {
        const int x = X0 + c;
        const int b_first = max(bx_lo, ceil_div_int(x - B + 1, S));
        const int b_last = min(bx_hi, x / S);
        for (int rr = rlt; rr < nr; rr += RT) {
          const int y = r0 + rr;
          float* cell = ring + (y % B) * J.tile_w + c;
          float sum = *cell;
          for (int bx = b_first; bx <= b_last; ++bx) {
            const int bi = bx - bx_lo;
            const int xx = x - bx * S;
            const int row = bi * BS + (y - oy) * KS;
            sum = sum + (xx < KS ? sre[row + xx] : sim[row + xx - KS]);
          }
          *cell = sum;
        }
      }
    }
    __syncthreads();
    // G: rows [oy, oy+S) are final: convert, store, clear their ring rows
    {
      const int r0 = max(oy, Y0), r1 = min(oy + S, Y1);
      const int nr = r1 - r0;
      for (int c = c_first; c < TW && rlt < RT; c += nthreads) {
        const int x = X0 + c, oxx = x - vx0;
        for (int rr = rlt; rr < nr; rr += RT) {
          const int y = r0 + rr;
          float* cell = ring + (y % B) * J.tile_w + c;
          const float a = *cell;
          *cell = 0.0f;
          bad |= !isfinite(a);
          const int oyy = y - vy0;
          if (J.acc_out) {
            J.acc_out[(long long)oyy * J.acc_stride + oxx] = a;
          } else {
            char* drow = static_cast<char*>(J.dst) + (long long)oyy * J.dst_stride;
            if (J.storage == kF32) {
              reinterpret_cast<float*>(drow)[oxx] = a * J.oscale;
            } else {
              const float v = (a * J.oscale) + 0.5f;
              const float cl = v < 0.0f ? 0.0f : (J.peak < v ? J.peak : v);
              if (J.storage == kU8) reinterpret_cast<uint8_t*>(drow)[oxx] = (uint8_t)cl;
              else reinterpret_cast<uint16_t*>(drow)[oxx] = (uint16_t)cl;
            }
          }
        }
      }
    }
    // The next block row's stages A..E touch neither the ring nor the rows
    // cleared above; stage F runs after four more barriers.
  }
  if (bad) atomicOr(flags, NF_FLAG_NONFINITE);
}

}  // namespace

template <int B>
__global__ __launch_bounds__(256) void dfttest_t1_kernel(const DftLaunch L) {
  int jid = 0;
  while (jid + 1 < L.njobs && (int)blockIdx.x >= L.tile_start[jid + 1]) ++jid;
  t1_tile<B>(L.jobs[jid], (int)blockIdx.x - L.tile_start[jid], L.flags);
}

// sbsize 32 holds two waves per SIMD; the explicit budget keeps the loop
// form free of spills
template <int B>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(B == 32 ? 2 : 1))) void dfttest_t1_list_kernel(
    const DftWorkList W) {
  for_each_work_item(W, [&](int j, int tile) { t1_tile<B>(list_job(W, j), tile, W.flags); });
}

size_t dfttest_t1_lds_bytes(int B, const DftJob& j) {
  return sizeof(float) * ((size_t)j.nbx_max * t1_block_stride(B) * 2 + (size_t)B * j.ww_max + (size_t)B * j.tile_w +
                          2 * (size_t)j.nbx_max + 2 * (size_t)B * (B + 1));
}

hipError_t launch_dfttest_t1(int B, const DftLaunch& L, int total_tiles, size_t lds, hipStream_t stream) {
  dim3 grid(total_tiles), block(256);
  switch (B) {
    case 8:
      hipLaunchKernelGGL(dfttest_t1_kernel<8>, grid, block, lds, stream, L);
      break;
    case 16:
      hipLaunchKernelGGL(dfttest_t1_kernel<16>, grid, block, lds, stream, L);
      break;
    case 32:
      hipLaunchKernelGGL(dfttest_t1_kernel<32>, grid, block, lds, stream, L);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_dfttest_t1_list(int B, const DftWorkList& W, int workgroups, size_t lds, hipStream_t stream) {
  dim3 grid(workgroups), block(256);
  switch (B) {
    case 8:
      hipLaunchKernelGGL(dfttest_t1_list_kernel<8>, grid, block, lds, stream, W);
      break;
    case 16:
      hipLaunchKernelGGL(dfttest_t1_list_kernel<16>, grid, block, lds, stream, W);
      break;
    case 32:
      hipLaunchKernelGGL(dfttest_t1_list_kernel<32>, grid, block, lds, stream, W);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

const void* dfttest_t1_list_function(int B) {
  switch (B) {
    case 8: return (const void*)dfttest_t1_list_kernel<8>;
    case 16: return (const void*)dfttest_t1_list_kernel<16>;
    case 32: return (const void*)dfttest_t1_list_kernel<32>;
    default: return nullptr;
  }
}

hipError_t dfttest_t1_set_lds_limit() {
  const void* fns[] = {(const void*)dfttest_t1_kernel<8>, (const void*)dfttest_t1_kernel<16>,
                       (const void*)dfttest_t1_kernel<32>};
  const void* lists[] = {(const void*)dfttest_t1_list_kernel<8>, (const void*)dfttest_t1_list_kernel<16>,
                         (const void*)dfttest_t1_list_kernel<32>};
  for (const void* f : fns) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
  }
  for (const void* f : lists) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kListLdsLimit);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace nf
