// nf_dfttest_mfma.hip -- DFTTest with the block DFTs as dense DFT-matrix
// products on the CDNA4 matrix cores (v_mfma_f32_16x16x16_f16), operands
// split into FP16 high and low parts with FP32 accumulation.
//
// This is the precision form beside the FP32 radix-2 path: the same filter
// (reference plan.cpp:538-640; gains scalar/spectral.cpp:5-53; temporal
// butterflies fft_backend.cpp:397-420), with every 16-point spatial
// transform evaluated as a 16x16x16 matrix product. A value v is carried as
// hi = fp16(v), lo = fp16(v - hi) (both rounded toward zero); a product A*B
// is hi*hi + lo*hi + hi*lo
// (the lo*lo term, 2^-22 relative, is dropped), accumulated in FP32. The
// DFT-matrix constants are held at 256x scale (|c| >= 0.38 for the non-zero
// entries, so their low parts stay normal FP16 numbers) and every product is
// brought back with an exact 1/256 in FP32. The magnitudes entering the FP16
// splits are checked: a block whose samples or spectrum would leave the FP16
// range raises NF_FLAG_RANGE instead of returning saturated values. Results
// are not bit-identical to the radix-2 path; their error against a binary64
// evaluation is measured by the tests.
//
// One wave transforms one 16x16 block as four chained products whose
// results stay in the accumulator layout (lane l holds column l%16, rows
// 4*(l/16)..+3), which is exactly the operand layout the next product needs:
//   R   = x  . Cf      rows y, columns kx   (row transform, kx < 9 kept)
//   X   = F  . R      rows ky, columns kx   (column transform)
//   Z^T = X^T . Fi/16 rows kx, columns yy   (inverse column transform)
//   y^T = M  . Z^T    rows xx, columns yy   (inverse row transform, Hermitian
//                                            weights 1,2,..,2,1 and 1/16)
// The temporal transform (tbsize > 1), the zmean template, the gain and the
// synthesis window run on the VALU between the products. Overlap-add uses
// the strip kernel's LDS ring in the canonical order.
#include "nf_device.hpp"
#include "nf_fft.hpp"
#include "nf_gain.hpp"
#include "nf_launch.hpp"
#include "nf_worklist.hpp"

namespace nf {

namespace {

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct Hl {
  h4 hi, lo;
};

// hi and lo come from one packed conversion each, so the hi that enters the
// product is the very value the residual was taken from (two separately
// emitted conversions of the same float could round a tie differently).
__device__ __forceinline__ Hl split(f4 v) {
  Hl s;
  const auto h01 = __builtin_amdgcn_cvt_pkrtz(v[0], v[1]);
  const auto h23 = __builtin_amdgcn_cvt_pkrtz(v[2], v[3]);
  const float r0 = v[0] - (float)h01[0], r1 = v[1] - (float)h01[1];
  const float r2 = v[2] - (float)h23[0], r3 = v[3] - (float)h23[1];
  const auto l01 = __builtin_amdgcn_cvt_pkrtz(r0, r1);
  const auto l23 = __builtin_amdgcn_cvt_pkrtz(r2, r3);
  s.hi[0] = h01[0];
  s.hi[1] = h01[1];
  s.hi[2] = h23[0];
  s.hi[3] = h23[1];
  s.lo[0] = l01[0];
  s.lo[1] = l01[1];
  s.lo[2] = l23[0];
  s.lo[3] = l23[1];
  return s;
}

__device__ __forceinline__ f4 mma(const Hl& a, const Hl& b, f4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.hi, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x16f16(a.lo, b.hi, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.lo, c, 0, 0, 0);
  return c;
}

constexpr float kConstScale = 256.0f, kUnscale = 1.0f / 256.0f;

__device__ __forceinline__ f4 unscale(f4 v) {
#pragma unroll
  for (int e = 0; e < 4; ++e) v[e] = v[e] * kUnscale;
  return v;
}

// Largest magnitude of a split operand: the FP16 maximum finite value.
constexpr float kHalfMax = 65504.0f;

__device__ __forceinline__ float abs_max(float m, f4 v) {
  return fmaxf(fmaxf(m, fmaxf(fabsf(v[0]), fabsf(v[1]))), fmaxf(fabsf(v[2]), fabsf(v[3])));
}

__device__ __forceinline__ Hl split_scaled(const float* v, float scale) {
  f4 t;
#pragma unroll
  for (int e = 0; e < 4; ++e) t[e] = v[e] * scale;
  return split(t);
}

// A: decoded, reflected source rows of every slot into win[TT][B][ww]; lanes
// own window columns (one column reflection each) and step over rows.
template <int B, int TT>
__device__ __forceinline__ void stage_sources(const DftJob& J, const DftTemporal& Q, float* win, int ww, int oy,
                                              int cx0, int WW, int tid, int nthreads, bool& bad) {
  const int R = nthreads >= WW ? nthreads / WW : 1;
  const int rl = R > 1 ? tid / WW : 0;
  for (int c = (R > 1 ? tid - rl * WW : tid); c < WW && rl < R; c += nthreads) {
    const int sx = reflect(cx0 + c, J.gx);
    for (int zr = rl; zr < TT * B; zr += R) {
      const int z = zr / B, rr = zr - z * B;
      const int f = Q.slot_frame[z];
      const int sy = reflect(oy + rr, J.gy);
      const float v = load_sample(Q.frames[f], Q.frame_stride[f], J.storage, sy, sx) * J.dscale;
      bad |= !isfinite(v);
      win[zr * ww + c] = v;
    }
  }
}

// F and G: overlap-add of the block outputs out[nbx][B][OS] of block row oy
// into the ring (block columns ascending), then conversion and store of the
// rows no later block row touches.
template <int B, int OS>
__device__ __forceinline__ void stage_overlap_store(const DftJob& J, const float* out, float* ring, int oy, int X0,
                                                    int X1, int Y0, int Y1, int bx_lo, int bx_hi, int S, int tid,
                                                    int nthreads, bool& bad) {
  const int vx0 = J.gx.offset, vy0 = J.gy.offset;
  const int TW = X1 - X0;
  // F: overlap-add into the ring, block columns in ascending order; lanes
  // own output columns
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
        for (int bx = b_first; bx <= b_last; ++bx)
          sum = sum + out[((bx - bx_lo) * B + (y - oy)) * OS + (x - bx * S)];
        *cell = sum;
      }
    }
  }
  __syncthreads();
  // G: rows [oy, oy+S) are final
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

// One output tile of job J with temporal inputs Q. tw: cos/sin of 2*pi*k/B,
// k = 0..B-1, as (cos, sin) pairs (binary64 -> float). ORs NF_FLAG_NONFINITE
// into *flags on a non-finite sample or intermediate.
//
// B = 8 packs four blocks into each 16x16 tile (tile rows 0-7 hold blocks
// u, u+1 side by side, rows 8-15 blocks u+2, u+3) and uses block-diagonal
// DFT matrices, so the same four products transform four blocks; the
// transposed products swap the off-diagonal quadrants, which the gain and
// output index maps follow.
template <int B, int TT, bool WIENER>
__device__ __forceinline__ void mfma_tile(const DftJob& J, const DftTemporal& Q, int tile, const float* tw,
                                          unsigned* flags) {
  // OS: row stride of the windows and block outputs, a multiple of four with
  // OS/4 odd, so the float4 accesses of eight lanes (rows r..r+7) fall on
  // distinct bank groups; the source window uses the same rule (ww).
  constexpr int K = B / 2 + 1, OS = mfma_out_stride(B), BINS2 = B * K;
  constexpr int BQ = 16 / B, PER_UNIT = BQ * BQ;  // blocks per tile side / per tile
  constexpr float kInvB = 1.0f / (float)B;
  // split-operand bounds: |R| <= B max|x| and |Z| <= 2 max(|Yr|, |Yi|)
  constexpr float kInLimit = kHalfMax / (float)B, kSpecLimit = 0.5f * kHalfMax;
  float xmax = 0.0f, ymax = 0.0f;
  extern __shared__ __align__(16) float smem[];
  const int tx = tile % J.tiles_x, ty = tile / J.tiles_x;
  const int S = J.S;
  const int tid = threadIdx.x, nthreads = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nwaves = nthreads >> 6;
  const int r = lane & 15, q = lane >> 4;

  const int vx0 = J.gx.offset, vx1 = J.gx.offset + J.gx.length;
  const int vy0 = J.gy.offset, vy1 = J.gy.offset + J.gy.length;
  const int X0 = max(vx0, (J.tile_x0 + tx) * J.tile_w), X1 = min(vx1, (J.tile_x0 + tx + 1) * J.tile_w);
  const int Y0 = max(vy0, (J.tile_y0 + ty) * J.tile_h), Y1 = min(vy1, (J.tile_y0 + ty + 1) * J.tile_h);
  const int bx_lo = ceil_div_int(X0 - B + 1, S), bx_hi = (X1 - 1) / S;
  const int by_lo = ceil_div_int(Y0 - B + 1, S), by_hi = (Y1 - 1) / S;
  const int nbx = bx_hi - bx_lo + 1;
  const int WW = (nbx - 1) * S + B;
  const int ww = mfma_window_stride(J.ww_max);
  const bool vec = (S & 3) == 0;  // block origins on float4 boundaries
  const int cx0 = bx_lo * S;
  const int TW = X1 - X0;
  const int target = TT / 2;

  float* win = smem;                              // [TT][B][ww]
  float* out = win + TT * B * ww;                 // [nbx_max][B][OS]
  float* ring = out + J.nbx_max * B * OS;         // [B][tile_w]
  float* hw = ring + B * J.tile_w;                // [TT][B][OS]
  float* hs = hw + TT * B * OS;                   // [B][OS]

  bool bad = false;
  for (int i = tid; i < TT * B * B; i += nthreads) hw[(i / B) * OS + i % B] = J.h[i];
  for (int i = tid; i < B * B; i += nthreads) hs[(i / B) * OS + i % B] = J.hsyn[target * B * B + i];
  for (int i = tid; i < B * J.tile_w; i += nthreads) ring[i] = 0.0f;

  // Per-lane DFT-matrix fragments; entry (r, m = 4q+e) is cos/sin(2 pi r m / B)
  // inside a diagonal B x B block of the tile, 0 outside.
  const int rb = r % B;
  float cs[4], sn[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int m = 4 * q + e, mb = m % B, idx = (rb * mb) & (B - 1);
    const bool diag = r / B == m / B;
    cs[e] = diag ? tw[2 * idx] : 0.0f;
    sn[e] = diag ? tw[2 * idx + 1] : 0.0f;
  }
  float t[4];
  // row transform (B operand, k = x = m, j = kx = r): cos, -sin, kx < K
#pragma unroll
  for (int e = 0; e < 4; ++e) t[e] = rb < K ? cs[e] : 0.0f;
  const Hl Cf = split_scaled(t, kConstScale);
#pragma unroll
  for (int e = 0; e < 4; ++e) t[e] = rb < K ? -sn[e] : 0.0f;
  const Hl Sf = split_scaled(t, kConstScale);
  // column transform (A operand, i = ky = r, k = y = m): cos, sin, -sin
  const Hl C2 = split_scaled(cs, kConstScale);
  const Hl S2 = split_scaled(sn, kConstScale);
  const Hl S2n = split_scaled(sn, -kConstScale);
  // inverse column transform (B operand, k = ky = m, j = yy = r), 1/B
  const Hl C3 = split_scaled(cs, kConstScale * kInvB);
  const Hl S3 = split_scaled(sn, kConstScale * kInvB);
  const Hl S3n = split_scaled(sn, -kConstScale * kInvB);
  // inverse row transform (A operand, i = xx = r, k = kx = m): Hermitian
  // weights, 1/B, kx < K
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int mb = (4 * q + e) % B;
    t[e] = mb < K ? ((mb == 0 || mb == B / 2) ? 1.0f : 2.0f) * cs[e] * kInvB : 0.0f;
  }
  const Hl Mc = split_scaled(t, kConstScale);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int mb = (4 * q + e) % B;
    t[e] = mb < K ? -((mb == 0 || mb == B / 2) ? 1.0f : 2.0f) * sn[e] * kInvB : 0.0f;
  }
  const Hl Ms = split_scaled(t, kConstScale);
  // index maps of this lane's four elements: input rows/columns (A layout,
  // row r, columns 4q+e), spectrum bins (D layout, rows 4q+e, column r) and
  // outputs (transposed D layout, rows xx = 4q+e, columns yy = r)
  const int in_blk = BQ * (r / B) + (4 * q) / B, in_y = rb, in_x0 = (4 * q) % B;
  // spectrum block of this lane: BQ * ((4 * q) / B) + r / B
  const int sp_ky0 = (4 * q) % B, sp_kx = rb;
  const int dc_lane = 16 * ((4 * q) / B) * (B / 4) + B * (r / B);
  const int out_blk = BQ * (r / B) + (4 * q) / B, out_yy = rb, out_xx0 = (4 * q) % B;
  const float tscale = TT > 1 ? 1.0f / (float)TT : 1.0f;
  const bool use_grid = J.grid && J.mean_scale != 0.0f && J.grid[0].x != 0.0f;
  __syncthreads();

  for (int by = by_lo; by <= by_hi; ++by) {
    const int oy = by * S;
    stage_sources<B, TT>(J, Q, win, ww, oy, cx0, WW, tid, nthreads, bad);
    __syncthreads();
    // M: each wave transforms NP tiles (of PER_UNIT blocks) at a time
    // (independent product chains interleave): forward products, gain,
    // inverse products
    constexpr int NP = TT >= 5 ? 1 : 2;
    const int nunits = (nbx + PER_UNIT - 1) / PER_UNIT;
    for (int b0 = wave * NP; b0 < nunits; b0 += nwaves * NP) {
      f4 Xr[NP][TT], Xi[NP][TT];
      const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int z = 0; z < TT; ++z) {
        Hl xs[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          const int bi = min((b0 + p) * PER_UNIT + in_blk, nbx - 1);
          f4 xv;
          if (vec) {
            const float4 a = *reinterpret_cast<const float4*>(win + (z * B + in_y) * ww + bi * S + in_x0);
            const float4 h = *reinterpret_cast<const float4*>(hw + (z * B + in_y) * OS + in_x0);
            xv[0] = a.x * h.x;
            xv[1] = a.y * h.y;
            xv[2] = a.z * h.z;
            xv[3] = a.w * h.w;
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int x = in_x0 + e;
              xv[e] = win[(z * B + in_y) * ww + bi * S + x] * hw[(z * B + in_y) * OS + x];
            }
          }
          xmax = abs_max(xmax, xv);
          xs[p] = split(xv);
        }
        Hl Rr[NP], Ri[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          Rr[p] = split(unscale(mma(xs[p], Cf, zero)));
          Ri[p] = split(unscale(mma(xs[p], Sf, zero)));
        }
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          Xr[p][z] = unscale(mma(S2, Ri[p], mma(C2, Rr[p], zero)));
          Xi[p][z] = unscale(mma(S2n, Rr[p], mma(C2, Ri[p], zero)));
        }
      }
      Hl Ar[NP], Ai[NP];
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        // temporal forward transform per bin (same lane, same element in every slot)
        if (TT > 1) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            Cpx v[TT > 1 ? TT : 2];
#pragma unroll
            for (int z = 0; z < TT; ++z) v[z] = Cpx{Xr[p][z][e], Xi[p][z][e]};
            temporal_butterfly(TT, v, true);
#pragma unroll
            for (int z = 0; z < TT; ++z) {
              Xr[p][z][e] = v[z].r;
              Xi[p][z][e] = v[z].i;
            }
          }
        }
        // zmean scale from the (3D) DC of this lane's block: element 0 of
        // the lane holding bin (0, 0) of the block
        const float dcr = __shfl(Xr[p][0][0], dc_lane, 64);
        const float gs = use_grid ? (J.mean_scale * dcr) / J.grid[0].x : 0.0f;
        // gain over the (3D) spectrum
#pragma unroll
        for (int z = 0; z < TT; ++z) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int ky = sp_ky0 + e;
            if (sp_kx < K) {
              const int k3 = z * BINS2 + ky * K + sp_kx;
              float re = Xr[p][z][e], im = Xi[p][z][e], mr = 0.0f, mi = 0.0f;
              if (J.grid) {
                const float2 gv = J.grid[k3];
                mr = gs * gv.x;
                mi = gs * gv.y;
              }
              apply_gain<WIENER>(J.g, k3, re, im, mr, mi, bad);
              Xr[p][z][e] = re;
              Xi[p][z][e] = im;
            } else {
              Xr[p][z][e] = 0.0f;
              Xi[p][z][e] = 0.0f;
            }
          }
        }
        // inverse temporal transform of the target slice
        f4 Yr = Xr[p][0], Yi = Xi[p][0];
        if (TT > 1) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            Cpx v[TT > 1 ? TT : 2];
#pragma unroll
            for (int z = 0; z < TT; ++z) v[z] = Cpx{Xr[p][z][e], Xi[p][z][e]};
            Cpx res;
            if (TT == 3 || TT == 5) {
              res = temporal_inverse_center(TT, v);
            } else {
              temporal_butterfly(TT, v, false);
              res = v[target];
            }
            Yr[e] = res.r * tscale;
            Yi[e] = res.i * tscale;
          }
        }
        ymax = abs_max(abs_max(ymax, Yr), Yi);
        Ar[p] = split(Yr);
        Ai[p] = split(Yi);
      }
      Hl Zr[NP], Zi[NP];
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        Zr[p] = split(unscale(mma(Ai[p], S3n, mma(Ar[p], C3, zero))));
        Zi[p] = split(unscale(mma(Ai[p], C3, mma(Ar[p], S3, zero))));
      }
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        const f4 yT = unscale(mma(Ms, Zi[p], mma(Mc, Zr[p], zero)));
        const int bi = (b0 + p) * PER_UNIT + out_blk;
        if (bi < nbx) {
          // yT: column yy, rows xx..xx+3 of block bi
          const float4 h = *reinterpret_cast<const float4*>(hs + out_yy * OS + out_xx0);
#pragma unroll
          for (int e = 0; e < 4; ++e) bad |= !isfinite(yT[e]);
          *reinterpret_cast<float4*>(out + (bi * B + out_yy) * OS + out_xx0) =
              make_float4(yT[0] * h.x, yT[1] * h.y, yT[2] * h.z, yT[3] * h.w);
        }
      }
    }
    __syncthreads();
    stage_overlap_store<B, OS>(J, out, ring, oy, X0, X1, Y0, Y1, bx_lo, bx_hi, S, tid, nthreads, bad);
  }
  if (bad) atomicOr(flags, NF_FLAG_NONFINITE);
  if (xmax > kInLimit || ymax > kSpecLimit) atomicOr(flags, NF_FLAG_RANGE);
}

// 32-point blocks (tbsize 1): one wave per block, the full complex spectrum
// through the even/odd splits of the 32-point DFT, so every product is a
// 16-point DFT matrix (the constants of sbsize 16) and one radix-2 stage
// per direction runs on the VALU:
//   rows    E, O = x[y][even x], x[y][odd x] . W16;  X[kx], X[kx+16] = E +- w^kx O
//   columns the row-parity tiles are the even/odd rows, so Y = W16 . X[even y]
//           +- w^ky (W16 . X[odd y])
//   gain    on all 32 x 32 bins; bins kx > 16 use the table entries and the
//           conjugate template of their mirror bin (32-ky mod 32, 32-kx)
//   inverse decimation in frequency: U = Y[k] + Y[k+16], V = (Y[k] -
//           Y[k+16]) w^-k, outputs 2m and 2m+1 from W16^-1 . U and . V, for
//           the columns (transposed product) and then the rows (real part)
// Tiles: x and X rows are y = 2i + p (i the tile row); Y rows ky (two
// tiles), columns kx (two tiles). The magnitudes entering every split are
// checked against the FP16 range.
template <bool WIENER>
__device__ __forceinline__ void mfma32_tile(const DftJob& J, const DftTemporal& Q, int tile, const float* tw,
                                            unsigned* flags) {
  constexpr int B = 32, K = 17, OS = mfma_out_stride(32);
  constexpr float kInv = kConstScale / 32.0f;
  float smax = 0.0f;
  extern __shared__ __align__(16) float smem[];
  const int tx = tile % J.tiles_x, ty = tile / J.tiles_x;
  const int S = J.S;
  const int tid = threadIdx.x, nthreads = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nwaves = nthreads >> 6;
  const int r = lane & 15, q = lane >> 4;

  const int vx0 = J.gx.offset, vx1 = J.gx.offset + J.gx.length;
  const int vy0 = J.gy.offset, vy1 = J.gy.offset + J.gy.length;
  const int X0 = max(vx0, (J.tile_x0 + tx) * J.tile_w), X1 = min(vx1, (J.tile_x0 + tx + 1) * J.tile_w);
  const int Y0 = max(vy0, (J.tile_y0 + ty) * J.tile_h), Y1 = min(vy1, (J.tile_y0 + ty + 1) * J.tile_h);
  const int bx_lo = ceil_div_int(X0 - B + 1, S), bx_hi = (X1 - 1) / S;
  const int by_lo = ceil_div_int(Y0 - B + 1, S), by_hi = (Y1 - 1) / S;
  const int nbx = bx_hi - bx_lo + 1;
  const int WW = (nbx - 1) * S + B;
  const int ww = mfma_window_stride(J.ww_max);
  const int cx0 = bx_lo * S;

  float* win = smem;                       // [B][ww]
  float* out = win + B * ww;               // [nbx_max][B][OS]
  float* ring = out + J.nbx_max * B * OS;  // [B][tile_w]
  float* hw = ring + B * J.tile_w;         // [B][OS]
  float* hs = hw + B * OS;                 // [B][OS]

  bool bad = false;
  for (int i = tid; i < B * B; i += nthreads) {
    hw[(i / B) * OS + i % B] = J.h[i];
    hs[(i / B) * OS + i % B] = J.hsyn[i];
  }
  for (int i = tid; i < B * J.tile_w; i += nthreads) ring[i] = 0.0f;

  // W16 fragment (symmetric, so one fragment serves the A and B layouts):
  // entry (r, 4q+e) = cos/sin(2 pi r (4q+e) / 16) = tw at 2 (r (4q+e) mod 16)
  float c16[4], s16[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int idx = 2 * ((r * (4 * q + e)) & 15);
    c16[e] = tw[2 * idx];
    s16[e] = tw[2 * idx + 1];
  }
  const Hl C = split_scaled(c16, kConstScale), Sp = split_scaled(s16, kConstScale);
  const Hl Sn = split_scaled(s16, -kConstScale);
  const Hl Ci = split_scaled(c16, kInv), Si = split_scaled(s16, kInv), Sin = split_scaled(s16, -kInv);
  // radix-2 stage twiddles: lane column r (row transforms), element rows
  // 4q+e (column transforms and the inverse stages); cos/sin(2 pi k / 32)
  const float lc = tw[2 * r], ls = tw[2 * r + 1];
  float ec[4], es[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    ec[e] = tw[2 * (4 * q + e)];
    es[e] = tw[2 * (4 * q + e) + 1];
  }
  const bool use_grid = J.grid && J.mean_scale != 0.0f && J.grid[0].x != 0.0f;
  const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
  __syncthreads();

  for (int by = by_lo; by <= by_hi; ++by) {
    const int oy = by * S;
    stage_sources<B, 1>(J, Q, win, ww, oy, cx0, WW, tid, nthreads, bad);
    __syncthreads();
    for (int bi = wave; bi < nbx; bi += nwaves) {
      // rows: tile row i = r is image row y = 2r + p; columns x = 8q + 2e (+1)
      f4 Xr[2][2], Xi[2][2];  // [p][kx tile]
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int y = 2 * r + p;
        const float* src = win + y * ww + bi * S + 8 * q;
        const float* h = hw + y * OS + 8 * q;
        f4 ve, vo;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          ve[e] = src[2 * e] * h[2 * e];
          vo[e] = src[2 * e + 1] * h[2 * e + 1];
        }
        smax = abs_max(abs_max(smax, ve), vo);
        const Hl xe = split(ve), xo = split(vo);
        const f4 er = unscale(mma(xe, C, zero)), ei = unscale(mma(xe, Sn, zero));
        const f4 orr = unscale(mma(xo, C, zero)), oi = unscale(mma(xo, Sn, zero));
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          // w^kx O with w^kx = (cos, -sin)
          const float vr = orr[e] * lc + oi[e] * ls, vi = oi[e] * lc - orr[e] * ls;
          Xr[p][0][e] = er[e] + vr;
          Xi[p][0][e] = ei[e] + vi;
          Xr[p][1][e] = er[e] - vr;
          Xi[p][1][e] = ei[e] - vi;
        }
      }
      // columns: Y = W16 . X[even rows] +- w^ky W16 . X[odd rows]
      f4 Yr[2][2], Yi[2][2];  // [ky tile][kx tile]
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        smax = abs_max(abs_max(abs_max(abs_max(smax, Xr[0][c]), Xi[0][c]), Xr[1][c]), Xi[1][c]);
        const Hl a_r = split(Xr[0][c]), a_i = split(Xi[0][c]), b_r = split(Xr[1][c]), b_i = split(Xi[1][c]);
        const f4 er = unscale(mma(Sp, a_i, mma(C, a_r, zero))), ei = unscale(mma(Sn, a_r, mma(C, a_i, zero)));
        const f4 orr = unscale(mma(Sp, b_i, mma(C, b_r, zero))), oi = unscale(mma(Sn, b_r, mma(C, b_i, zero)));
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float vr = orr[e] * ec[e] + oi[e] * es[e], vi = oi[e] * ec[e] - orr[e] * es[e];
          Yr[0][c][e] = er[e] + vr;
          Yi[0][c][e] = ei[e] + vi;
          Yr[1][c][e] = er[e] - vr;
          Yi[1][c][e] = ei[e] - vi;
        }
      }
      // gain over all bins; the zmean scale from the DC bin (lane 0, element 0)
      const float dcr = __shfl(Yr[0][0][0], 0, 64);
      const float gs = use_grid ? (J.mean_scale * dcr) / J.grid[0].x : 0.0f;
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int ky = 16 * t + 4 * q + e, kx = 16 * c + r;
            const bool mir = kx > K - 1;
            const int bin = (mir ? ((B - ky) & (B - 1)) : ky) * K + (mir ? B - kx : kx);
            float re = Yr[t][c][e], im = Yi[t][c][e], mr = 0.0f, mi = 0.0f;
            if (J.grid) {
              const float2 gv = J.grid[bin];
              mr = gs * gv.x;
              mi = gs * (mir ? -gv.y : gv.y);
            }
            apply_gain<WIENER>(J.g, bin, re, im, mr, mi, bad);
            Yr[t][c][e] = re;
            Yi[t][c][e] = im;
          }
      // inverse columns (transposed products: rows kx, columns m, y = 2m + p)
      f4 Zr[2][2], Zi[2][2];  // [kx tile][p]
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        f4 ur, ui, vr, vi;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          ur[e] = Yr[0][c][e] + Yr[1][c][e];
          ui[e] = Yi[0][c][e] + Yi[1][c][e];
          const float dr = Yr[0][c][e] - Yr[1][c][e], di = Yi[0][c][e] - Yi[1][c][e];
          // times w^-ky = (cos, +sin)
          vr[e] = dr * ec[e] - di * es[e];
          vi[e] = dr * es[e] + di * ec[e];
        }
        smax = abs_max(abs_max(abs_max(abs_max(smax, ur), ui), vr), vi);
        const Hl a_r = split(ur), a_i = split(ui), b_r = split(vr), b_i = split(vi);
        Zr[c][0] = unscale(mma(a_i, Sin, mma(a_r, Ci, zero)));
        Zi[c][0] = unscale(mma(a_i, Ci, mma(a_r, Si, zero)));
        Zr[c][1] = unscale(mma(b_i, Sin, mma(b_r, Ci, zero)));
        Zi[c][1] = unscale(mma(b_i, Ci, mma(b_r, Si, zero)));
      }
      // inverse rows (real part): outputs x = 2n and 2n + 1, n = 4q + e
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        f4 ur, ui, vr, vi;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          ur[e] = Zr[0][p][e] + Zr[1][p][e];
          ui[e] = Zi[0][p][e] + Zi[1][p][e];
          const float dr = Zr[0][p][e] - Zr[1][p][e], di = Zi[0][p][e] - Zi[1][p][e];
          vr[e] = dr * ec[e] - di * es[e];
          vi[e] = dr * es[e] + di * ec[e];
        }
        smax = abs_max(abs_max(abs_max(abs_max(smax, ur), ui), vr), vi);
        const f4 oe = unscale(mma(Sin, split(ui), mma(Ci, split(ur), zero)));
        const f4 oo = unscale(mma(Sin, split(vi), mma(Ci, split(vr), zero)));
        const int y = 2 * r + p;
        float* dst = out + (bi * B + y) * OS + 8 * q;
        const float* hsr = hs + y * OS + 8 * q;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          bad |= !isfinite(oe[e]) || !isfinite(oo[e]);
          *reinterpret_cast<float2*>(dst + 2 * e) = make_float2(oe[e] * hsr[2 * e], oo[e] * hsr[2 * e + 1]);
        }
      }
    }
    __syncthreads();
    stage_overlap_store<B, OS>(J, out, ring, oy, X0, X1, Y0, Y1, bx_lo, bx_hi, S, tid, nthreads, bad);
  }
  if (bad) atomicOr(flags, NF_FLAG_NONFINITE);
  if (smax > kHalfMax) atomicOr(flags, NF_FLAG_RANGE);
}

// The plain Wiener gain is selected once per tile (gain_is_wiener).
// G (32-point blocks): 1 = every job has the plain Wiener gain, 2 = general
// gains; the two are separate kernels there because one kernel holding both
// exceeds the register budget.
template <int B, int TT, int G>
__device__ __forceinline__ void mfma_tile_any(const DftJob& J, const DftTemporal& Q, int tile, const float* tw,
                                              unsigned* flags) {
  if constexpr (B == 32) {
    static_assert(TT == 1, "32-point blocks: tbsize 1");
    mfma32_tile<G == 1>(J, Q, tile, tw, flags);
  } else {
    // depth 5 keeps one path: two copies exceed its register budget
    if (TT < 5 && gain_is_wiener(J.g)) mfma_tile<B, TT, TT < 5>(J, Q, tile, tw, flags);
    else mfma_tile<B, TT, false>(J, Q, tile, tw, flags);
  }
}

}  // namespace

// Register budget: the compiler's default allocation for depths 3 and 5
// exceeds 256 VGPRs (one wave per SIMD); capped at two waves per SIMD (depth
// 1: four for the per-tile kernel, three for the work-list kernel, whose
// job lookup needs more registers) it fits without spills.
template <int B, int TT, int G = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TT == 1 && B < 32 ? 4 : 2))) void dfttest_mfma_kernel(
    const DftVolLaunch L, const float* tw) {
  int jid = 0;
  while (jid + 1 < L.njobs && (int)blockIdx.x >= L.tile_start[jid + 1]) ++jid;
  mfma_tile_any<B, TT, G>(L.jobs[jid], L.tq[jid], (int)blockIdx.x - L.tile_start[jid], tw, L.flags);
}

template <int B, int TT, int G = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TT == 1 && B < 32 ? 3 : 2))) void dfttest_mfma_list_kernel(
    const DftWorkList W, const float* tw) {
  for_each_work_item(W, [&](int j, int tile) {
    mfma_tile_any<B, TT, G>(list_job(W, j), list_temporal(W, j), tile, tw, W.flags);
  });
}

size_t dfttest_mfma_lds_bytes(int B, int T, int nbx_max, int ww, int tile_w) {
  const size_t OS = (size_t)mfma_out_stride(B);
  return sizeof(float) * ((size_t)T * B * mfma_window_stride(ww) + (size_t)nbx_max * B * OS + (size_t)B * tile_w +
                          (size_t)T * B * OS + (size_t)B * OS);
}

bool dfttest_mfma_supported(int B, int T) {
  return ((B == 8 || B == 16) && (T == 1 || T == 3 || T == 5)) || (B == 32 && T == 1);
}

namespace {
template <int B>
const void* mfma_function(int T, bool list, bool general) {
  if constexpr (B == 32) {
    if (T != 1) return nullptr;
    if (general)
      return list ? (const void*)dfttest_mfma_list_kernel<32, 1, 2> : (const void*)dfttest_mfma_kernel<32, 1, 2>;
    return list ? (const void*)dfttest_mfma_list_kernel<32, 1, 1> : (const void*)dfttest_mfma_kernel<32, 1, 1>;
  } else {
    switch (T) {
      case 1: return list ? (const void*)dfttest_mfma_list_kernel<B, 1> : (const void*)dfttest_mfma_kernel<B, 1>;
      case 3: return list ? (const void*)dfttest_mfma_list_kernel<B, 3> : (const void*)dfttest_mfma_kernel<B, 3>;
      case 5: return list ? (const void*)dfttest_mfma_list_kernel<B, 5> : (const void*)dfttest_mfma_kernel<B, 5>;
      default: return nullptr;
    }
  }
}
const void* mfma_function(int B, int T, bool list, bool general) {
  return B == 16 ? mfma_function<16>(T, list, general)
         : B == 8 ? mfma_function<8>(T, list, general)
         : B == 32 ? mfma_function<32>(T, list, general)
                   : nullptr;
}
}  // namespace

hipError_t launch_dfttest_mfma(int B, int T, const DftVolLaunch& L, int total_tiles, size_t lds, const float* tw,
                               hipStream_t stream, bool general) {
  const void* f = mfma_function(B, T, false, general);
  if (!f) return hipErrorInvalidValue;
  void* args[] = {const_cast<DftVolLaunch*>(&L), const_cast<const float**>(&tw)};
  return hipLaunchKernel(f, dim3(total_tiles), dim3(256), args, lds, stream);
}

const void* dfttest_mfma_list_function(int B, int T, bool general) { return mfma_function(B, T, true, general); }

hipError_t launch_dfttest_mfma_list(int B, int T, const DftWorkList& W, int workgroups, size_t lds, const float* tw,
                                    hipStream_t stream, bool general) {
  const void* f = dfttest_mfma_list_function(B, T, general);
  if (!f) return hipErrorInvalidValue;
  void* args[] = {const_cast<DftWorkList*>(&W), const_cast<const float**>(&tw)};
  return hipLaunchKernel(f, dim3(workgroups), dim3(256), args, lds, stream);
}

hipError_t dfttest_mfma_set_lds_limit() {
  for (int B : {8, 16, 32})
    for (int T : {1, 3, 5})
      for (bool general : {false, true}) {
        if (!mfma_function(B, T, false, general)) continue;
        hipError_t e = hipFuncSetAttribute(mfma_function(B, T, false, general),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute(mfma_function(B, T, true, general), hipFuncAttributeMaxDynamicSharedMemorySize,
                                kListLdsLimit);
        if (e != hipSuccess) return e;
      }
  return hipSuccess;
}

}  // namespace nf
