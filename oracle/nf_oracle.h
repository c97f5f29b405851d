/*
 * nf_oracle.h -- CPU restatement of neo-fft's blockwise spectral-filter core
 * (DFTTest / FFT3D), used ONLY as the checker of the GPU implementation.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library. The product path
 * (neo-fft_amd/) never links, loads or calls it.
 *
 * Arithmetic contract restated here (reference files cited per function in
 * nf_oracle.c):
 *   - own operators follow the reference's scalar operators (opt=1), which are
 *     compiled with -fno-fast-math -ffp-contract=off (CMakeLists.txt:119);
 *     this file is compiled the same way;
 *   - 8/16/32-point transforms follow the AVX2 codelets' radix-2 DIT butterfly
 *     order, float twiddles cosf/sinf(k*PI/8|16) and fused multiply-add in the
 *     complex twiddle multiply only (src/spectral/codelets/fft*_avx2.cpp);
 *   - other lengths use a direct DFT (the reference uses PocketFFT there; its
 *     bits are "parity unpinned", tolerance-pinned against a binary64 DFT).
 * All sample/plane pointers are host memory.
 */
#ifndef NF_ORACLE_H
#define NF_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int length, block, overlap, step, count, cover, offset;
} nfo_axis;

typedef struct {
  int ftype, sbsize, smode, sosize, tbsize, tmode, tosize, swin, twin;
  float sigma, sigma2, pmin, pmax, sbeta, tbeta, f0beta;
  int zmean, dither, dither_seed;
  /* optional frequency curves: (position, value) pairs, n = number of floats */
  const float *slocation, *ssx, *ssy, *sst;
  int n_slocation, n_ssx, n_ssy, n_sst, ssystem;
  /* sampled noise model (nlocation), T*B*(B/2+1) floats, or NULL */
  const float* noise_table;
} nfo_dft_cfg;

typedef struct {
  int bw, bh, ow, oh, bt, wintype;
  float sigma, beta, degrid, kratio;
  float sigma2, sigma3, sigma4;
  float sharpen, scutoff, svr, smin, smax, dehalo, hr, ht;
  /* pattern sampling: pfactor > 0 selects it; pattern = the published
   * per-bin model (bh*(bw/2+1) floats) from nfo_fft3d_pattern */
  float pfactor, pcutoff;
  int px, py, pshow;
  const float* pattern;
} nfo_fft3d_cfg;

const char* nfo_last_error(void);

/* geometry.hpp */
int nfo_fft3d_axis(int length, int block, int overlap, nfo_axis* out);
int nfo_dft_axis(int length, int block, int mode, int overlap, nfo_axis* out);
int nfo_reflect(int cover_coordinate, const nfo_axis* a);

/* windows.cpp */
int nfo_fft3d_window(int block, int overlap, int type, float* analysis, float* synthesis);
double nfo_dft_raw_window(int id, int j, int length, float beta);
int nfo_dft_window_3d(int tbsize, int block, int overlap, int mode, int swin, int twin, float sbeta,
                      float tbeta, int tmode, int tosize, float* h, float* wscale);

/* transforms; complex data interleaved (re, im) */
void nfo_cfft(int n, float* re, float* im, int forward);
void nfo_rfft2(int h, int w, const float* in, float* out);
void nfo_irfft2(int h, int w, const float* in, float* out, float fct);
void nfo_rfft3(int t, int h, int w, const float* in, float* out);
void nfo_irfft3_slice(int t, int h, int w, int slice, const float* in, float* out, float fct);

/* tables */
int nfo_fft3d_profile(int w, int h, const float* sigmas4, float* out);
int nfo_enhancement_tables(int w, int h, float factor, float sharpen, float scutoff, float svr,
                           float smin, float smax, float dehalo, float hr, float ht, float* sharpen_tab,
                           float* halo_tab, float* abc3);
int nfo_dft_profile(const nfo_dft_cfg* c, int T, int S, float sigma, float divisor, float* out);

/* per-bin operators; spectrum interleaved complex. enh_* may be NULL/0 */
int nfo_spectral(float* x, const float* grid, int count, float scale, int type, float a, float b,
                 float low, float high, float exponent, float floor_, const float* primary,
                 float sharpen, float dehalo, float ea, float eb, float ec, const float* sharpen_tab,
                 const float* halo_tab);
int nfo_fft3d_temporal(const float* const* spectra, int T, int c, int bins, float degrid,
                       const float* grid, float noise_uniform, const float* noise_table,
                       float noise_multiplier, float lower, float* out);
int nfo_kalman_step(const float* x, float* last, float* cov, float* proc, const float* pattern,
                    float uniform, float ratio2, int count);
uint32_t nfo_dither_hash(uint32_t seed, uint32_t n, uint32_t p, uint32_t y, uint32_t x);
int nfo_dither(const float* src, int w, int h, long long src_stride_floats, uint8_t* dst,
               long long dst_stride, int mode, int seed, int frame, int plane);

/* temporal_grid.hpp: returns count, fills starts/targets and slots[count][size] */
int nfo_temporal_blocks(int frame, int length, int size, int overlap, long long* starts,
                        int* targets, int* slots);

/* whole-plane processing (Plan::run). Sample storage: bits==8 -> u8,
 * floating -> f32, else u16. Strides in bytes. */
int nfo_dfttest_plane(const nfo_dft_cfg* c, int w, int h, int bits, int floating,
                      const void* const* srcs, const long long* strides, int nsrc,
                      const int* targets, int ntargets, void* dst, long long dst_stride,
                      int frame, int plane);
int nfo_fft3d_plane(const nfo_fft3d_cfg* c, int w, int h, int bits, int floating, int chroma,
                    const void* const* srcs, const long long* strides, int nsrc, void* dst,
                    long long dst_stride);
/* DFTNoise::prepare (dft_noise.cpp:5-41): rects[l*T+z] points at the top-left
 * sample of location l's S x S rectangle in frame (frame_l + z); strides in
 * bytes. alpha < 0 selects the default (5 for ftype 0, else 7). Writes the
 * calibrated power table, T*S*(S/2+1) floats. */
int nfo_dft_noise_model(const nfo_dft_cfg* c, float alpha, int bits, int floating, const void* const* rects,
                        const long long* strides, int nloc, float* out);
/* FFT3D pattern search and model (plan.cpp:142-160, 227-284): selected block
 * coordinates, the unscaled power of the selected block and the published
 * model (power * factor * weight, factor = 1 for Kalman plans else pfactor).
 * power/model may be NULL. */
int nfo_fft3d_pattern(const nfo_fft3d_cfg* c, int w, int h, int bits, int floating, int chroma, const void* src,
                      long long stride, int* sel_bx, int* sel_by, float* power, float* model);

/* Kalman (bt=0): state = last[2*bins] (complex), cov[bins], proc[bins] */
long long nfo_kalman_bins(const nfo_fft3d_cfg* c, int w, int h);
int nfo_kalman_init(const nfo_fft3d_cfg* c, int w, int h, int bits, int floating, float* last,
                    float* cov, float* proc);
int nfo_kalman_advance(const nfo_fft3d_cfg* c, int w, int h, int bits, int floating, int chroma,
                       const void* src, long long stride, float* last, float* cov, float* proc);
int nfo_kalman_render(const nfo_fft3d_cfg* c, int w, int h, int bits, int floating, int chroma,
                      const float* last, void* dst, long long dst_stride);
#ifdef __cplusplus
}
#endif
#endif
