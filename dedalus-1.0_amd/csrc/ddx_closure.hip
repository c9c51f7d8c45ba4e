// Learned sub-grid closure kernels (gfx950): BF16 MFMA 3x3 periodic
// convolutions with fused epilogues, and the data movement around them.
//
// Layouts
//   activations  [npix][Cp]        bf16 bit patterns, npix = B*H*W (b, y, x
//                                  order), channels padded to multiples of 16
//   weights      [Co][9][Ci]       bf16, tap t = 3*(dy+1) + (dx+1)
//   pre-activations / gradients    fp32, same pixel-major layout
// The domain is periodic in y and x, so taps wrap around.
//
// v_mfma_f32_16x16x16_bf16 per wavefront (D = A B, 16 x 16 x 16):
//   A (rows m, cols k): lane l holds A[l % 16][4*(l/16) + i], i < 4
//   B (rows k, cols n): lane l holds B[4*(l/16) + i][l % 16]
//   D:                  lane l holds D[4*(l/16) + i][l % 16]
// ddx_mfma_probe runs one tile with exactly these placements so the layout
// can be checked against a host product.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <type_traits>
#include <string>

#include "../../include/ddx.h"
#include "../../include/ddx_closure.h"

namespace ddx {
int set_error(int code, const std::string& msg);  // defined in ddx.hip
}

namespace {

// Workgroup barrier that orders LDS only: the release/acquire fences name the
// local address space, so global loads (the next tile's halo, issued before
// the current tile's compute) and stores stay in flight across it; a plain
// __syncthreads() also waits for every outstanding global access.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define CL_LAUNCHCHK(name)                                                                  \
  do {                                                                                      \
    hipError_t e_ = hipGetLastError();                                                      \
    if (e_ != hipSuccess)                                                                   \
      return ddx::set_error(DDX_ERR_HIP, std::string(name) + ": " + hipGetErrorString(e_)); \
  } while (0)

// fp32 -> bf16, round to nearest even (gfx950 v_cvt_pk_bf16_f32)
__device__ inline uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU in its tanh form, written as z sigma(s) with
// s = 2 sqrt(2/pi) (z + 0.044715 z^3): 0.5 z (1 + tanh(s/2)) = z sigma(s), so
// the negative tail has no cancellation.  sigma(s) = 1 / (1 + 2^u) with
// u = -s log2(e) = z (A + B z^2): the constants fold -2 sqrt(2/pi) log2(e),
// so one multiply, one fma and one multiply form the exponent of the
// hardware exp2, and one reciprocal finishes sigma (saturating to 0 and 1).
//   GELU'(z) = sigma(s) + z sigma(s) (1 - sigma(s)) s'(z),
//   s'(z) = 2 sqrt(2/pi) (1 + 3 * 0.044715 z^2)
constexpr float GELU_A = -2.3022082f;            // -2 sqrt(2/pi) log2(e)
constexpr float GELU_B = -2.3022082f * 0.044715f;
__device__ inline float gelu_sigma(float z) {
  const float u = z * __builtin_fmaf(z * z, GELU_B, GELU_A);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(u));
}
__device__ inline float gelu(float z) { return z * gelu_sigma(z); }
__device__ inline float gelu_grad(float z) {
  const float sg = gelu_sigma(z);
  const float ds = 1.5957691216057308f * (1.0f + 0.134145f * z * z);
  return sg + z * sg * (1.0f - sg) * ds;
}

__device__ inline bf16x4 load4(const uint16_t* p) {
  // four consecutive channels; channel offsets are multiples of 4, so 8-byte aligned
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  bf16x4 r;
  r[0] = (short)(v.x & 0xffffu);
  r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xffffu);
  r[3] = (short)(v.y >> 16);
  return r;
}

__device__ inline int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

struct ConvArgs {
  const uint16_t* x;  // [npix][Ci]
  const uint16_t* w;  // [Co][9][Ci]
  const float* bias;  // [Co] or null
  const float* aux;   // [npix][Co] pre-activation for DDX_CONV_GELU_GRAD
  float* out_f32;     // [npix][Co] or null
  uint16_t* out_bf16; // [npix][Co] or null
  int B, H, W, Ci, Co, mode;
};

// One wavefront: 16 consecutive pixels x all Co = 16*NCO output channels.
// Four wavefronts per block; no block-level synchronisation.
template <int NCO>
__global__ void __launch_bounds__(256) conv3x3_kernel(ConvArgs a) {
  const int lane = threadIdx.x & 63;
  const long tile = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const long npix = (long)a.B * a.H * a.W;
  if (tile * 16 >= npix) return;
  const long p0 = tile * 16;
  const int m = lane & 15, kq = lane >> 4;
  const long pm = p0 + m;
  const long hw = (long)a.H * a.W;
  const int b = (int)(pm / hw);
  const int rem = (int)(pm - (long)b * hw);
  const int py = rem / a.W, px = rem - py * a.W;
  f32x4 acc[NCO];
#pragma unroll
  for (int j = 0; j < NCO; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int t = 0; t < 9; ++t) {
    const int yy = wrapi(py + t / 3 - 1, a.H), xx = wrapi(px + t % 3 - 1, a.W);
    const uint16_t* xrow = a.x + (((long)b * a.H + yy) * a.W + xx) * a.Ci + 4 * kq;
    const uint16_t* wrow = a.w + ((long)m * 9 + t) * a.Ci + 4 * kq;
    const long wstride = 16L * 9 * a.Ci;  // next 16 output channels
    for (int kc = 0; kc < a.Ci; kc += 16) {
      const bf16x4 av = load4(xrow + kc);
#pragma unroll
      for (int j = 0; j < NCO; ++j) {
        const bf16x4 bv = load4(wrow + j * wstride + kc);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[j], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NCO; ++j) {
    const int c = j * 16 + m;
    const float bv = a.bias ? a.bias[c] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const long o = (p0 + 4 * kq + i) * a.Co + c;
      float v = acc[j][i];
      float act;
      if (a.mode == DDX_CONV_GELU_GRAD) {
        v *= gelu_grad(a.aux[o]);
        act = v;
      } else {
        v += bv;
        act = a.mode == DDX_CONV_GELU ? gelu(v) : v;
      }
      if (a.out_f32) a.out_f32[o] = v;
      if (a.out_bf16) a.out_bf16[o] = f2bf(act);
    }
  }
}

__device__ inline bf16x4 lds4(const uint16_t* p) {
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  bf16x4 r;
  r[0] = (short)(v.x & 0xffffu);
  r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xffffu);
  r[3] = (short)(v.y >> 16);
  return r;
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// eight consecutive bf16 from LDS as two 8-byte reads
__device__ inline bf16x8 lds8(const uint16_t* p) {
  const uint2 lo = *reinterpret_cast<const uint2*>(p);
  const uint2 hi = *reinterpret_cast<const uint2*>(p + 4);
  const u32x4 v = {lo.x, lo.y, hi.x, hi.y};
  return __builtin_bit_cast(bf16x8, v);
}

// LDS-tiled 3x3 convolution for CI, CO in {16, 32}, W % 64 == 0, H % 4 == 0.
// Output tile = 4 rows x 64 columns of one image; wavefront w owns row
// y0 + w as four 16-pixel fragments.  Blocks loop over tiles (tile c,
// c + gridDim.x, ...): the weights are staged in LDS once per block, and the
// next tile's 6 x 66 input halo is loaded into registers while the current
// tile computes.  Pixel stride and weight-row stride CI + 4 (bf16) place the
// 16 rows of a fragment read in distinct bank pairs.  CI = 32 uses the
// gfx950 16x16x32 bf16 MFMA, CI = 16 the 16x16x16 one.
//
// BF (bf16 output only, no fp32 output or GELU gradient: the inference
// pass): bias and activation are applied to the accumulators and the
// results staged as bf16, which halves the staging LDS; the kernel then fits
// three workgroups per CU instead of two.
template <int CI, int CO, bool BF>
__global__ void __launch_bounds__(256) conv3x3_tile_kernel(ConvArgs a, long ntiles) {
  constexpr int PS = CI + 4, WS = CI + 4;
  constexpr int HR = 6, HC = 66;
  constexpr int NCO = CO / 16;
  constexpr int CK = CI / 8;                       // 16-byte chunks per pixel
  constexpr int NH = HR * HC * CK, HPT = (NH + 255) / 256;
  constexpr int NW = CO * 9 * CK, WPT = (NW + 255) / 256;
  // the halo buffer also stages each wavefront's 64 x CO results (fp32, or
  // bf16 with BF)
  constexpr int ES = CO + 4;                                  // fp32 staging row stride (floats)
  constexpr int EB = CO + 8;                                  // bf16 staging row stride (bf16)
  constexpr int HALO_B = HR * HC * PS * 2, STAGE_B = BF ? 4 * 64 * EB * 2 : 4 * 64 * ES * 4;
  constexpr int BUF_B = HALO_B > STAGE_B ? HALO_B : STAGE_B;
  __shared__ __align__(16) uint16_t halo[BUF_B / 2];
  __shared__ __align__(16) uint16_t wl[CO * 9 * WS];
  const int tid = threadIdx.x;
  const int tiles_x = a.W / 64, tiles_y = a.H / 4;
  {
    uint4 wv[WPT];
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
      const int i = tid + k * 256;
      if (i < NW) wv[k] = *reinterpret_cast<const uint4*>(&a.w[(long)i * 8]);
    }
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
      const int i = tid + k * 256;
      if (i < NW) {
        const int row = i / CK, ch = i - row * CK;
        uint2* d = reinterpret_cast<uint2*>(&wl[row * WS + ch * 8]);
        d[0] = make_uint2(wv[k].x, wv[k].y);
        d[1] = make_uint2(wv[k].z, wv[k].w);
      }
    }
  }
  uint4 hv[HPT];
  auto fetch = [&](long tile) {
    long rest = tile;
    const int tx = (int)(rest % tiles_x);
    rest /= tiles_x;
    const int ty = (int)(rest % tiles_y);
    const int b = (int)(rest / tiles_y);
    const int x0 = tx * 64, y0 = ty * 4;
#pragma unroll
    for (int k = 0; k < HPT; ++k) {
      const int i = tid + k * 256;
      if (i < NH) {
        const int pix = i / CK, ch = i - pix * CK;
        const int r = pix / HC, c = pix - r * HC;
        const int yy = wrapi(y0 - 1 + r, a.H), xx = wrapi(x0 - 1 + c, a.W);
        hv[k] = *reinterpret_cast<const uint4*>(&a.x[(((long)b * a.H + yy) * a.W + xx) * CI + ch * 8]);
      }
    }
  };
  const int wave = tid >> 6, lane = tid & 63;
  const int m = lane & 15, kq = lane >> 4;
  const bool grad = a.mode == DDX_CONV_GELU_GRAD, gelu_on = a.mode == DDX_CONV_GELU;
  float bias4[4] = {0.f, 0.f, 0.f, 0.f};   // channels (lane % (CO/4)) * 4 + e (fp32 staging)
  float biasj[NCO];                         // channels j * 16 + m (BF)
#pragma unroll
  for (int j = 0; j < NCO; ++j) biasj[j] = 0.f;
  if (!grad && a.bias) {
    const int cb = (lane % (CO / 4)) * 4;
#pragma unroll
    for (int e = 0; e < 4; ++e) bias4[e] = a.bias[cb + e];
#pragma unroll
    for (int j = 0; j < NCO; ++j) biasj[j] = a.bias[j * 16 + m];
  }
  long tile = blockIdx.x;
  if (tile < ntiles) fetch(tile);
  for (; tile < ntiles; tile += gridDim.x) {
    lds_barrier();  // the previous tile's fragment reads are done
#pragma unroll
    for (int k = 0; k < HPT; ++k) {
      const int i = tid + k * 256;
      if (i < NH) {
        const int pix = i / CK, ch = i - pix * CK;
        uint2* d = reinterpret_cast<uint2*>(&halo[pix * PS + ch * 8]);
        d[0] = make_uint2(hv[k].x, hv[k].y);
        d[1] = make_uint2(hv[k].z, hv[k].w);
      }
    }
    lds_barrier();
    const long next = tile + gridDim.x;
    if (next < ntiles) fetch(next);  // in flight while this tile computes
    f32x4 acc[4][NCO];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int j = 0; j < NCO; ++j) acc[q][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int dy = t / 3, dx = t % 3;
      if constexpr (CI % 32 == 0) {
#pragma unroll
        for (int kc = 0; kc < CI / 32; ++kc) {
          bf16x8 bv[NCO];
#pragma unroll
          for (int j = 0; j < NCO; ++j) bv[j] = lds8(&wl[((j * 16 + m) * 9 + t) * WS + kc * 32 + 8 * kq]);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const bf16x8 av = lds8(&halo[((wave + dy) * HC + q * 16 + m + dx) * PS + kc * 32 + 8 * kq]);
#pragma unroll
            for (int j = 0; j < NCO; ++j)
              acc[q][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[j], acc[q][j], 0, 0, 0);
          }
        }
      } else {
#pragma unroll
        for (int kc = 0; kc < CI / 16; ++kc) {
          bf16x4 bv[NCO];
#pragma unroll
          for (int j = 0; j < NCO; ++j) bv[j] = lds4(&wl[((j * 16 + m) * 9 + t) * WS + kc * 16 + 4 * kq]);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const bf16x4 av = lds4(&halo[((wave + dy) * HC + q * 16 + m + dx) * PS + kc * 16 + 4 * kq]);
#pragma unroll
            for (int j = 0; j < NCO; ++j)
              acc[q][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv[j], acc[q][j], 0, 0, 0);
          }
        }
      }
    }
    long rest = tile;
    const int tx = (int)(rest % tiles_x);
    rest /= tiles_x;
    const int ty = (int)(rest % tiles_y);
    const int b = (int)(rest / tiles_y);
    const long prow = ((long)b * a.H + ty * 4 + wave) * a.W + tx * 64;
    lds_barrier();  // every wavefront is done with the halo
    if constexpr (BF) {
      // activations of the D fragments (pixel q*16 + 4kq + i, channel j*16 + m)
      // as [64][EB] bf16, then 4 channels (8 bytes) per lane and chunk to HBM
      uint16_t* sb = halo + wave * 64 * EB;
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int j = 0; j < NCO; ++j)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float v = acc[q][j][i] + biasj[j];
            sb[(q * 16 + 4 * kq + i) * EB + j * 16 + m] = f2bf(gelu_on ? gelu(v) : v);
          }
      lds_barrier();
      constexpr int C4 = CO / 4, NR = (64 * C4) / 64;
      const int c = (lane % C4) * 4;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const int px = (lane + 64 * r) / C4;
        *reinterpret_cast<uint2*>(&a.out_bf16[(prow + px) * CO + c]) =
            *reinterpret_cast<const uint2*>(&sb[px * EB + c]);
      }
      continue;
    }
    // stage D fragments (pixel q*16 + 4kq + i, channel j*16 + m) as [64][ES] fp32
    float* st = reinterpret_cast<float*>(halo) + wave * 64 * ES;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int j = 0; j < NCO; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) st[(q * 16 + 4 * kq + i) * ES + j * 16 + m] = acc[q][j][i];
    lds_barrier();
    // the wavefront's 64 pixels x CO channels are contiguous in HBM: 4 channels
    // per chunk; a lane's channels c..c+3 are the same in every chunk (C4
    // divides 64), so its bias was read once before the tile loop.  The aux
    // loads of all chunks are issued before any is used.
    constexpr int C4 = CO / 4, NR = (64 * C4) / 64;
    const int c = (lane % C4) * 4;
    float4 zaux[NR];
    if (grad) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const int px = (lane + 64 * r) / C4;
        zaux[r] = *reinterpret_cast<const float4*>(&a.aux[(prow + px) * CO + c]);
      }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int px = (lane + 64 * r) / C4;
      const float4 d = *reinterpret_cast<const float4*>(&st[px * ES + c]);
      float v[4] = {d.x, d.y, d.z, d.w}, act[4];
      const long o = (prow + px) * CO + c;
      if (grad) {
        const float zz[4] = {zaux[r].x, zaux[r].y, zaux[r].z, zaux[r].w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          v[e] *= gelu_grad(zz[e]);
          act[e] = v[e];
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          v[e] += bias4[e];
          act[e] = gelu_on ? gelu(v[e]) : v[e];
        }
      }
      if (a.out_f32) *reinterpret_cast<float4*>(&a.out_f32[o]) = make_float4(v[0], v[1], v[2], v[3]);
      if (a.out_bf16)
        *reinterpret_cast<uint2*>(&a.out_bf16[o]) =
            make_uint2((uint32_t)f2bf(act[0]) | ((uint32_t)f2bf(act[1]) << 16),
                       (uint32_t)f2bf(act[2]) | ((uint32_t)f2bf(act[3]) << 16));
    }
  }
}

// Weight-gradient partials over one chunk of pixels, all nine taps per wavefront:
// partial[chunk][co][t][ci] = sum_{p in chunk} dz[p][co] x[shift_t(p)][ci].
// grid.x = chunks, grid.y = (Co/16) * (Ci/16); one wavefront per block.
__global__ void __launch_bounds__(64) conv3x3_wgrad_kernel(const uint16_t* __restrict__ x,
                                                            const uint16_t* __restrict__ dz,
                                                            float* __restrict__ partial, int B,
                                                            int H, int W, int Ci, int Co,
                                                            int pch) {
  const int lane = threadIdx.x;
  const int m = lane & 15, kq = lane >> 4;
  const int nct = Ci / 16;
  const int co0 = (blockIdx.y / nct) * 16, ci0 = (blockIdx.y % nct) * 16;
  const long hw = (long)H * W;
  const long npix = (long)B * hw;
  const long pstart = (long)blockIdx.x * pch;
  const long pend = pstart + pch < npix ? pstart + pch : npix;
  f32x4 acc[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long pk = pstart; pk < pend; pk += 16) {
    bf16x4 av;
    int pb[4], py[4], px[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const long p = pk + 4 * kq + i;
      av[i] = (short)dz[p * Co + co0 + m];  // A[m = co][k = p]
      pb[i] = (int)(p / hw);
      const int rem = (int)(p - (long)pb[i] * hw);
      py[i] = rem / W;
      px[i] = rem - py[i] * W;
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      bf16x4 bv;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int yy = wrapi(py[i] + t / 3 - 1, H), xx = wrapi(px[i] + t % 3 - 1, W);
        bv[i] = (short)x[(((long)pb[i] * H + yy) * W + xx) * Ci + ci0 + m];  // B[k = p][n = ci]
      }
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int co = co0 + 4 * kq + i, ci = ci0 + m;
      partial[(((long)blockIdx.x * Co + co) * 9 + t) * Ci + ci] = acc[t][i];
    }
}

// LDS-tiled weight gradient for CI, CO in {16, 32}, W % 64 == 0, H % 4 == 0.
// Block c accumulates over output tiles c, c + gridDim.x, ... (tile = 4 rows
// x 64 columns of one image, as in the forward kernel) and writes
// partial[c][co][t][ci].  Per tile, dz (256 pixels x CO) and the input halo
// (6 x 66 pixels x CI) are staged channel-major, so the MFMA's K dimension
// (pixels) is contiguous: A[m = co][k = p] from dzT[co][p] (8-byte reads),
// B[k = p][n = ci] from haloT[ci][row][col].  Wavefront w owns the
// (co tile, ci tile) pairs w, w + 4, ... with all nine tap accumulators.
//
// With dzf (the fp32 dz), the block also sums dz over its pixels per output
// channel in double (the bias gradient), thread t taking channel t mod CO and
// every (256 / CO)-th pixel, combined in a fixed order into
// bpartial[block][co]: the column sums ride along with the tiles this
// kernel reads anyway, instead of a separate pass over dz.
template <int CI, int CO>
__global__ void __launch_bounds__(256) conv3x3_wgrad_tile_kernel(const uint16_t* __restrict__ x,
                                                                 const uint16_t* __restrict__ dz,
                                                                 float* __restrict__ partial,
                                                                 int B, int H, int W,
                                                                 const float* __restrict__ dzf = nullptr,
                                                                 double* __restrict__ bpartial = nullptr) {
  constexpr int DS = 260;            // dzT row stride (bf16): 256 pixels + 4
  constexpr int HRS = 68;            // halo row stride (66 columns + 2)
  constexpr int HS = 6 * HRS + 2;    // halo channel stride, odd in 4-byte words
  constexpr int NCT = CI / 16, NPAIR = (CO / 16) * (CI / 16);
  constexpr int NPW = (NPAIR + 3) / 4;  // pairs per wavefront
  __shared__ __align__(16) uint16_t dzT[CO * DS];
  __shared__ __align__(16) uint16_t haloT[CI * HS];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int m = lane & 15, kq = lane >> 4;
  const int tiles_x = W / 64, tiles_y = H / 4;
  const long ntiles = (long)B * tiles_y * tiles_x;
  f32x4 acc[NPW][9];
#pragma unroll
  for (int u = 0; u < NPW; ++u)
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[u][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  constexpr int PG = 256 / CO;              // pixel groups of the bias sums
  const int bco = tid % CO, bpg = tid / CO;
  double bsum = 0.0;
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    long rest = tile;
    const int tx = (int)(rest % tiles_x);
    rest /= tiles_x;
    const int ty = (int)(rest % tiles_y);
    const int b = (int)(rest / tiles_y);
    const int x0 = tx * 64, y0 = ty * 4;
    if (dzf) {
      // loads first, then the double adds in pixel order
      float bv[256 / PG];
#pragma unroll
      for (int k = 0; k < 256 / PG; ++k) {
        const int pix = bpg + PG * k;
        bv[k] = dzf[(((long)b * H + y0 + (pix >> 6)) * W + x0 + (pix & 63)) * CO + bco];
      }
#pragma unroll
      for (int k = 0; k < 256 / PG; ++k) bsum += (double)bv[k];
    }
    lds_barrier();  // the previous tile's reads are done
    // dz: 256 pixels x CO, 4 channels per load, stored transposed
    for (int i = tid; i < 256 * (CO / 4); i += 256) {
      const int pix = i / (CO / 4), ch = (i - pix * (CO / 4)) * 4;
      const int r = pix >> 6, c = pix & 63;
      const uint2 v = *reinterpret_cast<const uint2*>(
          &dz[(((long)b * H + y0 + r) * W + x0 + c) * CO + ch]);
      dzT[(ch + 0) * DS + pix] = (uint16_t)(v.x & 0xffffu);
      dzT[(ch + 1) * DS + pix] = (uint16_t)(v.x >> 16);
      dzT[(ch + 2) * DS + pix] = (uint16_t)(v.y & 0xffffu);
      dzT[(ch + 3) * DS + pix] = (uint16_t)(v.y >> 16);
    }
    // halo: 6 x 66 pixels x CI, stored transposed
    for (int i = tid; i < 6 * 66 * (CI / 4); i += 256) {
      const int pix = i / (CI / 4), ch = (i - pix * (CI / 4)) * 4;
      const int r = pix / 66, c = pix - r * 66;
      const int yy = wrapi(y0 - 1 + r, H), xx = wrapi(x0 - 1 + c, W);
      const uint2 v = *reinterpret_cast<const uint2*>(
          &x[(((long)b * H + yy) * W + xx) * CI + ch]);
      const int o = r * HRS + c;
      haloT[(ch + 0) * HS + o] = (uint16_t)(v.x & 0xffffu);
      haloT[(ch + 1) * HS + o] = (uint16_t)(v.x >> 16);
      haloT[(ch + 2) * HS + o] = (uint16_t)(v.y & 0xffffu);
      haloT[(ch + 3) * HS + o] = (uint16_t)(v.y >> 16);
    }
    lds_barrier();
#pragma unroll
    for (int u = 0; u < NPW; ++u) {
      const int pair = wave + 4 * u;
      if (pair >= NPAIR) break;
      const int co0 = (pair / NCT) * 16, ci0 = (pair % NCT) * 16;
      for (int ks = 0; ks < 16; ++ks) {       // 16 pixels per K step
        const int row = ks >> 2, col0 = (ks & 3) * 16;
        const bf16x4 av = lds4(&dzT[(co0 + m) * DS + ks * 16 + 4 * kq]);
        const uint16_t* hb = &haloT[(ci0 + m) * HS];
#pragma unroll
        for (int t = 0; t < 9; ++t) {
          const uint16_t* hp = hb + (row + t / 3) * HRS + col0 + 4 * kq + t % 3;
          bf16x4 bv;
          bv[0] = (short)hp[0];
          bv[1] = (short)hp[1];
          bv[2] = (short)hp[2];
          bv[3] = (short)hp[3];
          acc[u][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[u][t], 0, 0, 0);
        }
      }
    }
  }
  // D[row = co (4 kq + i)][col = ci (m)] of each tap
#pragma unroll
  for (int u = 0; u < NPW; ++u) {
    const int pair = wave + 4 * u;
    if (pair >= NPAIR) break;
    const int co0 = (pair / NCT) * 16, ci0 = (pair % NCT) * 16;
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        partial[(((long)blockIdx.x * CO + co0 + 4 * kq + i) * 9 + t) * CI + ci0 + m] = acc[u][t][i];
  }
  if (dzf) {
    // the PG partial sums of each channel, combined in group order
    lds_barrier();
    double* sh = reinterpret_cast<double*>(haloT);   // 256 doubles fit the halo buffer
    sh[tid] = bsum;
    lds_barrier();
    if (tid < CO) {
      double r = 0.0;
      for (int g = 0; g < PG; ++g) r += sh[g * CO + tid];
      bpartial[(long)blockIdx.x * CO + tid] = r;
    }
  }
}

// out[j] = sum over chunks of partial[c][j] in a fixed order (double):
// block = 64 consecutive j x SUM_GROUPS chunk groups (group q sums chunks q,
// q + SUM_GROUPS, ... with four loads in flight), the group sums are then
// added in group order.  Sixteen groups keep ~150 k threads loading (a
// 32 x 32 layer's 9216 weights x 512 chunks is 19 MB).
constexpr int SUM_GROUPS = 16;

__global__ void __launch_bounds__(64 * SUM_GROUPS) sum_chunks_kernel(const float* __restrict__ partial,
                                                                     float* __restrict__ out,
                                                                     int nchunk, long n) {
  __shared__ double sh[SUM_GROUPS][64];
  const int e = threadIdx.x & 63, q = threadIdx.x >> 6;
  const long j = (long)blockIdx.x * 64 + e;
  const long jc = j < n ? j : n - 1;   // in-range address for the spare lanes
  double s = 0.0;
  int c = q;
  for (; c + 3 * SUM_GROUPS < nchunk; c += 4 * SUM_GROUPS) {
    const float v0 = partial[(long)c * n + jc];
    const float v1 = partial[(long)(c + SUM_GROUPS) * n + jc];
    const float v2 = partial[(long)(c + 2 * SUM_GROUPS) * n + jc];
    const float v3 = partial[(long)(c + 3 * SUM_GROUPS) * n + jc];
    s += v0;
    s += v1;
    s += v2;
    s += v3;
  }
  for (; c < nchunk; c += SUM_GROUPS) s += partial[(long)c * n + jc];
  sh[q][e] = s;
  __syncthreads();
  if (q == 0 && j < n) {
    double t = sh[0][e];
#pragma unroll
    for (int g = 1; g < SUM_GROUPS; ++g) t += sh[g][e];
    out[j] = (float)t;
  }
}

// Column sums of a [npix][C] fp32 array (bias gradient), two stages with a
// fixed summation order.  Stage 1: block b sums pixels [b*chunk, (b+1)*chunk)
// of every column; the 256 threads read rows of C floats contiguously
// (thread t: column t % C, pixel offset t / C, G = 256 / C pixel groups).
__global__ void __launch_bounds__(256) colsum_partial_kernel(const float* __restrict__ a,
                                                             double* __restrict__ partial,
                                                             long npix, int C, long chunk) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  const int G = 256 / C;
  const int c = t % C, g = t / C;
  const long p0 = (long)blockIdx.x * chunk;
  const long p1 = p0 + chunk < npix ? p0 + chunk : npix;
  double s = 0.0;
  if (g < G)
    for (long p = p0 + g; p < p1; p += G) s += a[p * C + c];
  sh[t] = s;
  __syncthreads();
  if (t < C) {
    double r = 0.0;
    for (int k = 0; k < G; ++k) r += sh[k * C + t];
    partial[(long)blockIdx.x * C + t] = r;
  }
}

// Stage 2: out[c] = sum over blocks of partial[b][c]; one block per column,
// thread t sums blocks t, t+256, ..., then a fixed-order tree
__global__ void __launch_bounds__(256) colsum_final_kernel(const double* __restrict__ partial,
                                                           float* __restrict__ out, int nblk,
                                                           int C) {
  __shared__ double sh[256];
  const int c = blockIdx.x, t = threadIdx.x;
  double s = 0.0;
  for (int b = t; b < nblk; b += 256) s += partial[(long)b * C + c];
  sh[t] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (t < k) sh[t] += sh[t + k];
    __syncthreads();
  }
  if (t == 0) out[c] = (float)sh[0];
}

// fp32 weights in (co, ci, ky, kx) order -> padded bf16 forward and transposed
// (input-gradient) weights: wf[co][t][ci] = w[co][ci][t], wt[ci][t][co] = w[co][ci][8 - t]
__global__ void wprep_kernel(const float* __restrict__ w, int Co, int Ci, int Cop, int Cip,
                             uint16_t* __restrict__ wf, uint16_t* __restrict__ wt) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)Cop * 9 * Cip;
  if (j >= n) return;
  const int ci = (int)(j % Cip);
  const int t = (int)((j / Cip) % 9);
  const int co = (int)(j / (9L * Cip));
  const float v = (co < Co && ci < Ci) ? w[((long)co * Ci + ci) * 9 + t] : 0.f;
  const uint16_t h = f2bf(v);
  if (wf) wf[j] = h;
  if (wt) wt[((long)ci * 9 + (8 - t)) * Cop + co] = h;
}

// padded weight gradient [Cop][9][Cip] -> (co, ci, ky, kx) order
__global__ void wunpack_kernel(const float* __restrict__ dwp, int Co, int Ci, int Cop, int Cip,
                               float* __restrict__ dw) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)Co * Ci * 9;
  if (j >= n) return;
  const int t = (int)(j % 9);
  const int ci = (int)((j / 9) % Ci);
  const int co = (int)(j / (9L * Ci));
  dw[j] = dwp[((long)co * 9 + t) * Cip + ci];
}

// Gather channels: dst_bf16[p][c] = scale_c * src_c[p * pstride] (0 for c >= C)
struct PackArgs {
  const void* src[DDX_CONV_MAXCH];
  float scale[DDX_CONV_MAXCH];
  long pstride;
  long npix;
  int C, Cp, f64;
  uint16_t* dst_bf16;
  float* dst_f32;
};

__global__ void pack_kernel(PackArgs a) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.npix * a.Cp) return;
  const long p = j / a.Cp;
  const int c = (int)(j - p * a.Cp);
  float v = 0.f;
  if (c < a.C) {
    const long o = p * a.pstride;
    v = a.f64 ? (float)((const double*)a.src[c])[o] : ((const float*)a.src[c])[o];
    v *= a.scale[c];
  }
  if (a.dst_bf16) a.dst_bf16[j] = f2bf(v);
  if (a.dst_f32) a.dst_f32[j] = v;
}

// One pixel per thread (Cp % 8 == 0, 16-byte aligned outputs): the C <= CMAX
// source values are read once, unconditionally (channels past C read channel
// 0 and are zeroed), and the Cp padded channels written as 16-byte chunks.
template <bool F64, int CMAX>
__global__ void pack_pixel_kernel(PackArgs a) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.npix) return;
  float v[CMAX], x[CMAX];
  const long o = p * a.pstride;
#pragma unroll
  for (int c = 0; c < CMAX; ++c) {
    const int cc = c < a.C ? c : 0;
    x[c] = F64 ? (float)((const double*)a.src[cc])[o] : ((const float*)a.src[cc])[o];
  }
  // empty asm statements taking the loaded values as operands, after all the
  // loads: no load can be moved under the (uniform) channel test below, which
  // would serialise them
#pragma unroll
  for (int c = 0; c < CMAX; ++c) asm volatile("" : "+v"(x[c]));
#pragma unroll
  for (int c = 0; c < CMAX; ++c) v[c] = c < a.C ? x[c] * a.scale[c < a.C ? c : 0] : 0.f;
  for (int k = 0; k < a.Cp / 8; ++k) {
    float w[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      w[e] = 0.f;
#pragma unroll
      for (int d = 0; d < CMAX; ++d)
        if (d == 8 * k + e) w[e] = v[d];
    }
    if (a.dst_bf16) {
      uint4 q;
      q.x = (uint32_t)f2bf(w[0]) | ((uint32_t)f2bf(w[1]) << 16);
      q.y = (uint32_t)f2bf(w[2]) | ((uint32_t)f2bf(w[3]) << 16);
      q.z = (uint32_t)f2bf(w[4]) | ((uint32_t)f2bf(w[5]) << 16);
      q.w = (uint32_t)f2bf(w[6]) | ((uint32_t)f2bf(w[7]) << 16);
      *reinterpret_cast<uint4*>(&a.dst_bf16[p * a.Cp + 8 * k]) = q;
    }
    if (a.dst_f32) {
      float4* d = reinterpret_cast<float4*>(&a.dst_f32[p * a.Cp + 8 * k]);
      d[0] = make_float4(w[0], w[1], w[2], w[3]);
      d[1] = make_float4(w[4], w[5], w[6], w[7]);
    }
  }
}

// Scatter channels: dst_c[p * pstride] = scale_c * src[p][c] (c < C), fp32 or fp64
struct UnpackArgs {
  void* dst[DDX_CONV_MAXCH];
  const void* add[DDX_CONV_MAXCH];   // all null, or one per channel
  float scale[DDX_CONV_MAXCH];
  long pstride;
  long npix;
  int C, Cp, f64;
  const float* src;
};

__global__ void unpack_kernel(UnpackArgs a) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.npix) return;
  const long o = p * a.pstride;
  const float* __restrict__ sp = a.src + p * a.Cp;
  float v[DDX_CONV_MAXCH];
#pragma unroll
  for (int c = 0; c < DDX_CONV_MAXCH; ++c) v[c] = sp[c < a.C ? c : 0];   // loads first
#pragma unroll
  for (int c = 0; c < DDX_CONV_MAXCH; ++c) {
    if (c >= a.C) break;
    const float x = v[c] * a.scale[c];
    if (a.f64)
      ((double*)a.dst[c])[o] = (double)x + (a.add[0] ? ((const double*)a.add[c])[o] : 0.0);
    else
      ((float*)a.dst[c])[o] = x + (a.add[0] ? ((const float*)a.add[c])[o] : 0.f);
  }
}

// Box filter: dst[Y][X] = mean of the f x f block of src starting at (f Y, f X)
template <typename T>
__global__ void box_filter_kernel(const T* __restrict__ src, T* __restrict__ dst, int ny, int nx,
                                  int f) {
  const int mx = nx / f, my = ny / f;
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= (long)mx * my) return;
  const int Y = (int)(j / mx), X = (int)(j - (long)Y * mx);
  double s = 0.0;
  for (int a = 0; a < f; ++a) {
    const T* row = src + ((long)(Y * f + a)) * nx + (long)X * f;
    for (int b = 0; b < f; ++b) s += (double)row[b];
  }
  dst[j] = (T)(s / ((double)f * f));
}

// Layout probe, K = 32 (gfx950): lane l holds A[l % 16][8 (l / 16) + i] and
// B[8 (l / 16) + i][l % 16], i < 8; D as for K = 16.

__device__ inline bf16x8 to_bf16x8(const uint16_t (&h)[8]) {
  bf16x8 r;
#pragma unroll
  for (int i = 0; i < 8; ++i) r[i] = __builtin_bit_cast(__bf16, h[i]);
  return r;
}

__global__ void __launch_bounds__(64) mfma_probe32_kernel(const uint16_t* A, const uint16_t* Bm, float* D) {
  const int lane = threadIdx.x;
  const int m = lane & 15, kq = lane >> 4;
  uint16_t ha[8], hb[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    ha[i] = A[m * 32 + 8 * kq + i];
    hb[i] = Bm[(8 * kq + i) * 16 + m];
  }
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(to_bf16x8(ha), to_bf16x8(hb), acc, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 4; ++i) D[(4 * kq + i) * 16 + m] = acc[i];
}

// Layout probe: D = A B for one 16 x 16 x 16 tile; A, B bf16 row-major, D fp32 row-major.
__global__ void __launch_bounds__(64) mfma_probe_kernel(const uint16_t* A, const uint16_t* Bm, float* D) {
  const int lane = threadIdx.x;
  const int m = lane & 15, kq = lane >> 4;
  bf16x4 a, b;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    a[i] = (short)A[m * 16 + 4 * kq + i];
    b[i] = (short)Bm[(4 * kq + i) * 16 + m];
  }
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 4; ++i) D[(4 * kq + i) * 16 + m] = acc[i];
}

int check_conv(int B, int H, int W, int Ci, int Co) {
  if (B < 1 || H < 1 || W < 1) return ddx::set_error(DDX_ERR_ARG, "empty convolution");
  if (Ci < 16 || Co < 16 || Ci % 16 || Co % 16)
    return ddx::set_error(DDX_ERR_ARG, "padded channel counts must be positive multiples of 16");
  if (Co > 64) return ddx::set_error(DDX_ERR_UNSUPPORTED, "at most 64 output channels");
  if (((long)B * H * W) % 16) return ddx::set_error(DDX_ERR_ARG, "B*H*W must be a multiple of 16");
  if ((long)B * H * W * (Ci > Co ? Ci : Co) >= (1L << 40)) return ddx::set_error(DDX_ERR_ARG, "problem too large");
  return 0;
}

inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// The whole network in one launch (inference)
// ---------------------------------------------------------------------------
//
// NL layers of 3x3 periodic convolution: 16 (padded) input channels, HID
// hidden channels with GELU, 16 (padded) linear output channels.  A
// workgroup computes one TX x TY output tile of one image: it stages the
// tile's input with a halo of NL pixels in LDS, then each layer reads its
// input region from one LDS buffer and writes its output region, NL - 1 - l
// pixels of halo narrower, to the other.  Only the network's input and
// output touch HBM; the halo is recomputed by neighbouring tiles.
//
// Every layer is the same arithmetic as conv3x3_tile_kernel's (tap loop
// t = 0..8, K in the same chunks, bias then GELU, bf16 activations), with
// the MFMA operands swapped: A = weights (16 output channels x K), B = 16
// pixels' input channels, so a lane's D values are four consecutive output
// channels of one pixel (one 8-byte LDS write).
constexpr int NET_MAXL = 6;

struct NetArgs {
  const uint16_t* w[NET_MAXL];    // layer l: [Cop][9][Cip] bf16
  const float* bias[NET_MAXL];    // layer l: [Cop] fp32
  const uint16_t* x;              // [npix][16] bf16, or null: the fields below
  const double* src[DDX_CONV_MAXCH];
  float in_scale[DDX_CONV_MAXCH];
  int cin;
  float* out;                     // [npix][16] fp32, or null: the fields below
  double* dst[DDX_CONV_MAXCH];
  const double* add[DDX_CONV_MAXCH];  // all null, or one per output channel
  float out_scale[DDX_CONV_MAXCH];
  int cout;
  int B, H, W;
};

// LDS plan of one tile: buffer 0 holds the input and the outputs of layers
// 1, 3, ...; buffer 1 the outputs of layers 0, 2, ...  Pixel strides of 16 + 4
// and HID + 4 bf16 (8-byte aligned rows, consecutive pixels on distinct banks).
template <int NL, int HID, int TX, int TY>
struct NetPlan {
  static constexpr int PSI = 16 + 4, PSH = HID + 4;
  static constexpr int in_bytes() { return (TX + 2 * NL) * (TY + 2 * NL) * PSI * 2; }
  static constexpr int out_bytes(int l) {
    return (TX + 2 * (NL - 1 - l)) * (TY + 2 * (NL - 1 - l)) * PSH * 2;
  }
  static constexpr int buf_bytes(int b) {
    int m = b == 0 ? in_bytes() : 0;
    for (int l = 0; l < NL - 1; ++l)
      if ((l + 1) % 2 == b && out_bytes(l) > m) m = out_bytes(l);
    return (m + 15) / 16 * 16;
  }
  // the current layer's weights, staged as MFMA A operands (NetStage)
  static constexpr int wbuf_bytes() {
    int m = 0;
    for (int l = 0; l < NL; ++l) {
      const int ci = l == 0 ? 16 : HID, co = l == NL - 1 ? 16 : HID;
      const int b = 9 * (co / 16) * 64 * (ci == 32 ? 16 : 8);
      if (b > m) m = b;
    }
    return m;
  }
  static constexpr int lds_bytes() { return buf_bytes(0) + buf_bytes(1) + wbuf_bytes(); }
};

__device__ inline bf16x8 gld8(const uint16_t* p) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  const u32x4 u = {v.x, v.y, v.z, v.w};
  return __builtin_bit_cast(bf16x8, u);
}

__device__ inline bf16x4 gld4(const uint16_t* p) { return load4(p); }

__device__ inline int wrapm(int i, int n) {
  const int r = i % n;
  return r < 0 ? r + n : r;
}

// A layer's weights staged through registers into LDS, and its bias.  The
// LDS copy holds the MFMA A operands in lane order: block (t, j) (tap t,
// output channels 16 j ..) is 64 consecutive chunks, chunk l = output channel
// 16 j + l % 16, input channels (l / 16) KQ .. (KQ = 8 of 32 input channels,
// 4 of 16), so a wavefront's operand read of one block is one contiguous run
// of LDS.  The workgroup's threads share the chunks: the layer is read from
// global memory once per tile, not once per wavefront.  load() runs a layer
// ahead, under the previous layer's arithmetic; the bias stays in registers
// (a lane's four output channels of each 16).
__device__ inline bf16x8 lds16(const uint16_t* p) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  const u32x4 u = {v.x, v.y, v.z, v.w};
  return __builtin_bit_cast(bf16x8, u);
}

template <int CI, int CO, int NT>
struct NetStage {
  static constexpr int NCO = CO / 16, KQ = CI == 32 ? 8 : 4;
  static constexpr int NCH = 9 * NCO * 64, R = (NCH + NT - 1) / NT;
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  typedef typename std::conditional<CI == 32, u32x4, u32x2>::type chunk;
  chunk c[R];
  float b[NCO][4];
  __device__ __forceinline__ void load(const uint16_t* w_l, const float* b_l) {
    // unconditional loads (threads past the last chunk repeat it; store()
    // skips them), which keeps the chunks in registers
#pragma unroll
    for (int r = 0; r < R; ++r) {
      int q = (int)threadIdx.x + r * NT;
      q = q < NCH ? q : NCH - 1;
      const int blk = q >> 6, l = q & 63;
      const int t = blk / NCO, j = blk - t * NCO;
      c[r] = *reinterpret_cast<const chunk*>(w_l + ((j * 16 + (l & 15)) * 9 + t) * CI + (l >> 4) * KQ);
    }
    const int kq = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int j = 0; j < NCO; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) b[j][i] = b_l[j * 16 + 4 * kq + i];
  }
  __device__ __forceinline__ void store(uint16_t* wl) const {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = (int)threadIdx.x + r * NT;
      if (q < NCH) *reinterpret_cast<chunk*>(wl + q * KQ) = c[r];
    }
  }
};

// One layer over an output region of w x h pixels (row-major, flattened into
// 16-pixel fragments; the last fragment's spare lanes repeat the last pixel
// and store nothing).  in: (w + 2) x (h + 2) pixels with stride psi;
// LAST layers write the tile's pixels to global memory.
template <int CI, int CO, bool LAST, int NT>
__device__ __forceinline__ void net_layer(const NetArgs& a, const uint16_t* wl, const float (&bv)[CO / 16][4],
                                          const uint16_t* in, int psi, uint16_t* outb, int pso,
                                          int w, int h, int b, int y0, int x0) {
  constexpr int NCO = CO / 16;
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int npx = w * h, nfrag = (npx + 15) / 16;
  const int wi = w + 2;
  const int npairs = (nfrag + 1) / 2;
  // A pair of 16-pixel fragments per step: four independent accumulation
  // chains per wave.  The steps are software-pipelined over two accumulator
  // sets: the MFMAs of one pair are issued, then the epilogue (bias, GELU,
  // bf16 stores: VALU work) of the previous pair runs while the matrix cores
  // work through them.
  auto compute = [&](int pr, f32x4 (&acc)[2][NCO], int (&pp)[2], bool (&valid)[2]) {
    int base[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      int p = (2 * pr + u) * 16 + m;
      valid[u] = p < npx;
      if (!valid[u]) p = npx - 1;
      const int r = p / w, c = p - r * w;
      pp[u] = p;
      base[u] = r * wi + c;
    }
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int j = 0; j < NCO; ++j) acc[u][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int off = (t / 3) * wi + (t % 3);
      if constexpr (CI == 32) {
        bf16x8 wv[NCO];
#pragma unroll
        for (int j = 0; j < NCO; ++j) wv[j] = lds16(wl + ((t * NCO + j) * 64 + lane) * 8);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const bf16x8 xv = lds8(in + (base[u] + off) * psi + 8 * kq);
#pragma unroll
          for (int j = 0; j < NCO; ++j)
            acc[u][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wv[j], xv, acc[u][j], 0, 0, 0);
        }
      } else {
        bf16x4 wv[NCO];
#pragma unroll
        for (int j = 0; j < NCO; ++j) wv[j] = lds4(wl + ((t * NCO + j) * 64 + lane) * 4);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const bf16x4 xv = lds4(in + (base[u] + off) * psi + 4 * kq);
#pragma unroll
          for (int j = 0; j < NCO; ++j)
            acc[u][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wv[j], xv, acc[u][j], 0, 0, 0);
        }
      }
    }
  };
  auto finish = [&](const f32x4 (&acc)[2][NCO], const int (&pp)[2], const bool (&valid)[2]) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (!LAST) {
        // no branch on valid[u]: a spare lane computed the clamped last pixel
        // with the same operands as the lane that owns it, so its store
        // writes the same bits to the same LDS address
#pragma unroll
        for (int j = 0; j < NCO; ++j) {
          uint32_t h2[2];
#pragma unroll
          for (int e = 0; e < 2; ++e) {
            const float v0 = gelu(acc[u][j][2 * e] + bv[j][2 * e]);
            const float v1 = gelu(acc[u][j][2 * e + 1] + bv[j][2 * e + 1]);
            h2[e] = (uint32_t)f2bf(v0) | ((uint32_t)f2bf(v1) << 16);
          }
          *reinterpret_cast<uint2*>(&outb[pp[u] * pso + j * 16 + 4 * kq]) = make_uint2(h2[0], h2[1]);
        }
      } else {
        if (!valid[u]) continue;
        const int r = pp[u] / w, c = pp[u] - r * w;
        const long gp = ((long)b * a.H + y0 + r) * a.W + x0 + c;
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = acc[u][0][i] + bv[0][i];
        if (a.out) {
          *reinterpret_cast<float4*>(&a.out[gp * 16 + 4 * kq]) = make_float4(v[0], v[1], v[2], v[3]);
        } else {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int co = 4 * kq + i;
            if (co < a.cout) {
              const float x = v[i] * a.out_scale[co];
              a.dst[co][gp] = (double)x + (a.add[0] ? a.add[co][gp] : 0.0);
            }
          }
        }
      }
    }
  };
  f32x4 accA[2][NCO], accB[2][NCO];
  int ppA[2], ppB[2];
  bool vA[2], vB[2];
  bool pendA = false, pendB = false;
  for (int pr = wave; pr < npairs;) {
    compute(pr, accA, ppA, vA);
    if (pendB) finish(accB, ppB, vB);
    pendA = true;
    pendB = false;
    pr += NW;
    if (pr >= npairs) break;
    compute(pr, accB, ppB, vB);
    finish(accA, ppA, vA);
    pendA = false;
    pendB = true;
    pr += NW;
  }
  if (pendA) finish(accA, ppA, vA);
  if (pendB) finish(accB, ppB, vB);
}

template <int L, int NL, int HID, int NT>
struct NetLayerShape {
  static constexpr int CI = L == 0 ? 16 : HID;
  static constexpr int CO = L == NL - 1 ? 16 : HID;
  typedef NetStage<CI, CO, NT> W;
};

// Layer L with its weights in LDS (wl) and its bias in cur: the next
// layer's weights are loaded into registers first, then layer L runs; after
// a barrier (every wavefront done with layer L's input and weights) the next
// weights go to LDS, and after a second one the next layer runs.
template <int L, int NL, int HID, int TX, int TY, int NT>
__device__ __forceinline__ void net_layers(const NetArgs& a, uint16_t* buf0, uint16_t* buf1, uint16_t* wl,
                                           int b, int y0, int x0,
                                           const typename NetLayerShape<L, NL, HID, NT>::W& cur) {
  using P = NetPlan<NL, HID, TX, TY>;
  using S = NetLayerShape<L, NL, HID, NT>;
  constexpr int g = NL - 1 - L;  // output halo
  const uint16_t* in = (L % 2 == 0) ? buf0 : buf1;
  uint16_t* outb = (L % 2 == 0) ? buf1 : buf0;
  if constexpr (L < NL - 1) {
    typename NetLayerShape<L + 1, NL, HID, NT>::W nxt;
    nxt.load(a.w[L + 1], a.bias[L + 1]);
    net_layer<S::CI, S::CO, false, NT>(a, wl, cur.b, in, L == 0 ? P::PSI : P::PSH, outb, P::PSH,
                                       TX + 2 * g, TY + 2 * g, b, y0, x0);
    lds_barrier();
    nxt.store(wl);
    lds_barrier();
    net_layers<L + 1, NL, HID, TX, TY, NT>(a, buf0, buf1, wl, b, y0, x0, nxt);
  } else {
    net_layer<S::CI, S::CO, true, NT>(a, wl, cur.b, in, P::PSH, outb, P::PSH, TX + 2 * g, TY + 2 * g,
                                      b, y0, x0);
  }
}

template <int NL, int HID, int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) conv_net_kernel(NetArgs a) {
  using P = NetPlan<NL, HID, TX, TY>;
  extern __shared__ __align__(16) uint16_t netlds[];
  uint16_t* buf0 = netlds;
  uint16_t* buf1 = netlds + P::buf_bytes(0) / 2;
  uint16_t* wl = netlds + (P::buf_bytes(0) + P::buf_bytes(1)) / 2;
  const int tiles_x = a.W / TX, tiles_y = a.H / TY;
  int rest = blockIdx.x;
  const int tx = rest % tiles_x;
  rest /= tiles_x;
  const int ty = rest % tiles_y;
  const int b = rest / tiles_y;
  const int x0 = tx * TX, y0 = ty * TY;
  // the first layer's weights load while the input is staged
  typename NetLayerShape<0, NL, HID, NT>::W w0;
  w0.load(a.w[0], a.bias[0]);
  // stage the input region with a halo of NL pixels (periodic wrap)
  constexpr int RW = TX + 2 * NL, RH = TY + 2 * NL;
  for (int i = threadIdx.x; i < RW * RH; i += NT) {
    const int r = i / RW, c = i - r * RW;
    const int yy = wrapm(y0 - NL + r, a.H), xx = wrapm(x0 - NL + c, a.W);
    const long gp = ((long)b * a.H + yy) * a.W + xx;
    uint4 q0, q1;
    if (a.x) {
      q0 = *reinterpret_cast<const uint4*>(&a.x[gp * 16]);
      q1 = *reinterpret_cast<const uint4*>(&a.x[gp * 16 + 8]);
    } else {
      float v[16];
#pragma unroll
      for (int ch = 0; ch < 16; ++ch) {
        const int cc = ch < a.cin ? ch : 0;
        v[ch] = ch < a.cin ? (float)a.src[cc][gp] * a.in_scale[cc] : 0.f;
      }
      q0 = make_uint4((uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16),
                      (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16),
                      (uint32_t)f2bf(v[4]) | ((uint32_t)f2bf(v[5]) << 16),
                      (uint32_t)f2bf(v[6]) | ((uint32_t)f2bf(v[7]) << 16));
      q1 = make_uint4((uint32_t)f2bf(v[8]) | ((uint32_t)f2bf(v[9]) << 16),
                      (uint32_t)f2bf(v[10]) | ((uint32_t)f2bf(v[11]) << 16),
                      (uint32_t)f2bf(v[12]) | ((uint32_t)f2bf(v[13]) << 16),
                      (uint32_t)f2bf(v[14]) | ((uint32_t)f2bf(v[15]) << 16));
    }
    uint2* d = reinterpret_cast<uint2*>(&buf0[i * P::PSI]);
    d[0] = make_uint2(q0.x, q0.y);
    d[1] = make_uint2(q0.z, q0.w);
    d[2] = make_uint2(q1.x, q1.y);
    d[3] = make_uint2(q1.z, q1.w);
  }
  w0.store(wl);
  lds_barrier();
  net_layers<0, NL, HID, TX, TY, NT>(a, buf0, buf1, wl, b, y0, x0, w0);
}

template <int NL, int HID, int TX, int TY, int NT>
int launch_net(const NetArgs& a, hipStream_t s) {
  using P = NetPlan<NL, HID, TX, TY>;
  constexpr int lds = P::lds_bytes();
  static_assert(lds <= 160 * 1024, "tile does not fit the LDS");
  static bool attr = false;
  if (!attr) {
    const hipError_t e = hipFuncSetAttribute((const void*)conv_net_kernel<NL, HID, TX, TY, NT>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess)
      return ddx::set_error(DDX_ERR_HIP, std::string("conv_net LDS attribute: ") + hipGetErrorString(e));
    attr = true;
  }
  const long ntiles = (long)a.B * (a.H / TY) * (a.W / TX);
  hipLaunchKernelGGL((conv_net_kernel<NL, HID, TX, TY, NT>), dim3((unsigned)ntiles), dim3(NT), lds, s, a);
  CL_LAUNCHCHK("conv_net");
  return 0;
}

}  // namespace

extern "C" {

int ddx_conv3x3(void* stream, int B, int H, int W, int Ci, int Co, const void* x, const void* w,
                const float* bias, const float* aux, int mode, float* out_f32, void* out_bf16) {
  int rc = check_conv(B, H, W, Ci, Co);
  if (rc) return rc;
  if (!x || !w || (!out_f32 && !out_bf16)) return ddx::set_error(DDX_ERR_ARG, "null argument");
  if (mode != DDX_CONV_LINEAR && mode != DDX_CONV_GELU && mode != DDX_CONV_GELU_GRAD)
    return ddx::set_error(DDX_ERR_ARG, "unknown convolution mode");
  if (mode == DDX_CONV_GELU_GRAD && !aux)
    return ddx::set_error(DDX_ERR_ARG, "DDX_CONV_GELU_GRAD needs the pre-activation");
  ConvArgs a{(const uint16_t*)x, (const uint16_t*)w, bias, aux, out_f32, (uint16_t*)out_bf16,
             B, H, W, Ci, Co, mode};
  hipStream_t s = (hipStream_t)stream;
  auto al = [](const void* q, uintptr_t n) { return ((uintptr_t)q % n) == 0; };
  const bool aligned = al(x, 16) && al(w, 16) && al(out_f32, 16) && al(aux, 16) && al(out_bf16, 8);
  if (aligned && W % 64 == 0 && H % 4 == 0 && (Ci == 16 || Ci == 32) && (Co == 16 || Co == 32)) {
    // blocks loop over tiles; about four resident blocks per CU cover the chip
    const long ntiles = (long)B * (H / 4) * (W / 64);
    const dim3 tgrid((unsigned)(ntiles < 1024 ? ntiles : 1024));
    const int key = Ci * 100 + Co;
    // bf16-only output without the GELU gradient stages bf16 (BF)
    if (!out_f32 && mode != DDX_CONV_GELU_GRAD) {
      if (key == 1616) hipLaunchKernelGGL((conv3x3_tile_kernel<16, 16, true>), tgrid, dim3(256), 0, s, a, ntiles);
      else if (key == 1632) hipLaunchKernelGGL((conv3x3_tile_kernel<16, 32, true>), tgrid, dim3(256), 0, s, a, ntiles);
      else if (key == 3216) hipLaunchKernelGGL((conv3x3_tile_kernel<32, 16, true>), tgrid, dim3(256), 0, s, a, ntiles);
      else hipLaunchKernelGGL((conv3x3_tile_kernel<32, 32, true>), tgrid, dim3(256), 0, s, a, ntiles);
    } else {
      if (key == 1616) hipLaunchKernelGGL((conv3x3_tile_kernel<16, 16, false>), tgrid, dim3(256), 0, s, a, ntiles);
      else if (key == 1632) hipLaunchKernelGGL((conv3x3_tile_kernel<16, 32, false>), tgrid, dim3(256), 0, s, a, ntiles);
      else if (key == 3216) hipLaunchKernelGGL((conv3x3_tile_kernel<32, 16, false>), tgrid, dim3(256), 0, s, a, ntiles);
      else hipLaunchKernelGGL((conv3x3_tile_kernel<32, 32, false>), tgrid, dim3(256), 0, s, a, ntiles);
    }
    CL_LAUNCHCHK("conv3x3 tile");
    return 0;
  }
  const long tiles = (long)B * H * W / 16;
  const dim3 grid(nblk(tiles, 4));
  switch (Co / 16) {
    case 1: hipLaunchKernelGGL(conv3x3_kernel<1>, grid, dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL(conv3x3_kernel<2>, grid, dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL(conv3x3_kernel<3>, grid, dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL(conv3x3_kernel<4>, grid, dim3(256), 0, s, a); break;
  }
  CL_LAUNCHCHK("conv3x3");
  return 0;
}

int ddx_conv3x3_wgrad(void* stream, int B, int H, int W, int Ci, int Co, const void* x,
                      const void* dz, float* partial, int nchunk, float* dw) {
  int rc = check_conv(B, H, W, Ci, Co);
  if (rc) return rc;
  if (!x || !dz || !partial || !dw || nchunk < 1) return ddx::set_error(DDX_ERR_ARG, "null argument");
  const long npix = (long)B * H * W;
  hipStream_t s = (hipStream_t)stream;
  if (W % 64 == 0 && H % 4 == 0 && (Ci == 16 || Ci == 32) && (Co == 16 || Co == 32)) {
    const long ntiles = (long)B * (H / 4) * (W / 64);
    const int nch = (int)(ntiles < nchunk ? ntiles : nchunk);
    const int key = Ci * 100 + Co;
    if (key == 1616)
      hipLaunchKernelGGL((conv3x3_wgrad_tile_kernel<16, 16>), dim3(nch), dim3(256), 0, s,
                         (const uint16_t*)x, (const uint16_t*)dz, partial, B, H, W);
    else if (key == 1632)
      hipLaunchKernelGGL((conv3x3_wgrad_tile_kernel<16, 32>), dim3(nch), dim3(256), 0, s,
                         (const uint16_t*)x, (const uint16_t*)dz, partial, B, H, W);
    else if (key == 3216)
      hipLaunchKernelGGL((conv3x3_wgrad_tile_kernel<32, 16>), dim3(nch), dim3(256), 0, s,
                         (const uint16_t*)x, (const uint16_t*)dz, partial, B, H, W);
    else
      hipLaunchKernelGGL((conv3x3_wgrad_tile_kernel<32, 32>), dim3(nch), dim3(256), 0, s,
                         (const uint16_t*)x, (const uint16_t*)dz, partial, B, H, W);
    CL_LAUNCHCHK("conv3x3_wgrad tile");
    const long n = (long)Co * 9 * Ci;
    hipLaunchKernelGGL(sum_chunks_kernel, dim3(nblk(n, 64)), dim3(64 * SUM_GROUPS), 0, s, partial, dw, nch, n);
    CL_LAUNCHCHK("conv3x3_wgrad sum");
    return 0;
  }
  long pch = (npix + nchunk - 1) / nchunk;
  pch = (pch + 15) / 16 * 16;
  const int nch = (int)((npix + pch - 1) / pch);  // <= nchunk
  hipLaunchKernelGGL(conv3x3_wgrad_kernel, dim3(nch, (Co / 16) * (Ci / 16)), dim3(64), 0, s,
                     (const uint16_t*)x, (const uint16_t*)dz, partial, B, H, W, Ci, Co, (int)pch);
  CL_LAUNCHCHK("conv3x3_wgrad");
  const long n = (long)Co * 9 * Ci;
  hipLaunchKernelGGL(sum_chunks_kernel, dim3(nblk(n, 64)), dim3(64 * SUM_GROUPS), 0, s, partial, dw, nch, n);
  CL_LAUNCHCHK("conv3x3_wgrad sum");
  return 0;
}

int ddx_conv3x3_grads(void* stream, int B, int H, int W, int Ci, int Co, const void* x,
                      const void* dz, const float* dzf, float* partial, double* bpartial, int nchunk,
                      float* dw, float* db) {
  int rc = check_conv(B, H, W, Ci, Co);
  if (rc) return rc;
  if (!x || !dz || !dzf || !partial || !bpartial || !dw || !db || nchunk < 1)
    return ddx::set_error(DDX_ERR_ARG, "null argument");
  if (!(W % 64 == 0 && H % 4 == 0 && (Ci == 16 || Ci == 32) && (Co == 16 || Co == 32)))
    return ddx::set_error(DDX_ERR_UNSUPPORTED,
                          "fused gradients need W % 64 == 0, H % 4 == 0 and 16 or 32 channels");
  hipStream_t s = (hipStream_t)stream;
  const long ntiles = (long)B * (H / 4) * (W / 64);
  const int nch = (int)(ntiles < nchunk ? ntiles : nchunk);
  const int key = Ci * 100 + Co;
  const uint16_t* xx = (const uint16_t*)x;
  const uint16_t* dd = (const uint16_t*)dz;
  if (key == 1616)
    hipLaunchKernelGGL((conv3x3_wgrad_tile_kernel<16, 16>), dim3(nch), dim3(256), 0, s, xx, dd, partial,
                       B, H, W, dzf, bpartial);
  else if (key == 1632)
    hipLaunchKernelGGL((conv3x3_wgrad_tile_kernel<16, 32>), dim3(nch), dim3(256), 0, s, xx, dd, partial,
                       B, H, W, dzf, bpartial);
  else if (key == 3216)
    hipLaunchKernelGGL((conv3x3_wgrad_tile_kernel<32, 16>), dim3(nch), dim3(256), 0, s, xx, dd, partial,
                       B, H, W, dzf, bpartial);
  else
    hipLaunchKernelGGL((conv3x3_wgrad_tile_kernel<32, 32>), dim3(nch), dim3(256), 0, s, xx, dd, partial,
                       B, H, W, dzf, bpartial);
  CL_LAUNCHCHK("conv3x3 grads");
  const long n = (long)Co * 9 * Ci;
  hipLaunchKernelGGL(sum_chunks_kernel, dim3(nblk(n, 64)), dim3(64 * SUM_GROUPS), 0, s, partial, dw, nch, n);
  CL_LAUNCHCHK("conv3x3 grads: weight sum");
  hipLaunchKernelGGL(colsum_final_kernel, dim3(Co), dim3(256), 0, s, bpartial, db, nch, Co);
  CL_LAUNCHCHK("conv3x3 grads: bias sum");
  return 0;
}

int ddx_conv_colsum(void* stream, const float* a, int64_t npix, int C, double* partial, int nchunk,
                    float* out) {
  if (!a || !out || !partial || npix < 1 || C < 1 || C > 256 || nchunk < 1)
    return ddx::set_error(DDX_ERR_ARG, "bad column-sum arguments");
  const long chunk = (npix + nchunk - 1) / nchunk;
  const int nblk = (int)((npix + chunk - 1) / chunk);  // <= nchunk
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(colsum_partial_kernel, dim3(nblk), dim3(256), 0, s, a, partial, (long)npix, C,
                     chunk);
  CL_LAUNCHCHK("conv_colsum");
  hipLaunchKernelGGL(colsum_final_kernel, dim3(C), dim3(256), 0, s, partial, out, nblk, C);
  CL_LAUNCHCHK("conv_colsum final");
  return 0;
}

int ddx_conv_wprep(void* stream, const float* w, int Co, int Ci, int Cop, int Cip, void* wf, void* wt) {
  if (!w || (!wf && !wt) || Co < 1 || Ci < 1 || Cop < Co || Cip < Ci)
    return ddx::set_error(DDX_ERR_ARG, "bad weight-prep arguments");
  const long n = (long)Cop * 9 * Cip;
  hipLaunchKernelGGL(wprep_kernel, dim3(nblk(n, 256)), dim3(256), 0, (hipStream_t)stream, w, Co, Ci,
                     Cop, Cip, (uint16_t*)wf, (uint16_t*)wt);
  CL_LAUNCHCHK("conv_wprep");
  return 0;
}

int ddx_conv_wunpack(void* stream, const float* dwp, int Co, int Ci, int Cop, int Cip, float* dw) {
  if (!dwp || !dw || Co < 1 || Ci < 1 || Cop < Co || Cip < Ci)
    return ddx::set_error(DDX_ERR_ARG, "bad weight-unpack arguments");
  const long n = (long)Co * Ci * 9;
  hipLaunchKernelGGL(wunpack_kernel, dim3(nblk(n, 256)), dim3(256), 0, (hipStream_t)stream, dwp, Co,
                     Ci, Cop, Cip, dw);
  CL_LAUNCHCHK("conv_wunpack");
  return 0;
}

int ddx_conv_pack(void* stream, int64_t npix, int C, int Cp, const void* const* src, int64_t pstride,
                  int src_f64, const float* scale, void* dst_bf16, float* dst_f32) {
  if (!src || npix < 1 || C < 1 || C > DDX_CONV_MAXCH || Cp < C || pstride < 1 || (!dst_bf16 && !dst_f32))
    return ddx::set_error(DDX_ERR_ARG, "bad pack arguments");
  PackArgs a{};
  for (int c = 0; c < C; ++c) {
    if (!src[c]) return ddx::set_error(DDX_ERR_ARG, "null channel source");
    a.src[c] = src[c];
    a.scale[c] = scale ? scale[c] : 1.0f;
  }
  a.pstride = pstride;
  a.npix = npix;
  a.C = C;
  a.Cp = Cp;
  a.f64 = src_f64 ? 1 : 0;
  a.dst_bf16 = (uint16_t*)dst_bf16;
  a.dst_f32 = dst_f32;
  const bool al = ((uintptr_t)dst_bf16 % 16) == 0 && ((uintptr_t)dst_f32 % 16) == 0;
  const dim3 pg(nblk(npix, 256));
  hipStream_t ps = (hipStream_t)stream;
  if (Cp % 8 == 0 && al) {
    if (a.f64 && C <= 4) hipLaunchKernelGGL((pack_pixel_kernel<true, 4>), pg, dim3(256), 0, ps, a);
    else if (a.f64) hipLaunchKernelGGL((pack_pixel_kernel<true, DDX_CONV_MAXCH>), pg, dim3(256), 0, ps, a);
    else if (C <= 4) hipLaunchKernelGGL((pack_pixel_kernel<false, 4>), pg, dim3(256), 0, ps, a);
    else hipLaunchKernelGGL((pack_pixel_kernel<false, DDX_CONV_MAXCH>), pg, dim3(256), 0, ps, a);
  } else
    hipLaunchKernelGGL(pack_kernel, dim3(nblk(npix * Cp, 256)), dim3(256), 0, (hipStream_t)stream, a);
  CL_LAUNCHCHK("conv_pack");
  return 0;
}

int ddx_conv_unpack(void* stream, int64_t npix, int C, int Cp, const float* src, void* const* dst,
                    int64_t pstride, int dst_f64, const float* scale, const void* const* add) {
  if (!src || !dst || npix < 1 || C < 1 || C > DDX_CONV_MAXCH || Cp < C || pstride < 1)
    return ddx::set_error(DDX_ERR_ARG, "bad unpack arguments");
  UnpackArgs a{};
  for (int c = 0; c < C; ++c) {
    if (!dst[c]) return ddx::set_error(DDX_ERR_ARG, "null channel destination");
    if (add && !add[c]) return ddx::set_error(DDX_ERR_ARG, "null channel addend");
    a.dst[c] = dst[c];
    a.add[c] = add ? add[c] : nullptr;
    a.scale[c] = scale ? scale[c] : 1.0f;
  }
  a.pstride = pstride;
  a.npix = npix;
  a.C = C;
  a.Cp = Cp;
  a.f64 = dst_f64 ? 1 : 0;
  a.src = src;
  hipLaunchKernelGGL(unpack_kernel, dim3(nblk(npix, 256)), dim3(256), 0, (hipStream_t)stream, a);
  CL_LAUNCHCHK("conv_unpack");
  return 0;
}

int ddx_box_filter(void* stream, int ny, int nx, int factor, int f64, const void* src, void* dst) {
  if (!src || !dst || ny < 1 || nx < 1 || factor < 1 || ny % factor || nx % factor)
    return ddx::set_error(DDX_ERR_ARG, "box filter: extents must be positive multiples of the factor");
  const long n = (long)(ny / factor) * (nx / factor);
  hipStream_t s = (hipStream_t)stream;
  if (f64)
    hipLaunchKernelGGL(box_filter_kernel<double>, dim3(nblk(n, 256)), dim3(256), 0, s,
                       (const double*)src, (double*)dst, ny, nx, factor);
  else
    hipLaunchKernelGGL(box_filter_kernel<float>, dim3(nblk(n, 256)), dim3(256), 0, s,
                       (const float*)src, (float*)dst, ny, nx, factor);
  CL_LAUNCHCHK("box_filter");
  return 0;
}

int ddx_conv_net(void* stream, int B, int H, int W, int nl, const int* chans,
                 const void* const* w, const float* const* bias, const void* x, float* out,
                 const void* const* src, const float* in_scale, void* const* dst,
                 const void* const* add, const float* out_scale) {
  if (B < 1 || H < 1 || W < 1) return ddx::set_error(DDX_ERR_ARG, "bad image shape");
  if (!chans || !w || !bias) return ddx::set_error(DDX_ERR_ARG, "null argument");
  if ((!x && !src) || (!out && !dst)) return ddx::set_error(DDX_ERR_ARG, "no input or no output");
  if (nl < 2 || nl > NET_MAXL) return ddx::set_error(DDX_ERR_UNSUPPORTED, "network depth");
  const int cin = chans[0], cout = chans[nl], hid = chans[1];
  if (cin < 1 || cin > 16 || cout < 1 || cout > 16)
    return ddx::set_error(DDX_ERR_UNSUPPORTED, "network input and output need <= 16 channels");
  for (int l = 1; l < nl; ++l)
    if (chans[l] != hid) return ddx::set_error(DDX_ERR_UNSUPPORTED, "hidden layers of one width");
  NetArgs a;
  memset(&a, 0, sizeof(a));
  for (int l = 0; l < nl; ++l) {
    if (!w[l] || !bias[l]) return ddx::set_error(DDX_ERR_ARG, "null weights or bias");
    if ((uintptr_t)w[l] % 16) return ddx::set_error(DDX_ERR_ARG, "weights must be 16-byte aligned");
    a.w[l] = (const uint16_t*)w[l];
    a.bias[l] = bias[l];
  }
  if (x && (uintptr_t)x % 16) return ddx::set_error(DDX_ERR_ARG, "x must be 16-byte aligned");
  if (out && (uintptr_t)out % 16) return ddx::set_error(DDX_ERR_ARG, "out must be 16-byte aligned");
  a.x = (const uint16_t*)x;
  a.out = out;
  a.cin = cin;
  a.cout = cout;
  if (!x) {
    for (int c = 0; c < cin; ++c) {
      if (!src[c]) return ddx::set_error(DDX_ERR_ARG, "null input field");
      a.src[c] = (const double*)src[c];
      a.in_scale[c] = in_scale ? in_scale[c] : 1.f;
    }
  }
  if (!out) {
    for (int c = 0; c < cout; ++c) {
      if (!dst[c] || (add && !add[c])) return ddx::set_error(DDX_ERR_ARG, "null output field");
      a.dst[c] = (double*)dst[c];
      a.add[c] = add ? (const double*)add[c] : nullptr;
      a.out_scale[c] = out_scale ? out_scale[c] : 1.f;
    }
  }
  a.B = B;
  a.H = H;
  a.W = W;
  hipStream_t s = (hipStream_t)stream;
  // large tiles (less recomputed halo) when they alone fill the chip, small
  // tiles (more workgroups) for a few images
  const long big = (H % 16 == 0 && W % 32 == 0) ? (long)B * (H / 16) * (W / 32) : 0;
  const bool small_ok = H % 8 == 0 && W % 16 == 0;
  if (nl == 5 && hid == 32) {
    if (big >= 256 || (big > 0 && !small_ok)) return launch_net<5, 32, 32, 16, 768>(a, s);
    // a few images: small tiles, whose layers are a few short dependent
    // phases each (one 64^2 image, a single coarse run: 4 x 4 tiles, one
    // per CU, 9.4 us per pass against 10.0 us with 8 x 8 tiles, though the
    // executed MFMA work grows from 3.0x to 5.5x the network's: the pass
    // is a chain of latencies, not of arithmetic)
    if ((long)B * (H / 4) * (W / 4) <= 256 && H % 4 == 0 && W % 4 == 0)
      return launch_net<5, 32, 4, 4, 256>(a, s);
    // then 8 x 8 tiles: each layer's region at most one fragment pair per
    // wavefront, so the five layers are five short phases rather than rounds
    // of pairs
    const long small = small_ok ? (long)B * (H / 8) * (W / 16) : 0;
    if (small < 128 && H % 8 == 0 && W % 8 == 0) return launch_net<5, 32, 8, 8, 512>(a, s);
    if (small_ok) return launch_net<5, 32, 16, 8, 256>(a, s);
  } else if (nl == 3 && hid == 16) {
    if (big >= 256 || (big > 0 && !small_ok)) return launch_net<3, 16, 32, 16, 256>(a, s);
    if (small_ok) return launch_net<3, 16, 16, 8, 256>(a, s);
  }
  return ddx::set_error(DDX_ERR_UNSUPPORTED,
                        "fused network: depth 5 width 32 on images of multiples of 8 x 8 "
                        "pixels, or depth 3 width 16 on multiples of 8 x 16");
}

int ddx_mfma_probe(void* stream, int k, const void* A, const void* Bm, float* D) {
  if (!A || !Bm || !D) return ddx::set_error(DDX_ERR_ARG, "null argument");
  if (k != 16 && k != 32) return ddx::set_error(DDX_ERR_ARG, "probe K must be 16 or 32");
  if (k == 16)
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream,
                       (const uint16_t*)A, (const uint16_t*)Bm, D);
  else
    hipLaunchKernelGGL(mfma_probe32_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream,
                       (const uint16_t*)A, (const uint16_t*)Bm, D);
  CL_LAUNCHCHK("mfma_probe");
  return 0;
}

}  // extern "C"
