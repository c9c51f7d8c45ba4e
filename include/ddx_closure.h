/*
 * ddx_closure.h -- C ABI of the learned sub-grid closure kernels in libddx.so.
 *
 * The reference has no learned closure.  Its plug point is the velocity
 * forcing hook (dedalus/physics/physics.py:524-525, applied at :569-572):
 * the closure's predicted tendency enters the step as that forcing, through
 * ddx_rhs / ddx_step (ddx.h).  These entry points are the network's
 * building blocks: BF16 MFMA 3x3 periodic convolutions with FP32
 * accumulation and fused epilogues, weight preparation, channel
 * packing between the solver's x-space fields and the network's
 * pixel-major activations, and the coarse-graining box filter.
 *
 * Conventions
 *   - Return 0 or a negative DDX_ERR_* code (ddx.h); message in
 *     ddx_last_error().
 *   - stream is a hipStream_t (NULL = the null stream); every call is
 *     asynchronous on it.
 *   - Activations are pixel-major: [npix][C] with npix = B*H*W in (b, y, x)
 *     order; bf16 arrays hold raw 16-bit patterns; channel counts Ci, Co of
 *     the convolutions are padded to multiples of 16 (Co <= 64).
 *   - Convolution weights are [Co][9][Ci] bf16 with tap t = 3*ky + kx, i.e.
 *     out[y][x] = sum w[co][3*ky+kx][ci] in[y+ky-1][x+kx-1], indices
 *     periodic; ddx_conv_wprep builds them from fp32 (co, ci, ky, kx).
 */
#ifndef DDX_CLOSURE_H
#define DDX_CLOSURE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DDX_CONV_LINEAR = 0,    /* out = conv + bias */
  DDX_CONV_GELU = 1,      /* out_f32 = conv + bias; out_bf16 = GELU(conv + bias) */
  DDX_CONV_GELU_GRAD = 2  /* out = conv * GELU'(aux): the input gradient through
                             the previous layer's activation */
};

enum { DDX_CONV_MAXCH = 16 }; /* channels per pack / unpack call */

/* 3x3 periodic convolution on B images of H x W pixels.  x [npix][Ci] bf16,
 * w [Co][9][Ci] bf16, bias [Co] fp32 or NULL (ignored in GELU_GRAD mode),
 * aux [npix][Co] fp32 (GELU_GRAD only).  Writes out_f32 [npix][Co] and/or
 * out_bf16 [npix][Co] (either may be NULL, not both).  With the transposed
 * weights of ddx_conv_wprep, the same call gives the input gradient.
 * Widths W % 64 == 0, heights H % 4 == 0 and 16 or 32 channels with 16-byte
 * aligned x, w, out_f32, aux (8-byte aligned out_bf16) take the LDS-tiled
 * kernel; other shapes a per-fragment kernel with the same results. */
int ddx_conv3x3(void* stream, int B, int H, int W, int Ci, int Co, const void* x, const void* w,
                const float* bias, const float* aux, int mode, float* out_f32, void* out_bf16);

/* Weight gradient dw[co][t][ci] = sum_p dz[p][co] x[shift_t(p)][ci] (fp32,
 * padded layout).  partial is scratch of nchunk * Co * 9 * Ci floats; the
 * pixel range is split into at most nchunk chunks whose partial sums are
 * added in a fixed order (deterministic). */
int ddx_conv3x3_wgrad(void* stream, int B, int H, int W, int Ci, int Co, const void* x,
                      const void* dz, float* partial, int nchunk, float* dw);

/* Both gradients of one layer in one pass over dz: ddx_conv3x3_wgrad's
 * dw [Co][9][Ci] and ddx_conv_colsum's db[co] = sum_p dzf[p][co] (dzf the
 * fp32 dz, dz its bf16 rounding), the column sums taken per block of the
 * weight-gradient kernel in double and added in a fixed order.  partial:
 * nchunk * Co * 9 * Ci floats, bpartial: nchunk * Co doubles of scratch.
 * Widths W % 64 == 0, heights H % 4 == 0 and 16 or 32 channels only
 * (DDX_ERR_UNSUPPORTED otherwise: call the two separately). */
int ddx_conv3x3_grads(void* stream, int B, int H, int W, int Ci, int Co, const void* x,
                      const void* dz, const float* dzf, float* partial, double* bpartial, int nchunk,
                      float* dw, float* db);

/* out[c] = sum_p a[p][c] (bias gradient), c < C <= 256.  partial is scratch
 * of nchunk * C doubles; pixels are split into at most nchunk chunks summed
 * in a fixed order (deterministic). */
int ddx_conv_colsum(void* stream, const float* a, int64_t npix, int C, double* partial,
                    int nchunk, float* out);

/* fp32 weights (co, ci, ky, kx) of Co x Ci -> padded bf16 forward weights
 * wf [Cop][9][Cip] and transposed weights wt [Cip][9][Cop] with
 * wt[ci][t][co] = w[co][ci][8 - t] (either output may be NULL). */
int ddx_conv_wprep(void* stream, const float* w, int Co, int Ci, int Cop, int Cip, void* wf,
                   void* wt);

/* Padded weight gradient [Cop][9][Cip] -> (co, ci, ky, kx) fp32 of Co x Ci. */
int ddx_conv_wunpack(void* stream, const float* dwp, int Co, int Ci, int Cop, int Cip, float* dw);

/* Gather C channels into pixel-major padded activations:
 * dst[p][c] = scale[c] * src[c][p * pstride] for c < C, 0 for C <= c < Cp.
 * src entries are float64 (src_f64 = 1) or float32; scale may be NULL (1).
 * Writes dst_bf16 and/or dst_f32. */
int ddx_conv_pack(void* stream, int64_t npix, int C, int Cp, const void* const* src,
                  int64_t pstride, int src_f64, const float* scale, void* dst_bf16,
                  float* dst_f32);

/* Scatter: dst[c][p * pstride] = scale[c] * src[p][c] (+ add[c][p * pstride])
 * for c < C (src has Cp channels per pixel); dst and add entries float64
 * (dst_f64 = 1) or float32; add may be NULL.  With add, a constant x-space
 * forcing joins the network's tendency before its forward transform. */
int ddx_conv_unpack(void* stream, int64_t npix, int C, int Cp, const float* src,
                    void* const* dst, int64_t pstride, int dst_f64, const float* scale,
                    const void* const* add);

/* Coarse-graining: dst[Y][X] = mean of src over the factor x factor block at
 * (factor*Y, factor*X); src is ny x nx (float64 if f64, else float32). */
int ddx_box_filter(void* stream, int ny, int nx, int factor, int f64, const void* src, void* dst);

/* The whole network's forward pass in one launch (inference): nl 3x3
 * periodic convolution layers with chans[0..nl] channels (padded to 16 in
 * the weights and activations), GELU after every layer but the last, bf16
 * activations and fp32 accumulation -- the arithmetic of nl ddx_conv3x3
 * calls (GELU mode, then LINEAR), without the intermediate activations
 * leaving the chip.  w[l] = layer l's [Cop][9][Cip] bf16 weights (16-byte
 * aligned), bias[l] its [Cop] fp32 bias.  Input: x [npix][16] bf16, or (x
 * NULL) chans[0] float64 fields src[c][p] scaled by in_scale[c] and rounded
 * to bf16 as ddx_conv_pack does.  Output: out [npix][16] fp32, or (out NULL)
 * chans[nl] float64 fields dst[c][p] = out_scale[c] * y_c[p] (+ add[c][p])
 * as ddx_conv_unpack does.  Supported: depth 5 with 32 hidden channels (H
 * and W multiples of 8) and depth 3 with 16 (H a multiple of 8, W of 16),
 * inputs and outputs of at most 16 channels; otherwise DDX_ERR_UNSUPPORTED
 * and the layer-by-layer calls apply.  The tile size follows the image
 * count: 32 x 16 output tiles when they alone fill the chip, 4 x 4 when at
 * most 256 of them cover the images (one 64^2 coarse run), 8 x 8 for a few
 * images, 16 x 8 in between.  in_scale, out_scale, add may be NULL (1, 1,
 * nothing). */
int ddx_conv_net(void* stream, int B, int H, int W, int nl, const int* chans,
                 const void* const* w, const float* const* bias, const void* x, float* out,
                 const void* const* src, const float* in_scale, void* const* dst,
                 const void* const* add, const float* out_scale);

/* Layout check: D = A B for one 16 x 16 x k bf16 tile (k = 16 or 32) on one
 * wavefront, with the lane placement the convolution kernels use (A 16 x k,
 * B k x 16 row-major bf16, D 16 x 16 row-major fp32). */
int ddx_mfma_probe(void* stream, int k, const void* A, const void* Bm, float* D);

#ifdef __cplusplus
}
#endif

#endif /* DDX_CLOSURE_H */
