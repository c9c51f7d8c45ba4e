// Device kernels of libddx: column FFT passes, x-row passes fused with the
// physics products, k-space epilogues (projection, curl, buoyancy), ETD
// updates, layout conversion.
//
// Data layout in HBM
//   x space   : (Nz, Ny, Nx) real, x fastest (Nz = 1 in 2D)
//   compact k : (Ky, Kz, Kxp) complex, kx fastest, only 2/3-kept modes;
//               row index of a full axis j <-> frequency j (j <= c) or
//               j + N - K (j > c); pad columns kx >= Kx are zero
//   mixed     : (Ky, Nz, Kxp) after the z pass, (Nz, Ny, Kxp) after the y pass
#pragma once

#include "fft_core.h"

namespace ddx {

#define DDX_MAXF 16

// One term of a derived field: coef * (i k_dim)^[dim >= 0] * source[src]
struct Term {
  int src;
  int dim;  // -1: none, 0: x, 1: y, 2: z
  double coef;
};
struct Recipe {
  int nterm;
  Term t[2];
};

// ---------------------------------------------------------------------------
// column pass: batched FFT along a strided axis, W columns per workgroup
// ---------------------------------------------------------------------------

template <typename T>
struct ColArgs {
  const cpx<T>* in[DDX_MAXF];
  cpx<T>* out[DDX_MAXF];
  const cpx<T>* src[DDX_MAXF];  // sources for the spectral prologue
  Recipe rec[DDX_MAXF];
  FftPlan fp;
  const cpx<T>* tw;
  const double* k_row;   // compact wavenumbers along the row axis (prologue)
  const double* k_slab;  // compact wavenumbers along the slab axis (prologue)
  const double* k_col;   // kx table (prologue)
  long in_rs, in_ss, out_rs, out_ss;  // row / slab strides (complex elements)
  int n;                 // FFT length
  int in_rows, in_c;     // in_c < 0: rows are the full axis
  int out_rows, out_c;   // out_c < 0: rows are the full axis
  int ncols;             // columns per slab
  int W;                 // tile width (columns per workgroup)
  int row_dim, slab_dim; // physical axis (0 x, 1 y, 2 z) of rows / slabs, -1 none
  int nslabs, nfields;   // grid extents (used by the remapped 1D grids)
  int ntiles;            // column tiles per slab
  int xcd_group;         // 1D grid on XCDs: 1 the fields of a (slab, tile), 2 whole slabs
  int group_hint;        // set by the caller: group the fields per XCD (spectral passes)
  int wide;              // set by the caller: double-width tiles (n = 256)
  // the row axis's wavenumbers from the index, k = ((ms * inv) * 2) * kny
  // with ms the signed frequency index (some size-specialised kernels)
  double kr_inv, kr_kny;
  int nt;                // set by the caller: nontemporal output stores
  T scale;
};

template <typename T, int D, bool SPECTRAL>
__global__ void __launch_bounds__(256) col_fft_kernel(const ColArgs<T> a) {
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem_raw);
  const int W = a.W, n = a.n;
  cpx<T>* B = A + n * W;
  const int f = blockIdx.z;
  const long slab = blockIdx.y;
  const int c0 = blockIdx.x * W;
  const int tid = threadIdx.x, nthr = blockDim.x;

  for (int e = tid; e < n * W; e += nthr) {
    const int m = e / W, c = e - m * W;
    const int col = c0 + c;
    cpx<T> v = mk(T(0), T(0));
    if (col < a.ncols) {
      int row;
      if (a.in_c < 0) row = m;
      else if (m <= a.in_c) row = m;
      else if (m >= n - a.in_c) row = m - (n - a.in_rows);
      else row = -1;
      if (row >= 0) {
        const long off = slab * a.in_ss + (long)row * a.in_rs + col;
        if (SPECTRAL) {
          const int nterm = a.rec[f].nterm;
#pragma unroll
          for (int t = 0; t < 2; ++t) {
            if (t < nterm) {
              const Term tm = a.rec[f].t[t];
              cpx<T> x = a.src[tm.src][off];
              if (tm.dim >= 0) {
                const double kk = (tm.dim == 0) ? a.k_col[col]
                                  : (tm.dim == a.row_dim ? a.k_row[row] : a.k_slab[slab]);
                x = mul_is(x, T(kk));
              }
              v = v + scale(x, T(tm.coef));
            }
          }
        } else {
          v = a.in[f][off];
        }
      }
    }
    A[m * W + c] = v;
  }
  __syncthreads();
  cpx<T>* R = lds_fft<T, T, D>(A, B, a.fp, W, W, 1, true, a.tw, tid, nthr);
  for (int e = tid; e < a.out_rows * W; e += nthr) {
    const int j = e / W, c = e - j * W;
    const int col = c0 + c;
    if (col >= a.ncols) continue;
    const int m = (a.out_c < 0 || j <= a.out_c) ? j : j + (n - a.out_rows);
    a.out[f][slab * a.out_ss + (long)j * a.out_rs + col] = scale(R[m * W + c], a.scale);
  }
}

// ---------------------------------------------------------------------------
// x-row pass: c2r of complex half-spectrum rows (two real rows per complex
// FFT), x-space physics, r2c back to half-spectrum rows
// ---------------------------------------------------------------------------

enum { ROW_C2R = 0, ROW_R2C = 1, ROW_PHYS = 2 };
enum { PH_HYDRO = 0, PH_BOUSS = 1, PH_MHD = 2 };

template <typename T>
struct RowArgs {
  const cpx<T>* in[DDX_MAXF];
  cpx<T>* out[DDX_MAXF];
  const T* xin[DDX_MAXF];
  T* xout[DDX_MAXF];
  FftPlan fp;
  const cpx<T>* tw;
  int nin, nout;
  int nx, kx, kxp;
  int mode, physics, ndim;
  T fpr_inv;  // 1 / (4 pi rho0)
  int nt;     // nontemporal output stores (x-row products)
  // ROW_PHYS: x-space arrays added to product f before its r2c (a forcing
  // given in x space: the forward transform then carries it), or null
  const T* xadd[DDX_MAXF];
};

// x-space products.  Inputs per physics (ndim 3 / 2):
//   HYDRO : u[3], w[3]            | u[2], w
//   BOUSS : u[3], w[3], T, dT[3]  | u[2], w, T, dT[2]
//   MHD   : u[3], w[3], B[3], J[3]| u[2], w, B[2], J
// Outputs: velocity u x w (+ J x B / (4 pi rho0)); BOUSS adds -u.grad T;
// MHD adds the electric field u x B (3D) or its z component (2D).
template <typename T>
__device__ inline void row_physics(int physics, int ndim, T fi, const T* q, T* o) {
  if (ndim == 3) {
    const T u0 = q[0], u1 = q[1], u2 = q[2], w0 = q[3], w1 = q[4], w2 = q[5];
    o[0] = u1 * w2 - u2 * w1;
    o[1] = u2 * w0 - u0 * w2;
    o[2] = u0 * w1 - u1 * w0;
    if (physics == PH_BOUSS) {
      o[3] = -(u0 * q[7] + u1 * q[8] + u2 * q[9]);
    } else if (physics == PH_MHD) {
      const T b0 = q[6], b1 = q[7], b2 = q[8], j0 = q[9], j1 = q[10], j2 = q[11];
      o[0] += fi * (j1 * b2 - j2 * b1);
      o[1] += fi * (j2 * b0 - j0 * b2);
      o[2] += fi * (j0 * b1 - j1 * b0);
      o[3] = u1 * b2 - u2 * b1;
      o[4] = u2 * b0 - u0 * b2;
      o[5] = u0 * b1 - u1 * b0;
    }
  } else {
    const T u0 = q[0], u1 = q[1], w = q[2];
    o[0] = u1 * w;
    o[1] = -u0 * w;
    if (physics == PH_BOUSS) {
      o[2] = -(u0 * q[4] + u1 * q[5]);
    } else if (physics == PH_MHD) {
      const T b0 = q[3], b1 = q[4], j = q[5];
      o[0] -= fi * j * b1;
      o[1] += fi * j * b0;
      o[2] = u0 * b1 - u1 * b0;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) row_kernel(const RowArgs<T> a) {
  extern __shared__ unsigned char smem_raw[];
  const int nx = a.nx;
  const int nsin = (a.nin + 1) / 2, nsout = (a.nout + 1) / 2;
  const int nseq = nsin > nsout ? nsin : nsout;
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* B = A + nseq * nx;
  const long r = blockIdx.x;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int cx = a.kx - 1;
  cpx<T>* X;

  if (a.mode == ROW_R2C) {
    for (int e = tid; e < nsout * nx; e += nthr) {
      const int q = e / nx, m = e - q * nx;
      T re = a.xin[2 * q][r * nx + m];
      T im = (2 * q + 1 < a.nout) ? a.xin[2 * q + 1][r * nx + m] : T(0);
      A[q * nx + m] = mk(re, im);
    }
    __syncthreads();
    X = A;
  } else {
    // Z[k] = A[k] + i B[k] over the full spectrum from the half spectra
    for (int e = tid; e < nsin * nx; e += nthr) {
      const int q = e / nx, m = e - q * nx;
      const bool hasb = (2 * q + 1 < a.nin);
      cpx<T> va = mk(T(0), T(0)), vb = mk(T(0), T(0));
      if (m <= cx) {
        va = a.in[2 * q][r * a.kxp + m];
        if (hasb) vb = a.in[2 * q + 1][r * a.kxp + m];
        if (m == 0) { va.y = T(0); vb.y = T(0); }  // c2r keeps the real part of kx = 0
      } else if (m >= nx - cx) {
        va = conj(a.in[2 * q][r * a.kxp + (nx - m)]);
        if (hasb) vb = conj(a.in[2 * q + 1][r * a.kxp + (nx - m)]);
      }
      A[q * nx + m] = mk(va.x - vb.y, va.y + vb.x);
    }
    __syncthreads();
    X = lds_fft<T, T, 1>(A, B, a.fp, nsin, 1, nx, false, a.tw, tid, nthr);
    if (a.mode == ROW_C2R) {
      for (int e = tid; e < a.nin * nx; e += nthr) {
        const int f = e / nx, m = e - f * nx;
        const cpx<T> z = X[(f >> 1) * nx + m];
        a.xout[f][r * nx + m] = (f & 1) ? z.y : z.x;
      }
      return;
    }
    // physics on the real values, result packed into the other buffer
    cpx<T>* Y = (X == A) ? B : A;
    for (int m = tid; m < nx; m += nthr) {
      T q[12], o[6];
#pragma unroll
      for (int s = 0; s < 6; ++s) {
        cpx<T> z = mk(T(0), T(0));
        if (s < nsin) z = X[s * nx + m];
        q[2 * s] = z.x;
        q[2 * s + 1] = z.y;
      }
#pragma unroll
      for (int f = 0; f < 6; ++f) o[f] = T(0);
      row_physics<T>(a.physics, a.ndim, a.fpr_inv, q, o);
#pragma unroll
      for (int f = 0; f < 6; ++f)
        if (f < a.nout && a.xadd[f]) o[f] += a.xadd[f][r * nx + m];
#pragma unroll
      for (int s = 0; s < 3; ++s)
        if (s < nsout) Y[s * nx + m] = mk(o[2 * s], o[2 * s + 1]);
    }
    __syncthreads();
    X = Y;
  }
  cpx<T>* Bf = (X == A) ? B : A;
  cpx<T>* Z = lds_fft<T, T, -1>(X, Bf, a.fp, nsout, 1, nx, false, a.tw, tid, nthr);
  for (int e = tid; e < nsout * a.kxp; e += nthr) {
    const int q = e / a.kxp, k = e - q * a.kxp;
    cpx<T> ra = mk(T(0), T(0)), rb = mk(T(0), T(0));
    if (k < a.kx) {
      const cpx<T> zk = Z[q * nx + k];
      const cpx<T> zm = conj(Z[q * nx + ((nx - k) % nx)]);
      const cpx<T> s = zk + zm, d = zk - zm;
      ra = mk(T(0.5) * s.x, T(0.5) * s.y);
      rb = mk(T(0.5) * d.y, T(-0.5) * d.x);
    }
    a.out[2 * q][r * a.kxp + k] = ra;
    if (2 * q + 1 < a.nout) a.out[2 * q + 1][r * a.kxp + k] = rb;
  }
}

// ---------------------------------------------------------------------------
// compact-k pointwise kernels
// ---------------------------------------------------------------------------

struct KGeom {
  int Ky, Kz, Kx, Kxp;
  const double* kx;
  const double* ky;
  const double* kz;
};

template <typename T>
struct EpiArgs {
  const cpx<T>* U[8];            // forward-transformed x-space products (unscaled)
  const cpx<T>* state[DDX_MAXF];
  const cpx<T>* forcing[DDX_MAXF];
  cpx<T>* deriv[DDX_MAXF];
  KGeom g;
  long nk;
  int physics, ndim, bdir;
  double npts;    // N = prod(xshape); the epilogue multiplies by 1 / N
  double galpha;  // g * alpha_t
  double beta;
  // fused ETD sweep (upd_kind >= 0): upd_out = ETD(upd_y, F1, F2) with
  // F1 = this right-hand side (ETD1, Euler) or upd_f1 (ETD2RK1/2) and F2 =
  // this right-hand side; deriv is written only when write_deriv is set
  int upd_kind, upd_order, write_deriv;
  double upd_dt;
  double upd_coeff[DDX_MAXF];
  const cpx<T>* upd_y[DDX_MAXF];
  const cpx<T>* upd_f1[DDX_MAXF];
  cpx<T>* upd_out[DDX_MAXF];
  const cpx<T>* tw_y;   // twiddles of length Ny (fused forward-y epilogue)
  int nt;               // nontemporal stores of deriv and upd_out
};

// store, nontemporal (streaming) when nt is set (uniform per launch)
template <typename T>
__device__ __forceinline__ void st_nt(cpx<T>* p, cpx<T> v, bool nt) {
  if (nt) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    v2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<v2*>(p));
  } else {
    *p = v;
  }
}

template <typename T>
__device__ inline void kvec(const KGeom& g, long idx, int& col, T& kx, T& ky, T& kz) {
  col = (int)(idx % g.Kxp);
  const long rs = idx / g.Kxp;
  const int row = (int)(rs % g.Kz);
  const int slab = (int)(rs / g.Kz);
  kx = T(g.kx[col]);
  ky = T(g.ky[slab]);
  kz = T(g.kz[row]);
}

template <typename T>
__device__ inline cpx<double> c64(cpx<T> v) { return mk(double(v.x), double(v.y)); }
template <typename T>
__device__ inline cpx<T> cT(cpx<double> v) { return mk(T(v.x), T(v.y)); }

// k-space epilogue of the RHS, computed in FP64 whatever the storage type
// (projection and curl in double precision).
// phi functions with the reference's branches (forward_step_cy_2d.pyx:123-148,
// forward_step_cy_3d.pyx:54-59): Taylor series to degree 14 for |Z| < 0.5
// (phi0 by Taylor in 2D, by exp in 3D), closed forms otherwise.
// 1 / j! for j = 0..14, evaluated at compile time
constexpr double inv_fact(int j) {
  double f = 1.0;
  for (int i = 2; i <= j; ++i) f *= i;
  return 1.0 / f;
}

// phi_0, phi_1, phi_2 of z.  For |z| < 0.5 the reference's Taylor series
// (the exponential's series truncated after z^14 / 14!, every phi_k from
// the same terms): phi_2 = sum_{j=2..14} z^(j-2) / j! by Horner, then
// phi_1 = 1 + z phi_2 and phi_0 = 1 + z phi_1 (identities of the truncated
// sums), or exp(z) for phi_0 when taylor0 is 0.  Otherwise the closed forms
// phi_0 = e^z, phi_1 = (phi_0 - 1) / z, phi_2 = (phi_1 - 1) / z.
__device__ inline void phi012(double z, int taylor0, double& p0, double& p1, double& p2) {
  if (z < 0.5 && z > -0.5) {
    constexpr double c[15] = {inv_fact(0),  inv_fact(1),  inv_fact(2),  inv_fact(3),
                              inv_fact(4),  inv_fact(5),  inv_fact(6),  inv_fact(7),
                              inv_fact(8),  inv_fact(9),  inv_fact(10), inv_fact(11),
                              inv_fact(12), inv_fact(13), inv_fact(14)};
    double s = c[14];
#pragma unroll
    for (int j = 13; j >= 2; --j) s = fma(s, z, c[j]);
    p2 = s;
    p1 = fma(z, p2, 1.0);
    p0 = taylor0 ? fma(z, p1, 1.0) : exp(z);
  } else {
    p0 = exp(z);
    const double r = 1.0 / z;
    p1 = (p0 - 1.0) * r;
    p2 = (p1 - 1.0) * r;
  }
}

template <typename T>
struct UpdArgs {
  const cpx<T>* y[DDX_MAXF];
  cpx<T>* out[DDX_MAXF];
  const cpx<T>* F1[DDX_MAXF];
  const cpx<T>* F2[DDX_MAXF];
  double coeff[DDX_MAXF];
  // per-mode integrating factor (compact layout, FP64), or null: coeff |k|^(2 order)
  const double* ifac;
  KGeom g;
  long nk;
  int ncomp, kind, order, ndim;
  double dt;
};

enum { UPD_EULER = 0, UPD_ETD1 = 1, UPD_ETD2RK1 = 2, UPD_ETD2RK2 = 3 };

// One ETD / Euler update of a mode (forward_step_cy_2d.pyx:21-120,
// forward_step_cy_3d.pyx:17-124), Z = -coeff |k|^(2 order) dt with its phi values
__device__ inline cpx<double> etd_combine(int kind, cpx<double> y, cpx<double> f1, cpx<double> f2,
                                          double Z, double p0, double p1, double p2, double dt) {
  if (Z == 0.0) {
    if (kind == UPD_ETD2RK1) return y + scale(f2 - f1, dt / 2.0);
    if (kind == UPD_ETD2RK2) return y + scale(f2, dt);
    return y + scale(f1, dt);
  }
  if (kind == UPD_ETD1) return scale(y, p0) + scale(f1, p1 * dt);
  if (kind == UPD_ETD2RK1) return y + scale(f2 - f1, p2 * dt);
  return scale(y, p0) + scale(f2 - f1, 2.0 * p2 * dt) + scale(f1, p1 * dt);
}

// One thread per compact mode, all components: the phi functions depend on
// the mode and the component's coefficient only, so components sharing a
// coefficient (all velocity components; all field components) share them.
template <typename T>
__global__ void update_kernel(const UpdArgs<T> a) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.nk) return;
  int col;
  double kx, ky, kz;
  kvec<double>(a.g, idx, col, kx, ky, kz);
  if (col >= a.g.Kx) {
    for (int comp = 0; comp < a.ncomp; ++comp) a.out[comp][idx] = mk(T(0), T(0));
    return;
  }
  const double dt = a.dt;
  const double k2 = kx * kx + ky * ky + kz * kz;
  double kp;
  if (a.order == 1) kp = k2;
  else if (a.order == 2) kp = k2 * k2;
  else kp = pow(k2, (double)a.order);
  double cprev = 0.0, Z = 0.0, p0 = 1.0, p1 = 1.0, p2 = 0.5;
  bool have = false;
  if (a.ifac) {
    // an integrating-factor array: the reference passes -IF to the sweep,
    // which forms Z = (-IF) dt (forward_step_cy_3d.pyx:54)
    Z = a.kind != UPD_EULER ? (-a.ifac[idx]) * dt : 0.0;
    if (Z != 0.0) phi012(Z, a.ndim == 2, p0, p1, p2);
    have = true;
  }
  for (int comp = 0; comp < a.ncomp; ++comp) {
    const cpx<double> y = c64(a.y[comp][idx]);
    const cpx<double> f1 = c64(a.F1[comp][idx]);
    const cpx<double> f2 = (a.kind >= UPD_ETD2RK1) ? c64(a.F2[comp][idx]) : mk(0.0, 0.0);
    const double c = a.coeff[comp];
    if (!a.ifac && (!have || c != cprev)) {
      Z = (c != 0.0 && a.kind != UPD_EULER) ? (-(c * kp)) * dt : 0.0;
      if (Z != 0.0) phi012(Z, a.ndim == 2, p0, p1, p2);
      cprev = c;
      have = true;
    }
    a.out[comp][idx] = cT<T>(etd_combine(a.kind, y, f1, f2, Z, p0, p1, p2, dt));
  }
}

// Spherical 2/3 rule (representations.py:410-416): zero, in nf compact
// arrays, the modes with sqrt(|k|^2) >= thr, |k|^2 summed in the reference's
// axis order ((kx^2 + ky^2) in 2D, (ky^2 + kz^2) + kx^2 in 3D) so the
// comparison matches its rounding.  The kept cube of the compact layout
// holds the sphere; pad columns stay as they are (zero).
template <typename T>
struct MaskArgs {
  cpx<T>* f[DDX_MAXF];
  KGeom g;
  long nk;
  int nf, ndim;
  double thr;
};

template <typename T>
__global__ void sphere_mask_kernel(const MaskArgs<T> a) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.nk) return;
  int col;
  double kx, ky, kz;
  kvec<double>(a.g, idx, col, kx, ky, kz);
  if (col >= a.g.Kx) return;
  const double k2 = a.ndim == 2 ? kx * kx + ky * ky : (ky * ky + kz * kz) + kx * kx;
  if (sqrt(k2) < a.thr) return;
  for (int f = 0; f < a.nf; ++f) a.f[f][idx] = mk(T(0), T(0));
}

// Per-mode pieces of the k-space epilogue, shared by the epilogue kernel and
// the fused forward-y epilogue of small 2D grids.  The ETD sweep of
// components v0 <= v < v1 of one mode (update_kernel's arithmetic): phi is
// formed once per distinct coefficient.  The caller has read every state
// entry of this mode that it still needs, so in-place outputs are safe.
template <typename T>
__device__ __forceinline__ void epi_etd(const EpiArgs<T>& a, long idx, double k2true, int v0, int v1,
                                        const cpx<double> (&D)[6], const cpx<T>* yv = nullptr,
                                        const cpx<T>* fv = nullptr) {
  const int nd = a.ndim;
  double kp;
  if (a.upd_order == 1) kp = k2true;
  else if (a.upd_order == 2) kp = k2true * k2true;
  else kp = pow(k2true, (double)a.upd_order);
  const double dt = a.upd_dt;
  double cprev = 0.0, Z = 0.0, p0 = 1.0, p1 = 1.0, p2 = 0.5;
  bool have = false;
#pragma unroll
  for (int v = 0; v < 6; ++v) {
    if (v < v0 || v >= v1) continue;
    const double c = a.upd_coeff[v];
    if (!have || c != cprev) {
      Z = (c != 0.0 && a.upd_kind != UPD_EULER) ? (-(c * kp)) * dt : 0.0;
      if (Z != 0.0) phi012(Z, nd == 2, p0, p1, p2);
      cprev = c;
      have = true;
    }
    const cpx<double> dv = c64(cT<T>(D[v]));  // the value an unfused sweep would read back
    // yv / fv: the mode's y and F1, read by the caller before any write
    const cpx<double> y = c64(yv ? yv[v] : a.upd_y[v][idx]);
    const cpx<double> f1 = (a.upd_kind >= UPD_ETD2RK1) ? c64(fv ? fv[v] : a.upd_f1[v][idx]) : dv;
    st_nt(&a.upd_out[v][idx], cT<T>(etd_combine(a.upd_kind, y, f1, dv, Z, p0, p1, p2, dt)), a.nt);
  }
}

// All components of one mode: F = NL / N + forcing (+ g alpha T along the
// buoyancy axis), projected (F <- F - k (k.F) / |k|^2, |k|^2 = 0 replaced by
// 1); Boussinesq T; MHD induction i k x E.  n0..n2: the unscaled forward
// products of the velocity; e0..e2: those of T (e0), of E (3D) or of the
// scalar E (2D, e0); tval, bval: the state's T and its velocity along the
// buoyancy axis at this mode (Boussinesq), read by the caller before any
// write of the mode.
template <typename T>
__device__ __forceinline__ void epi_mode(const EpiArgs<T>& a, long idx, double kx, double ky,
                                         double kz, cpx<T> n0, cpx<T> n1, cpx<T> n2, cpx<T> e0_,
                                         cpx<T> e1_, cpx<T> e2_, cpx<T> tval, cpx<T> bval,
                                         const cpx<T>* yv = nullptr, const cpx<T>* fv = nullptr) {
  const int nd = a.ndim;
  const int nvar = nd + (a.physics == PH_BOUSS ? 1 : (a.physics == PH_MHD ? nd : 0));
  const bool upd = a.upd_kind >= 0;
  const bool wd = !upd || a.write_deriv;
  const double s = 1.0 / a.npts;
  const double kk[3] = {kx, ky, kz};
  const cpx<T> nl[3] = {n0, n1, n2};
  cpx<double> F[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    F[i] = mk(0.0, 0.0);
    if (i < nd) {
      F[i] = scale(c64(nl[i]), s);
      if (a.forcing[i]) F[i] = F[i] + c64(a.forcing[i][idx]);
      if (a.physics == PH_BOUSS && i == a.bdir) F[i] = F[i] + scale(c64(tval), a.galpha);
    }
  }
  // projection F <- F - k (k.F) / |k|^2, |k|^2 = 0 replaced by 1
  double k2 = kx * kx + ky * ky + kz * kz;
  const double k2true = k2;
  if (k2 == 0.0) k2 = 1.0;
  cpx<double> kf = mk(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < 3; ++i) kf = kf + scale(F[i], kk[i]);
  cpx<double> D[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) D[i] = mk(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (i < nd) D[i] = F[i] - scale(kf, kk[i] / k2);
  if (a.physics == PH_BOUSS) {
    const cpx<T>* fT = nd == 2 ? a.forcing[2] : a.forcing[3];
    cpx<double> ft = scale(c64(e0_), s) - scale(c64(bval), a.beta);
    if (fT) ft = ft + c64(fT[idx]);
    if (nd == 2) D[2] = ft;
    else D[3] = ft;
  } else if (a.physics == PH_MHD) {
    if (nd == 3) {
      const cpx<double> e0 = scale(c64(e0_), s), e1 = scale(c64(e1_), s), e2 = scale(c64(e2_), s);
      // i k x E
      D[3] = mul_is(e2, ky) - mul_is(e1, kz);
      D[4] = mul_is(e0, kz) - mul_is(e2, kx);
      D[5] = mul_is(e1, kx) - mul_is(e0, ky);
    } else {
      const cpx<double> e = scale(c64(e0_), s);
      D[2] = mul_is(e, ky);
      D[3] = mk(0.0, 0.0) - mul_is(e, kx);
    }
  }
  if (wd) {
#pragma unroll
    for (int v = 0; v < 6; ++v)
      if (v < nvar) st_nt(&a.deriv[v][idx], cT<T>(D[v]), a.nt);
  }
  if (upd) epi_etd(a, idx, k2true, 0, nvar, D, yv, fv);
}

// Zeros for the pad columns kx >= Kx of one mode (all components).
template <typename T>
__device__ __forceinline__ void epi_pad(const EpiArgs<T>& a, long idx, int nvar) {
  const bool upd = a.upd_kind >= 0;
  const bool wd = !upd || a.write_deriv;
#pragma unroll
  for (int v = 0; v < 6; ++v) {
    if (v >= nvar) break;
    if (wd) a.deriv[v][idx] = mk(T(0), T(0));
    if (upd) a.upd_out[v][idx] = mk(T(0), T(0));
  }
}

// k-space epilogue of the fused right-hand side: 1/N, forcing, buoyancy,
// projection, induction curl; optionally the ETD sweep that follows it.
// PRE (3D MHD with a fused ETD sweep): the sweep's y and F1 are read with
// the products, so that every load of a mode is issued before its first
// store; otherwise the compiler, which cannot rule out that the deriv and
// out arrays alias y and F1, issues them after the deriv stores and between
// the per-component out stores.
template <typename T, bool PRE>
__global__ void epilogue_kernel(const EpiArgs<T> a) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.nk) return;
  int col;
  double kx, ky, kz;
  kvec<double>(a.g, idx, col, kx, ky, kz);
  const int nd = a.ndim;
  const int nvar = nd + (a.physics == PH_BOUSS ? 1 : (a.physics == PH_MHD ? nd : 0));
  if (col >= a.g.Kx) {
    epi_pad(a, idx, nvar);
    return;
  }
  const cpx<T> z = mk(T(0), T(0));
  // arrays picked by ndim / direction without dynamic indexing (keeps the
  // argument arrays and locals out of scratch memory)
  cpx<T> tval = z, bval = z;
  if (a.physics == PH_BOUSS) {
    tval = (nd == 2 ? a.state[2] : a.state[3])[idx];
    bval = (a.bdir == 0 ? a.state[0] : (a.bdir == 1 ? a.state[1] : a.state[2]))[idx];
  }
  const cpx<T> n0 = a.U[0][idx], n1 = a.U[1][idx], n2 = nd == 3 ? a.U[2][idx] : z;
  cpx<T> e0 = z, e1 = z, e2 = z;
  if (nvar > nd) {
    e0 = (nd == 2 ? a.U[2] : a.U[3])[idx];
    if (a.physics == PH_MHD && nd == 3) {
      e1 = a.U[4][idx];
      e2 = a.U[5][idx];
    }
  }
  if (PRE) {
    // F1 exists for the ETD2RK sweeps only; the other sweeps read y twice
    // rather than branch around a load
    const bool hasf = a.upd_kind >= UPD_ETD2RK1;
    cpx<T> yv[6], fv[6];
#pragma unroll
    for (int v = 0; v < 6; ++v) {
      yv[v] = a.upd_y[v][idx];
      fv[v] = (hasf ? a.upd_f1[v] : a.upd_y[v])[idx];
    }
    epi_mode(a, idx, kx, ky, kz, n0, n1, n2, e0, e1, e2, tval, bval, yv, fv);
  } else {
    epi_mode(a, idx, kx, ky, kz, n0, n1, n2, e0, e1, e2, tval, bval);
  }
}

// layout conversion: kref (reference transposed half spectrum) <-> compact
struct LayoutArgs {
  int ndim;
  int nx, ny, nz;       // x-space extents
  int Kx, Kxp, Ky, Kz;
  int cy, cz;
  long nk;              // compact elements
  long nref;            // reference elements
};

// 2D: fz is the member of a batched plan (0 otherwise); members are stacked
// (kx, ky) arrays
__device__ inline long ref_index(const LayoutArgs& L, int m, int fy, int fz) {
  const int hx = L.nx / 2 + 1;
  if (L.ndim == 2) return ((long)fz * hx + m) * L.ny + fy;
  return ((long)fy * L.nz + fz) * hx + m;
}

template <typename T>
__global__ void pack_kernel(const LayoutArgs L, const cpx<double>* ref, cpx<T>* kc) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= L.nk) return;
  const int col = (int)(idx % L.Kxp);
  const long rs = idx / L.Kxp;
  const int row = (int)(rs % L.Kz);
  const int slab = (int)(rs / L.Kz);
  if (col >= L.Kx) { kc[idx] = mk(T(0), T(0)); return; }
  const int fy = slab <= L.cy ? slab : slab + (L.ny - L.Ky);
  const int fz = row <= L.cz ? row : row + (L.nz - L.Kz);
  const cpx<double> v = ref[ref_index(L, col, fy, fz)];
  kc[idx] = mk(T(v.x), T(v.y));
}

template <typename T>
__global__ void unpack_kernel(const LayoutArgs L, const cpx<T>* kc, cpx<double>* ref) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= L.nref) return;
  const int hx = L.nx / 2 + 1;
  int m, fy, fz;
  if (L.ndim == 2) {
    const long per = (long)hx * L.ny;
    fz = (int)(idx / per);
    const long r = idx - (long)fz * per;
    m = (int)(r / L.ny);
    fy = (int)(r % L.ny);
  }
  else { m = (int)(idx % hx); long t = idx / hx; fz = (int)(t % L.nz); fy = (int)(t / L.nz); }
  cpx<double> v = mk(0.0, 0.0);
  int row = -1, slab = -1;
  if (fy <= L.cy) slab = fy; else if (fy >= L.ny - L.cy) slab = fy - (L.ny - L.Ky);
  if (fz <= L.cz) row = fz; else if (fz >= L.nz - L.cz) row = fz - (L.nz - L.Kz);
  if (m < L.Kx && slab >= 0 && row >= 0) {
    const cpx<T> u = kc[((long)slab * L.Kz + row) * L.Kxp + m];
    v = mk(double(u.x), double(u.y));
  }
  ref[idx] = v;
}

template <typename T>
__global__ void deriv_kernel(KGeom g, long nk, int dim, const cpx<T>* in, cpx<T>* out) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nk) return;
  int col;
  T kx, ky, kz;
  kvec<T>(g, idx, col, kx, ky, kz);
  if (col >= g.Kx) { out[idx] = mk(T(0), T(0)); return; }
  const T k = dim == 0 ? kx : (dim == 1 ? ky : kz);
  out[idx] = mul_is(in[idx], k);
}

template <typename T>
struct ProjArgs {
  cpx<T>* c[3];
  KGeom g;
  long nk;
  int ndim;
};

template <typename T>
__global__ void project_kernel(const ProjArgs<T> a) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.nk) return;
  int col;
  T kx, ky, kz;
  kvec<T>(a.g, idx, col, kx, ky, kz);
  if (col >= a.g.Kx) return;
  const double kk[3] = {double(kx), double(ky), double(kz)};
  double k2 = kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2];
  if (k2 == 0.0) k2 = 1.0;
  cpx<double> kf = mk(0.0, 0.0);
  for (int i = 0; i < a.ndim; ++i) kf = kf + scale(c64(a.c[i][idx]), kk[i]);
  for (int i = 0; i < a.ndim; ++i) a.c[i][idx] = cT<T>(c64(a.c[i][idx]) - scale(kf, kk[i] / k2));
}

// out (+)= sum_t coef_t * op_t(k) * src_t on compact k arrays.
// op: -1 identity, 0/1/2 multiply by i k_x / i k_y / i k_z, 3 divide by |k|^2
// with |k|^2 = 0 replaced by 1.
#define DDX_MAXT 4
template <typename T>
struct KCombArgs {
  const cpx<T>* src[DDX_MAXT];
  double coef[DDX_MAXT];
  int op[DDX_MAXT];
  int nterm;
  int accumulate;
  cpx<T>* out;
  KGeom g;
  long nk;
};

template <typename T>
__global__ void kcombine_kernel(const KCombArgs<T> a) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.nk) return;
  int col;
  T kx, ky, kz;
  kvec<T>(a.g, idx, col, kx, ky, kz);
  if (col >= a.g.Kx) { a.out[idx] = mk(T(0), T(0)); return; }
  cpx<T> acc = a.accumulate ? a.out[idx] : mk(T(0), T(0));
#pragma unroll
  for (int t = 0; t < DDX_MAXT; ++t) {
    if (t < a.nterm) {
      cpx<T> v = a.src[t][idx];
      const int op = a.op[t];
      if (op == 0) v = mul_is(v, kx);
      else if (op == 1) v = mul_is(v, ky);
      else if (op == 2) v = mul_is(v, kz);
      else if (op == 3) {
        T k2 = kx * kx + ky * ky + kz * kz;
        if (k2 == T(0)) k2 = T(1);
        v = mk(v.x / k2, v.y / k2);
      }
      acc = acc + scale(v, T(a.coef[t]));
    }
  }
  a.out[idx] = acc;
}

// out (+)= sum_t coef_t * a_t * b_t on x-space arrays (b_t == NULL: factor 1)
template <typename T>
struct XCombArgs {
  const T* a[DDX_MAXT];
  const T* b[DDX_MAXT];
  double coef[DDX_MAXT];
  int nterm;
  int accumulate;
  T* out;
  long n;
};

template <typename T>
__global__ void xcombine_kernel(const XCombArgs<T> a) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.n) return;
  T acc = a.accumulate ? a.out[idx] : T(0);
#pragma unroll
  for (int t = 0; t < DDX_MAXT; ++t) {
    if (t < a.nterm) {
      T v = a.a[t][idx];
      if (a.b[t]) v *= a.b[t][idx];
      acc += T(a.coef[t]) * v;
    }
  }
  a.out[idx] = acc;
}

// Block partial reductions; the host finishes the sum / max over blocks.
// op 0: max of a^2, 1: sum of a, 2: sum of a^2 (x space, real)
template <typename T>
__global__ void xreduce_kernel(const T* a, long n, int op, double* partial) {
  __shared__ double sh[256];
  double acc = (op == 0) ? 0.0 : 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double v = (double)a[i];
    if (op == 0) acc = fmax(acc, v * v);
    else if (op == 1) acc += v;
    else acc += v * v;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      sh[threadIdx.x] = (op == 0) ? fmax(sh[threadIdx.x], sh[threadIdx.x + s]) : sh[threadIdx.x] + sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// Hermitian-weighted sums over compact k arrays (weight 1 on kx = 0, 2 elsewhere),
// i.e. x-space means by Parseval.  op 0: sum |src|^2; 1: sum |i k x src|^2 summed
// over the 2D curl of a (src0, src1) pair (enstrophy-like); 2: Re(src) at weight.
// weighted contribution of compact mode i to a k-space reduction
template <typename T>
__device__ inline double kred_term(const cpx<T>* s0, const cpx<T>* s1, const KGeom& g, long i, int op) {
  int col;
  T kx, ky, kz;
  kvec<T>(g, i, col, kx, ky, kz);
  if (col >= g.Kx) return 0.0;
  const double w = col == 0 ? 1.0 : 2.0;
  if (op == 0) {
    const cpx<T> v = s0[i];
    return w * ((double)v.x * v.x + (double)v.y * v.y);
  }
  if (op == 1) {
    // w_z = i kx s1 - i ky s0
    const cpx<T> v = mul_is(s1[i], kx) - mul_is(s0[i], ky);
    return w * ((double)v.x * v.x + (double)v.y * v.y);
  }
  return w * (double)s0[i].x;
}

template <typename T>
__global__ void kreduce_kernel(const cpx<T>* s0, const cpx<T>* s1, KGeom g, long nk, int op,
                               double* partial) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nk; i += (long)gridDim.x * blockDim.x)
    acc += kred_term(s0, s1, g, i, op);
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// The same reduction per member of a batched plan: block b sums the modes
// (ky, b, kx) of member b (the member is the z slab of the compact layout)
// in a fixed order and writes out[b].
template <typename T>
__global__ void __launch_bounds__(256) kreduce_members_kernel(const cpx<T>* s0, const cpx<T>* s1,
                                                              KGeom g, int op, double* out) {
  __shared__ double sh[256];
  const int b = blockIdx.x;
  const long per = (long)g.Ky * g.Kxp;
  double acc = 0.0;
  for (long j = threadIdx.x; j < per; j += 256) {
    const long ky = j / g.Kxp, c = j - ky * g.Kxp;
    acc += kred_term(s0, s1, g, (ky * g.Kz + b) * g.Kxp + c, op);
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b] = sh[0];
}

// real array conversion (x-space staging between the FP64 host boundary and FP32 plans)
template <typename TI, typename TO>
__global__ void convert_kernel(const TI* in, TO* out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = TO(in[i]);
}

__global__ void twiddle_kernel(cpx<double>* tw, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double s, c;
  sincospi(2.0 * (double)j / (double)n, &s, &c);
  tw[j] = mk(c, -s);
}

__global__ void twiddle_f32_kernel(const cpx<double>* tw64, cpx<float>* tw32, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  tw32[j] = mk(float(tw64[j].x), float(tw64[j].y));
}

__global__ void zero_kernel(uint32_t* p, size_t nwords) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < nwords; i += stride) p[i] = 0u;
}

__global__ void copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

__global__ void copy1_kernel(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void zero_bytes_kernel(uint8_t* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace ddx
