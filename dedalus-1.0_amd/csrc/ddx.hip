// libddx: C ABI of the MI355X pseudospectral step (declarations in include/ddx.h).
//
// Host orchestration of the kernels in ddx_kernels.h.  One plan per grid;
// every launch goes to the plan's stream; argument and shape checks happen on
// the host before any launch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/ddx.h"
#include "ddx_kernels.h"
#ifndef DDX_COL_XCD_GROUP
#define DDX_COL_XCD_GROUP 1
#endif
#ifndef DDX_COL_NT
#define DDX_COL_NT 1
#endif
#ifndef DDX_ROW_NT
#define DDX_ROW_NT 1
#endif
#ifndef DDX_EPI_NT
#define DDX_EPI_NT 1
#endif
#ifndef DDX_EPI_PRELOAD
#define DDX_EPI_PRELOAD 1
#endif
#ifndef DDX_YSPEC_GROUP
#define DDX_YSPEC_GROUP 1
#endif
#ifndef DDX_YSPEC_WIDE
#define DDX_YSPEC_WIDE 1
#endif
#include "ddx_fast.h"

using namespace ddx;

// Device scratch of ddx_step_noalloc: one copy per GPU, resident from the
// library's first use on that GPU until the process ends.  Host code never
// passes `scratch` itself to a kernel (on the host it names a stand-in, not
// a device address); it uses the address hipGetSymbolAddress returns for
// each device (scratch_base below).
#define DDX_SCRATCH_BYTES (64u << 20)
static __device__ __align__(256) unsigned char scratch[DDX_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;
int g_device = -1;

// per-device address of `scratch`, resolved by ddx_init with the device current
constexpr int kMaxDevices = 64;
constexpr size_t kScratchBytes = DDX_SCRATCH_BYTES;
void* g_scratch[kMaxDevices] = {};

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(DDX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

#define LAUNCHCHK()                                                                        \
  do {                                                                                     \
    hipError_t e_ = hipGetLastError();                                                     \
    if (e_ != hipSuccess)                                                                  \
      return fail(DDX_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_));    \
  } while (0)

// Per-axis wavenumbers, computed exactly as the reference's numpy does
// (representations.py:204-226): k = fftfreq(n)[m] * 2 * kny with
// kny = pi n / L and the Nyquist entry (even n) sign-flipped to +kny; the
// 2/3 rule keeps |k| < (2/3) kny (representations.py:384-397).
struct AxisInfo {
  int n = 1;
  double L = 1.0;
  int c = 0;  // largest kept |frequency index|
  int K = 1;  // kept count: 2c+1 on full axes, c+1 on the x axis
  std::vector<double> kfull;  // k for frequency index m in [0, n)
  // k(m) = ((ms * inv) * 2) * kny with ms the signed index: the factors, for
  // kernels that form k from the index instead of reading the table
  double inv = 1.0, kny = 0.0;
};

// Kept set of one axis: |k| < thr (the 2/3 rules, thr = (2/3) kny or, for
// the spherical rule, (2/3) min(kny) of all axes), or with nyq_only every
// mode but the Nyquist one (|k| == kny, zero_nyquist).
int setup_axis(int n, double L, bool half, AxisInfo* ax, std::string* why, double thr_sph = 0.0,
               bool nyq_only = false) {
  ax->n = n;
  ax->L = L;
  const double kny = (M_PI * (double)n) / L;
  const double inv = 1.0 / (double)n;
  const double thr = thr_sph > 0.0 ? thr_sph : 2.0 / 3.0 * kny;
  ax->inv = inv;
  ax->kny = kny;
  ax->kfull.resize(n);
  const int npos = (n - 1) / 2 + 1;  // numpy fftfreq split
  std::vector<int> keep(n, 0);
  for (int m = 0; m < n; ++m) {
    const int ms = (m < npos) ? m : m - n;
    double k = ((double)ms * inv) * 2.0 * kny;
    if (n % 2 == 0 && m == n / 2) k = -k;
    ax->kfull[m] = k;
    keep[m] = nyq_only ? !(std::fabs(k) == kny) : std::fabs(k) < thr;
  }
  int c = -1;
  while (c + 1 <= n / 2 && keep[c + 1]) ++c;
  if (c < 0) {
    *why = "2/3 rule removes the mean mode";
    return -1;
  }
  // check the kept set is {0..c} (x axis: within the half spectrum) or {0..c} U {n-c..n-1}
  const int hmax = half ? n / 2 : n - 1;
  for (int m = 0; m <= hmax; ++m) {
    bool exp_keep = (m <= c) || (!half && m >= n - c);
    if (!!keep[m] != exp_keep) {
      *why = "2/3 kept set is not symmetric";
      return -1;
    }
  }
  ax->c = c;
  ax->K = half ? c + 1 : 2 * c + 1;
  return 0;
}

}  // namespace

struct ddx_plan {
  int device = 0;
  int ndim = 2;
  int precision = DDX_F64;
  AxisInfo ax, ay, az;  // az: n = 1 in 2D
  int Kx = 1, Kxp = 1, Ky = 1, Kz = 1;
  long nk = 0;          // Ky*Kz*Kxp
  long npts = 0;        // nx*ny*nz
  long nref = 0;        // reference k elements
  FftPlan fx, fy, fz;
  double* dkx = nullptr;
  double* dky = nullptr;
  double* dkz = nullptr;
  std::map<int, void*> tw;  // n -> twiddle table (precision of the plan)
  std::vector<void*> owned;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  void* wsA = nullptr;
  void* wsB = nullptr;
  size_t ws_bytes = 0;
  // bumped whenever the workspace is reallocated (part of the step graph's key)
  unsigned long ws_gen = 0;
  // second-stage derivative of ddx_step when the caller passes no deriv2 and
  // the ETD sweeps run as separate launches (DDX_NO_FUSED_UPDATE=1)
  void* f2buf = nullptr;
  size_t f2_bytes = 0;
  // kernel instrumentation (HIP events around each launch when enabled)
  double* partial = nullptr;  // reduction partials (device)
  std::vector<double> hpartial;
  bool generic_only = false;  // DDX_FORCE_GENERIC=1: size-generic kernels everywhere
  // one RK step captured as a HIP graph and replayed (DDX_NO_GRAPH=1 disables)
  bool use_graph = true;
  // z transforms run (3D); a batched 2D plan keeps members on the z axis
  // untransformed: compact (Ky, nbatch, Kxp), x space (nbatch, Ny, Nx)
  bool zfft = false;
  // ddx_step applies each ETD sweep inside the preceding RHS epilogue
  // (DDX_NO_FUSED_UPDATE=1 runs them as separate launches)
  bool fuse_update = true;
  // single 2D grids with 64-point y axes run the forward y pass and the
  // k-space epilogue as one kernel (DDX_NO_FUSED_YEPI=1 runs them as two)
  bool fuse_yepi = true;
  // FP64 3D right-hand sides transform along z only the arrays whose
  // derivatives need k_z, and form the other derived fields in the y pass
  // (DDX_NO_ZSPLIT=1 forms every derived field before the z pass)
  bool zsplit = true;
  // ddx_plan_set_tuning choices, -1: the built-in rule
  int tune_wide = -1, tune_xcd = -1, tune_yepi = -1, tune_radix = -1;
  // dealiasing rule (DDX_DEALIAS_*); the spherical rule's radius (2/3) min(kny)
  int dealias = DDX_DEALIAS_23;
  double sph_thr = 0.0;
  int nbatch = 1;
  long nper = 0;  // points of one transform (the 1/N of the forward transform)
  hipGraphExec_t step_exec = nullptr;
  std::vector<unsigned char> step_key;
  // key of the last single step that ran eagerly because it missed the
  // graph's key: a second call with the same key captures a new graph, so a
  // dt that changes every step (CFL) never pays for a capture
  std::vector<unsigned char> miss_key;
  bool prof = false;
  struct Rec { int cat; hipEvent_t a, b; double bytes; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> evfree;
};

namespace {

size_t cpx_bytes(const ddx_plan* p) { return p->precision == DDX_F64 ? 16 : 8; }
size_t real_bytes(const ddx_plan* p) { return p->precision == DDX_F64 ? 8 : 4; }

// kernel classes reported by ddx_prof_read
enum ProfCat {
  PC_COL_BWD_Z_SPECTRAL, PC_COL_BWD_Y_SPECTRAL, PC_COL_BWD_Z, PC_COL_BWD_Y,
  PC_COL_FWD_Y, PC_COL_FWD_Z, PC_ROW_PHYS, PC_ROW_C2R, PC_ROW_R2C,
  PC_EPILOGUE, PC_UPDATE, PC_LAYOUT, PC_POINTWISE, PC_COPY, PC_YEPI, PC_NCAT
};
const char* const kCatNames[PC_NCAT] = {
  "col_bwd_z_spectral", "col_bwd_y_spectral", "col_bwd_z", "col_bwd_y",
  "col_fwd_y", "col_fwd_z", "row_phys", "row_c2r", "row_r2c",
  "epilogue", "update", "layout", "pointwise", "copy", "col_fwd_y_epilogue"};

hipEvent_t take_event(ddx_plan* p) {
  if (!p->evfree.empty()) {
    hipEvent_t e = p->evfree.back();
    p->evfree.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Records a start event on construction and a stop event on destruction
// (around one kernel launch) when instrumentation is on.
struct ProfScope {
  ddx_plan* p;
  int cat;
  double bytes;
  hipEvent_t a = nullptr;
  ProfScope(ddx_plan* p_, int cat_, double bytes_) : p(p_), cat(cat_), bytes(bytes_) {
    if (p->prof) {
      a = take_event(p);
      if (a && hipEventRecord(a, p->stream) != hipSuccess) a = nullptr;
    }
  }
  ~ProfScope() {
    if (!a) return;
    hipEvent_t b = take_event(p);
    if (b && hipEventRecord(b, p->stream) == hipSuccess) p->recs.push_back({cat, a, b, bytes});
  }
};

KGeom kgeom(const ddx_plan* p) {
  KGeom g;
  g.Ky = p->Ky;
  g.Kz = p->Kz;
  g.Kx = p->Kx;
  g.Kxp = p->Kxp;
  g.kx = p->dkx;
  g.ky = p->dky;
  g.kz = p->dkz;
  return g;
}

int ensure_ws(ddx_plan* p, int nfields) {
  // both buffers hold nfields arrays of the largest mixed layout
  const long ny = p->ay.n, nz = p->az.n;
  long per = std::max((long)p->Ky * nz, nz * ny) * p->Kxp;
  per = std::max(per, (long)p->nk);
  const size_t need = (size_t)nfields * per * cpx_bytes(p);
  if (need <= p->ws_bytes) return 0;
  // the captured step graph launches kernels on the old buffers: drain the
  // stream and drop it before they are freed (a new block may start at the
  // old address, so the generation counter, not the address, keys the graph)
  HIPCHK(hipStreamSynchronize(p->stream));
  if (p->step_exec) {
    HIPCHK(hipGraphExecDestroy(p->step_exec));
    p->step_exec = nullptr;
    p->step_key.clear();
  }
  ++p->ws_gen;
  if (p->wsA) HIPCHK(hipFree(p->wsA));
  if (p->wsB) HIPCHK(hipFree(p->wsB));
  p->wsA = p->wsB = nullptr;
  p->ws_bytes = 0;
  if (hipMalloc(&p->wsA, need) != hipSuccess || hipMalloc(&p->wsB, need) != hipSuccess)
    return fail(DDX_ERR_NOMEM, "workspace allocation failed");
  p->ws_bytes = need;
  return 0;
}

// ns arrays of nk complex for the second-stage derivative of ddx_step when
// the caller passes no deriv2 buffers and the sweeps run as separate launches
int ensure_f2(ddx_plan* p, int ns) {
  const size_t need = (size_t)ns * p->nk * cpx_bytes(p);
  if (need <= p->f2_bytes) return 0;
  HIPCHK(hipStreamSynchronize(p->stream));
  if (p->step_exec) {
    HIPCHK(hipGraphExecDestroy(p->step_exec));
    p->step_exec = nullptr;
    p->step_key.clear();
  }
  ++p->ws_gen;
  if (p->f2buf) HIPCHK(hipFree(p->f2buf));
  p->f2buf = nullptr;
  p->f2_bytes = 0;
  if (hipMalloc(&p->f2buf, need) != hipSuccess) return fail(DDX_ERR_NOMEM, "deriv2 scratch allocation failed");
  p->f2_bytes = need;
  return 0;
}

// Device address of the static scratch on `device` (set up by ddx_init).
int scratch_base(int device, void** out) {
  if (device < 0 || device >= kMaxDevices || !g_scratch[device])
    return fail(DDX_ERR_NODEV, "ddx_init has not been called for this plan's device");
  *out = g_scratch[device];
  return 0;
}

size_t scratch_round(size_t n) { return (n + 255) / 256 * 256; }

// Device scratch of one right-hand side (two buffers of the mixed layouts)
// and, for ddx_step without deriv2 buffers and unfused sweeps, the second
// stage's derivative.
struct Ws {
  void* A = nullptr;
  void* B = nullptr;
  void* F2 = nullptr;
};

long ws_field_stride(const ddx_plan* p) {
  const long ny = p->ay.n, nz = p->az.n;
  long per = std::max((long)p->Ky * nz, nz * ny) * p->Kxp;
  return std::max(per, (long)p->nk);
}

int tile_width(int n, size_t cb) {
  // two ping-pong tiles of n x W complex within 64 KiB of LDS
  int W = 8;
  while (W > 1 && (size_t)2 * n * W * cb > 65536) W /= 2;
  return W;
}

// ---------------------------------------------------------------------------
// column passes
// ---------------------------------------------------------------------------

enum ColKind { COL_PLAIN, COL_SPECTRAL };

template <typename T, int N, bool WIDE, int V>
void col_fast_go(ddx_plan* p, const ColArgs<T>& a, dim3 grid, int dir, bool spectral) {
  const dim3 block(ColPlanV<N, V>::LANES * col_width<N, WIDE, V>());
  // spectral prologues belong to backward passes only (launch_col_fast)
  if (spectral) hipLaunchKernelGGL((col_fft_fast<T, 1, N, true, WIDE, V>), grid, block, 0, p->stream, a);
  else if (dir > 0) hipLaunchKernelGGL((col_fft_fast<T, 1, N, false, WIDE, V>), grid, block, 0, p->stream, a);
  else hipLaunchKernelGGL((col_fft_fast<T, -1, N, false, WIDE, V>), grid, block, 0, p->stream, a);
}

// radix variant of the size-specialised passes (ddx_plan_set_tuning)
int radix_variant(const ddx_plan* p) { return p->tune_radix == 1 ? 1 : 0; }

// Size-specialised column kernel for n in {64, 128, 256, 512}
template <typename T, int N>
void col_fast_n(ddx_plan* p, ColArgs<T>& a, int nslabs, int nfields, int dir, bool spectral) {
  // double-width tiles at n = 256: FP32, the plain FP64 passes, and the
  // spectral passes whose caller asks for them
  const bool wide = N == 256 && (p->tune_wide >= 0 ? p->tune_wide == 1
                                                   : (sizeof(T) == 4 || !spectral || a.wide));
  const int V = radix_variant(p);
  const int W = V ? (wide ? col_width<N, true, 1>() : col_width<N, false, 1>())
                  : (wide ? col_width<N, true>() : col_width<N, false>());
  const int ntiles = (a.ncols + W - 1) / W;
  dim3 grid(ntiles, nslabs, nfields);
  a.nslabs = nslabs;
  a.nfields = nfields;
  a.ntiles = ntiles;
  // The 3D spectral z pass reads each source for up to three fields: group
  // them per XCD (11 % faster at 256^3; no gain measured on the 2D y pass).
  // The 3D spectral y pass (group_hint) walks whole slabs per XCD, so that
  // its lanes past the last column find the previous tile in L2 too (67 MB
  // less fetch per launch at 256^3); on the z pass that order was 5 % slower.
  a.xcd_group = (spectral && nfields > 1 && (a.row_dim == 2 || a.group_hint) && DDX_COL_XCD_GROUP)
                    ? (a.group_hint ? 2 : 1) : 0;
  if (spectral && p->tune_xcd >= 0) a.xcd_group = p->tune_xcd;
  if (a.xcd_group == 2) {
    grid = dim3((unsigned)(((nslabs + 7) / 8) * 8 * (long)ntiles * nfields), 1, 1);
  } else if (a.xcd_group) {
    const long units = (long)ntiles * nslabs;
    grid = dim3((unsigned)(((units + 7) / 8) * 8 * nfields), 1, 1);
  }
  if constexpr (N == 256) {
    if (wide) {
      if (V) col_fast_go<T, N, true, 1>(p, a, grid, dir, spectral);
      else col_fast_go<T, N, true, 0>(p, a, grid, dir, spectral);
      return;
    }
  }
  if (V) col_fast_go<T, N, false, 1>(p, a, grid, dir, spectral);
  else col_fast_go<T, N, false, 0>(p, a, grid, dir, spectral);
}

template <typename T>
bool launch_col_fast(ddx_plan* p, ColArgs<T>& a, int nslabs, int nfields, int dir, bool spectral) {
  if (spectral && dir < 0) return false;   // not a pass of this library: the generic kernel
  switch (a.n) {
    case 64: col_fast_n<T, 64>(p, a, nslabs, nfields, dir, spectral); return true;
    case 128: col_fast_n<T, 128>(p, a, nslabs, nfields, dir, spectral); return true;
    case 256: col_fast_n<T, 256>(p, a, nslabs, nfields, dir, spectral); return true;
    case 512: col_fast_n<T, 512>(p, a, nslabs, nfields, dir, spectral); return true;
    default: return false;
  }
}

template <typename T>
int launch_col(ddx_plan* p, ColArgs<T>& a, int nslabs, int nfields, int dir, bool spectral,
               int cat, int nsrc = 0) {
  if (nfields < 1 || nfields > DDX_MAXF) return fail(DDX_ERR_ARG, "bad field count");
  a.W = tile_width(a.n, sizeof(cpx<T>));
  const size_t lds = (size_t)2 * a.n * a.W * sizeof(cpx<T>);
  if (lds > 65536) return fail(DDX_ERR_UNSUPPORTED, "column FFT too long for LDS");
  dim3 grid((a.ncols + a.W - 1) / a.W, nslabs, nfields);
  dim3 block(256);
  // algorithmic bytes: every valid input element read once, output written once
  const double per_in = (double)nslabs * a.in_rows * p->Kx;
  const double per_out = (double)nslabs * a.out_rows * p->Kx;
  const double bytes = ((spectral ? nsrc : nfields) * per_in + nfields * per_out) * sizeof(cpx<T>);
  ProfScope prof(p, cat, bytes);
  if (!p->generic_only && launch_col_fast<T>(p, a, nslabs, nfields, dir, spectral)) {
    LAUNCHCHK();
    return 0;
  }
  if (dir > 0) {
    if (spectral) hipLaunchKernelGGL((col_fft_kernel<T, 1, true>), grid, block, lds, p->stream, a);
    else hipLaunchKernelGGL((col_fft_kernel<T, 1, false>), grid, block, lds, p->stream, a);
  } else {
    if (spectral) hipLaunchKernelGGL((col_fft_kernel<T, -1, true>), grid, block, lds, p->stream, a);
    else hipLaunchKernelGGL((col_fft_kernel<T, -1, false>), grid, block, lds, p->stream, a);
  }
  LAUNCHCHK();
  return 0;
}

template <typename T>
void col_defaults(ColArgs<T>& a) {
  memset(&a, 0, sizeof(a));
  a.in_c = -1;
  a.out_c = -1;
  a.row_dim = -1;
  a.slab_dim = -1;
  a.scale = T(1);
}

// Backward z pass (3D): compact (Ky, Kz, Kxp) -> (Ky, Nz, Kxp)
template <typename T>
void geom_bz(const ddx_plan* p, ColArgs<T>& a) {
  a.fp = p->fz;
  a.tw = (const cpx<T>*)p->tw.at(p->az.n);
  a.n = p->az.n;
  a.in_rows = p->Kz; a.in_c = p->az.c;
  a.out_rows = p->az.n; a.out_c = -1;
  a.ncols = p->Kxp;
  a.in_rs = p->Kxp; a.in_ss = (long)p->Kz * p->Kxp;
  a.out_rs = p->Kxp; a.out_ss = (long)p->az.n * p->Kxp;
  a.k_row = p->dkz; a.row_dim = 2;
  a.k_slab = p->dky; a.slab_dim = 1;
  a.k_col = p->dkx;
  a.kr_inv = p->az.inv; a.kr_kny = p->az.kny;
}

// Backward y pass.  3D: (Ky, Nz, Kxp) -> (Nz, Ny, Kxp); 2D: (Ky, Kxp) -> (Ny, Kxp)
template <typename T>
void geom_by(const ddx_plan* p, ColArgs<T>& a) {
  const long nz = p->az.n;
  a.fp = p->fy;
  a.tw = (const cpx<T>*)p->tw.at(p->ay.n);
  a.n = p->ay.n;
  a.in_rows = p->Ky; a.in_c = p->ay.c;
  a.out_rows = p->ay.n; a.out_c = -1;
  a.ncols = p->Kxp;
  a.in_rs = nz * p->Kxp; a.in_ss = p->Kxp;
  a.out_rs = p->Kxp; a.out_ss = (long)p->ay.n * p->Kxp;
  a.k_row = p->dky; a.row_dim = 1;
  a.k_slab = nullptr; a.slab_dim = -1;
  a.k_col = p->dkx;
  a.kr_inv = p->ay.inv; a.kr_kny = p->ay.kny;
}

// Forward y pass.  3D: (Nz, Ny, Kxp) -> (Ky, Nz, Kxp); 2D: (Ny, Kxp) -> (Ky, Kxp)
template <typename T>
void geom_fy(const ddx_plan* p, ColArgs<T>& a) {
  const long nz = p->az.n;
  a.fp = p->fy;
  a.tw = (const cpx<T>*)p->tw.at(p->ay.n);
  a.n = p->ay.n;
  a.in_rows = p->ay.n; a.in_c = -1;
  a.out_rows = p->Ky; a.out_c = p->ay.c;
  a.ncols = p->Kxp;
  a.in_rs = p->Kxp; a.in_ss = (long)p->ay.n * p->Kxp;
  a.out_rs = nz * p->Kxp; a.out_ss = p->Kxp;
}

// Forward z pass (3D): (Ky, Nz, Kxp) -> (Ky, Kz, Kxp)
template <typename T>
void geom_fz(const ddx_plan* p, ColArgs<T>& a) {
  a.fp = p->fz;
  a.tw = (const cpx<T>*)p->tw.at(p->az.n);
  a.n = p->az.n;
  a.in_rows = p->az.n; a.in_c = -1;
  a.out_rows = p->Kz; a.out_c = p->az.c;
  a.ncols = p->Kxp;
  a.in_rs = p->Kxp; a.in_ss = (long)p->az.n * p->Kxp;
  a.out_rs = p->Kxp; a.out_ss = (long)p->Kz * p->Kxp;
}

template <typename T, int N, int PH, int ND>
void row_fast_go(ddx_plan* p, const RowArgs<T>& a, long nrows) {
  if constexpr (ND == 2) {
    if (a.xin[0]) {
      // inputs given in x space (rhs_impl checks that this kernel applies)
      constexpr int RPB = 64 / RowPlan<N>::LANES;
      hipLaunchKernelGGL((row_phys_fast<T, N, PH, ND, 0, true, true>), dim3((nrows + RPB - 1) / RPB),
                         dim3(64), 0, p->stream, a, nrows);
      return;
    }
  }
  bool xf = false;
  for (int f = 0; f < a.nout; ++f) xf = xf || a.xadd[f] != nullptr;
  if (xf) {
    // x-space forcing: the base radix plan with the forcing loads
    constexpr int RPB = 64 / RowPlan<N>::LANES;
    hipLaunchKernelGGL((row_phys_fast<T, N, PH, ND, 0, true>), dim3((nrows + RPB - 1) / RPB),
                       dim3(64), 0, p->stream, a, nrows);
    return;
  }
  if (radix_variant(p)) {
    constexpr int RPB = 64 / RowPlanV<N, 1>::LANES;
    hipLaunchKernelGGL((row_phys_fast<T, N, PH, ND, 1>), dim3((nrows + RPB - 1) / RPB), dim3(64), 0,
                       p->stream, a, nrows);
    return;
  }
  constexpr int RPB = 64 / RowPlan<N>::LANES;
  hipLaunchKernelGGL((row_phys_fast<T, N, PH, ND>), dim3((nrows + RPB - 1) / RPB), dim3(64), 0,
                     p->stream, a, nrows);
}

template <typename T, int N>
bool row_fast_n(ddx_plan* p, const RowArgs<T>& a, long nrows) {
  const int key = a.physics * 10 + a.ndim;
  switch (key) {
    case PH_HYDRO * 10 + 2: row_fast_go<T, N, PH_HYDRO, 2>(p, a, nrows); return true;
    case PH_HYDRO * 10 + 3: row_fast_go<T, N, PH_HYDRO, 3>(p, a, nrows); return true;
    case PH_BOUSS * 10 + 2: row_fast_go<T, N, PH_BOUSS, 2>(p, a, nrows); return true;
    case PH_BOUSS * 10 + 3: row_fast_go<T, N, PH_BOUSS, 3>(p, a, nrows); return true;
    case PH_MHD * 10 + 2: row_fast_go<T, N, PH_MHD, 2>(p, a, nrows); return true;
    case PH_MHD * 10 + 3: row_fast_go<T, N, PH_MHD, 3>(p, a, nrows); return true;
    default: return false;
  }
}

template <typename T>
bool launch_row_fast(ddx_plan* p, const RowArgs<T>& a, long nrows) {
  if (a.kx > a.nx / 3 + 1 || a.nin > 12 || a.nout > 6) return false;  // LDS staging bound
  switch (a.nx) {
    case 64: return row_fast_n<T, 64>(p, a, nrows);
    case 128: return row_fast_n<T, 128>(p, a, nrows);
    case 256: return row_fast_n<T, 256>(p, a, nrows);
    case 512: return row_fast_n<T, 512>(p, a, nrows);
    default: return false;
  }
}

// Size-specialised single-field c2r / r2c row passes (two rows per team);
// false if none applies.
template <typename T, int N, int V>
void row_plain_nv(ddx_plan* p, const RowArgs<T>& a, long nrows) {
  constexpr int RPB = 64 / RowPlanV<N, V>::LANES;
  const long teams = (nrows + 1) / 2;
  const int nf = a.mode == ROW_C2R ? a.nin : a.nout;   // grid.y: the field
  const dim3 grid((unsigned)((teams + RPB - 1) / RPB), (unsigned)nf);
  if (a.mode == ROW_C2R) hipLaunchKernelGGL((row_c2r_fast<T, N, V>), grid, dim3(64), 0, p->stream, a, nrows);
  else hipLaunchKernelGGL((row_r2c_fast<T, N, V>), grid, dim3(64), 0, p->stream, a, nrows);
}

template <typename T, int N>
void row_plain_n(ddx_plan* p, const RowArgs<T>& a, long nrows) {
  if (radix_variant(p)) row_plain_nv<T, N, 1>(p, a, nrows);
  else row_plain_nv<T, N, 0>(p, a, nrows);
}

template <typename T>
bool launch_row_plain_fast(ddx_plan* p, const RowArgs<T>& a, long nrows) {
  if (a.kx > a.nx / 3 + 1) return false;
  if (a.mode == ROW_C2R ? (a.nin < 1 || a.nout != 0) : (a.nout < 1 || a.nin != 0)) return false;
  switch (a.nx) {
    case 64: row_plain_n<T, 64>(p, a, nrows); return true;
    case 128: row_plain_n<T, 128>(p, a, nrows); return true;
    case 256: row_plain_n<T, 256>(p, a, nrows); return true;
    case 512: row_plain_n<T, 512>(p, a, nrows); return true;
    default: return false;
  }
}

template <typename T>
int launch_row(ddx_plan* p, RowArgs<T>& a) {
  const long nrows = (long)p->ay.n * p->az.n;
  const double cb = sizeof(cpx<T>), rb = sizeof(T);
  double bytes;
  int cat;
  if (a.mode == ROW_PHYS) { cat = PC_ROW_PHYS; bytes = nrows * (double)(a.nin + a.nout) * p->Kx * cb; }
  else if (a.mode == ROW_C2R) { cat = PC_ROW_C2R; bytes = nrows * (double)a.nin * (p->Kx * cb + a.nx * rb); }
  else { cat = PC_ROW_R2C; bytes = nrows * (double)a.nout * (p->Kx * cb + a.nx * rb); }
  // the size-specialised kernels keep rows in registers (any field count)
  if (!p->generic_only) {
    ProfScope prof(p, cat, bytes);
    const bool fast = a.mode == ROW_PHYS ? launch_row_fast<T>(p, a, nrows)
                                         : launch_row_plain_fast<T>(p, a, nrows);
    if (fast) {
      LAUNCHCHK();
      return 0;
    }
    if (prof.a) p->evfree.push_back(prof.a);   // nothing launched: no record
    prof.a = nullptr;
  }
  // the generic kernel stages whole rows of every field in LDS
  const int nseq = std::max((a.nin + 1) / 2, (a.nout + 1) / 2);
  const size_t lds = (size_t)2 * nseq * a.nx * sizeof(cpx<T>);
  if (lds > 65536 && a.mode != ROW_PHYS && std::max(a.nin, a.nout) > 1) {
    // several plain fields over the LDS of the generic kernel: one at a time
    const int nf = std::max(a.nin, a.nout);
    for (int f = 0; f < nf; ++f) {
      RowArgs<T> b = a;
      if (a.mode == ROW_C2R) { b.nin = 1; b.in[0] = a.in[f]; b.xout[0] = a.xout[f]; }
      else { b.nout = 1; b.xin[0] = a.xin[f]; b.out[0] = a.out[f]; }
      int rc = launch_row(p, b);
      if (rc) return rc;
    }
    return 0;
  }
  if (lds > 65536) return fail(DDX_ERR_UNSUPPORTED, "x rows too long for LDS");
  ProfScope prof(p, cat, bytes);
  hipLaunchKernelGGL((row_kernel<T>), dim3(nrows), dim3(256), lds, p->stream, a);
  LAUNCHCHK();
  return 0;
}

template <typename T>
void row_defaults(const ddx_plan* p, RowArgs<T>& a) {
  memset(&a, 0, sizeof(a));
  a.fp = p->fx;
  a.tw = (const cpx<T>*)p->tw.at(p->ax.n);
  a.nx = p->ax.n;
  a.kx = p->Kx;
  a.kxp = p->Kxp;
  a.ndim = p->ndim;
}

// ---------------------------------------------------------------------------
// transforms
// ---------------------------------------------------------------------------

// The fused right-hand side forms the nonlinear terms in rotational form
// (u x omega, and the induction and Lorentz terms likewise), which equals
// the reference's convective forms after projection only when the products
// are dealiased by a 2/3 rule; without dealiasing the aliasing errors of the
// two forms differ.
const char* const kNoFusedAliased =
    "the fused right-hand side needs a 2/3 dealiasing rule (its rotational form equals the "
    "reference's convective form only for dealiased products); with dealiasing 'None' use "
    "the per-operation kernels";

// The spherical rule's zeroing of the modes outside the sphere, in place on
// nf compact arrays (nothing for the other rules, whose compact layout holds
// only kept modes).
template <typename T>
int sphere_mask(ddx_plan* p, int nf, cpx<T>* const* f) {
  if (p->dealias != DDX_DEALIAS_SPHERICAL || nf < 1) return 0;
  if (nf > DDX_MAXF) return fail(DDX_ERR_ARG, "too many fields");
  MaskArgs<T> a;
  memset(&a, 0, sizeof(a));
  for (int i = 0; i < nf; ++i) a.f[i] = f[i];
  a.g = kgeom(p);
  a.nk = p->nk;
  a.nf = nf;
  a.ndim = p->ndim;
  a.thr = p->sph_thr;
  hipLaunchKernelGGL((sphere_mask_kernel<T>), dim3((p->nk + 255) / 256), dim3(256), 0, p->stream, a);
  LAUNCHCHK();
  return 0;
}

// nf fields at once: every pass is one launch over all the fields
template <typename T>
int fwd_impl(ddx_plan* p, int nf, const T* const* x, cpx<T>* const* kc) {
  int rc = ensure_ws(p, nf);
  if (rc) return rc;
  const long fs = ws_field_stride(p);
  cpx<T>* A = (cpx<T>*)p->wsA;
  cpx<T>* B = (cpx<T>*)p->wsB;
  RowArgs<T> r;
  row_defaults(p, r);
  r.mode = ROW_R2C;
  r.nin = 0; r.nout = nf;
  for (int f = 0; f < nf; ++f) { r.xin[f] = x[f]; r.out[f] = A + f * fs; }
  if ((rc = launch_row(p, r))) return rc;
  ColArgs<T> c;
  col_defaults(c);
  geom_fy(p, c);
  for (int f = 0; f < nf; ++f) c.in[f] = A + f * fs;
  if (!p->zfft) {
    for (int f = 0; f < nf; ++f) c.out[f] = kc[f];
    c.scale = T(1.0 / (double)p->nper);
    if ((rc = launch_col(p, c, p->az.n, nf, -1, false, PC_COL_FWD_Y))) return rc;
    return sphere_mask(p, nf, kc);
  }
  for (int f = 0; f < nf; ++f) c.out[f] = B + f * fs;
  if ((rc = launch_col(p, c, p->az.n, nf, -1, false, PC_COL_FWD_Y))) return rc;
  col_defaults(c);
  geom_fz(p, c);
  for (int f = 0; f < nf; ++f) { c.in[f] = B + f * fs; c.out[f] = kc[f]; }
  c.scale = T(1.0 / (double)p->nper);
  if ((rc = launch_col(p, c, p->Ky, nf, -1, false, PC_COL_FWD_Z))) return rc;
  // forward() dealiases its output (representations.py:344)
  return sphere_mask(p, nf, kc);
}

template <typename T>
int bwd_impl(ddx_plan* p, int nf, const cpx<T>* const* kc, T* const* x) {
  int rc = ensure_ws(p, nf);
  if (rc) return rc;
  // backward() dealiases its input in place first (representations.py:353)
  if ((rc = sphere_mask(p, nf, const_cast<cpx<T>* const*>(kc)))) return rc;
  const long fs = ws_field_stride(p);
  cpx<T>* A = (cpx<T>*)p->wsA;
  cpx<T>* B = (cpx<T>*)p->wsB;
  ColArgs<T> c;
  if (p->zfft) {
    col_defaults(c);
    geom_bz(p, c);
    for (int f = 0; f < nf; ++f) { c.in[f] = kc[f]; c.out[f] = A + f * fs; }
    if ((rc = launch_col(p, c, p->Ky, nf, 1, false, PC_COL_BWD_Z))) return rc;
    col_defaults(c);
    geom_by(p, c);
    for (int f = 0; f < nf; ++f) { c.in[f] = A + f * fs; c.out[f] = B + f * fs; }
    if ((rc = launch_col(p, c, p->az.n, nf, 1, false, PC_COL_BWD_Y))) return rc;
  } else {
    col_defaults(c);
    geom_by(p, c);
    for (int f = 0; f < nf; ++f) { c.in[f] = kc[f]; c.out[f] = B + f * fs; }
    if ((rc = launch_col(p, c, p->az.n, nf, 1, false, PC_COL_BWD_Y))) return rc;
  }
  RowArgs<T> r;
  row_defaults(p, r);
  r.mode = ROW_C2R;
  r.nin = nf; r.nout = 0;
  for (int f = 0; f < nf; ++f) { r.in[f] = B + f * fs; r.xout[f] = x[f]; }
  return launch_row(p, r);
}

// Backward transforms of linear combinations of compact k arrays: the
// combination (recipes, as the RHS's derived fields) is formed in the first
// column pass, so derivative fields need no k-space pass of their own.
template <typename T>
int bwd_combine_impl(ddx_plan* p, int nsrc, const cpx<T>* const* kc, int nout, const Recipe* rec,
                     T* const* x) {
  int rc = ensure_ws(p, nout);
  if (rc) return rc;
  // backward() dealiases what it transforms in place (representations.py:353)
  if ((rc = sphere_mask(p, nsrc, const_cast<cpx<T>* const*>(kc)))) return rc;
  const long fs = ws_field_stride(p);
  cpx<T>* A = (cpx<T>*)p->wsA;
  cpx<T>* B = (cpx<T>*)p->wsB;
  ColArgs<T> c;
  col_defaults(c);
  for (int s = 0; s < nsrc; ++s) c.src[s] = kc[s];
  for (int q = 0; q < nout; ++q) c.rec[q] = rec[q];
  if (p->zfft) {
    geom_bz(p, c);
    for (int q = 0; q < nout; ++q) c.out[q] = A + q * fs;
    if ((rc = launch_col(p, c, p->Ky, nout, 1, true, PC_COL_BWD_Z_SPECTRAL, nsrc))) return rc;
    col_defaults(c);
    geom_by(p, c);
    for (int q = 0; q < nout; ++q) { c.in[q] = A + q * fs; c.out[q] = B + q * fs; }
    if ((rc = launch_col(p, c, p->az.n, nout, 1, false, PC_COL_BWD_Y))) return rc;
  } else {
    geom_by(p, c);
    for (int q = 0; q < nout; ++q) c.out[q] = B + q * fs;
    if ((rc = launch_col(p, c, p->az.n, nout, 1, true, PC_COL_BWD_Y_SPECTRAL, nsrc))) return rc;
  }
  RowArgs<T> r;
  row_defaults(p, r);
  r.mode = ROW_C2R;
  r.nin = nout;
  r.nout = 0;
  for (int q = 0; q < nout; ++q) { r.in[q] = B + q * fs; r.xout[q] = x[q]; }
  return launch_row(p, r);
}

// ---------------------------------------------------------------------------
// fused RHS
// ---------------------------------------------------------------------------

struct PhysLayout {
  int nstate;
  int nq;  // derived x-space inputs
  int np;  // x-space outputs
  Recipe rec[DDX_MAXF];
};

Recipe one(int src, int dim = -1, double coef = 1.0) {
  Recipe r;
  memset(&r, 0, sizeof(r));
  r.nterm = 1;
  r.t[0].src = src; r.t[0].dim = dim; r.t[0].coef = coef;
  return r;
}

Recipe two(int s0, int d0, double c0, int s1, int d1, double c1) {
  Recipe r = one(s0, d0, c0);
  r.nterm = 2;
  r.t[1].src = s1; r.t[1].dim = d1; r.t[1].coef = c1;
  return r;
}

// curl of the vector starting at state index b (3D): (dy vz - dz vy, dz vx - dx vz, dx vy - dy vx)
void add_curl3(PhysLayout& L, int b) {
  L.rec[L.nq++] = two(b + 2, 1, 1.0, b + 1, 2, -1.0);
  L.rec[L.nq++] = two(b + 0, 2, 1.0, b + 2, 0, -1.0);
  L.rec[L.nq++] = two(b + 1, 0, 1.0, b + 0, 1, -1.0);
}

int phys_layout(int physics, int nd, PhysLayout* L) {
  memset(L, 0, sizeof(*L));
  // velocity and vorticity
  for (int i = 0; i < nd; ++i) L->rec[L->nq++] = one(i);
  if (nd == 3) add_curl3(*L, 0);
  else L->rec[L->nq++] = two(1, 0, 1.0, 0, 1, -1.0);
  if (physics == DDX_HYDRO) {
    L->nstate = nd;
    L->np = nd;
  } else if (physics == DDX_BOUSSINESQ) {
    L->nstate = nd + 1;
    L->rec[L->nq++] = one(nd);
    for (int d = 0; d < nd; ++d) L->rec[L->nq++] = one(nd, d, 1.0);
    L->np = nd + 1;
  } else if (physics == DDX_MHD) {
    L->nstate = 2 * nd;
    for (int i = 0; i < nd; ++i) L->rec[L->nq++] = one(nd + i);
    if (nd == 3) add_curl3(*L, nd);
    else L->rec[L->nq++] = two(nd + 1, 0, 1.0, nd + 0, 1, -1.0);
    L->np = (nd == 3) ? 6 : 3;
  } else {
    return -1;
  }
  return 0;
}

// The backward z pass of a 3D right-hand side splits the derived fields: a
// derivative along z must be taken before the z transform, but one along x or
// y can wait for the y pass.  Fields with a d/dz term, and plain copies, are
// formed in the z pass; the others wait for the y pass and are formed there
// from plain z-pass outputs.  MHD: the z pass forms u, wx, wy, B, Jx, Jy (10
// arrays instead of 12), the y pass copies those and forms wz and Jz; hydro
// transforms 5 arrays instead of 6, Boussinesq 7 instead of 10.  Returns the
// number of z-pass outputs.
int split_z(const PhysLayout& L, Recipe* zrec, Recipe* yrec) {
  int nz = 0;
  // index of the z-pass output that is source s, plain (added if missing)
  auto plain = [&](int s) {
    int j = 0;
    while (j < nz && !(zrec[j].nterm == 1 && zrec[j].t[0].src == s && zrec[j].t[0].dim < 0 &&
                       zrec[j].t[0].coef == 1.0)) ++j;
    if (j == nz) zrec[nz++] = one(s);
    return j;
  };
  int zi[DDX_MAXF];
  for (int q = 0; q < L.nq; ++q) {
    const Recipe& r = L.rec[q];
    bool dz = false;
    for (int t = 0; t < r.nterm; ++t) dz = dz || r.t[t].dim == 2;
    const bool copy = r.nterm == 1 && r.t[0].dim < 0 && r.t[0].coef == 1.0;
    zi[q] = -1;
    if (copy) zi[q] = plain(r.t[0].src);
    else if (dz) { zi[q] = nz; zrec[nz++] = r; }
  }
  for (int q = 0; q < L.nq; ++q) {
    if (zi[q] >= 0) { yrec[q] = one(zi[q]); continue; }
    yrec[q] = L.rec[q];
    for (int t = 0; t < L.rec[q].nterm; ++t) yrec[q].t[t].src = plain(L.rec[q].t[t].src);
  }
  return nz;
}

int nstate_of(const ddx_params* prm, int nd) {
  PhysLayout L;
  if (phys_layout(prm->physics, nd, &L)) return -1;
  return L.nstate;
}

// An ETD sweep to fuse into the epilogue of a right-hand side
template <typename T>
struct FusedUpdate {
  int kind;                  // DDX_ETD1 / DDX_ETD2RK1 / DDX_ETD2RK2
  const cpx<T>* const* y;
  const cpx<T>* const* f1;   // ETD2RK*: the first stage's derivative
  cpx<T>* const* out;
  const double* coeff;
  int order;
  double dt;
  bool write_deriv;
};

// Arguments of the k-space epilogue of one right-hand side (and of the ETD
// sweep fused into it when fu is given); the products U are the caller's.
template <typename T>
EpiArgs<T> epi_args(const ddx_plan* p, const ddx_params* prm, int nstate, const cpx<T>* const* state,
                    cpx<T>* const* deriv, const cpx<T>* const* forcing, const FusedUpdate<T>* fu) {
  EpiArgs<T> e;
  memset(&e, 0, sizeof(e));
  for (int s = 0; s < nstate; ++s) {
    e.state[s] = state[s];
    e.deriv[s] = deriv ? deriv[s] : nullptr;
    e.forcing[s] = forcing ? forcing[s] : nullptr;
  }
  e.g = kgeom(p);
  e.nk = p->nk;
  e.physics = prm->physics;
  e.ndim = p->ndim;
  e.bdir = prm->bdir;
  e.npts = (double)p->nper;
  e.galpha = prm->g * prm->alpha_t;
  e.beta = prm->beta;
  e.upd_kind = -1;
  if (fu) {
    e.upd_kind = fu->kind;
    e.upd_order = fu->order;
    e.upd_dt = fu->dt;
    e.write_deriv = fu->write_deriv ? 1 : 0;
    for (int s = 0; s < nstate; ++s) {
      e.upd_y[s] = fu->y[s];
      e.upd_f1[s] = fu->f1 ? fu->f1[s] : nullptr;
      e.upd_out[s] = fu->out[s];
      e.upd_coeff[s] = fu->coeff[s];
    }
  }
  return e;
}

// Workspace fields a right-hand side of this physics needs in each buffer
int rhs_ws_fields(int physics, int nd) {
  PhysLayout L;
  if (phys_layout(physics, nd, &L)) return -1;
  return std::max(L.nq, L.np);
}

// The right-hand side on the workspace ws (ws.A, ws.B hold rhs_ws_fields
// arrays of ws_field_stride each).  deriv may be null when a fused sweep
// that does not write it (fu->write_deriv false) follows.
// xin (2D): the x-space inputs of the x-row kernel (the recipe fields of
// phys_layout, as ddx_bwd_combine writes them), which then replace the
// backward column pass; used where the size-specialised x-row kernel applies
// to a cube-rule plan, otherwise ignored (the pass runs).
template <typename T>
bool xin_applies(const ddx_plan* p, int nq) {
  return p->ndim == 2 && !p->generic_only && p->dealias == DDX_DEALIAS_23 && nq <= 12 &&
         p->Kx <= p->ax.n / 3 + 1 &&
         (p->ax.n == 64 || p->ax.n == 128 || p->ax.n == 256 || p->ax.n == 512);
}

template <typename T>
int rhs_impl(ddx_plan* p, const ddx_params* prm, const cpx<T>* const* state, cpx<T>* const* deriv,
             const cpx<T>* const* forcing, const FusedUpdate<T>* fu, const Ws& ws,
             const T* const* xforcing = nullptr, const T* const* xin = nullptr) {
  const int nd = p->ndim;
  PhysLayout L;
  if (phys_layout(prm->physics, nd, &L)) return fail(DDX_ERR_ARG, "unknown physics");
  if (prm->physics == DDX_BOUSSINESQ && (prm->bdir < 0 || prm->bdir >= nd))
    return fail(DDX_ERR_ARG, "boussinesq direction outside the grid's dimensions");
  const bool need_deriv = !fu || fu->write_deriv;
  for (int s = 0; s < L.nstate; ++s)
    if (!state[s] || (need_deriv && (!deriv || !deriv[s])))
      return fail(DDX_ERR_ARG, "null state or deriv pointer");
  if (!ws.A || !ws.B) return fail(DDX_ERR_ARG, "no workspace");
  int rc;
  const long fs = ws_field_stride(p);
  cpx<T>* A = (cpx<T>*)ws.A;
  cpx<T>* B = (cpx<T>*)ws.B;
  // the reference's first backward transforms of the state dealias it in
  // place (representations.py:353)
  if ((rc = sphere_mask(p, L.nstate, const_cast<cpx<T>* const*>(state)))) return rc;

  // backward: spectral prologue in the first column pass
  ColArgs<T> c;
  col_defaults(c);
  for (int s = 0; s < L.nstate; ++s) c.src[s] = state[s];
  for (int q = 0; q < L.nq; ++q) c.rec[q] = L.rec[q];
  cpx<T>* mixed;  // (Nz, Ny, Kxp) per field
  // FP64 only: at 256^3 MHD the z pass gets 26 % faster (10 outputs
  // instead of 12) and the y pass 1 % faster: 2.5 % per step.  FP32 at
  // 128^3 measured 1-2 % slower per step with the split (its passes are
  // short of issue slots, not of bytes).
  Recipe zrec[DDX_MAXF], yrec[DDX_MAXF];
  const int nzs = p->zfft && p->zsplit && sizeof(T) == 8 ? split_z(L, zrec, yrec) : L.nq;
  // 3D grids whose arrays dwarf the caches (>= 2^22 complex per field): the
  // column passes, the x rows and the epilogue store nontemporally.  A pass
  // that stores so gets slower, and the kernel reading its output faster
  // (256^3 MHD per 40 launches: z pass 6.7 -> 7.3, forward y 6.8 -> 7.3 with
  // column stores alone; y pass 14.1 -> 13.3, x rows 13.9 -> 13.3, forward z
  // 4.7 -> 4.4, epilogue 8.1 -> 7.5): 2 % per step for the column passes,
  // 1 % more with the x rows and the epilogue.  Only the y pass and the
  // forward z pass storing so was no faster, and at 128^3, where the
  // intermediates fit the last-level cache, it was slower.
  const bool big_nt = DDX_COL_NT && p->zfft && (long)p->az.n * p->ay.n * p->Kxp >= (1L << 22);
  const bool use_xin = xin && xin_applies<T>(p, L.nq);
  if (use_xin) {
    // the caller's x-space inputs replace the backward pass
  } else if (p->zfft && nzs < L.nq) {
    geom_bz(p, c);
    for (int j = 0; j < nzs; ++j) { c.rec[j] = zrec[j]; c.out[j] = A + j * fs; }
    c.nt = big_nt;
    if ((rc = launch_col(p, c, p->Ky, nzs, 1, true, PC_COL_BWD_Z_SPECTRAL, L.nstate))) return rc;
    col_defaults(c);
    geom_by(p, c);
    for (int j = 0; j < nzs; ++j) c.src[j] = A + j * fs;
    for (int q = 0; q < L.nq; ++q) { c.rec[q] = yrec[q]; c.out[q] = B + q * fs; }
    c.group_hint = DDX_YSPEC_GROUP;
    c.wide = DDX_YSPEC_WIDE;
    c.nt = big_nt;
    if ((rc = launch_col(p, c, p->az.n, L.nq, 1, true, PC_COL_BWD_Y_SPECTRAL, nzs))) return rc;
  } else if (p->zfft) {
    geom_bz(p, c);
    for (int q = 0; q < L.nq; ++q) c.out[q] = A + q * fs;
    c.nt = big_nt;
    if ((rc = launch_col(p, c, p->Ky, L.nq, 1, true, PC_COL_BWD_Z_SPECTRAL, L.nstate))) return rc;
    col_defaults(c);
    geom_by(p, c);
    for (int q = 0; q < L.nq; ++q) { c.in[q] = A + q * fs; c.out[q] = B + q * fs; }
    c.nt = big_nt;
    if ((rc = launch_col(p, c, p->az.n, L.nq, 1, false, PC_COL_BWD_Y))) return rc;
  } else {
    geom_by(p, c);
    for (int q = 0; q < L.nq; ++q) c.out[q] = B + q * fs;
    if ((rc = launch_col(p, c, p->az.n, L.nq, 1, true, PC_COL_BWD_Y_SPECTRAL, L.nstate))) return rc;
  }
  mixed = B;

  // x rows: c2r, products, r2c
  RowArgs<T> r;
  row_defaults(p, r);
  r.mode = ROW_PHYS;
  r.physics = prm->physics;
  r.nin = L.nq;
  r.nout = L.np;
  r.fpr_inv = T(1.0 / (4.0 * M_PI * (prm->physics == DDX_MHD ? prm->rho0 : 1.0)));
  for (int q = 0; q < L.nq; ++q) r.in[q] = mixed + q * fs;
  if (use_xin)
    for (int q = 0; q < L.nq; ++q) r.xin[q] = xin[q];
  for (int o = 0; o < L.np; ++o) r.out[o] = A + o * fs;
  r.nt = big_nt && DDX_ROW_NT;
  // an x-space forcing of the velocity (and of T) joins its product: the
  // product outputs 0 .. nd-1 are the velocity's, output nd is T's
  if (xforcing) {
    const int nforce = nd + (prm->physics == DDX_BOUSSINESQ ? 1 : 0);
    for (int o = 0; o < nforce; ++o) r.xadd[o] = xforcing[o];
  }
  if ((rc = launch_row(p, r))) return rc;

  // forward column passes (single 2D grids with 64-point y axes: the y pass
  // runs inside the epilogue; batched plans keep the two launches, which
  // measured as fast for 256 members)
  const bool yepi = !p->zfft && nd == 2 && (p->tune_yepi >= 0 ? p->tune_yepi == 1 : p->nbatch == 1) &&
                    p->fuse_yepi && !p->generic_only && radix_variant(p) == 0 &&
                    p->dealias != DDX_DEALIAS_SPHERICAL &&
                    p->ay.n == 64 && p->Ky <= 2 * ((64 + 2) / 3) + 1 && L.np <= 3;
  cpx<T>* Uc;
  col_defaults(c);
  geom_fy(p, c);
  for (int o = 0; o < L.np; ++o) c.in[o] = A + o * fs;
  if (p->zfft) {
    for (int o = 0; o < L.np; ++o) c.out[o] = B + o * fs;
    c.nt = big_nt;
    if ((rc = launch_col(p, c, p->az.n, L.np, -1, false, PC_COL_FWD_Y))) return rc;
    col_defaults(c);
    geom_fz(p, c);
    for (int o = 0; o < L.np; ++o) { c.in[o] = B + o * fs; c.out[o] = A + o * fs; }
    c.nt = big_nt;
    if ((rc = launch_col(p, c, p->Ky, L.np, -1, false, PC_COL_FWD_Z))) return rc;
    Uc = A;
  } else if (yepi) {
    Uc = A;   // the x-row products, (members, 64, Kxp) per field
  } else {
    for (int o = 0; o < L.np; ++o) c.out[o] = B + o * fs;
    if ((rc = launch_col(p, c, p->az.n, L.np, -1, false, PC_COL_FWD_Y))) return rc;
    Uc = B;
  }

  // the products' forward transforms dealias them (representations.py:344);
  // forcing, added in the epilogue, is not
  if (p->dealias == DDX_DEALIAS_SPHERICAL) {
    cpx<T>* uf[DDX_MAXF];
    for (int o = 0; o < L.np; ++o) uf[o] = Uc + o * fs;
    if ((rc = sphere_mask(p, L.np, uf))) return rc;
  }

  // k-space epilogue
  EpiArgs<T> e = epi_args<T>(p, prm, L.nstate, state, deriv, forcing, fu);
  for (int o = 0; o < L.np; ++o) e.U[o] = Uc + o * fs;
  e.nt = big_nt && DDX_EPI_NT;
  int upd_arrays = 0;
  if (fu) {
    // y read, out written, F1 read (ETD2RK*), deriv not written (stage 2)
    upd_arrays = 2 + (fu->kind >= DDX_ETD2RK1 ? 1 : 0) - (fu->write_deriv ? 0 : 1);
  }
  const int bs = 256;
  int nforce = 0;
  for (int s = 0; s < L.nstate; ++s) nforce += e.forcing[s] ? 1 : 0;
  const double kvalid = (double)p->Ky * p->Kz * p->Kx;
  const int extra = prm->physics == DDX_BOUSSINESQ ? 2 : 0;
  if (yepi) {
    // the y pass's reads of the x-row products replace the epilogue's reads
    // of its output
    const double yin = (double)p->az.n * p->ay.n * p->Kx * L.np;
    ProfScope prof(p, PC_YEPI,
                   (yin + kvalid * (L.nstate + extra + nforce + upd_arrays * L.nstate)) * sizeof(cpx<T>));
    e.tw_y = (const cpx<T>*)p->tw.at(p->ay.n);
    const dim3 grid((p->Kxp + YEPI_W - 1) / YEPI_W, p->az.n, YEPI_SPLIT);
    hipLaunchKernelGGL((yepi64_kernel<T>), grid, dim3(64 * YEPI_WAVES), 0, p->stream, e);
    LAUNCHCHK();
    return 0;
  }
  ProfScope prof(p, PC_EPILOGUE,
                 kvalid * (L.np + L.nstate + extra + nforce + upd_arrays * L.nstate) * sizeof(cpx<T>));
  if (DDX_EPI_PRELOAD && fu && prm->physics == DDX_MHD && nd == 3)
    hipLaunchKernelGGL((epilogue_kernel<T, true>), dim3((p->nk + bs - 1) / bs), dim3(bs), 0, p->stream, e);
  else
    hipLaunchKernelGGL((epilogue_kernel<T, false>), dim3((p->nk + bs - 1) / bs), dim3(bs), 0, p->stream, e);
  LAUNCHCHK();
  return 0;
}

template <typename T>
int update_impl(ddx_plan* p, int kind, int ncomp, const cpx<T>* const* y, cpx<T>* const* out,
                const cpx<T>* const* F1, const cpx<T>* const* F2, const double* coeff, int order,
                double dt, const double* ifac = nullptr) {
  if (ncomp < 1 || ncomp > DDX_MAXF) return fail(DDX_ERR_ARG, "bad component count");
  if (kind < DDX_EULER || kind > DDX_ETD2RK2) return fail(DDX_ERR_ARG, "bad update kind");
  UpdArgs<T> u;
  memset(&u, 0, sizeof(u));
  for (int i = 0; i < ncomp; ++i) {
    if (!y[i] || !out[i] || !F1[i] || (kind >= DDX_ETD2RK1 && !F2[i]))
      return fail(DDX_ERR_ARG, "null array in update");
    u.y[i] = y[i];
    u.out[i] = out[i];
    u.F1[i] = F1[i];
    u.F2[i] = F2 ? F2[i] : nullptr;
    u.coeff[i] = coeff[i];
  }
  u.ifac = ifac;
  u.g = kgeom(p);
  u.nk = p->nk;
  u.ncomp = ncomp;
  u.kind = kind;
  u.order = order;
  u.ndim = p->ndim;
  u.dt = dt;
  const int bs = 256;
  const double kvalid = (double)p->Ky * p->Kz * p->Kx;
  ProfScope prof(p, PC_UPDATE, kvalid * (ncomp * (kind >= DDX_ETD2RK1 ? 4 : 3) * sizeof(cpx<T>) +
                                         (ifac ? sizeof(double) : 0)));
  hipLaunchKernelGGL((update_kernel<T>), dim3((p->nk + bs - 1) / bs), dim3(bs), 0, p->stream, u);
  LAUNCHCHK();
  return 0;
}

void comp_coeffs(const ddx_params* prm, int nd, int nstate, double* coeff) {
  for (int i = 0; i < nstate; ++i) coeff[i] = prm->nu;
  if (prm->physics == DDX_BOUSSINESQ) coeff[nd] = prm->kappa;
  if (prm->physics == DDX_MHD)
    for (int i = nd; i < 2 * nd; ++i) coeff[i] = prm->eta;
}

template <typename T>
int step_impl(ddx_plan* p, int integrator, const ddx_params* prm, cpx<T>* const* y,
              cpx<T>* const* stage, cpx<T>* const* d1, cpx<T>* const* d2_in,
              const cpx<T>* const* forcing, int nsteps, double dt, const Ws* fixed_ws = nullptr,
              const T* const* xforcing = nullptr, const T* const* xin = nullptr) {
  const int nd = p->ndim;
  const int ns = nstate_of(prm, nd);
  // xin describes the state before the first step only
  if (xin && nsteps != 1) return fail(DDX_ERR_ARG, "x-space inputs apply to a single step");
  if (ns < 0) return fail(DDX_ERR_ARG, "unknown physics");
  if (integrator != DDX_RK2MID && integrator != DDX_RK2TRAP)
    return fail(DDX_ERR_ARG, "unknown integrator");
  for (int i = 0; i < ns; ++i) {
    if (!y[i] || !d1[i] || (d2_in && !d2_in[i]) || (integrator == DDX_RK2MID && !stage[i]))
      return fail(DDX_ERR_ARG, "null buffer in step");
  }
  int rc;
  Ws ws;
  if (fixed_ws) {
    ws = *fixed_ws;
  } else {
    if ((rc = ensure_ws(p, rhs_ws_fields(prm->physics, nd)))) return rc;
    if (!d2_in && !p->fuse_update && (rc = ensure_f2(p, ns))) return rc;
    ws.A = p->wsA;
    ws.B = p->wsB;
    ws.F2 = p->f2buf;
  }
  // the second stage's derivative: the caller's deriv2, or (unfused sweeps)
  // scratch, or not stored at all (fused sweeps)
  cpx<T>* d2v[DDX_MAXF] = {};
  for (int i = 0; i < ns; ++i) {
    if (d2_in) d2v[i] = d2_in[i];
    else if (!p->fuse_update) d2v[i] = (cpx<T>*)ws.F2 + (long)i * p->nk;
  }
  if (!d2_in && !p->fuse_update && !ws.F2) return fail(DDX_ERR_ARG, "no deriv2 scratch");
  cpx<T>* const* d2 = d2v;
  const bool keep_d2 = d2_in != nullptr;
  double coeff[DDX_MAXF];
  comp_coeffs(prm, nd, ns, coeff);
  const int vo = prm->viscosity_order;
  auto one_step = [&]() -> int {
    int rc;
    if (p->fuse_update) {
      // the same two RHS + two ETD sweeps, each sweep inside the preceding
      // epilogue; deriv2 is written when the caller passed it
      const cpx<T>* const* yc = (const cpx<T>* const*)y;
      const cpx<T>* const* d1c = (const cpx<T>* const*)d1;
      if (integrator == DDX_RK2MID) {
        // time_step.py:285-305
        FusedUpdate<T> u1{DDX_ETD1, yc, nullptr, stage, coeff, vo, dt / 2.0, true};
        if ((rc = rhs_impl<T>(p, prm, y, d1, forcing, &u1, ws, xforcing, xin))) return rc;
        FusedUpdate<T> u2{DDX_ETD2RK2, yc, d1c, y, coeff, vo, dt, keep_d2};
        return rhs_impl<T>(p, prm, stage, keep_d2 ? d2 : nullptr, forcing, &u2, ws, xforcing);
      }
      // time_step.py:369-389
      FusedUpdate<T> u1{DDX_ETD1, yc, nullptr, y, coeff, vo, dt, true};
      if ((rc = rhs_impl<T>(p, prm, y, d1, forcing, &u1, ws, xforcing, xin))) return rc;
      FusedUpdate<T> u2{DDX_ETD2RK1, yc, d1c, y, coeff, vo, dt, keep_d2};
      return rhs_impl<T>(p, prm, y, keep_d2 ? d2 : nullptr, forcing, &u2, ws, xforcing);
    }
    if (integrator == DDX_RK2MID) {
      // time_step.py:285-305
      if ((rc = rhs_impl<T>(p, prm, y, d1, forcing, nullptr, ws, xforcing, xin))) return rc;
      if ((rc = update_impl<T>(p, DDX_ETD1, ns, y, stage, d1, nullptr, coeff, vo, dt / 2.0))) return rc;
      if ((rc = rhs_impl<T>(p, prm, stage, d2, forcing, nullptr, ws, xforcing))) return rc;
      if ((rc = update_impl<T>(p, DDX_ETD2RK2, ns, y, y, d1, d2, coeff, vo, dt))) return rc;
    } else {
      // time_step.py:369-389
      if ((rc = rhs_impl<T>(p, prm, y, d1, forcing, nullptr, ws, xforcing, xin))) return rc;
      if ((rc = update_impl<T>(p, DDX_ETD1, ns, y, y, d1, nullptr, coeff, vo, dt))) return rc;
      if ((rc = rhs_impl<T>(p, prm, y, d2, forcing, nullptr, ws, xforcing))) return rc;
      if ((rc = update_impl<T>(p, DDX_ETD2RK1, ns, y, y, d1, d2, coeff, vo, dt))) return rc;
    }
    return 0;
  };
  auto eager = [&]() -> int {
    for (int s = 0; s < nsteps; ++s)
      if ((rc = one_step())) return rc;
    return 0;
  };
  if (fixed_ws || !p->use_graph || p->prof || nsteps < 1) return eager();
  // Capture one step as a graph keyed by everything it depends on; replay it.
  std::vector<unsigned char> key;
  auto add = [&key](const void* q, size_t n) {
    const unsigned char* c = (const unsigned char*)q;
    key.insert(key.end(), c, c + n);
  };
  add(&integrator, sizeof(integrator));
  add(prm, sizeof(*prm));
  add(&dt, sizeof(dt));
  add(&keep_d2, sizeof(keep_d2));
  add(y, sizeof(void*) * ns);
  add(stage, sizeof(void*) * ns);
  add(d1, sizeof(void*) * ns);
  add(d2v, sizeof(void*) * ns);
  for (int i = 0; i < ns; ++i) {
    const void* f = forcing ? (const void*)forcing[i] : nullptr;
    const void* xf = xforcing ? (const void*)xforcing[i] : nullptr;
    add(&f, sizeof(f));
    add(&xf, sizeof(xf));
  }
  {
    PhysLayout L;
    const int nq = phys_layout(prm->physics, nd, &L) ? 0 : L.nq;
    for (int q = 0; q < nq; ++q) {
      const void* xi = xin ? (const void*)xin[q] : nullptr;
      add(&xi, sizeof(xi));
    }
  }
  add(&p->ws_gen, sizeof(p->ws_gen));
  add(&ws.A, sizeof(void*));
  add(&ws.B, sizeof(void*));
  add(&ws.F2, sizeof(void*));
  add(&p->stream, sizeof(void*));
  if (!p->step_exec || key != p->step_key) {
    // A single step whose key was not seen on the previous call runs
    // eagerly: a capture pays only once the same step repeats.
    if (nsteps == 1 && key != p->miss_key) {
      p->miss_key = key;
      return eager();
    }
    if (p->step_exec) {
      // the previous graph may still be running on the stream
      HIPCHK(hipStreamSynchronize(p->stream));
      HIPCHK(hipGraphExecDestroy(p->step_exec));
      p->step_exec = nullptr;
      p->step_key.clear();
    }
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    const int crc = one_step();
    const hipError_t ec = hipStreamEndCapture(p->stream, &g);
    if (crc) {
      if (g) (void)hipGraphDestroy(g);
      return crc;
    }
    if (ec != hipSuccess) return fail(DDX_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&p->step_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      p->step_exec = nullptr;
      return fail(DDX_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
    }
    p->step_key = key;
  }
  p->miss_key.clear();
  for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(p->step_exec, p->stream));
  return 0;
}

LayoutArgs layout_args(const ddx_plan* p) {
  LayoutArgs L;
  L.ndim = p->ndim;
  L.nx = p->ax.n;
  L.ny = p->ay.n;
  L.nz = p->az.n;
  L.Kx = p->Kx;
  L.Kxp = p->Kxp;
  L.Ky = p->Ky;
  L.Kz = p->Kz;
  L.cy = p->ay.c;
  L.cz = p->az.c;
  L.nk = p->nk;
  L.nref = p->nref;
  return L;
}

int check_plan(const ddx_plan* p) {
  if (!p) return fail(DDX_ERR_ARG, "null plan");
  return 0;
}

}  // namespace

namespace ddx {
// error reporting shared with the closure kernels' entry points
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace ddx

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

const char* ddx_last_error(void) { return g_err.c_str(); }

int ddx_version(void) { return 1; }

int ddx_init(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(DDX_ERR_NODEV, "no HIP device available");
  if (device < 0 || device >= n) return fail(DDX_ERR_ARG, "device index out of range");
  HIPCHK(hipSetDevice(device));
  if (device < kMaxDevices && !g_scratch[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(DDX_ERR_HIP, std::string("scratch symbol address: ") + hipGetErrorString(e));
    g_scratch[device] = ptr;
  }
  g_device = device;
  return 0;
}

}  // extern "C"

namespace {

// nbatch > 0: a 2D plan for nbatch independent members (z axis untransformed)
int plan_create(int ndim, const int64_t* xshape, const double* length, int precision, int nbatch,
                ddx_plan** out, int dealias = DDX_DEALIAS_23) {
  if (!out || !xshape || !length) return fail(DDX_ERR_ARG, "null argument");
  *out = nullptr;
  if (dealias != DDX_DEALIAS_23 && dealias != DDX_DEALIAS_SPHERICAL && dealias != DDX_DEALIAS_NONE)
    return fail(DDX_ERR_ARG, "Specified dealiasing method not implemented.");
  if (g_device < 0) return fail(DDX_ERR_NODEV, "ddx_init has not been called");
  if (ndim != 2 && ndim != 3) return fail(DDX_ERR_ARG, "Must use either 2 or 3 dimensions.");
  if (precision != DDX_F64 && precision != DDX_F32) return fail(DDX_ERR_ARG, "bad precision");
  for (int i = 0; i < ndim; ++i) {
    if (xshape[i] < 1 || xshape[i] > (1 << 20)) return fail(DDX_ERR_ARG, "bad grid extent");
    if (!(length[i] > 0.0)) return fail(DDX_ERR_ARG, "box lengths must be positive");
  }
  const int nx = (int)xshape[ndim - 1];
  if (nx % 2) return fail(DDX_ERR_UNSUPPORTED, "x extent must be even");
  ddx_plan* p = new ddx_plan();
  p->device = g_device;
  {
    const char* env = getenv("DDX_FORCE_GENERIC");
    p->generic_only = env && env[0] == '1';
    const char* ng = getenv("DDX_NO_GRAPH");
    p->use_graph = !(ng && ng[0] == '1');
    const char* nf = getenv("DDX_NO_FUSED_UPDATE");
    p->fuse_update = !(nf && nf[0] == '1');
    const char* ny = getenv("DDX_NO_FUSED_YEPI");
    p->fuse_yepi = !(ny && ny[0] == '1');
    const char* nz = getenv("DDX_NO_ZSPLIT");
    p->zsplit = !(nz && nz[0] == '1');
  }
  p->ndim = ndim;
  p->precision = precision;
  p->dealias = dealias;
  std::string why;
  int ny = (int)xshape[ndim - 2], nz = ndim == 3 ? (int)xshape[0] : 1;
  double lx = length[ndim - 1], ly = length[ndim - 2], lz = ndim == 3 ? length[0] : 1.0;
  if (dealias == DDX_DEALIAS_SPHERICAL) {
    // 2 / 3. * np.min(self.kny), kny = pi N / L per axis
    double kmin = (M_PI * (double)nx) / lx;
    kmin = std::min(kmin, (M_PI * (double)ny) / ly);
    if (ndim == 3) kmin = std::min(kmin, (M_PI * (double)nz) / lz);
    p->sph_thr = 2.0 / 3.0 * kmin;
  }
  const bool nyq = dealias == DDX_DEALIAS_NONE;
  if (setup_axis(nx, lx, true, &p->ax, &why, p->sph_thr, nyq) ||
      setup_axis(ny, ly, false, &p->ay, &why, p->sph_thr, nyq)) {
    delete p;
    return fail(DDX_ERR_UNSUPPORTED, why);
  }
  if (ndim == 3) {
    if (setup_axis(nz, lz, false, &p->az, &why, p->sph_thr, nyq)) {
      delete p;
      return fail(DDX_ERR_UNSUPPORTED, why);
    }
  } else if (nbatch > 0) {
    p->az.n = nbatch; p->az.L = 1.0; p->az.c = nbatch; p->az.K = nbatch;
    p->az.kfull.assign(nbatch, 0.0);
  } else {
    p->az.n = 1; p->az.L = 1.0; p->az.c = 0; p->az.K = 1; p->az.kfull.assign(1, 0.0);
  }
  p->zfft = ndim == 3;
  p->nbatch = nbatch > 0 ? nbatch : 1;
  if (make_fft_plan(p->ax.n, &p->fx) || make_fft_plan(p->ay.n, &p->fy) ||
      make_fft_plan(p->zfft ? p->az.n : 1, &p->fz)) {
    delete p;
    return fail(DDX_ERR_UNSUPPORTED, "grid extents must factor into 2, 3 and 5");
  }
  p->Kx = p->ax.K;
  p->Kxp = (p->Kx + 7) / 8 * 8;
  p->Ky = p->ay.K;
  p->Kz = p->az.K;
  p->nk = (long)p->Ky * p->Kz * p->Kxp;
  p->npts = (long)nx * ny * p->az.n;
  p->nper = (long)nx * ny * (p->zfft ? p->az.n : 1);
  p->nref = (long)(nx / 2 + 1) * ny * p->az.n;

  // compact k tables
  std::vector<double> hx(p->Kxp, 0.0), hy(p->Ky), hz(p->Kz);
  for (int j = 0; j < p->Kx; ++j) hx[j] = p->ax.kfull[j];
  for (int j = 0; j < p->Ky; ++j) hy[j] = p->ay.kfull[j <= p->ay.c ? j : j + p->ay.n - p->Ky];
  for (int j = 0; j < p->Kz; ++j) hz[j] = p->az.kfull[j <= p->az.c ? j : j + p->az.n - p->Kz];
  auto upload = [&](const std::vector<double>& h, double** d) -> int {
    HIPCHK(hipMalloc((void**)d, h.size() * sizeof(double)));
    p->owned.push_back(*d);
    HIPCHK(hipMemcpy(*d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return 0;
  };
  int rc;
  if ((rc = upload(hx, &p->dkx)) || (rc = upload(hy, &p->dky)) || (rc = upload(hz, &p->dkz))) {
    ddx_plan_destroy(p);
    return rc;
  }
  HIPCHK(hipStreamCreate(&p->own_stream));
  p->stream = p->own_stream;
  // FP64 twiddles generated on the device (sincospi); FP32 plans round them
  const int ns[3] = {p->ax.n, p->ay.n, p->zfft ? p->az.n : 1};
  for (int i = 0; i < 3; ++i) {
    const int n = ns[i];
    if (p->tw.count(n)) continue;
    void* t64 = nullptr;
    if (hipMalloc(&t64, (size_t)n * 16) != hipSuccess) {
      ddx_plan_destroy(p);
      return fail(DDX_ERR_NOMEM, "twiddle allocation failed");
    }
    p->owned.push_back(t64);
    hipLaunchKernelGGL(twiddle_kernel, dim3((n + 255) / 256), dim3(256), 0, p->stream,
                       (cpx<double>*)t64, n);
    if (precision == DDX_F64) {
      p->tw[n] = t64;
    } else {
      void* t32 = nullptr;
      if (hipMalloc(&t32, (size_t)n * 8) != hipSuccess) {
        ddx_plan_destroy(p);
        return fail(DDX_ERR_NOMEM, "twiddle allocation failed");
      }
      p->owned.push_back(t32);
      hipLaunchKernelGGL(twiddle_f32_kernel, dim3((n + 255) / 256), dim3(256), 0, p->stream,
                         (const cpx<double>*)t64, (cpx<float>*)t32, n);
      p->tw[n] = t32;
    }
  }
  if (hipStreamSynchronize(p->stream) != hipSuccess) {
    ddx_plan_destroy(p);
    return fail(DDX_ERR_HIP, "plan setup failed");
  }
  *out = p;
  return 0;
}

}  // namespace

extern "C" {

int ddx_plan_create(int ndim, const int64_t* xshape, const double* length, int precision,
                    ddx_plan** out) {
  return plan_create(ndim, xshape, length, precision, 0, out);
}

int ddx_plan_create_ex(int ndim, const int64_t* xshape, const double* length, int precision,
                       int dealias, ddx_plan** out) {
  return plan_create(ndim, xshape, length, precision, 0, out, dealias);
}

int ddx_plan_create_batched(int nbatch, int64_t ny, int64_t nx, const double* length, int precision,
                            ddx_plan** out) {
  if (nbatch < 1 || nbatch > (1 << 20)) return fail(DDX_ERR_ARG, "bad member count");
  const int64_t xs[2] = {ny, nx};
  return plan_create(2, xs, length, precision, nbatch, out);
}

int ddx_plan_destroy(ddx_plan* p) {
  if (!p) return 0;
  if (p->own_stream) (void)hipStreamSynchronize(p->own_stream);
  for (void* q : p->owned) (void)hipFree(q);
  if (p->wsA) (void)hipFree(p->wsA);
  if (p->wsB) (void)hipFree(p->wsB);
  if (p->partial) (void)hipFree(p->partial);
  if (p->step_exec) (void)hipGraphExecDestroy(p->step_exec);
  for (auto& r : p->recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (hipEvent_t e : p->evfree) (void)hipEventDestroy(e);
  if (p->own_stream) (void)hipStreamDestroy(p->own_stream);
  delete p;
  return 0;
}

int ddx_plan_kshape(const ddx_plan* p, int64_t* kref, int64_t* kc, int64_t* cut, int64_t* nkc) {
  if (check_plan(p)) return DDX_ERR_ARG;
  const int hx = p->ax.n / 2 + 1;
  if (kref) {
    if (p->ndim == 2) { kref[0] = hx; kref[1] = p->ay.n; kref[2] = p->nbatch; }
    else { kref[0] = p->ay.n; kref[1] = p->az.n; kref[2] = hx; }
  }
  if (kc) { kc[0] = p->Ky; kc[1] = p->Kz; kc[2] = p->Kxp; }
  if (cut) { cut[0] = p->ax.c; cut[1] = p->ay.c; cut[2] = p->az.c; }
  if (nkc) *nkc = p->nk;
  return 0;
}

int ddx_plan_ktable(const ddx_plan* p, int axis, double* out, int64_t n) {
  if (check_plan(p)) return DDX_ERR_ARG;
  const AxisInfo* a = axis == 0 ? &p->ax : axis == 1 ? &p->ay : axis == 2 ? &p->az : nullptr;
  if (!a || !out) return fail(DDX_ERR_ARG, "bad axis");
  if (n != a->K) return fail(DDX_ERR_ARG, "ktable length mismatch");
  for (int j = 0; j < a->K; ++j) out[j] = a->kfull[(axis == 0 || j <= a->c) ? j : j + a->n - a->K];
  return 0;
}

int ddx_plan_set_graph(ddx_plan* p, int on) {
  if (check_plan(p)) return DDX_ERR_ARG;
  if (p->step_exec) {
    HIPCHK(hipStreamSynchronize(p->stream));
    HIPCHK(hipGraphExecDestroy(p->step_exec));
    p->step_exec = nullptr;
  }
  p->use_graph = on != 0;
  return 0;
}

int ddx_plan_set_tuning(ddx_plan* p, int knob, int value) {
  if (check_plan(p)) return DDX_ERR_ARG;
  const int top = knob == DDX_TUNE_XCD_ORDER ? 2 : 1;
  if (knob < DDX_TUNE_WIDE_TILES || knob > DDX_TUNE_RADIX || value < -1 || value > top)
    return fail(DDX_ERR_ARG, "unknown tuning knob or value");
  if (p->step_exec) {
    HIPCHK(hipStreamSynchronize(p->stream));
    HIPCHK(hipGraphExecDestroy(p->step_exec));
    p->step_exec = nullptr;
  }
  if (knob == DDX_TUNE_WIDE_TILES) p->tune_wide = value;
  else if (knob == DDX_TUNE_XCD_ORDER) p->tune_xcd = value;
  else if (knob == DDX_TUNE_FUSED_YEPI) p->tune_yepi = value;
  else p->tune_radix = value;
  return 0;
}

int ddx_plan_set_stream(ddx_plan* p, void* stream) {
  if (check_plan(p)) return DDX_ERR_ARG;
  p->stream = stream ? (hipStream_t)stream : p->own_stream;
  return 0;
}

int ddx_alloc(size_t bytes, void** ptr) {
  if (!ptr) return fail(DDX_ERR_ARG, "null pointer");
  if (g_device < 0) return fail(DDX_ERR_NODEV, "ddx_init has not been called");
  *ptr = nullptr;
  if (hipMalloc(ptr, bytes ? bytes : 16) != hipSuccess) return fail(DDX_ERR_NOMEM, "hipMalloc failed");
  return 0;
}

int ddx_free(void* ptr) {
  if (ptr) HIPCHK(hipFree(ptr));
  return 0;
}

int ddx_h2d(ddx_plan* p, void* dst, const void* src, size_t bytes) {
  if (check_plan(p) || !dst || !src) return fail(DDX_ERR_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return 0;
}

int ddx_d2h(ddx_plan* p, void* dst, const void* src, size_t bytes) {
  if (check_plan(p) || !dst || !src) return fail(DDX_ERR_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return 0;
}

int ddx_zero(ddx_plan* p, void* dst, size_t bytes) {
  if (check_plan(p) || !dst) return fail(DDX_ERR_ARG, "null argument");
  if (bytes == 0) return 0;
  if (bytes % 4 == 0 && ((uintptr_t)dst % 4) == 0) {
    const size_t nw = bytes / 4;
    const size_t blocks = std::min<size_t>((nw + 255) / 256, 65536);
    hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, p->stream, (uint32_t*)dst, nw);
  } else {
    hipLaunchKernelGGL(zero_bytes_kernel, dim3((bytes + 255) / 256), dim3(256), 0, p->stream,
                       (uint8_t*)dst, bytes);
  }
  LAUNCHCHK();
  return 0;
}

int ddx_copy(ddx_plan* p, void* dst, const void* src, size_t bytes) {
  if (check_plan(p) || !dst || !src) return fail(DDX_ERR_ARG, "null argument");
  if (bytes == 0) return 0;
  ProfScope prof(p, PC_COPY, 2.0 * (double)bytes);
  if (bytes % 16 == 0 && ((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0) {
    // one 16-byte chunk per thread: the fastest plain copy measured here
    // (6.2 TB/s read + write for 1 GiB; a grid of 8192 blocks looping over
    // the chunks reached 4.5)
    const size_t n = bytes / 16;
    const size_t blocks = std::min<size_t>((n + 255) / 256, (size_t)1 << 30);
    hipLaunchKernelGGL(copy16_kernel, dim3(blocks), dim3(256), 0, p->stream, (const uint4*)src,
                       (uint4*)dst, n);
  } else {
    hipLaunchKernelGGL(copy1_kernel, dim3((bytes + 255) / 256), dim3(256), 0, p->stream,
                       (const uint8_t*)src, (uint8_t*)dst, bytes);
  }
  LAUNCHCHK();
  return 0;
}

int ddx_sync(ddx_plan* p) {
  if (check_plan(p)) return DDX_ERR_ARG;
  HIPCHK(hipStreamSynchronize(p->stream));
  return 0;
}

#define DDX_DISPATCH(p, fn, ...) \
  ((p)->precision == DDX_F64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

extern "C++" {
template <typename T>
static int pack_t(ddx_plan* p, const void* kref, void* kc) {
  LayoutArgs L = layout_args(p);
  ProfScope prof(p, PC_LAYOUT, (double)p->nk * (16 + sizeof(cpx<T>)));
  hipLaunchKernelGGL((pack_kernel<T>), dim3((p->nk + 255) / 256), dim3(256), 0, p->stream, L,
                     (const cpx<double>*)kref, (cpx<T>*)kc);
  LAUNCHCHK();
  return 0;
}
}  // extern "C++"

int ddx_pack(ddx_plan* p, const void* kref, void* kc) {
  if (check_plan(p) || !kref || !kc) return fail(DDX_ERR_ARG, "null argument");
  return DDX_DISPATCH(p, pack_t, p, kref, kc);
}

extern "C++" {
template <typename T>
static int unpack_t(ddx_plan* p, const void* kc, void* kref) {
  LayoutArgs L = layout_args(p);
  ProfScope prof(p, PC_LAYOUT, (double)p->nref * 16 + (double)p->nk * sizeof(cpx<T>));
  hipLaunchKernelGGL((unpack_kernel<T>), dim3((p->nref + 255) / 256), dim3(256), 0, p->stream,
                     L, (const cpx<T>*)kc, (cpx<double>*)kref);
  LAUNCHCHK();
  return 0;
}
}  // extern "C++"

int ddx_unpack(ddx_plan* p, const void* kc, void* kref) {
  if (check_plan(p) || !kref || !kc) return fail(DDX_ERR_ARG, "null argument");
  return DDX_DISPATCH(p, unpack_t, p, kc, kref);
}

int ddx_fwd(ddx_plan* p, const void* x, void* kc) {
  if (check_plan(p) || !x || !kc) return fail(DDX_ERR_ARG, "null argument");
  return ddx_fwd_many(p, 1, &x, &kc);
}

int ddx_bwd(ddx_plan* p, const void* kc, void* x) {
  if (check_plan(p) || !x || !kc) return fail(DDX_ERR_ARG, "null argument");
  return ddx_bwd_many(p, 1, &kc, &x);
}

int ddx_fwd_many(ddx_plan* p, int nf, const void* const* x, void* const* kc) {
  if (check_plan(p) || !x || !kc) return fail(DDX_ERR_ARG, "null argument");
  if (nf < 1 || nf > DDX_MAXF) return fail(DDX_ERR_ARG, "1..16 fields");
  for (int f = 0; f < nf; ++f)
    if (!x[f] || !kc[f]) return fail(DDX_ERR_ARG, "null field pointer");
  if (p->precision == DDX_F64)
    return fwd_impl<double>(p, nf, (const double* const*)x, (cpx<double>* const*)kc);
  return fwd_impl<float>(p, nf, (const float* const*)x, (cpx<float>* const*)kc);
}

int ddx_bwd_many(ddx_plan* p, int nf, const void* const* kc, void* const* x) {
  if (check_plan(p) || !x || !kc) return fail(DDX_ERR_ARG, "null argument");
  if (nf < 1 || nf > DDX_MAXF) return fail(DDX_ERR_ARG, "1..16 fields");
  for (int f = 0; f < nf; ++f)
    if (!x[f] || !kc[f]) return fail(DDX_ERR_ARG, "null field pointer");
  if (p->precision == DDX_F64)
    return bwd_impl<double>(p, nf, (const cpx<double>* const*)kc, (double* const*)x);
  return bwd_impl<float>(p, nf, (const cpx<float>* const*)kc, (float* const*)x);
}

int ddx_bwd_combine(ddx_plan* p, int nsrc, const void* const* kc, int nout, const int* nterm,
                    const int* src, const int* dim, const double* coef, void* const* x) {
  if (check_plan(p) || !kc || !x || !nterm || !src || !dim || !coef)
    return fail(DDX_ERR_ARG, "null argument");
  if (nsrc < 1 || nsrc > DDX_MAXF || nout < 1 || nout > DDX_MAXF)
    return fail(DDX_ERR_ARG, "1..16 sources and outputs");
  for (int s = 0; s < nsrc; ++s)
    if (!kc[s]) return fail(DDX_ERR_ARG, "null source");
  Recipe rec[DDX_MAXF];
  memset(rec, 0, sizeof(rec));
  for (int q = 0; q < nout; ++q) {
    if (!x[q]) return fail(DDX_ERR_ARG, "null output");
    if (nterm[q] < 1 || nterm[q] > 2) return fail(DDX_ERR_ARG, "1 or 2 terms per output");
    rec[q].nterm = nterm[q];
    for (int t = 0; t < nterm[q]; ++t) {
      const int i = 2 * q + t;
      if (src[i] < 0 || src[i] >= nsrc) return fail(DDX_ERR_ARG, "term source out of range");
      if (dim[i] < -1 || dim[i] >= p->ndim) return fail(DDX_ERR_ARG, "term dimension out of range");
      rec[q].t[t].src = src[i];
      rec[q].t[t].dim = dim[i];
      rec[q].t[t].coef = coef[i];
    }
  }
  if (p->precision == DDX_F64)
    return bwd_combine_impl<double>(p, nsrc, (const cpx<double>* const*)kc, nout, rec, (double* const*)x);
  return bwd_combine_impl<float>(p, nsrc, (const cpx<float>* const*)kc, nout, rec, (float* const*)x);
}

int ddx_dealias(ddx_plan* p, void* kc) {
  if (check_plan(p) || !kc) return fail(DDX_ERR_ARG, "null argument");
  if (p->precision == DDX_F64) {
    cpx<double>* f = (cpx<double>*)kc;
    return sphere_mask<double>(p, 1, &f);
  }
  cpx<float>* f = (cpx<float>*)kc;
  return sphere_mask<float>(p, 1, &f);
}

int ddx_deriv(ddx_plan* p, const void* kc, void* out, int dim) {
  if (check_plan(p) || !kc || !out) return fail(DDX_ERR_ARG, "null argument");
  if (dim < 0 || dim >= p->ndim) return fail(DDX_ERR_ARG, "bad derivative dimension");
  if (p->precision == DDX_F64)
    hipLaunchKernelGGL((deriv_kernel<double>), dim3((p->nk + 255) / 256), dim3(256), 0, p->stream,
                       kgeom(p), p->nk, dim, (const cpx<double>*)kc, (cpx<double>*)out);
  else
    hipLaunchKernelGGL((deriv_kernel<float>), dim3((p->nk + 255) / 256), dim3(256), 0, p->stream,
                       kgeom(p), p->nk, dim, (const cpx<float>*)kc, (cpx<float>*)out);
  LAUNCHCHK();
  return 0;
}

extern "C++" {
template <typename T>
static int project_t(ddx_plan* p, void* const* comps) {
  ProjArgs<T> a;
  memset(&a, 0, sizeof(a));
  for (int i = 0; i < p->ndim; ++i) {
    if (!comps[i]) return fail(DDX_ERR_ARG, "null component");
    a.c[i] = (cpx<T>*)comps[i];
  }
  a.g = kgeom(p);
  a.nk = p->nk;
  a.ndim = p->ndim;
  hipLaunchKernelGGL((project_kernel<T>), dim3((p->nk + 255) / 256), dim3(256), 0, p->stream, a);
  LAUNCHCHK();
  return 0;
}
}  // extern "C++"

int ddx_project(ddx_plan* p, void* const* comps) {
  if (check_plan(p) || !comps) return fail(DDX_ERR_ARG, "null argument");
  return DDX_DISPATCH(p, project_t, p, comps);
}

int ddx_update(ddx_plan* p, int kind, const void* y, void* out, const void* F1, const void* F2,
               double coeff, int order, double dt) {
  if (check_plan(p)) return DDX_ERR_ARG;
  if (p->precision == DDX_F64) {
    const cpx<double>* yy = (const cpx<double>*)y;
    cpx<double>* oo = (cpx<double>*)out;
    const cpx<double>* f1 = (const cpx<double>*)F1;
    const cpx<double>* f2 = (const cpx<double>*)F2;
    return update_impl<double>(p, kind, 1, &yy, &oo, &f1, &f2, &coeff, order, dt);
  }
  const cpx<float>* yy = (const cpx<float>*)y;
  cpx<float>* oo = (cpx<float>*)out;
  const cpx<float>* f1 = (const cpx<float>*)F1;
  const cpx<float>* f2 = (const cpx<float>*)F2;
  return update_impl<float>(p, kind, 1, &yy, &oo, &f1, &f2, &coeff, order, dt);
}

int ddx_update_if(ddx_plan* p, int kind, const void* y, void* out, const void* F1, const void* F2,
                  const double* intfactor, double dt) {
  if (check_plan(p)) return DDX_ERR_ARG;
  if (!intfactor) return fail(DDX_ERR_ARG, "null integrating factor");
  const double coeff = 0.0;
  if (p->precision == DDX_F64) {
    const cpx<double>* yy = (const cpx<double>*)y;
    cpx<double>* oo = (cpx<double>*)out;
    const cpx<double>* f1 = (const cpx<double>*)F1;
    const cpx<double>* f2 = (const cpx<double>*)F2;
    return update_impl<double>(p, kind, 1, &yy, &oo, &f1, &f2, &coeff, 1, dt, intfactor);
  }
  const cpx<float>* yy = (const cpx<float>*)y;
  cpx<float>* oo = (cpx<float>*)out;
  const cpx<float>* f1 = (const cpx<float>*)F1;
  const cpx<float>* f2 = (const cpx<float>*)F2;
  return update_impl<float>(p, kind, 1, &yy, &oo, &f1, &f2, &coeff, 1, dt, intfactor);
}

int ddx_rhs(ddx_plan* p, const ddx_params* prm, const void* const* state, void* const* deriv,
            const void* const* forcing) {
  if (check_plan(p) || !prm || !state || !deriv) return fail(DDX_ERR_ARG, "null argument");
  if (p->dealias == DDX_DEALIAS_NONE) return fail(DDX_ERR_UNSUPPORTED, kNoFusedAliased);
  const int nf = rhs_ws_fields(prm->physics, p->ndim);
  if (nf < 0) return fail(DDX_ERR_ARG, "unknown physics");
  int rc = ensure_ws(p, nf);
  if (rc) return rc;
  Ws ws;
  ws.A = p->wsA;
  ws.B = p->wsB;
  if (p->precision == DDX_F64)
    return rhs_impl<double>(p, prm, (const cpx<double>* const*)state, (cpx<double>* const*)deriv,
                            (const cpx<double>* const*)forcing, nullptr, ws);
  return rhs_impl<float>(p, prm, (const cpx<float>* const*)state, (cpx<float>* const*)deriv,
                         (const cpx<float>* const*)forcing, nullptr, ws);
}

}  // extern "C"

namespace {

int step_dispatch(ddx_plan* p, int integrator, const ddx_params* prm, void* const* state,
                  void* const* stage, void* const* deriv1, void* const* deriv2,
                  const void* const* forcing, int nsteps, double dt, const Ws* fixed_ws,
                  const void* const* xforcing = nullptr, const void* const* xin = nullptr) {
  if (check_plan(p) || !prm || !state || !deriv1) return fail(DDX_ERR_ARG, "null argument");
  if (p->dealias == DDX_DEALIAS_NONE) return fail(DDX_ERR_UNSUPPORTED, kNoFusedAliased);
  if (integrator == DDX_RK2MID && !stage) return fail(DDX_ERR_ARG, "RK2mid needs stage buffers");
  if (nsteps < 0) return fail(DDX_ERR_ARG, "negative step count");
  void* const* st = stage ? stage : state;
  if (p->precision == DDX_F64)
    return step_impl<double>(p, integrator, prm, (cpx<double>* const*)state, (cpx<double>* const*)st,
                             (cpx<double>* const*)deriv1, (cpx<double>* const*)deriv2,
                             (const cpx<double>* const*)forcing, nsteps, dt, fixed_ws,
                             (const double* const*)xforcing, (const double* const*)xin);
  return step_impl<float>(p, integrator, prm, (cpx<float>* const*)state, (cpx<float>* const*)st,
                          (cpx<float>* const*)deriv1, (cpx<float>* const*)deriv2,
                          (const cpx<float>* const*)forcing, nsteps, dt, fixed_ws,
                          (const float* const*)xforcing, (const float* const*)xin);
}

}  // namespace

extern "C" {

int ddx_step(ddx_plan* p, int integrator, const ddx_params* prm, void* const* state,
             void* const* stage, void* const* deriv1, void* const* deriv2,
             const void* const* forcing, int nsteps, double dt) {
  return step_dispatch(p, integrator, prm, state, stage, deriv1, deriv2, forcing, nsteps, dt, nullptr);
}

int ddx_step_xforced(ddx_plan* p, int integrator, const ddx_params* prm, void* const* state,
                     void* const* stage, void* const* deriv1, void* const* deriv2,
                     const void* const* forcing, const void* const* xforcing, int nsteps,
                     double dt) {
  if (check_plan(p) || !prm) return fail(DDX_ERR_ARG, "null argument");
  if (xforcing) {
    const int ns = nstate_of(prm, p->ndim);
    if (ns < 0) return fail(DDX_ERR_ARG, "unknown physics");
    const int nforce = p->ndim + (prm->physics == DDX_BOUSSINESQ ? 1 : 0);
    for (int i = nforce; i < ns; ++i)
      if (xforcing[i]) return fail(DDX_ERR_ARG, "x-space forcing applies to the velocity and T only");
  }
  return step_dispatch(p, integrator, prm, state, stage, deriv1, deriv2, forcing, nsteps, dt, nullptr,
                       xforcing);
}

int ddx_step_xin(ddx_plan* p, int integrator, const ddx_params* prm, void* const* state,
                 void* const* stage, void* const* deriv1, void* const* deriv2,
                 const void* const* forcing, const void* const* xforcing, const void* const* xin,
                 double dt) {
  if (check_plan(p) || !prm || !xin) return fail(DDX_ERR_ARG, "null argument");
  PhysLayout L;
  if (phys_layout(prm->physics, p->ndim, &L)) return fail(DDX_ERR_ARG, "unknown physics");
  if (p->ndim != 2) return fail(DDX_ERR_UNSUPPORTED, "x-space step inputs are for 2D grids");
  for (int q = 0; q < L.nq; ++q)
    if (!xin[q]) return fail(DDX_ERR_ARG, "null x-space input");
  if (xforcing) {
    const int ns = L.nstate;
    const int nforce = p->ndim + (prm->physics == DDX_BOUSSINESQ ? 1 : 0);
    for (int i = nforce; i < ns; ++i)
      if (xforcing[i]) return fail(DDX_ERR_ARG, "x-space forcing applies to the velocity and T only");
  }
  return step_dispatch(p, integrator, prm, state, stage, deriv1, deriv2, forcing, 1, dt, nullptr,
                       xforcing, xin);
}

int ddx_step_noalloc(ddx_plan* p, int integrator, const ddx_params* prm, void* const* state,
                     void* const* stage, void* const* deriv1, void* const* deriv2,
                     const void* const* forcing, int nsteps, double dt) {
  if (check_plan(p) || !prm) return fail(DDX_ERR_ARG, "null argument");
  const int nf = rhs_ws_fields(prm->physics, p->ndim);
  const int ns = nstate_of(prm, p->ndim);
  if (nf < 0 || ns < 0) return fail(DDX_ERR_ARG, "unknown physics");
  void* base = nullptr;
  int rc = scratch_base(p->device, &base);
  if (rc) return rc;
  // A | B | F2 (F2 only when the sweeps run unfused and no deriv2 is given)
  const size_t half = scratch_round((size_t)nf * ws_field_stride(p) * cpx_bytes(p));
  const size_t f2 = (!deriv2 && !p->fuse_update) ? scratch_round((size_t)ns * p->nk * cpx_bytes(p)) : 0;
  const size_t need = 2 * half + f2;
  if (need > kScratchBytes)
    return fail(DDX_ERR_ARG, "ddx_step_noalloc: this grid and physics need " + std::to_string(need) +
                                 " bytes of scratch; the static scratch holds " +
                                 std::to_string(kScratchBytes) + " (use ddx_step)");
  Ws ws;
  ws.A = base;
  ws.B = (unsigned char*)base + half;
  ws.F2 = f2 ? (unsigned char*)base + 2 * half : nullptr;
  // per-launch timing would create HIP events: off for this call
  const bool prof = p->prof;
  p->prof = false;
  rc = step_dispatch(p, integrator, prm, state, stage, deriv1, deriv2, forcing, nsteps, dt, &ws);
  p->prof = prof;
  return rc;
}

extern "C++" {
template <typename T>
static int kcombine_t(ddx_plan* p, void* out, int nterm, const void* const* src, const double* coef,
                      const int* op, int accumulate) {
  KCombArgs<T> a;
  memset(&a, 0, sizeof(a));
  for (int t = 0; t < nterm; ++t) {
    if (!src[t]) return fail(DDX_ERR_ARG, "null source");
    if (op[t] < -1 || op[t] > 3 || (op[t] >= 0 && op[t] <= 2 && op[t] >= p->ndim))
      return fail(DDX_ERR_ARG, "bad operator");
    a.src[t] = (const cpx<T>*)src[t];
    a.coef[t] = coef[t];
    a.op[t] = op[t];
  }
  a.nterm = nterm;
  a.accumulate = accumulate;
  a.out = (cpx<T>*)out;
  a.g = kgeom(p);
  a.nk = p->nk;
  ProfScope prof(p, PC_POINTWISE,
                 (double)p->Ky * p->Kz * p->Kx * (nterm + 1 + accumulate) * sizeof(cpx<T>));
  hipLaunchKernelGGL((kcombine_kernel<T>), dim3((p->nk + 255) / 256), dim3(256), 0, p->stream, a);
  LAUNCHCHK();
  return 0;
}
}  // extern "C++"

int ddx_kcombine(ddx_plan* p, void* out, int nterm, const void* const* src, const double* coef,
                 const int* op, int accumulate) {
  if (check_plan(p) || !out || (nterm > 0 && (!src || !coef || !op)))
    return fail(DDX_ERR_ARG, "null argument");
  if (nterm < 0 || nterm > DDX_MAXT) return fail(DDX_ERR_ARG, "1..4 terms");
  return DDX_DISPATCH(p, kcombine_t, p, out, nterm, src, coef, op, accumulate);
}

extern "C++" {
template <typename T>
static int xcombine_t(ddx_plan* p, void* out, int nterm, const void* const* a_, const void* const* b_,
                      const double* coef, int accumulate) {
  XCombArgs<T> a;
  memset(&a, 0, sizeof(a));
  int nb = 0;
  for (int t = 0; t < nterm; ++t) {
    if (!a_[t]) return fail(DDX_ERR_ARG, "null operand");
    a.a[t] = (const T*)a_[t];
    a.b[t] = b_ ? (const T*)b_[t] : nullptr;
    nb += a.b[t] ? 1 : 0;
    a.coef[t] = coef[t];
  }
  a.nterm = nterm;
  a.accumulate = accumulate;
  a.out = (T*)out;
  a.n = p->npts;
  ProfScope prof(p, PC_POINTWISE, (double)p->npts * (nterm + nb + 1 + accumulate) * sizeof(T));
  hipLaunchKernelGGL((xcombine_kernel<T>), dim3((p->npts + 255) / 256), dim3(256), 0, p->stream, a);
  LAUNCHCHK();
  return 0;
}
}  // extern "C++"

int ddx_xcombine(ddx_plan* p, void* out, int nterm, const void* const* a_, const void* const* b_,
                 const double* coef, int accumulate) {
  if (check_plan(p) || !out || (nterm > 0 && (!a_ || !coef))) return fail(DDX_ERR_ARG, "null argument");
  if (nterm < 0 || nterm > DDX_MAXT) return fail(DDX_ERR_ARG, "1..4 terms");
  return DDX_DISPATCH(p, xcombine_t, p, out, nterm, a_, b_, coef, accumulate);
}

static int finish_reduce(ddx_plan* p, int nblk, int op, double* out) {
  p->hpartial.resize(nblk);
  HIPCHK(hipMemcpyAsync(p->hpartial.data(), p->partial, nblk * sizeof(double), hipMemcpyDeviceToHost,
                        p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  double r = 0.0;
  for (int i = 0; i < nblk; ++i) r = (op == 0) ? std::max(r, p->hpartial[i]) : r + p->hpartial[i];
  *out = r;
  return 0;
}

static int ensure_partial(ddx_plan* p) {
  if (!p->partial && hipMalloc((void**)&p->partial, 1024 * sizeof(double)) != hipSuccess)
    return fail(DDX_ERR_NOMEM, "reduction buffer");
  return 0;
}

int ddx_xreduce(ddx_plan* p, const void* x, int op, double* out) {
  if (check_plan(p) || !x || !out) return fail(DDX_ERR_ARG, "null argument");
  if (op < 0 || op > 2) return fail(DDX_ERR_ARG, "bad reduction");
  int rc = ensure_partial(p);
  if (rc) return rc;
  const int nblk = (int)std::min<long>(1024, (p->npts + 255) / 256);
  if (p->precision == DDX_F64)
    hipLaunchKernelGGL((xreduce_kernel<double>), dim3(nblk), dim3(256), 0, p->stream, (const double*)x,
                       p->npts, op, p->partial);
  else
    hipLaunchKernelGGL((xreduce_kernel<float>), dim3(nblk), dim3(256), 0, p->stream, (const float*)x,
                       p->npts, op, p->partial);
  LAUNCHCHK();
  return finish_reduce(p, nblk, op, out);
}

int ddx_kreduce(ddx_plan* p, const void* k0, const void* k1, int op, double* out) {
  if (check_plan(p) || !k0 || !out || (op == 1 && !k1)) return fail(DDX_ERR_ARG, "null argument");
  if (op < 0 || op > 2) return fail(DDX_ERR_ARG, "bad reduction");
  int rc = ensure_partial(p);
  if (rc) return rc;
  const int nblk = (int)std::min<long>(1024, (p->nk + 255) / 256);
  if (p->precision == DDX_F64)
    hipLaunchKernelGGL((kreduce_kernel<double>), dim3(nblk), dim3(256), 0, p->stream,
                       (const cpx<double>*)k0, (const cpx<double>*)k1, kgeom(p), p->nk, op, p->partial);
  else
    hipLaunchKernelGGL((kreduce_kernel<float>), dim3(nblk), dim3(256), 0, p->stream,
                       (const cpx<float>*)k0, (const cpx<float>*)k1, kgeom(p), p->nk, op, p->partial);
  LAUNCHCHK();
  return finish_reduce(p, nblk, 1, out);
}

int ddx_kreduce_members(ddx_plan* p, const void* k0, const void* k1, int op, double* out) {
  if (check_plan(p) || !k0 || !out || (op == 1 && !k1)) return fail(DDX_ERR_ARG, "null argument");
  if (op < 0 || op > 2) return fail(DDX_ERR_ARG, "bad reduction");
  if (p->ndim != 2 || p->zfft) return fail(DDX_ERR_ARG, "per-member reductions need a 2D plan");
  const int nb = p->az.n;
  if (p->precision == DDX_F64)
    hipLaunchKernelGGL((kreduce_members_kernel<double>), dim3(nb), dim3(256), 0, p->stream,
                       (const cpx<double>*)k0, (const cpx<double>*)k1, kgeom(p), op, out);
  else
    hipLaunchKernelGGL((kreduce_members_kernel<float>), dim3(nb), dim3(256), 0, p->stream,
                       (const cpx<float>*)k0, (const cpx<float>*)k1, kgeom(p), op, out);
  LAUNCHCHK();
  return 0;
}

int ddx_prof_enable(ddx_plan* p, int on) {
  if (check_plan(p)) return DDX_ERR_ARG;
  HIPCHK(hipStreamSynchronize(p->stream));
  for (auto& r : p->recs) {
    p->evfree.push_back(r.a);
    p->evfree.push_back(r.b);
  }
  p->recs.clear();
  p->prof = on != 0;
  return 0;
}

int ddx_prof_read(ddx_plan* p, int max_cat, int* ncat, char* names, int64_t* launches,
                  double* ms_total, double* bytes_total) {
  if (check_plan(p) || !ncat || !names || !launches || !ms_total || !bytes_total)
    return fail(DDX_ERR_ARG, "null argument");
  if (max_cat < PC_NCAT) return fail(DDX_ERR_ARG, "max_cat too small");
  HIPCHK(hipStreamSynchronize(p->stream));
  for (int c = 0; c < PC_NCAT; ++c) {
    strncpy(names + 32 * c, kCatNames[c], 31);
    names[32 * c + 31] = 0;
    launches[c] = 0;
    ms_total[c] = 0.0;
    bytes_total[c] = 0.0;
  }
  for (auto& r : p->recs) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
    launches[r.cat] += 1;
    ms_total[r.cat] += ms;
    bytes_total[r.cat] += r.bytes;
    p->evfree.push_back(r.a);
    p->evfree.push_back(r.b);
  }
  p->recs.clear();
  *ncat = PC_NCAT;
  return 0;
}

}  // extern "C"
