// MI355X (gfx950) kernels, host launchers and plan state of the orbital-free DFT hot path.
//
// One evaluation of E[rho] and V[rho] is five launches on the plan stream:
//   rows_fwd   : pre-pointwise (rho, sqrt(rho), rho^beta ...) fused with the
//                axis-2 real-to-complex transform (two real rows per complex
//                sequence), writes F half spectra
//   axis_c2c   : axis-1 complex transform, W-column tiles staged in LDS
//   pencil     : axis-0 forward transform, reciprocal multiplier
//                (4 pi/G^2, G^2, WT kernel), axis-0 inverse transform -- one
//                launch, the spectrum is never written in between
//   axis_c2c   : axis-1 inverse transform
//   rows_inv   : axis-2 complex-to-real transform fused with the
//                post-pointwise functional (TF, vW, WT, LDA, Hartree, V_loc),
//                writes V and per-block FP64 energy partial sums
// plus a tiny reduction launch.  All fields of one evaluation (up to three
// spectra) travel through every launch together.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <atomic>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/ofdft_amd.h"
#include "fft_lds.h"
#include "fft_reg.h"

using namespace ofdft;

// last error of the calling thread (defined once, in ofdft_api.hip)
namespace ofdft_shared {
extern thread_local std::string g_err;
}

namespace {

using ofdft_shared::g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(OFDFT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
  } while (0)

#define LAUNCHCHK()                                                                      \
  do {                                                                                   \
    hipError_t e_ = hipGetLastError();                                                   \
    if (e_ != hipSuccess)                                                                \
      return fail(OFDFT_EHIP, std::string("kernel launch: ") + hipGetErrorString(e_));   \
  } while (0)

constexpr double kPi = 3.14159265358979323846;
constexpr double kZero = 1e-30;  // constants.py:85
constexpr int kRedThreads = 1024;
constexpr size_t kMaxLdsBytes = 160 * 1024;  // LDS a gfx950 workgroup may allocate

// ---------------------------------------------------------------------------
// host-side pass schedule and twiddles
// ---------------------------------------------------------------------------
// `asc`: passes in ascending radix order (FP32 plans), as make_sched.
bool make_pass(int n, Pass1D* pl, bool asc) {
  pl->n = n;
  pl->npass = 0;
  int m = n;
  auto push = [&](int r) {
    if (pl->npass >= kMaxPass) return false;
    pl->radix[pl->npass++] = r;
    return true;
  };
  int c8 = 0, c4 = 0, c2 = 0, c3 = 0, c5 = 0, c7 = 0, c11 = 0, c13 = 0;
  while (m % 8 == 0) { ++c8; m /= 8; }
  while (m % 4 == 0) { ++c4; m /= 4; }
  while (m % 2 == 0) { ++c2; m /= 2; }
  while (m % 3 == 0) { ++c3; m /= 3; }
  while (m % 5 == 0) { ++c5; m /= 5; }
  while (m % 7 == 0) { ++c7; m /= 7; }
  while (m % 11 == 0) { ++c11; m /= 11; }
  while (m % 13 == 0) { ++c13; m /= 13; }
  // 7, 11 and 13 (the other factors of the reference's max_prime = 13
  // grids) run last, after the order of the 2/3/5 passes
  const int order[8] = {asc ? 3 : 8, asc ? 5 : 4, 2, asc ? 4 : 3, asc ? 8 : 5, 7, 11, 13};
  for (int r : order) {
    const int c = r == 8 ? c8 : r == 4 ? c4 : r == 2 ? c2 : r == 3 ? c3 : r == 5 ? c5 : r == 7 ? c7 : r == 11 ? c11 : c13;
    for (int i = 0; i < c; ++i)
      if (!push(r)) return false;
  }
  return m == 1;
}

// the schedule holds a radix 11 or 13 pass (runtime kernels with larger
// butterfly registers)
bool has_large(const Pass1D& pl) {
  for (int i = 0; i < pl.npass; ++i)
    if (pl.radix[i] > 8) return true;
  return false;
}

int max_butterflies(const Pass1D& pl) {
  int mx = 0;
  for (int i = 0; i < pl.npass; ++i) mx = std::max(mx, pl.n / pl.radix[i]);
  return mx;
}

struct FFTLaunch {
  int threads;
  int kb;
};

FFTLaunch fft_launch(const Pass1D& pl, int ntr, int min_threads = 64) {
  long total = (long)ntr * max_butterflies(pl);
  long th = std::max<long>(min_threads, ((total + 63) / 64) * 64);
  if (th > 1024) th = 1024;
  int kb = (int)((total + th - 1) / th);
  if (kb == 3) kb = 4;
  return {(int)th, kb};
}

}  // namespace

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
struct ofdft_plan_s {
  int n[3];
  int nh;            // n2/2 + 1 logical half-spectrum length
  int nhp;           // row stride of the internal spectrum (nh rounded up to a multiple of 16)
  long long N, Nh, Nhp;
  int idx32;  // every field and the 3-field spectrum fit 2^31 bytes: 32-bit element offsets
  int prec;
  int full;   // reference full=True semantics (Hermitian-symmetrised multipliers)
  size_t rs;  // bytes per real
  double lat[9], bg[9], volume, dV;
  hipStream_t stream;
  Pass1D pl[3];
  void* tw[3];      // twiddle tables, plan precision complex
  void* spec;       // 3*Nh complex workspace
  void* wtk;        // WT kernel table, Nh reals (plan precision)
  void* wtks;       // the same table in split rows (n2 even): n0*n1 rows of n2/2, then the Nyquist plane [k0][i1]
  void* vloc;       // N reals
  double* partials; // per-block partial sums
  long long npartials;
  double* red;      // device, kMaxSum doubles
  double* hred;     // pinned host mirror
  double* dhred;    // device address of hred (reduction kernels write final sums there)
  void* tnst;       // device TnState of ofdft_tn_direction
  void* htnst;      // pinned host mirror of it
  void* dhtnst;     // device address of htnst: the control kernels publish the state there
  const int* skip;  // while set, FFT and reduction kernels return at once when *skip != 0
  ofdft_functional fn;
  int have_fn, have_vloc, have_wtk;
  // ion state of the last ofdft_vloc_build (local PP forces reuse it)
  int nspecies, nat, pme, order;
  double* d_knots;   // all species' knots, concatenated
  double* d_coefs;   // all species' coefficients, concatenated
  void* d_splines;   // SplineDev[nspecies]
  double* d_frac;    // nat x 3 wrapped fractional coordinates
  double* d_pos;     // nat x 3 Cartesian positions (bohr)
  int* d_species;    // nat
  double* d_bm;      // PME b-moduli: complex n0 + n1 + nbm2
  double* d_work;    // N doubles: PME Q array / rhoPB
  void* d_gga;       // 4 N (plan precision): GGA pot1 and vector field
  // LWT / LMGP: kernel tables (knots and coefficients on the device), the
  // kf node list, and the workspace B | Acc (spectra) | pot1 | pot3 (fields)
  ofdft_lwt_table lwt_tab[2];
  double* d_lwttab;
  void* d_lwt;
  // per-kernel-class event timing (ofdft_plan_set_profiling): classes in
  // prof_on, one launch in every prof_stride of a class timed
  int prof_on;
  int prof_stride = 1;
  long long prof_seq[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<hipEvent_t> ev_pool;  // reused across profile windows
  size_t ev_used = 0;
  struct ProfRec { int cls; long long bytes; int e0, e1, tag; };
  std::vector<ProfRec> prof_pending;
  int prof_tag = -1;  // CG iteration a launch belongs to (-1: none)
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;  // events of the launch being profiled
  double prof_ms[8];
  long long prof_count[8];
  double prof_bytes[8];
  // hipGraph of one device CG iteration of ofdft_tn_direction
  // (ofdft_plan_set_graph): replayed once per queued iteration in place of
  // its launches; rebuilt when the iteration's buffers change
  // columns per workgroup of the register column kernels (axis-1 pass and
  // axis-0 pencils): 8, 16 or 32 (FP32), 8 or 16 (FP64); chosen per grid
  // length and precision by the autotuner (ofdft_plan_set_tiles)
  int tile_axis = 16, tile_pencil = 16;
  // spectra of the TN inner loop's Hessian products (ofdft_plan_set_cg_spectra):
  // 0 plan precision, 1 bf16 (FP32 plans of compile-time lengths, see
  // cg_bf16_tiles), with tile_bf_* columns per workgroup of the axis pass
  // and of the pencils (0: automatic)
  int cg_spectra = 0;
  int tile_bf_axis = 0, tile_bf_pencil = 0;
  int use_graph = 0;
  hipGraphExec_t tn_exec = nullptr;
  hipEvent_t tn_ev = nullptr;  // the TN loop's check point, one iteration behind the queue
  const void* tn_key[8] = {};
  hipStream_t cap_stream = nullptr;  // capture happens here (the null stream cannot capture)
};

// A group of plans whose TN inner loops run in one persistent launch
// (ofdft_batch_create / ofdft_batch_tn_direction).
struct ofdft_batch_s {
  std::vector<ofdft_plan_s*> plans;
  int prec = 0;
  int use_graph = 1;            // 1: graph form (per-length stage launches), 0: one cooperative launch
  void* d_cells = nullptr;      // per-plan descriptors
  long long* d_pref = nullptr;  // work-item prefix sums of the stages
  int* d_idx = nullptr;         // graph form: the cells of each length class
  size_t cells_cap = 0, pref_cap = 0, idx_cap = 0;
  void* h_states = nullptr;     // pinned, mapped: every cell's TnState after a call
  void* d_states = nullptr;
  size_t states_cap = 0;
  int* h_live = nullptr;        // pinned, mapped: cells still iterating (graph form)
  int* d_live = nullptr;
  hipStream_t cap_stream = nullptr;  // graph form: capture stream
  hipGraphExec_t exec = nullptr;     // graph form: one CG iteration of every cell
  std::vector<long long> key;        // launch shapes and buffers the graph was captured with
  int last_iters = 0;                // longest loop of the previous call (first chunk of the next)
};

namespace {

constexpr int kMaxSum = 16;
// row pairs per rows_* workgroup
template <typename T>
constexpr int kRowPB = 2;
constexpr int kRowPBMin = 2;
constexpr int kColTileC = 16;   // W: contiguous columns per axis-1 / axis-0 tile

// kernel classes reported by ofdft_plan_get_profile (OFDFT_NPROF of them):
// one per kernel variant that a solve runs
enum { PC_ROWS_FWD = 0, PC_ROWS_FWD_HVP = 1, PC_AXIS = 2, PC_PENCIL = 3, PC_ROWS_INV = 4, PC_ROWS_INV_HVP = 5,
       PC_VEC = 6, PC_NCLS = 7 };

// Event bracket around one launch (profiling only).
// Times one kernel launch: the launch made inside the scope passes
// P->ev_start / P->ev_stop to hipExtLaunchKernel, which stamps them from the
// dispatch itself (the kernel's own start and end, no gap around it).
struct ProfScope {
  ofdft_plan_s* P;
  int cls;
  long long bytes;
  int e0 = -1;
  ProfScope(ofdft_plan_s* p, int c, long long b) : P(p), cls(c), bytes(b) {
    if (!(P->prof_on & (1 << c))) return;
    if (P->prof_seq[c]++ % P->prof_stride != 0) return;  // a sampled launch of the class
    while (P->ev_pool.size() < P->ev_used + 2) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      P->ev_pool.push_back(e);
    }
    e0 = (int)P->ev_used;
    P->ev_used += 2;
    P->ev_start = P->ev_pool[e0];
    P->ev_stop = P->ev_pool[e0 + 1];
  }
  ~ProfScope() {
    if (e0 < 0) return;
    P->ev_start = P->ev_stop = nullptr;
    P->prof_pending.push_back({cls, bytes, e0, e0 + 1, P->prof_tag});
  }
};

// launch on the plan stream, timed when a ProfScope is open
#define OFDFT_LAUNCH(K, G, B, S, P, ...) \
  hipExtLaunchKernelGGL(K, G, B, S, (P)->stream, (P)->ev_start, (P)->ev_stop, 0u, __VA_ARGS__)

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T dev_cbrt(T x);
template <>
__device__ __forceinline__ double dev_cbrt<double>(double x) { return cbrt(x); }
// FP32 transcendentals use the hardware's 1-ulp approximations (v_log_f32 =
// log2, v_exp_f32 = 2^x, v_sqrt_f32, v_rcp_f32): the FP32 path accumulates
// every energy in FP64 and is refined in FP64, and the row transforms that
// evaluate these functions are bound by instruction issue.  FP64 keeps the
// correctly rounded library functions.
template <>
__device__ __forceinline__ float dev_cbrt<float>(float x) {
  return x > 0.f ? __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.f / 3.f)) : cbrtf(x);
}

template <typename T>
__device__ __forceinline__ T dev_pow(T x, T y);
template <>
__device__ __forceinline__ double dev_pow<double>(double x, double y) { return pow(x, y); }
template <>
__device__ __forceinline__ float dev_pow<float>(float x, float y) {
  return x > 0.f ? __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)) : powf(x, y);
}

template <typename T>
__device__ __forceinline__ T dev_log(T x);
template <>
__device__ __forceinline__ double dev_log<double>(double x) { return log(x); }
template <>
__device__ __forceinline__ float dev_log<float>(float x) {
  return x > 0.f ? __builtin_amdgcn_logf(x) * 0.693147180559945309f : logf(x);
}

template <typename T>
__device__ __forceinline__ T dev_sqrt(T x);
template <>
__device__ __forceinline__ double dev_sqrt<double>(double x) { return sqrt(x); }
template <>
__device__ __forceinline__ float dev_sqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

// a / b: IEEE division in FP64, a * rcp(b) in FP32
template <typename T>
__device__ __forceinline__ T dev_div(T a, T b);
template <>
__device__ __forceinline__ double dev_div<double>(double a, double b) { return a / b; }
template <>
__device__ __forceinline__ float dev_div<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// sign(x) |x|^(2/3) in FP32 for x != 0: phi rho^(beta-1) at beta = 5/6 with
// rho = phi^2, two transcendental instructions instead of the five of
// sqrt, cbrt and the reciprocals.
__device__ __forceinline__ float sgn_pow23(float x) {
  const float a = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(__builtin_fabsf(x)) * (2.f / 3.f));
  return x < 0.f ? -a : a;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sums of NS doubles; result written by thread 0 to out[0..NS).
template <int NS>
__device__ void block_sums(double (&acc)[NS], double* __restrict__ out) {
  __shared__ double red[16][NS];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double v = wave_sum(acc[s]);
    if (lane == 0) red[wave][s] = v;
  }
  __syncthreads();
  if (threadIdx.x < NS) {
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += red[w][threadIdx.x];
    out[threadIdx.x] = t;
  }
}

// Element i of a field.  With a 32-bit index the byte offset is formed in
// 32 bits, so that the access takes the scalar-base + 32-bit vector-offset
// form of global memory instructions instead of a 64-bit address per
// lane; used where the plan guarantees every field and the spectrum fit in
// 2^31 bytes (ofdft_plan_s::idx32).
template <typename T>
__device__ __forceinline__ T& at(T* p, long long i) {
  return p[i];
}
template <typename T>
__device__ __forceinline__ T& at(T* p, unsigned i) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + i * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T& at(const T* p, unsigned i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + i * (unsigned)sizeof(T));
}

// bf16 storage of spectra (the TN inner loop's Hessian products on FP32
// plans, when selected): one complex value in 32 bits, the real part in the
// low half and the imaginary part in the high half, each rounded to nearest
// even from FP32 by one v_cvt_pk_bf16_f32; the transforms compute in FP32.
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned pack_bf16(cplx<float> v) {
  const f32x2_t f = {v.x, v.y};
  return __builtin_bit_cast(unsigned, __builtin_convertvector(f, bf16x2_t));
}
__device__ __forceinline__ cplx<float> unpack_bf16(unsigned u) {
  return cplx<float>{__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u)};
}

// A spectrum value as stored in layout SL: SL = 0 (padded rows) and SL = 2
// (split rows) keep the plan precision, SL = 1 (split rows) a bf16 pair.
template <typename T, int SL>
struct SpecElem {
  using E = cplx<T>;
  static __device__ __forceinline__ E put(cplx<T> v) { return v; }
  static __device__ __forceinline__ cplx<T> get(E e) { return e; }
};
template <>
struct SpecElem<float, 1> {
  using E = unsigned;
  static __device__ __forceinline__ E put(cplx<float> v) { return pack_bf16(v); }
  static __device__ __forceinline__ cplx<float> get(E e) { return unpack_bf16(e); }
};

// Select one of three per-point values without a dynamically indexed array.
template <typename T>
__device__ __forceinline__ T pick3(const T* c, int f) {
  return f == 0 ? c[0] : (f == 1 ? c[1] : c[2]);
}

template <typename T>
__device__ __forceinline__ void put3(T* out, int f, T v) {
  if (f == 0) out[0] = v;
  else if (f == 1) out[1] = v;
  else if (f == 2) out[2] = v;
}

// ---------------------------------------------------------------------------
// pointwise functionals (reference formulas, see each citation)
// ---------------------------------------------------------------------------
struct FnParams {
  int tf, vw, wt, lda, hartree, pseudo;
  int beta56;  // beta == 5/6: rho^beta = sqrt(rho) cbrt(rho), no pow
  double x, y, alpha, beta;
  int fH, fV, fW;  // spectrum slot of each convolution field (-1 if absent)
  int F;
};

// Spectrum slots and the beta = 5/6 flag of a Hessian-product functor.
// LAY = 1: the fixed layout of the full functional (Hartree, vW and WT
// fields in slots 0, 1, 2, beta = 5/6), so that every slot test and
// register select folds at compile time; LAY = 0: read from FnParams.
template <int LAY>
struct FieldSlots {
  __device__ __forceinline__ static int fH(const FnParams& p) { return LAY ? 0 : p.fH; }
  __device__ __forceinline__ static int fV(const FnParams& p) { return LAY ? 1 : p.fV; }
  __device__ __forceinline__ static int fW(const FnParams& p) { return LAY ? 2 : p.fW; }
  __device__ __forceinline__ static bool b56(const FnParams& p) { return LAY ? true : p.beta56 != 0; }
};

inline bool fixed_layout(const FnParams& p) {
  return p.F == 3 && p.fH == 0 && p.fV == 1 && p.fW == 2 && p.beta56;
}

// LDA exchange-correlation energy density per electron and potential:
// xc/semilocal_xc.py:287-320.
template <typename T>
__device__ __forceinline__ void lda_point(T rho, T rcbrt, T& ex, T& vxc) {
  const T a = T(0.0311), b = T(-0.048), c = T(0.0020), d = T(-0.0116);
  const T g = T(-0.1423), b1 = T(1.0529), b2 = T(0.3334);
  T rc = rcbrt;
  if (rc < T(1e-30)) rc = T(1e-30);
  const T rs = dev_div(T(0.62035049089940001667), rc);  // cbrt(3/(4 pi)) / cbrt(rho)
  ex = T(-0.73855876638202240588) * rc;         // -3/4 cbrt(3/pi)
  vxc = T(-0.98474502184269654118) * rc;        // cbrt(-3/pi)
  if (rs < T(1)) {
    const T lr = dev_log(rs);
    ex += a * lr + b + c * rs * lr + d * rs;
    vxc += lr * (a + T(2.0 / 3.0) * c * rs) + b - T(1.0 / 3.0) * a + T(1.0 / 3.0) * (T(2) * d - c) * rs;
  } else {
    const T sr = dev_sqrt(rs);
    const T den = T(1) + b1 * sr + b2 * rs;
    ex += dev_div(g, den);
    vxc += dev_div(g + (T(7.0 / 6.0) * g * b1) * sr + (T(4.0 / 3.0) * g * b2 * rs), den * den);
  }
}

// d V_xc / d rho: xc/semilocal_xc.py:321-333.
template <typename T>
__device__ __forceinline__ T lda_v2(T rho, T rcbrt) {
  const T a = T(0.0311), c = T(0.0020), d = T(-0.0116);
  const T g = T(-0.1423), b1 = T(1.0529), b2 = T(0.3334);
  T rc0 = rcbrt;
  T rc = rc0 < T(1e-30) ? T(1e-30) : rc0;
  const T rs = T(0.62035049089940001667) / rc;
  const T fx = T(-0.98474502184269654118) / T(3) * rc0 / rho;
  T fc;
  if (rs < T(1)) {
    fc = -a / T(3) - (c / T(9) * (dev_log(rs) * T(2) + T(1)) + d * T(2.0 / 9.0)) * rs;
  } else {
    const T ta = b1 * dev_sqrt(rs), tb = b2 * rs;
    const T den = T(1) + ta + tb;
    fc = g / T(36) * (T(5) * ta + T(7) * ta * ta + T(8) * tb + T(16) * tb * tb + T(21) * ta * tb) / den / den / den;
  }
  return fx + fc / rho;
}

// Inverse Lindhard function with the reference branch precedence
// (kedf/kernel.py:14-136): large > |eta-1| small > eta small > normal.
__device__ double lindhard_dev(double eta, double lbda, double mu) {
  const double tol = 1e-10, tol2 = 1e-11;
  if (eta > 3.65) {
    const double t = 1.0 / (eta * eta);
    const double c[15] = {-0.13714285714285712,     -6.39999999999999875e-2, -3.77825602968460128e-2,
                          -2.51824061652633074e-2,  -1.80879839616166146e-2, -1.36715733124818332e-2,
                          -1.07236045520990083e-2,  -8.65192783339199453e-3, -7.1372762502456763e-3,
                          -5.9945117538835746e-3,   -5.10997527675418131e-3, -4.41060829979912465e-3,
                          -3.84763737842981233e-3,  -3.38745061493813488e-3, -3.00624946457977689e-3};
    double acc = c[14];
    for (int i = 13; i >= 0; --i) acc = c[i] + t * acc;
    return 3.0 * (1.0 - mu) * eta * eta - lbda - 0.6 + t * acc;
  }
  if (fabs(eta - 1.0) < tol) return 2.0 - lbda - 3.0 * mu + 20.0 * (eta - 1.0);
  if (eta < tol) return 1.0 - lbda + eta * eta * (1.0 / 3.0 - 3.0 * mu);
  if (eta > tol2 && fabs(eta - 1.0) > tol2)
    return 1.0 / (0.5 + 0.25 * (1.0 - eta * eta) * log((1.0 + eta) / fabs(1.0 - eta)) / eta) - 3.0 * mu * eta * eta -
           lbda;
  return eta;
}

struct GridGeom {
  int n0, n1, n2, nh;
  int nhp;   // row stride of the internal spectrum
  int full;  // 1: frequencies of the full complex grid (grid.py:584-620, full=True)
  double bg[9];
  float bgf[9];
};

__device__ __forceinline__ double gsq_m(const GridGeom& g, double m0, double m1, double m2) {
  const double gx = m0 * g.bg[0] + m1 * g.bg[3] + m2 * g.bg[6];
  const double gy = m0 * g.bg[1] + m1 * g.bg[4] + m2 * g.bg[7];
  const double gz = m0 * g.bg[2] + m1 * g.bg[5] + m2 * g.bg[8];
  return gx * gx + gy * gy + gz * gz;
}

__device__ __forceinline__ int fftfreq_i(int k, int n) { return (k <= (n - 1) / 2) ? k : k - n; }

// |G|^2 at half-spectrum index (k0, k1, k2).  full=0: rfftfreq on axis 2.
// full=1: fftfreq on every axis; partner=1 gives |G|^2 at the index -k
// (mod n), whose frequency equals -m except on Nyquist planes.
__device__ __forceinline__ double gsq(const GridGeom& g, int k0, int k1, int k2, int partner = 0) {
  int m0 = fftfreq_i(k0, g.n0), m1 = fftfreq_i(k1, g.n1);
  int m2 = g.full ? fftfreq_i(k2, g.n2) : k2;
  if (partner) {
    m0 = fftfreq_i((g.n0 - k0) % g.n0, g.n0);
    m1 = fftfreq_i((g.n1 - k1) % g.n1, g.n1);
    m2 = fftfreq_i((g.n2 - k2) % g.n2, g.n2);
  }
  return gsq_m(g, m0, m1, m2);
}

// ---------------------------------------------------------------------------
// pre-pointwise functors (feed rows_fwd)
// ---------------------------------------------------------------------------
template <typename T>
struct PreCopy {
  static constexpr int kProf = PC_ROWS_FWD;
  const T* __restrict__ in;
  __host__ __device__ int nin() const { return 1; }
  struct In {
    T v;
  };
  __device__ __forceinline__ In load(long long i) const { return {in[i]}; }
  __device__ __forceinline__ void apply(const In& x, T* out) const { out[0] = x.v; }
  __device__ __forceinline__ void operator()(long long i, T* out) const { apply(load(i), out); }
  __device__ __forceinline__ void apply(const In& x, T* out, long long) const { apply(x, out); }
};

// rho (mode 0) or phi(t) = c*phi + s*p (mode 1), then the convolution inputs
// rho (Hartree), sqrt(max(rho,ZERO)) (vW, vw.py:34-44), max(rho,ZERO)^beta
// (WT, wt.py:15-19).
template <typename T>
struct PreEval {
  static constexpr int kProf = PC_ROWS_FWD;
  const T* __restrict__ rho;
  const T* __restrict__ phi;
  const T* __restrict__ p;
  T c, s;
  FnParams fp;
  __host__ __device__ int nin() const { return rho ? 1 : (p ? 2 : 1); }
  // raw loads only (no arithmetic on them), so that callers can issue the
  // loads of several points before the first use
  struct In {
    T a, b;
  };
  template <class I>
  __device__ __forceinline__ In load(I i) const {
    if (rho) return {at(rho, i), T(0)};
    return {at(phi, i), p ? at(p, i) : T(0)};
  }
  __device__ __forceinline__ T density(const In& x) const {
    if (rho) return x.a;
    const T ph = p ? c * x.a + s * x.b : x.a;
    return ph * ph;
  }
  __device__ __forceinline__ void operator()(long long i, T* out) const { apply(load(i), out); }
  template <class I>
  __device__ __forceinline__ void apply(const In& x, T* out, I) const { apply(x, out); }
  __device__ __forceinline__ void apply(const In& x, T* out) const {
    const T r = density(x);
    if (fp.fH >= 0) put3(out, fp.fH, r);
    const T s = dev_sqrt(r > T(0) ? r : T(kZero));
    if (fp.fV >= 0) put3(out, fp.fV, s);
    if (fp.fW >= 0) {
      const T rz = r < T(kZero) ? T(kZero) : r;
      const T rb = (fp.beta56 && r >= T(kZero)) ? s * dev_cbrt(rz) : dev_pow(rz, T(fp.beta));
      put3(out, fp.fW, rb);
    }
  }
};

// HVP inputs for direction p at phi: drho = 2 phi p (Hartree), ds = drho/(2s)
// (vW), beta*rho^(beta-1)*drho (WT).
template <typename T, int LAY = 0>
struct PreHvp {
  using S = FieldSlots<LAY>;
  static constexpr int kProf = PC_ROWS_FWD_HVP;
  const T* __restrict__ phi;
  const T* __restrict__ p;
  FnParams fp;
  __host__ __device__ int nin() const { return 2; }
  struct In {
    T ph, pp;
  };
  template <class I>
  __device__ __forceinline__ In load(I i) const { return {at(phi, i), at(p, i)}; }
  __device__ __forceinline__ void operator()(long long i, T* out) const { apply(load(i), out); }
  template <class I>
  __device__ __forceinline__ void apply(const In& x, T* out, I) const { apply(x, out); }
  __device__ __forceinline__ void apply(const In& x, T* out) const {
    const T ph = x.ph, pp = x.pp;
    const T r = ph * ph;
    const T dr = T(2) * ph * pp;
    if constexpr (sizeof(T) == 4) {
      // the same fields with sqrt(rho) = |phi| cancelled: drho / (2 sqrt(rho))
      // = sign(phi) p, beta rho^(beta-1) drho = 2 beta sign(phi) |phi|^(2/3) p
      if (S::fH(fp) >= 0) put3(out, S::fH(fp), dr);
      if (S::fV(fp) >= 0) put3(out, S::fV(fp), r > T(0) ? (ph > T(0) ? pp : -pp) : T(0));
      if (S::fW(fp) >= 0) {
        const T w = S::b56(fp) ? T(2 * fp.beta) * sgn_pow23(ph) * pp
                               : T(fp.beta) * dev_pow(r, T(fp.beta) - T(1)) * dr;
        put3(out, S::fW(fp), r < T(kZero) ? T(0) : w);
      }
      return;
    }
    const T s = dev_sqrt(r);
    if (S::fH(fp) >= 0) put3(out, S::fH(fp), dr);
    if (S::fV(fp) >= 0) put3(out, S::fV(fp), r > T(0) ? dev_div(dr, T(2) * s) : T(0));
    if (S::fW(fp) >= 0) {
      const T rb1 = S::b56(fp) ? dev_div(dev_cbrt(r), s) : dev_pow(r, T(fp.beta) - T(1));  // rho^(beta-1)
      put3(out, S::fW(fp), r < T(kZero) ? T(0) : T(fp.beta) * rb1 * dr);
    }
  }
};

// ---------------------------------------------------------------------------
// post-pointwise functors (consume rows_inv); c[f] are the normalised
// convolution results of the F fields at point i.
// ---------------------------------------------------------------------------
template <typename T>
struct PostStore {
  static constexpr int NS = 1;
  static constexpr int kProf = PC_ROWS_INV;
  T* __restrict__ out;
  __host__ __device__ int nio() const { return 1; }
  struct In {};
  __device__ __forceinline__ In load(long long) const { return {}; }
  __device__ __forceinline__ void apply(long long i, const In&, const T* c, double*) const { out[i] = c[0]; }
  __device__ __forceinline__ void operator()(long long i, const T* c, double* acc) const { apply(i, In{}, c, acc); }
};

// rho^a (numpy power: rho ** a) as the row transform's input, for the WT
// stress transforms (wt.py:86-87)
template <typename T>
struct PrePow {
  static constexpr int kProf = PC_ROWS_FWD;
  const T* __restrict__ rho;
  T a;
  __host__ __device__ int nin() const { return 1; }
  struct In {
    T v;
  };
  __device__ __forceinline__ In load(long long i) const { return {rho[i]}; }
  __device__ __forceinline__ void apply(const In& x, T* out) const { out[0] = pow(x.v, a); }
  __device__ __forceinline__ void apply(const In& x, T* out, long long) const { apply(x, out); }
  __device__ __forceinline__ void operator()(long long i, T* out) const { apply(load(i), out); }
};

// the six products d_a rho d_b rho / rho of the three gradient fields
// (vonWeizsackerStress, kedf/vw.py:97-112); c[a] = d_a rho at point i
template <typename T>
struct PostStressVW {
  static constexpr int NS = 6;
  static constexpr int kProf = PC_ROWS_INV;
  const T* __restrict__ rho;
  __host__ __device__ int nio() const { return 1; }
  struct In {
    T r;
  };
  __device__ __forceinline__ In load(long long i) const { return {rho[i]}; }
  __device__ __forceinline__ void apply(long long, const In& in, const T* c, double* acc) const {
    const double inv = 1.0 / (double)in.r;
    acc[0] += (double)c[0] * c[0] * inv;
    acc[1] += (double)c[0] * c[1] * inv;
    acc[2] += (double)c[0] * c[2] * inv;
    acc[3] += (double)c[1] * c[1] * inv;
    acc[4] += (double)c[1] * c[2] * inv;
    acc[5] += (double)c[2] * c[2] * inv;
  }
  __device__ __forceinline__ void operator()(long long i, const T* c, double* acc) const { apply(i, load(i), c, acc); }
};

// ---------------------------------------------------------------------------
// GGA kinetic functionals (kedf/gga.py): enhancement factor F(s) and
// dFds2 = (1/s) dF/ds of the reduced gradient s = |grad rho| / rho^(4/3);
// the factors marked "ss" take ss = s / TKF0 and return dFds2 / TKF0^2
// (s_with_tkf0, kedf/gga.py:18-40).  FP64 in every precision.
// ---------------------------------------------------------------------------
constexpr double kCTF = 2.87123400018819181594;   // constants.py:82
constexpr double kTKF0 = 6.18733545256027186194;  // constants.py:83
constexpr double kCbrt2 = 1.25992104989487316477; // constants.py:84

struct GgaSpec {
  int id;  // OFDFT_GGA_*
  int np;
  double p[8];
};

__device__ void gga_enhance(const GgaSpec& g, double s, double& F, double& d) {
  const double* p = g.p;
  bool tk = true;
  double x = s / kTKF0;  // ss
  switch (g.id) {
    case OFDFT_GGA_LKTVW: {  // :42-55
      const double e1 = exp(-p[0] * x), e2 = exp(-2.0 * p[0] * x);
      F = 2.0 * e1 / (1.0 + e2);
      d = -p[0] * (1.0 - e2) / (1.0 + e2) * F / x;
      break;
    }
    case OFDFT_GGA_LKT: {  // :547-575
      const double s2 = x * x;
      if (x > 100.0) {
        F = 5.0 / 3.0 * s2 * p[1];
        d = 10.0 / 3.0 * p[1];
      } else if (x < 1e-5) {
        const double a2 = p[0] * p[0];
        F = 1.0 + (5.0 / 3.0 * p[1] - 0.5 * a2) * s2 + 5.0 / 24.0 * a2 * a2 * s2 * s2;
        d = 10.0 / 3.0 * p[1] - a2 + 5.0 / 6.0 * a2 * a2 * s2 - 61.0 / 120.0 * a2 * a2 * a2 * s2 * s2;
      } else {
        const double ch = cosh(p[0] * x);
        F = 1.0 / ch + 5.0 / 3.0 * s2 * p[1];
        d = 10.0 / 3.0 * p[1] - p[0] * sinh(p[0] * x) / (ch * ch) / x;
      }
      break;
    }
    case OFDFT_GGA_DK: {  // :58-74 (s)
      tk = false;
      const double t = s * s / (72.0 * kCTF);
      const double Fa = 9.0 * p[5] * t * t * t * t + p[2] * t * t * t + p[1] * t * t + p[0] * t + 1.0;
      const double Fb = p[5] * t * t * t + p[4] * t * t + p[3] * t + 1.0;
      F = Fa / Fb;
      d = ((36.0 * p[5] * t * t * t + 3 * p[2] * t * t + 2 * p[1] * t + p[0]) / Fb -
           Fa / (Fb * Fb) * (3.0 * p[5] * t * t + 2.0 * p[4] * t + p[3])) / (36.0 * kCTF);
      break;
    }
    case OFDFT_GGA_LLP: {  // :77-95 (s)
      tk = false;
      const double bs = kCbrt2 * s, bs2 = bs * bs;
      const double Fa = p[0] * bs2, Fb = 1.0 + p[1] * bs * asinh(bs);
      F = 1.0 + Fa / Fb;
      d = (2.0 * p[0] / Fb - (p[0] * p[1] * bs * asinh(bs) + Fa * p[1] / sqrt(1.0 + bs2)) / (Fb * Fb)) * kCbrt2 * kCbrt2;
      break;
    }
    case OFDFT_GGA_OL: {  // :98-112 (s)
      tk = false;
      F = 1.0 + s * s / 72.0 / kCTF + p[0] / kCTF * s / (1 + p[1] * 4 * s);
      d = 1.0 / 36.0 / kCTF;
      if (s > 1.0e-8) {
        const double q = 1 + p[1] * 4 * s;
        d += p[0] / kCTF / (q * q) / s;
      }
      break;
    }
    case OFDFT_GGA_THAK: {  // :115-142 (s)
      tk = false;
      const double bs = kCbrt2 * s, bs2 = bs * bs;
      const double c53 = 3.1748021039363989495;  // 2^(5/3)
      F = 1.0 + p[0] * bs2 / (1.0 + p[1] * bs * asinh(bs)) - p[2] * bs / (1.0 + c53 * bs);
      const double fb = 1.0 + p[1] * bs * asinh(bs);
      d = (-(p[0] * p[1] * bs2) / sqrt(1 + bs2) + (p[0] * p[1] * bs * asinh(bs)) + 2.0 * p[0]) / (fb * fb);
      if (s > 1.0e-8) {
        const double q = 1.0 + c53 * bs;
        d -= p[2] / (q * q) / bs;
      }
      d *= kCbrt2 * kCbrt2;
      break;
    }
    case OFDFT_GGA_B86A: {  // :145-159 (s)
      tk = false;
      const double bs2 = kCbrt2 * s * kCbrt2 * s, Fb = 1.0 + p[1] * bs2;
      F = 1.0 + p[0] * bs2 / Fb;
      d = 2 * p[0] / (Fb * Fb) * (kCbrt2 * kCbrt2);
      break;
    }
    case OFDFT_GGA_B86B: {  // :162-176 (s)
      tk = false;
      const double bs2 = kCbrt2 * s * kCbrt2 * s, q = 1.0 + p[1] * bs2, Fb = pow(q, 4.0 / 5.0);
      F = 1.0 + p[0] * bs2 / Fb;
      d = (2 * p[0] * (p[1] * bs2 + 5.0)) / (5 * q * Fb) * (kCbrt2 * kCbrt2);
      break;
    }
    case OFDFT_GGA_DK87: {  // :179-195 (s)
      tk = false;
      const double bs = kCbrt2 * s, bs2 = bs * bs;
      const double Fa = p[0] * bs2 * (1 + p[1] * bs), Fb = 1.0 + p[2] * bs2;
      F = 1.0 + Fa / Fb;
      d = p[0] * (2.0 + 3.0 * p[1] * bs + p[1] * p[2] * bs2 * bs) / (Fb * Fb) * (kCbrt2 * kCbrt2);
      break;
    }
    case OFDFT_GGA_PW86: {  // :198-213 (ss)
      const double s2 = x * x, Fa = 1.0 + p[0] * s2 + p[1] * s2 * s2 + p[2] * s2 * s2 * s2;
      F = pow(Fa, 1.0 / 15.0);
      d = 2.0 / 15.0 * (p[0] + 2 * p[1] * s2 + 3 * p[2] * s2 * s2) / pow(Fa, 14.0 / 15.0);
      break;
    }
    case OFDFT_GGA_PW91O: {  // :216-237 (ss)
      const double s2 = x * x, s4 = s2 * s2, ex = exp(-p[3] * s2);
      const double Fa = (p[1] - p[2] * ex) * s2 - p[5] * s4;
      const double Fb = 1.0 + p[0] * x * asinh(p[4] * x) + p[5] * s4;
      const double Fa_s2 = (p[1] - p[2] * ex) - p[5] * s2;
      F = 1.0 + Fa / Fb;
      d = 2.0 * (p[1] + (p[3] * s2 - 1) * p[2] * ex - 4.0 * p[5] * s2) / Fb -
          Fa_s2 * x / (Fb * Fb) * ((p[0] * p[4] * x) / (p[4] * p[4] * s2 + 1) + p[0] * asinh(p[4] * x) + 4.0 * p[5] * s2);
      break;
    }
    case OFDFT_GGA_PW91: {  // :240-277 (ss); dFds2 as the reference writes it
      const double s2 = x * x, ex = exp(-p[3] * s2), ash = asinh(p[4] * x);
      const double Fa = 1.0 + p[0] * x * ash + (p[1] - p[2] * ex) * s2;
      const double Fb = 1.0 + p[0] * x * ash + p[5] * s2 * s2;
      const double sq = sqrt(p[4] * p[4] * s2 + 1);
      const double Fa_s = p[0] * p[4] * x / sq + p[0] * ash + 2.0 * x * (p[1] + 2.0 * p[2] * p[3] * s2 * ex - 2.0 * p[2] * ex);
      F = Fa / Fb;
      d = Fa_s / Fb - Fa / (Fb * Fb) * (p[0] * p[4] * x / sq + (p[0] * ash + 4.0 * p[5] * s2 * x));
      if (x > 1.0e-8) d /= x;
      break;
    }
    case OFDFT_GGA_LG94: {  // :280-316 (ss)
      const double s2 = x * x, s4 = s2 * s2, s6 = s4 * s2, s8 = s4 * s4, s10 = s4 * s6, s12 = s6 * s6;
      const double Fa = 1.0 + p[0] * s2 + p[1] * s4 + p[2] * s6 + p[3] * s8 + p[4] * s10 + p[5] * s12;
      const double Fb = 1.0 + 1e-8 * s2;
      F = pow(Fa / Fb, p[6]);
      d = -2 * p[6] * F *
          (1e-8 * Fa - Fb * (p[0] + 2 * p[1] * s2 + 3 * p[2] * s4 + 4 * p[3] * s6 + 5 * p[4] * s8 + 6 * p[5] * s10)) /
          (Fa * Fb);
      break;
    }
    case OFDFT_GGA_P92: {  // :319-335 (ss)
      const double s2 = x * x, Fa = p[0] + p[1] * s2 + p[2] * s2 * s2, Fb = p[0] + p[3] * s2;
      F = Fa / Fb;
      d = (2.0 * p[1] + 4.0 * p[2] * s2) / Fb - (2.0 * p[3] * Fa) / (Fb * Fb);
      break;
    }
    case OFDFT_GGA_PBE: {  // :338-368 (ss), 2-4 parameters
      const double s2 = x * x, Fb = 1.0 + p[1] / p[0] * s2, Fb2 = Fb * Fb;
      F = 1.0 + p[1] * s2 / Fb;
      d = 2.0 * p[1] / Fb2;
      if (g.np > 2) {
        const double Fb3 = Fb2 * Fb;
        F += p[2] * s2 * s2 / Fb2;
        d += 4 * p[2] * s2 / Fb3;
        if (g.np > 3) {
          F += p[3] * s2 * s2 * s2 / Fb3;
          d += 6 * p[3] * s2 * s2 / (Fb3 * Fb);
        }
      }
      break;
    }
    case OFDFT_GGA_P82: {  // :371-385 (ss)
      const double s2 = x * x, s6 = s2 * s2 * s2, Fb = 1 + s6;
      F = 1.0 + p[0] * s2 / Fb;
      d = p[0] * (2.0 / Fb - 6.0 * s6 / (Fb * Fb));
      break;
    }
    case OFDFT_GGA_VJKS00: {  // :388-404 (ss)
      const double s2 = x * x, s4 = s2 * s2, s6 = s4 * s2;
      const double Fa = 1.0 + p[0] * s2 - p[2] * s6, Fb = 1.0 + p[1] * s2 + p[2] * s4;
      F = Fa / Fb;
      d = (2.0 * p[0] - 6.0 * p[2] * s4) / Fb - (2.0 * p[1] + 4.0 * p[2] * s2) * Fa / (Fb * Fb);
      break;
    }
    case OFDFT_GGA_VT84F: {  // :407-433 (ss), s^2 floored at 1e-8
      double s2 = x * x;
      if (s2 < 1.0e-8) s2 = 1.0e-8;
      const double s4 = s2 * s2, e1 = exp(-p[1] * s2), e2 = exp(-p[1] * s4), q = 1 + p[0] * s2;
      F = 1.0 + 5.0 / 3.0 * s2 + p[0] * s2 * e1 / q + (1 - e2) * (1.0 / s2 - 1.0);
      d = 10.0 / 3.0 + 2.0 * p[0] * e1 * (1.0 - p[1] * s2) / q - 2.0 * p[0] * s2 * e1 / (q * q) +
          4.0 * p[1] * s2 * (1.0 / s2 - 1.0) * e2 - 2.0 * (1 - e2) / s4;
      break;
    }
    case OFDFT_GGA_TFVW: {  // :436-447 (ss)
      F = p[0] + 5.0 / 3.0 * p[1] * x * x;
      d = 10.0 / 3.0 * p[1];
      break;
    }
    case OFDFT_GGA_STV: {  // :450-462 (ss)
      const double s2 = x * x, Fb = p[3] + p[2] * s2;
      F = p[0] + 5.0 / 3.0 * p[1] * s2 / Fb;
      d = 5.0 / 3.0 * p[1] * (2.0 / Fb - 2.0 * s2 * p[2] / (Fb * Fb));
      break;
    }
    case OFDFT_GGA_PBE2M: {  // :465-478 (ss)
      const double s2 = x * x, Fb = p[0] + p[1] * s2;
      F = 1.0 + p[2] * s2 / Fb;
      d = 2.0 * p[2] / (Fb * Fb);
      break;
    }
    default: {  // OFDFT_GGA_PG, :481-490 (ss)
      const double s2 = x * x, Fa = exp(-p[0] * s2);
      F = Fa + 5.0 / 3.0 * p[1] * s2;
      d = -2.0 * p[0] * Fa + 10.0 / 3.0 * p[1];
    }
  }
  if (tk) d /= kTKF0 * kTKF0;
}

// rho (mode rho) or (c phi + s p)^2 as a single-field row-transform input
template <typename T>
struct PreRho {
  static constexpr int kProf = PC_ROWS_FWD;
  const T* __restrict__ rho;
  const T* __restrict__ phi;
  const T* __restrict__ p;
  T c, s;
  __host__ __device__ int nin() const { return rho ? 1 : (p ? 2 : 1); }
  struct In {
    T a, b;
  };
  __device__ __forceinline__ In load(long long i) const {
    if (rho) return {rho[i], T(0)};
    return {phi[i], p ? p[i] : T(0)};
  }
  __device__ __forceinline__ double dens(const In& x) const {
    if (rho) return (double)x.a;
    const double ph = p ? (double)c * x.a + (double)s * x.b : (double)x.a;
    return ph * ph;
  }
  __device__ __forceinline__ void apply(const In& x, T* out) const { out[0] = T(dens(x)); }
  __device__ __forceinline__ void apply(const In& x, T* out, long long) const { apply(x, out); }
  __device__ __forceinline__ void operator()(long long i, T* out) const { apply(load(i), out); }
};

// three stored fields as the row-transform inputs (w[f*N + i])
template <typename T>
struct PreCopy3 {
  static constexpr int kProf = PC_ROWS_FWD;
  const T* __restrict__ w;
  long long N;
  __host__ __device__ int nin() const { return 3; }
  struct In {
    T a, b, c;
  };
  __device__ __forceinline__ In load(long long i) const { return {w[i], w[N + i], w[2 * N + i]}; }
  __device__ __forceinline__ void apply(const In& x, T* out) const {
    out[0] = x.a;
    out[1] = x.b;
    out[2] = x.c;
  }
  __device__ __forceinline__ void apply(const In& x, T* out, long long) const { apply(x, out); }
  __device__ __forceinline__ void operator()(long long i, T* out) const { apply(load(i), out); }
};

// First GGA pass, on the gradient c[0..2] of rho (GGA, kedf/gga.py:1240-1283,
// :1299-1311): energy density tf F, pot1, the vector field
// tf dFds2 grad rho / rho^(8/3) (stored for the divergence) and the stress
// sums.  acc: [sum tf F, sum pot1 rho, 6 x sum vkin2 d_a rho d_b rho].
template <typename T>
struct PostGga1 {
  static constexpr int NS = 8;
  static constexpr int kProf = PC_ROWS_INV;
  PreRho<T> r;
  GgaSpec g;
  T* __restrict__ W;  // [pot1 | w_x | w_y | w_z], 4 N
  long long N;
  __host__ __device__ int nio() const { return r.nin() + (W ? 4 : 0); }
  using In = typename PreRho<T>::In;
  __device__ __forceinline__ In load(long long i) const { return r.load(i); }
  __device__ __forceinline__ void apply(long long i, const In& in, const T* c, double* acc) const {
    const double rho = r.dens(in);
    const double rhom = rho < 1e-16 ? 1e-16 : rho;
    const double rho23 = cbrt(rhom * rhom);
    const double tf = kCTF * rho23 * rhom;
    const double rho43 = rho23 * rho23, rho83 = rho43 * rho43;
    const double g0 = c[0], g1 = c[1], g2 = c[2];
    double s = sqrt(g0 * g0 + g1 * g1 + g2 * g2) / rho43;
    if (s < 1e-16) s = 1e-16;
    double F, d;
    gga_enhance(g, s, F, d);
    const double pot1 = 5.0 / 3.0 * kCTF * rho23 * F - 4.0 / 3.0 * tf * d * s * s / rhom;
    const double vk = tf * d / rho83;
    acc[0] += tf * F;
    acc[1] += pot1 * rho;
    acc[2] += vk * g0 * g0;
    acc[3] += vk * g0 * g1;
    acc[4] += vk * g0 * g2;
    acc[5] += vk * g1 * g1;
    acc[6] += vk * g1 * g2;
    acc[7] += vk * g2 * g2;
    if (W) {
      W[i] = T(pot1);
      W[N + i] = T(vk * g0);
      W[2 * N + i] = T(vk * g1);
      W[3 * N + i] = T(vk * g2);
    }
  }
  __device__ __forceinline__ void operator()(long long i, const T* c, double* acc) const { apply(i, load(i), c, acc); }
};

// Second GGA pass: V_gga = pot1 - div (c[0]), added to V; acc: [sum V_gga rho,
// 2 sum V_gga phi(t) dphi/dt] (the terms the evaluation's mu and line-search
// derivative sums get from this functional).
template <typename T>
struct PostGga2 {
  static constexpr int NS = 2;
  static constexpr int kProf = PC_ROWS_INV;
  PreRho<T> r;
  const T* __restrict__ pot1;
  T* __restrict__ V;
  __host__ __device__ int nio() const { return r.nin() + 3; }
  struct In {
    typename PreRho<T>::In x;
    T p1, v;
  };
  __device__ __forceinline__ In load(long long i) const { return {r.load(i), pot1[i], V[i]}; }
  __device__ __forceinline__ void apply(long long i, const In& in, const T* c, double* acc) const {
    const double vg = (double)in.p1 - (double)c[0];
    V[i] = T((double)in.v + vg);
    acc[0] += vg * r.dens(in.x);
    if (!r.rho && r.p) {
      const double ph = (double)r.c * in.x.a + (double)r.s * in.x.b;
      const double dph = (double)r.c * in.x.b - (double)r.s * in.x.a;
      acc[1] += 2.0 * vg * ph * dph;
    }
  }
  __device__ __forceinline__ void operator()(long long i, const T* c, double* acc) const { apply(i, load(i), c, acc); }
};

template <typename T>
struct PostEval {
  static constexpr int NS = OFDFT_NSUM;
  static constexpr int kProf = PC_ROWS_INV;
  PreEval<T> pre;
  const T* __restrict__ vloc;
  T* __restrict__ V;
  T* __restrict__ aux;  // [0,N) vW raw potential, [N,2N) WT convolution
  long long N;
  // fields read (density inputs, V_loc) and written (V, aux)
  __host__ __device__ int nio() const {
    return pre.nin() + (pre.fp.pseudo ? 1 : 0) + (V ? 1 : 0) + (aux ? (pre.fp.fV >= 0) + (pre.fp.fW >= 0) : 0);
  }
  struct In {
    typename PreEval<T>::In d;
    T vl;
  };
  template <class I>
  __device__ __forceinline__ In load(I i) const { return {pre.load(i), pre.fp.pseudo ? at(vloc, i) : T(0)}; }
  __device__ __forceinline__ void operator()(long long i, const T* c, double* acc) const { apply(i, load(i), c, acc); }
  template <class I>
  __device__ __forceinline__ void apply(I i, const In& in, const T* c, double* acc) const {
    const FnParams& fp = pre.fp;
    T ph = T(0), dph = T(0);
    T r;
    if (pre.rho) {
      r = in.d.a;
    } else {
      const T p0 = in.d.a;
      const T pp = in.d.b;
      ph = pre.c * p0 + pre.s * pp;
      dph = pre.c * pp - pre.s * p0;
      r = ph * ph;
    }
    T v = T(0);
    // one cube root and one square root serve TF, LDA, vW and WT (beta = 5/6)
    const T cb = dev_cbrt(r);
    T s = dev_sqrt(r > T(0) ? r : T(kZero));
    if (s < T(kZero)) s = T(kZero);
    if (fp.tf) {
      // TF: energy density cbrt(rho^5) = rho cbrt(rho)^2, potential (5/3)C_TF cbrt(rho)^2 (tf.py:13-35)
      acc[OFDFT_E_TF] += (double)(r * cb * cb);
      v += T(4.7853900003136530265) * (cb * cb) * T(fp.x);
    }
    if (fp.fV >= 0) {
      // vW: 0.5 * IFFT[G^2 FFT(s)] / s, energy over rho > 0 (vw.py:29-53, :93)
      const T vraw = dev_div(T(0.5) * pick3(c, fp.fV), s);
      if (r > T(0)) acc[OFDFT_E_VW] += (double)(r * vraw);
      v += vraw * T(fp.y);
      if (aux) at(aux, i) = vraw;
    }
    if (fp.fW >= 0) {
      // WT: 2 alpha rho^beta / rho * conv, energy density V rho/(2 alpha) (wt.py:12-28, :125)
      const T rz = r < T(kZero) ? T(kZero) : r;
      const T rbm1 = (fp.beta56 && r >= T(kZero)) ? dev_div(cb, s) : dev_div(dev_pow(rz, T(fp.beta)), rz);
      const T vw = T(2 * fp.alpha) * rbm1 * pick3(c, fp.fW);
      acc[OFDFT_E_NL] += (double)(vw * r / T(2 * fp.alpha));
      v += vw;
      if (aux) at(aux + N, i) = pick3(c, fp.fW);
    }
    if (fp.lda) {
      T ex, vx;
      lda_point(r, cb, ex, vx);
      acc[OFDFT_E_XC] += (double)(ex * r);
      v += vx;
    }
    if (fp.fH >= 0) {
      const T vh = pick3(c, fp.fH);
      acc[OFDFT_E_H] += (double)(vh * r);
      v += vh;
    }
    if (fp.pseudo) {
      const T vl = in.vl;
      acc[OFDFT_E_PP] += (double)(vl * r);
      v += vl;
    }
    acc[OFDFT_S_VRHO] += (double)(v * r);
    acc[OFDFT_S_RHO] += (double)r;
    if (!pre.rho) acc[OFDFT_S_DTHETA] += (double)(v * ph * dph);
    if (V) at(V, i) = v;
  }
};

// Ap = (V - mu) p + phi * dV[drho]; acc[0] = sum p*Ap.
template <typename T>
struct PostHvp {
  static constexpr int NS = 1;
  static constexpr int kProf = PC_ROWS_INV_HVP;
  const T* __restrict__ phi;
  const T* __restrict__ p;
  const T* __restrict__ V;
  const T* __restrict__ aux;
  T* __restrict__ Ap;
  T mu;
  long long N;
  FnParams fp;
  int raw = 0;  // 1: store dV (the potential's response to drho = 2 phi p) instead of Ap
  __host__ __device__ int nio() const { return 4 + (fp.fV >= 0) + (fp.fW >= 0); }
  struct In {
    T ph, pp, v, a0, a1;
  };
  __device__ __forceinline__ In load(long long i) const {
    return {phi[i], p[i], V[i], fp.fV >= 0 ? aux[i] : T(0), fp.fW >= 0 ? aux[N + i] : T(0)};
  }
  __device__ __forceinline__ void operator()(long long i, const T* c, double* acc) const { apply(i, load(i), c, acc); }
  __device__ __forceinline__ void apply(long long i, const In& in, const T* c, double* acc) const {
    const T ph = in.ph, pp = in.pp;
    const T r = ph * ph;
    const T dr = T(2) * ph * pp;
    const T cb = dev_cbrt(r);
    const T s = dev_sqrt(r);
    T dv = T(0);
    if (fp.tf && r > T(0)) dv += T(fp.x) * T(3.1902600002091020177) * dr / cb;  // (2/3)(5/3)C_TF
    if (fp.lda && r > T(0)) dv += lda_v2(r, cb) * dr;
    if (fp.fH >= 0) dv += pick3(c, fp.fH);
    if (fp.fV >= 0 && r > T(0)) {
      const T ds = dr / (T(2) * s);
      dv += T(fp.y) * (T(0.5) * pick3(c, fp.fV) / s - in.a0 * ds / s);
    }
    if (fp.fW >= 0 && r >= T(kZero)) {
      const T rb1 = fp.beta56 ? cb / s : dev_pow(r, T(fp.beta) - T(1));
      dv += T(2 * fp.alpha) * ((T(fp.beta) - T(1)) * rb1 / r * dr * in.a1 + rb1 * pick3(c, fp.fW));
    }
    const T ap = (in.v - mu) * pp + ph * dv;
    Ap[i] = raw ? dv : ap;
    acc[0] += (double)(pp * ap);
  }
};

// Ap = W0 p + phi * (H[drho] + y/2 GG[ds]/s + 2 alpha rho^(beta-1) K[beta rho^(beta-1) drho]),
// with W0 = V - mu + 2 rho a(rho) precomputed by k_hvp_prepare (a: the local
// second derivative of TF, LDA, vW and WT); acc[0] = sum p*Ap.
template <typename T, int LAY = 0>
struct PostHvpFast {
  using S = FieldSlots<LAY>;
  static constexpr int NS = 1;
  static constexpr int kProf = PC_ROWS_INV_HVP;
  const T* __restrict__ phi;
  const T* __restrict__ p;
  const T* __restrict__ W0;
  T* __restrict__ Ap;
  FnParams fp;
  __host__ __device__ int nio() const { return 4; }
  struct In {
    T ph, pp, w0;
  };
  template <class I>
  __device__ __forceinline__ In load(I i) const { return {at(phi, i), at(p, i), at(W0, i)}; }
  __device__ __forceinline__ void operator()(long long i, const T* c, double* acc) const { apply(i, load(i), c, acc); }
  // (H p)(i) from the point's phi, p, W0 and the convolutions c
  __device__ __forceinline__ T ap_of(const In& in, const T* c) const {
    const T ph = in.ph, pp = in.pp;
    const T r = ph * ph;
    if constexpr (sizeof(T) == 4) {
      // phi * dv with sqrt(rho) = |phi| cancelled: phi (y/2) c_vW / sqrt(rho)
      // = (y/2) sign(phi) c_vW, phi 2 alpha rho^(beta-1) = 2 alpha sign(phi) |phi|^(2/3)
      T ap = in.w0 * pp;
      if (S::fH(fp) >= 0) ap += ph * pick3(c, S::fH(fp));
      if (S::fV(fp) >= 0 && r > T(0)) {
        const T cv = T(0.5 * fp.y) * pick3(c, S::fV(fp));
        ap += ph > T(0) ? cv : -cv;
      }
      if (S::fW(fp) >= 0 && r >= T(kZero)) {
        const T g = S::b56(fp) ? sgn_pow23(ph) : ph * dev_pow(r, T(fp.beta) - T(1));
        ap += T(2 * fp.alpha) * g * pick3(c, S::fW(fp));
      }
      return ap;
    }
    T dv = T(0);
    if (S::fH(fp) >= 0) dv += pick3(c, S::fH(fp));
    if (S::fV(fp) >= 0 || S::fW(fp) >= 0) {
      const T s = dev_sqrt(r);
      if (S::fV(fp) >= 0 && r > T(0)) dv += dev_div(T(fp.y) * T(0.5) * pick3(c, S::fV(fp)), s);
      if (S::fW(fp) >= 0 && r >= T(kZero)) {
        const T rb1 = S::b56(fp) ? dev_div(dev_cbrt(r), s) : dev_pow(r, T(fp.beta) - T(1));
        dv += T(2 * fp.alpha) * rb1 * pick3(c, S::fW(fp));
      }
    }
    return in.w0 * pp + ph * dv;
  }
  template <class I>
  __device__ __forceinline__ void apply(I i, const In& in, const T* c, double* acc) const {
    const T ap = ap_of(in, c);
    at(Ap, i) = ap;
    acc[0] += (double)(in.pp * ap);
  }
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// Residual (V - mu) phi (optimization.py:323, :469) and the HVP diagonal
// W0 = V - mu + 2 rho a(rho) at phi; partial sums of res^2.
template <typename T>
__global__ void k_hvp_prepare(const T* __restrict__ phi, const T* __restrict__ V, const T* __restrict__ aux, T mu,
                              T* __restrict__ res, T* __restrict__ W0, long long N, FnParams fp,
                              double* __restrict__ partials) {
  double acc[1] = {0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
    const T ph = phi[i];
    const T r = ph * ph;
    const T vm = V[i] - mu;
    const T rr = vm * ph;
    res[i] = rr;
    acc[0] += (double)rr * (double)rr;
    T a = T(0);
    if (r > T(0)) {
      const T cb = dev_cbrt(r);
      if (fp.tf) a += T(fp.x) * T(3.1902600002091020177) / cb;  // (2/3)(5/3) C_TF rho^(-1/3)
      if (fp.lda) a += lda_v2(r, cb);
      if (fp.fV >= 0) a -= T(fp.y) * T(0.5) * aux[i] / r;   // -V_vW / (2 rho)
      if (fp.fW >= 0 && r >= T(kZero)) {
        const T s = dev_sqrt(r);
        const T rb1 = fp.beta56 ? cb / s : dev_pow(r, T(fp.beta) - T(1));
        a += T(2 * fp.alpha * (fp.beta - 1.0)) * rb1 / r * aux[N + i];
      }
    }
    W0[i] = vm + T(2) * r * a;
  }
  block_sums<1>(acc, partials + blockIdx.x);
}

// Axis-2 real-to-complex: block handles PB row pairs x F fields.  Rows r and
// r+1 of field f share one complex sequence z = a + i b; after the FFT,
// A[k] = (Z[k] + conj Z[-k])/2, B[k] = (Z[k] - conj Z[-k])/(2i).
// transform length of a compile-time FFT policy (1 for the run-time one)
template <class FFT>
constexpr int fixed_len() {
  if constexpr (FFT::kFixed) return FFT::n();
  else return 1;
}

// FC > 0: the field count F fixed at compile time (index arithmetic by F
// becomes constant); FC = 0: F taken from Frt.
// One workgroup's share of the forward row pass (row pairs bx*PB ..):
// the body of k_rows_fwd, also run by the persistent batched TN kernel.
// SL = 1 or 2 (compile-time even length; split rows): columns 0 .. n2/2-1
// of field f and row r at element f*Nh + r*nhp + k (the caller passes Nh =
// n0*n1*n2/2 and nhp = n2/2), the Nyquist column apart at element 3*Nh +
// f*nrows + r; SL = 1 stores bf16 pairs (FP32 plans), SL = 2 the plan
// precision.
template <typename T, class FFT, class Pre, int FC, int SL = 0>
__device__ __forceinline__ void rows_fwd_block(const Pre& pre, const FFT& fft, cplx<T>* __restrict__ spec,
                                               long long nrows, long long Nh, int nhp, int Frt,
                                               const cplx<T>* __restrict__ tw, long long bx, cplx<T>* lds) {
  const int F = FC > 0 ? FC : Frt;
  constexpr int PB = kRowPB<T>;
  const int n2 = fft.n();
  const int nh = n2 / 2 + 1;
  const int th = fft.th();
  const long long pair0 = bx * PB;
  const int ld = n2;
  if constexpr (FFT::kFixed) {
    // every load of the thread's points is issued before the first use;
    // compile-time lengths run on plans whose fields fit 32-bit offsets
    constexpr int NN = FFT::n();
    constexpr int IT = (PB * NN + FFT::TH - 1) / FFT::TH;
    const unsigned nr = (unsigned)nrows, p0 = (unsigned)pair0;
    typename Pre::In xa[IT], xb[IT];
#pragma unroll
    for (int u = 0; u < IT; ++u) {
      const int t = threadIdx.x + u * FFT::TH;
      const int lp = t / NN, j = t - lp * NN;
      const unsigned r0 = 2u * (p0 + lp), r1 = r0 + 1;
      xa[u] = {};
      xb[u] = {};
      if (t < PB * NN && r0 < nr) xa[u] = pre.load(r0 * NN + j);
      if (t < PB * NN && r1 < nr) xb[u] = pre.load(r1 * NN + j);
    }
#pragma unroll
    for (int u = 0; u < IT; ++u) {
      const int t = threadIdx.x + u * FFT::TH;
      if (t < PB * NN) {
        const int lp = t / NN, j = t - lp * NN;
        const unsigned r0 = 2u * (p0 + lp), r1 = r0 + 1;
        T a[3] = {T(0), T(0), T(0)}, b[3] = {T(0), T(0), T(0)};
        if (r0 < nr) pre.apply(xa[u], a, r0 * NN + j);
        if (r1 < nr) pre.apply(xb[u], b, r1 * NN + j);
#pragma unroll
        for (int f = 0; f < 3; ++f)
          if (f < F) lds[(lp * F + f) * ld + j] = cplx<T>{a[f], b[f]};
      }
    }
  } else {
    for (int t = threadIdx.x; t < PB * n2; t += th) {
      const int lp = t / n2, j = t - lp * n2;
      const long long r0 = 2 * (pair0 + lp), r1 = r0 + 1;
      T a[3] = {T(0), T(0), T(0)}, b[3] = {T(0), T(0), T(0)};
      if (r0 < nrows) pre(r0 * n2 + j, a);
      if (r1 < nrows) pre(r1 * n2 + j, b);
#pragma unroll
      for (int f = 0; f < 3; ++f)
        if (f < F) lds[(lp * F + f) * ld + j] = cplx<T>{a[f], b[f]};
    }
  }
  cplx<T>* ltw = lds + PB * F * ld;
  for (int t = threadIdx.x; t < n2; t += th) ltw[t] = tw[t];
  __syncthreads();
  fft.run(lds, ld, PB * F, ltw, false);
  for (int t = threadIdx.x; t < PB * F * nh; t += th) {
    const int tr = t / nh, k = t - tr * nh;
    const int lp = tr / F, f = tr - lp * F;
    const cplx<T> zk = lds[tr * ld + k];
    const cplx<T> zm = lds[tr * ld + (k == 0 ? 0 : n2 - k)];
    const cplx<T> A{T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
    const cplx<T> B{T(0.5) * (zk.y + zm.y), T(0.5) * (zm.x - zk.x)};
    if constexpr (SL >= 1) {
      using SE = SpecElem<T, SL>;
      auto* sb = reinterpret_cast<typename SE::E*>(spec);
      const unsigned r0 = 2u * ((unsigned)pair0 + lp), r1 = r0 + 1;
      const unsigned nb = (unsigned)nhp;
      const unsigned o = (unsigned)k < nb ? (unsigned)f * (unsigned)Nh + r0 * nb + k
                                          : 3u * (unsigned)Nh + (unsigned)f * (unsigned)nrows + r0;
      const unsigned o1 = (unsigned)k < nb ? o + nb : o + 1u;
      if (r0 < (unsigned)nrows) at(sb, o) = SE::put(A);
      if (r1 < (unsigned)nrows) at(sb, o1) = SE::put(B);
    } else if constexpr (FFT::kFixed) {
      const unsigned r0 = 2u * ((unsigned)pair0 + lp), r1 = r0 + 1;
      const unsigned o = (unsigned)f * (unsigned)Nh + r0 * (unsigned)nhp + k;
      if (r0 < (unsigned)nrows) at(spec, o) = A;
      if (r1 < (unsigned)nrows) at(spec, o + (unsigned)nhp) = B;
    } else {
      const long long r0 = 2 * (pair0 + lp), r1 = r0 + 1;
      if (r0 < nrows) spec[f * Nh + r0 * nhp + k] = A;
      if (r1 < nrows) spec[f * Nh + r1 * nhp + k] = B;
    }
  }
}

template <typename T, class FFT, class Pre, int FC, int SL = 0>
__global__ __launch_bounds__(FFT::TH) void k_rows_fwd(Pre pre, FFT fft, cplx<T>* __restrict__ spec,
                                                      long long nrows, long long Nh, int nhp, int Frt,
                                                      const cplx<T>* __restrict__ tw,
                                                      const int* __restrict__ skip) {
  if (skip && *skip) return;  // a finished device-side CG loop skips the rest of its queue
  extern __shared__ unsigned char smem_raw[];
  rows_fwd_block<T, FFT, Pre, FC, SL>(pre, fft, spec, nrows, Nh, nhp, Frt, tw, blockIdx.x,
                                      reinterpret_cast<cplx<T>*>(smem_raw));
}

// Axis-2 complex-to-real fused with the post functor.  Two half spectra A, B
// are combined into Z = A + iB on the full axis (Hermitian extension, DC and
// Nyquist imaginary parts dropped as irfft does); after the inverse FFT the
// real part is row r, the imaginary part row r+1.  For compile-time lengths
// the post functor's own inputs are loaded first, so that their latency
// overlaps the spectrum loads and the transform.
// One workgroup's share of the inverse row pass: the body of k_rows_inv,
// also run by the persistent batched TN kernel; the workgroup's sums go to
// partials[bx * NS ..] (none when partials is null).
// SL = 1, 2: the split-row spectrum layouts of rows_fwd_block.
template <typename T, class FFT, class Post, int FC, int SL = 0>
__device__ __forceinline__ void rows_inv_block(const Post& post, const FFT& fft, const cplx<T>* __restrict__ spec,
                                               long long nrows, long long Nh, int nhp, int Frt,
                                               const cplx<T>* __restrict__ tw, double* __restrict__ partials,
                                               long long bx, cplx<T>* lds) {
  const int F = FC > 0 ? FC : Frt;
  constexpr int PB = kRowPB<T>;
  const int n2 = fft.n();
  const int nh = n2 / 2 + 1;
  const int th = fft.th();
  const long long pair0 = bx * PB;
  const int ld = n2;
  const bool even = (n2 % 2) == 0;
  constexpr int NN = fixed_len<FFT>();
  constexpr int IT = FFT::kFixed ? (PB * NN + FFT::TH - 1) / FFT::TH : 1;
  typename Post::In xa[IT], xb[IT];
  const unsigned nr = (unsigned)nrows, p0 = (unsigned)pair0;
  if constexpr (FFT::kFixed) {
    // compile-time lengths run on plans whose fields fit 32-bit offsets
#pragma unroll
    for (int u = 0; u < IT; ++u) {
      const int t = threadIdx.x + u * FFT::TH;
      const int lp = t / NN, j = t - lp * NN;
      const unsigned r0 = 2u * (p0 + lp), r1 = r0 + 1;
      xa[u] = {};
      xb[u] = {};
      if (t < PB * NN && r0 < nr) xa[u] = post.load(r0 * NN + j);
      if (t < PB * NN && r1 < nr) xb[u] = post.load(r1 * NN + j);
    }
  }
  for (int t = threadIdx.x; t < PB * F * nh; t += th) {
    const int tr = t / nh, k = t - tr * nh;
    const int lp = tr / F, f = tr - lp * F;
    cplx<T> A{T(0), T(0)}, B{T(0), T(0)};
    if constexpr (SL >= 1) {
      using SE = SpecElem<T, SL>;
      const auto* sb = reinterpret_cast<const typename SE::E*>(spec);
      const unsigned r0 = 2u * (p0 + lp), r1 = r0 + 1;
      const unsigned nb = (unsigned)nhp;
      const unsigned o = (unsigned)k < nb ? (unsigned)f * (unsigned)Nh + r0 * nb + k
                                          : 3u * (unsigned)Nh + (unsigned)f * nr + r0;
      const unsigned o1 = (unsigned)k < nb ? o + nb : o + 1u;
      if (r0 < nr) A = SE::get(at(sb, o));
      if (r1 < nr) B = SE::get(at(sb, o1));
    } else if constexpr (FFT::kFixed) {
      const unsigned r0 = 2u * (p0 + lp), r1 = r0 + 1;
      const unsigned o = (unsigned)f * (unsigned)Nh + r0 * (unsigned)nhp + k;
      if (r0 < nr) A = at(spec, o);
      if (r1 < nr) B = at(spec, o + (unsigned)nhp);
    } else {
      const long long r0 = 2 * (pair0 + lp), r1 = r0 + 1;
      if (r0 < nrows) A = spec[f * Nh + r0 * nhp + k];
      if (r1 < nrows) B = spec[f * Nh + r1 * nhp + k];
    }
    if (k == 0 || (even && k == n2 / 2)) {
      A.y = T(0);
      B.y = T(0);
    }
    lds[tr * ld + k] = cplx<T>{A.x - B.y, A.y + B.x};
    if (k != 0 && n2 - k != k) lds[tr * ld + (n2 - k)] = cplx<T>{A.x + B.y, B.x - A.y};
  }
  cplx<T>* ltw = lds + PB * F * ld;
  for (int t = threadIdx.x; t < n2; t += th) ltw[t] = tw[t];
  __syncthreads();
  fft.run(lds, ld, PB * F, ltw, true);
  double acc[Post::NS];
#pragma unroll
  for (int s = 0; s < Post::NS; ++s) acc[s] = 0.0;
  if constexpr (FFT::kFixed) {
#pragma unroll
    for (int u = 0; u < IT; ++u) {
      const int t = threadIdx.x + u * FFT::TH;
      if (t < PB * NN) {
        const int lp = t / NN, j = t - lp * NN;
        const unsigned r0 = 2u * (p0 + lp), r1 = r0 + 1;
        T ca[3] = {T(0), T(0), T(0)}, cb[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          if (f < F) {
            const cplx<T> z = lds[(lp * F + f) * ld + j];
            ca[f] = z.x;
            cb[f] = z.y;
          }
        }
        if (r0 < nr) post.apply(r0 * NN + j, xa[u], ca, acc);
        if (r1 < nr) post.apply(r1 * NN + j, xb[u], cb, acc);
      }
    }
  } else {
    for (int t = threadIdx.x; t < PB * n2; t += th) {
      const int lp = t / n2, j = t - lp * n2;
      const long long r0 = 2 * (pair0 + lp), r1 = r0 + 1;
      T ca[3] = {T(0), T(0), T(0)}, cb[3] = {T(0), T(0), T(0)};
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        if (f < F) {
          const cplx<T> z = lds[(lp * F + f) * ld + j];
          ca[f] = z.x;
          cb[f] = z.y;
        }
      }
      if (r0 < nrows) post(r0 * n2 + j, ca, acc);
      if (r1 < nrows) post(r1 * n2 + j, cb, acc);
    }
  }
  if (partials) block_sums<Post::NS>(acc, partials + bx * Post::NS);
}

template <typename T, class FFT, class Post, int FC, int SL = 0>
__global__ __launch_bounds__(FFT::TH) void k_rows_inv(Post post, FFT fft, const cplx<T>* __restrict__ spec,
                                                      long long nrows, long long Nh, int nhp, int Frt,
                                                      const cplx<T>* __restrict__ tw,
                                                      double* __restrict__ partials, const int* __restrict__ skip) {
  if (skip && *skip) return;  // a finished device-side CG loop skips the rest of its queue
  extern __shared__ unsigned char smem_raw[];
  rows_inv_block<T, FFT, Post, FC, SL>(post, fft, spec, nrows, Nh, nhp, Frt, tw, partials, blockIdx.x,
                                       reinterpret_cast<cplx<T>*>(smem_raw));
}

// Strided complex transform along one axis: blockIdx.y selects the outer
// index, blockIdx.x a tile of W contiguous columns, blockIdx.z the field.
template <typename T, class FFT>
__global__ __launch_bounds__(FFT::TH) void k_axis_c2c(FFT fft, cplx<T>* __restrict__ spec, long long Nh,
                                                      long long stride, long long ostride, int ncol,
                                                      const cplx<T>* __restrict__ tw, int inv,
                                                      const int* __restrict__ skip) {
  if (skip && *skip) return;  // a finished device-side CG loop skips the rest of its queue
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  constexpr int W = kColTileC;
  const int n = fft.n();
  const int th = fft.th();
  cplx<T>* base = spec + (long long)blockIdx.z * Nh + (long long)blockIdx.y * ostride;
  const int c0 = blockIdx.x * W;
  const int ld = n + 1;
  for (int t = threadIdx.x; t < n * W; t += th) {
    const int i = t / W, w = t - i * W;
    const int c = c0 + w;
    lds[w * ld + i] = c < ncol ? base[i * stride + c] : cplx<T>{T(0), T(0)};
  }
  cplx<T>* ltw = lds + W * ld;
  for (int t = threadIdx.x; t < n; t += th) ltw[t] = tw[t];
  __syncthreads();
  fft.run(lds, ld, W, ltw, inv != 0);
  for (int t = threadIdx.x; t < n * W; t += th) {
    const int i = t / W, w = t - i * W;
    const int c = c0 + w;
    if (c < ncol) base[i * stride + c] = lds[w * ld + i];
  }
}

struct ConvKinds {
  int kind[3];
  double scale;    // 1/N
  double coef[3];  // weight of field f in the spectral part of p.Ap (TN inner loop)
};

// |G|^2 in the transform precision (FP32 plans use the FP32 lattice copy).
template <typename T>
__device__ __forceinline__ T gsq_t(const GridGeom& g, int k0, int k1, int k2, int partner) {
  if constexpr (sizeof(T) == 8) {
    return gsq(g, k0, k1, k2, partner);
  } else {
    int m0 = fftfreq_i(k0, g.n0), m1 = fftfreq_i(k1, g.n1);
    int m2 = g.full ? fftfreq_i(k2, g.n2) : k2;
    if (partner) {
      m0 = fftfreq_i((g.n0 - k0) % g.n0, g.n0);
      m1 = fftfreq_i((g.n1 - k1) % g.n1, g.n1);
      m2 = fftfreq_i((g.n2 - k2) % g.n2, g.n2);
    }
    const float a = (float)m0, b = (float)m1, c = (float)m2;
    const float gx = a * g.bgf[0] + b * g.bgf[3] + c * g.bgf[6];
    const float gy = a * g.bgf[1] + b * g.bgf[4] + c * g.bgf[7];
    const float gz = a * g.bgf[2] + b * g.bgf[5] + c * g.bgf[8];
    return gx * gx + gy * gy + gz * gz;
  }
}

// Reciprocal multiplier M(G)/N of one convolution kind at (k0, k1, c).
template <typename T>
__device__ __forceinline__ T conv_mult(const GridGeom& geo, int kind, int k0, int k1, int c, T scale,
                                       const T* __restrict__ wtk) {
  T m;
  if (kind == OFDFT_CONV_HARTREE) {
    const T g2 = gsq_t<T>(geo, k0, k1, c, 0);
    m = g2 > T(0) ? T(4.0 * kPi) / g2 : T(0);
    if (geo.full) {
      const T h2 = gsq_t<T>(geo, k0, k1, c, 1);
      m = T(0.5) * (m + (h2 > T(0) ? T(4.0 * kPi) / h2 : T(0)));
    }
  } else if (kind == OFDFT_CONV_GG) {
    m = gsq_t<T>(geo, k0, k1, c, 0);
    if (geo.full) m = T(0.5) * (m + gsq_t<T>(geo, k0, k1, c, 1));
  } else if (kind == OFDFT_CONV_WT) {
    m = wtk[((long long)k0 * geo.n1 + k1) * geo.nh + c];
  } else {
    m = T(1);
  }
  return m * scale;
}

// Axis-0 pencils: forward FFT, multiply by M_f(G)/N, inverse FFT.
template <typename T, class FFT>
__global__ __launch_bounds__(FFT::TH) void k_pencil_conv(FFT fft, cplx<T>* __restrict__ spec, long long Nh,
                                                         GridGeom geo, const cplx<T>* __restrict__ tw,
                                                         ConvKinds ck, const T* __restrict__ wtk,
                                                         const int* __restrict__ skip) {
  if (skip && *skip) return;  // a finished device-side CG loop skips the rest of its queue
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  constexpr int W = kColTileC;
  const int n = fft.n();
  const int th = fft.th();
  const int i1 = blockIdx.y;
  const int f = blockIdx.z;
  const long long stride = (long long)geo.n1 * geo.nhp;
  cplx<T>* base = spec + (long long)f * Nh + (long long)i1 * geo.nhp;
  const int c0 = blockIdx.x * W;
  const int ld = n + 1;
  for (int t = threadIdx.x; t < n * W; t += th) {
    const int i = t / W, w = t - i * W;
    const int c = c0 + w;
    lds[w * ld + i] = c < geo.nh ? base[i * stride + c] : cplx<T>{T(0), T(0)};
  }
  cplx<T>* ltw = lds + W * ld;
  for (int t = threadIdx.x; t < n; t += th) ltw[t] = tw[t];
  __syncthreads();
  fft.run(lds, ld, W, ltw, false);
  const int kind = ck.kind[f];
  for (int t = threadIdx.x; t < n * W; t += th) {
    const int w = t / n, k0 = t - w * n;
    const int c = c0 + w;
    if (c >= geo.nh) continue;
    const T mm = conv_mult<T>(geo, kind, k0, i1, c, T(ck.scale), wtk);
    cplx<T> z = lds[w * ld + k0];
    lds[w * ld + k0] = cplx<T>{z.x * mm, z.y * mm};
  }
  __syncthreads();
  fft.run(lds, ld, W, ltw, true);
  for (int t = threadIdx.x; t < n * W; t += th) {
    const int i = t / W, w = t - i * W;
    const int c = c0 + w;
    if (c < geo.nh) base[i * stride + c] = lds[w * ld + i];
  }
}


// LDS stride of one column's N1 x (N2 + 1) block in the register column
// kernels.  Thread (w, q) of a W-column tile writes element w*S + k1*(N2+1)
// + q and reads w*S + q*(N2+1) + n2.  The stride is the smallest S >=
// N1*(N2+1) in the residue class that a gfx950 sweep over every split, tile
// width and S (SQ_LDS_BANK_CONFLICT per LDS instruction of both accesses,
// 32 consecutive S each) found conflict-free or least conflicting:
//   8-byte elements (FP32 complex): W >= 16 odd S (about 1 conflict cycle per
//   instruction against 7-13 for even S), W = 8 S = 2 mod 4 (about 1
//   against 7-13 for odd S or S = 0 mod 4);
//   16-byte elements (FP64 complex): odd S (W = 8: 6-12 against 11-19; W =
//   16: 9-12, the least of any S).
template <typename T, int N1, int N2, int W>
constexpr int col_block_stride() {
  int s = N1 * (N2 + 1);
  if (sizeof(T) == 4 && W == 8) {
    while (s % 4 != 2) ++s;
  } else {
    while (s % 2 != 1) ++s;
  }
  return s;
}

// One workgroup's share of the four-step axis transform (the body of
// k_axis_reg, also run by the persistent batched TN kernel).
template <typename T, int N1, int N2, int W>
__device__ __forceinline__ void axis_reg_block(cplx<T>* __restrict__ spec, long long Nh, long long stride,
                                               long long ostride, int ncol, const cplx<T>* __restrict__ tw, int inv,
                                               int bx, int by, int bz, cplx<T>* lds) {
  constexpr int S = col_block_stride<T, N1, N2, W>();
  // threads beyond the W x N2 of the split idle (the persistent kernel runs
  // every length with one workgroup size); they still reach the barrier
  const int t = threadIdx.x, w = t % W, q = t / W;
  const bool act = t < W * N2;
  cplx<T>* base = spec + (long long)bz * Nh + (long long)by * ostride;
  const int c = bx * W + w;
  const bool ok = act && c < ncol;
  const bool iv = inv != 0;
  cplx<T> v[N2];
#pragma unroll
  for (int n1 = 0; n1 < N1; ++n1)
    v[n1] = ok ? at(base, (unsigned)(N2 * n1 + q) * (unsigned)stride + c) : cplx<T>{T(0), T(0)};
  if (act) {
    dftreg<T, N1>(v, iv);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
      cplx<T> wv = tw[q * k1];
      if (iv) wv.y = -wv.y;
      v[k1] = cmul(v[k1], wv);
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) lds[w * S + k1 * (N2 + 1) + q] = v[k1];
  }
  __syncthreads();
  if (act && q < N1) {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) v[n2] = lds[w * S + q * (N2 + 1) + n2];
    dftreg<T, N2>(v, iv);
    if (ok) {
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) at(base, (unsigned)(q + N1 * k2) * (unsigned)stride + c) = v[k2];
    }
  }
}

// The axis-1 transform of axis_reg_block over the split-row spectra (SL =
// 1 bf16, SL = 2 plan precision; rows_fwd_block): Nh = n0*n1*nb elements per
// field, stride = nb, ostride = n1*nb, ncol = nb = n2/2 (a multiple of W),
// nouter = n0.  Tiles bx < nb/W take W columns of plane by; the last tile
// slot, bx = nb/W, takes the Nyquist columns of the W planes by*W ..
// by*W+W-1, stored contiguously along the axis at element 3*Nh + (bz*n0 +
// plane)*N (workgroups with by*W >= n0 have none), so that no workgroup
// transforms a mostly empty tile.  The stored elements are loaded first and
// converted after, so that every load of the thread is in flight at once.
template <typename T, int N1, int N2, int W, int SL>
__device__ __forceinline__ void axis_split_block(cplx<T>* __restrict__ spec, long long Nh, long long stride,
                                              long long ostride, int ncol, const cplx<T>* __restrict__ tw, int inv,
                                              int bx, int by, int bz, cplx<T>* lds, int nouter) {
  constexpr int S = col_block_stride<T, N1, N2, W>();
  constexpr unsigned NA = (unsigned)(N1 * N2);
  using SE = SpecElem<T, SL>;
  using E = typename SE::E;
  E* sp = reinterpret_cast<E*>(spec);
  const int t = threadIdx.x, w = t % W, q = t / W;
  const bool act = t < W * N2;
  const bool iv = inv != 0;
  const unsigned n0 = (unsigned)nouter;
  const bool nyq = bx * W >= ncol;  // uniform over the workgroup
  if (nyq && (unsigned)by * W >= n0) return;  // the whole workgroup: no plane left
  const unsigned pl = (unsigned)(by * W + w);
  const bool ok = act && (!nyq || pl < n0);
  const unsigned off = nyq ? 3u * (unsigned)Nh + ((unsigned)bz * n0 + pl) * NA
                           : (unsigned)bz * (unsigned)Nh + (unsigned)by * (unsigned)ostride + (unsigned)(bx * W + w);
  const unsigned js = nyq ? 1u : (unsigned)stride;
  E raw[N1];
#pragma unroll
  for (int n1 = 0; n1 < N1; ++n1) raw[n1] = ok ? at(sp, off + (unsigned)(N2 * n1 + q) * js) : E{};
  cplx<T> v[N2];
#pragma unroll
  for (int n1 = 0; n1 < N1; ++n1) v[n1] = SE::get(raw[n1]);
  if (act) {
    dftreg<T, N1>(v, iv);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
      cplx<T> wv = tw[q * k1];
      if (iv) wv.y = -wv.y;
      v[k1] = cmul(v[k1], wv);
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) lds[w * S + k1 * (N2 + 1) + q] = v[k1];
  }
  __syncthreads();
  if (act && q < N1) {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) v[n2] = lds[w * S + q * (N2 + 1) + n2];
    dftreg<T, N2>(v, iv);
    if (ok) {
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) at(sp, off + (unsigned)(q + N1 * k2) * js) = SE::put(v[k2]);
    }
  }
}

// Four-step register transform of W columns along a strided axis (one
// workgroup, W*N2 threads): x -> X in place, natural order in and out.
// SL = 1, 2: split-row spectra (axis_split_block); nouter: the number of planes.
template <typename T, int N1, int N2, int W, int SL = 0>
__global__ __launch_bounds__(W * N2) void k_axis_reg(cplx<T>* __restrict__ spec, long long Nh, long long stride,
                                                     long long ostride, int ncol, const cplx<T>* __restrict__ tw,
                                                     int inv, const int* __restrict__ skip, int nouter) {
  if (skip && *skip) return;  // a finished device-side CG loop skips the rest of its queue
  __shared__ cplx<T> lds[W * col_block_stride<T, N1, N2, W>()];
  if constexpr (SL >= 1)
    axis_split_block<T, N1, N2, W, SL>(spec, Nh, stride, ostride, ncol, tw, inv, blockIdx.x, blockIdx.y, blockIdx.z,
                                       lds, nouter);
  else
    axis_reg_block<T, N1, N2, W>(spec, Nh, stride, ostride, ncol, tw, inv, blockIdx.x, blockIdx.y, blockIdx.z, lds);
}

// The multiplier step of an axis-0 pencil: thread (w, q) holds the points
// k0 = q + N1*k2 (k2 < N2) of column (i1, c) after the forward transform;
// each becomes M_f(G)/N times itself (live == false: zero), and the return
// value is this thread's sum of M |X|^2 (its share of p.Ap in the TN loop).
// wm: the WT multipliers of the points, loaded by the caller (wt: WT field).
template <typename T, int N1, int N2>
__device__ __forceinline__ double pencil_multiply(cplx<T>* v, const GridGeom& geo, int kind, double cscale,
                                                  const T* __restrict__ wtk, const T* wm, bool wt, bool live, int i1,
                                                  int c, int q) {
  const T scale = T(cscale);
  double pap = 0.0;
  if constexpr (sizeof(T) == 4) {
    if (live && !wt && !geo.full && (kind == OFDFT_CONV_HARTREE || kind == OFDFT_CONV_GG)) {
      // FP32 Hartree / vW multipliers along the pencil: the (i1, c) part of
      // G is fixed per thread, only the axis-0 frequency changes
      const float m1 = (float)fftfreq_i(i1, geo.n1), m2 = (float)c;
      const float px = m1 * geo.bgf[3] + m2 * geo.bgf[6];
      const float py = m1 * geo.bgf[4] + m2 * geo.bgf[7];
      const float pz = m1 * geo.bgf[5] + m2 * geo.bgf[8];
      const bool hartree = kind == OFDFT_CONV_HARTREE;
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const float m0 = (float)fftfreq_i(q + N1 * k2, N1 * N2);
        const float gx = m0 * geo.bgf[0] + px, gy = m0 * geo.bgf[1] + py, gz = m0 * geo.bgf[2] + pz;
        const float g2 = gx * gx + gy * gy + gz * gz;
        float m = hartree ? (g2 > 0.f ? float(4.0 * kPi) * __builtin_amdgcn_rcpf(g2) : 0.f) : g2;
        m *= scale;
        pap += (double)m * ((double)v[k2].x * v[k2].x + (double)v[k2].y * v[k2].y);
        v[k2] = cplx<T>{v[k2].x * m, v[k2].y * m};
      }
    } else {
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const T m = !live ? T(0) : (wt ? wm[k2] * scale : conv_mult<T>(geo, kind, q + N1 * k2, i1, c, scale, wtk));
        pap += (double)m * ((double)v[k2].x * v[k2].x + (double)v[k2].y * v[k2].y);
        v[k2] = cplx<T>{v[k2].x * m, v[k2].y * m};
      }
    }
  } else {
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
      const T m = !live ? T(0) : (wt ? wm[k2] * scale : conv_mult<T>(geo, kind, q + N1 * k2, i1, c, scale, wtk));
      pap += (double)m * ((double)(v[k2].x * v[k2].x) + (double)(v[k2].y * v[k2].y));
      v[k2] = cplx<T>{v[k2].x * m, v[k2].y * m};
    }
  }
  return pap;
}

// Four-step axis-0 pencils of one workgroup: forward transform, multiplier,
// inverse transform with two LDS transposes and no intermediate global
// traffic.  (by, bz) = (axis-1 index, field); cgp: the workgroup's share of
// the spectral part of p.Ap goes to cgp[pslot] (TN inner loop).
template <typename T, int N1, int N2, int W>
__device__ __forceinline__ void pencil_reg_block(cplx<T>* __restrict__ spec, long long Nh, const GridGeom& geo,
                                                 const cplx<T>* __restrict__ tw, const ConvKinds& ck,
                                                 const T* __restrict__ wtk, double* __restrict__ cgp,
                                                 long long pslot, int bx, int by, int bz, cplx<T>* lds) {
  constexpr int S = col_block_stride<T, N1, N2, W>();
  const int t = threadIdx.x, w = t % W, q = t / W;
  const bool act = t < W * N2;
  const int i1 = by;
  const int f = bz;
  const long long stride = (long long)geo.n1 * geo.nhp;
  cplx<T>* base = spec + (long long)f * Nh + (long long)i1 * geo.nhp;
  const int c = bx * W + w;
  // padded columns (c >= nh) are neither read nor written: no kernel reads
  // them, and their transform would only move zeros
  const bool ok = act && c < geo.nh;
  const bool live = act && c < geo.nh;
  const int kind = ck.kind[f];
  cplx<T> v[N2];
#pragma unroll
  for (int n1 = 0; n1 < N1; ++n1)
    v[n1] = ok ? at(base, (unsigned)(N2 * n1 + q) * (unsigned)stride + c) : cplx<T>{T(0), T(0)};
  // WT multipliers of this thread's points, loaded with the spectrum so that
  // their latency is not exposed between the two transforms
  T wm[N2];
  const bool wt = kind == OFDFT_CONV_WT && q < N1 && live;
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2)
    wm[k2] = wt ? at(wtk, ((unsigned)(q + N1 * k2) * (unsigned)geo.n1 + i1) * (unsigned)geo.nh + c) : T(0);
  if (act) {
    dftreg<T, N1>(v, false);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], tw[q * k1]);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) lds[w * S + k1 * (N2 + 1) + q] = v[k1];
  }
  __syncthreads();
  // sum over this thread's points of m |X|^2: the field's share of p.Ap in
  // the TN inner loop (Hermitian weight 2 off the c = 0 and Nyquist planes)
  double pap = 0.0;
  if (act && q < N1) {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) v[n2] = lds[w * S + q * (N2 + 1) + n2];
    dftreg<T, N2>(v, false);
    pap = pencil_multiply<T, N1, N2>(v, geo, kind, ck.scale, wtk, wm, wt, live, i1, c, q);
    dftreg<T, N2>(v, true);
#pragma unroll
    for (int n2 = 1; n2 < N2; ++n2) {
      cplx<T> wv = tw[n2 * q];
      wv.y = -wv.y;
      v[n2] = cmul(v[n2], wv);
    }
    // this thread re-reads nothing it overwrites: row (w, q) is private to it
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) lds[w * S + q * (N2 + 1) + n2] = v[n2];
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) v[k1] = lds[w * S + k1 * (N2 + 1) + q];
    dftreg<T, N1>(v, true);
    if (ok) {
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) at(base, (unsigned)(N2 * n1 + q) * (unsigned)stride + c) = v[n1];
    }
  }
  if (cgp) {
    const bool edge = c == 0 || ((geo.n2 & 1) == 0 && c == geo.n2 / 2);
    double acc[1] = {live ? pap * ck.coef[f] * (edge ? 1.0 : 2.0) : 0.0};
    block_sums<1>(acc, cgp + pslot);
  }
}

// The axis-0 pencils of pencil_reg_block over the split-row spectra (SL =
// 1 bf16, SL = 2 plan precision; rows_fwd_block; nb = n2/2 a multiple of W,
// NB = n0*n1*nb elements per field).  Tiles bx < nb/W take columns bx*W + w
// at axis-1 index by; the last tile slot, bx = nb/W, takes the Nyquist
// columns of axis-1 indices by*W .. by*W+W-1 (element 3*NB + bz*n0*n1 +
// k0*n1 + i1; workgroups with by*W >= n1 have none and leave a zero p.Ap
// partial sum).  The stored elements are loaded first and converted after.
template <typename T, int N1, int N2, int W, int SL>
__device__ __forceinline__ void pencil_split_block(cplx<T>* __restrict__ spec, const GridGeom& geo,
                                                const cplx<T>* __restrict__ tw, const ConvKinds& ck,
                                                const T* __restrict__ wtk, double* __restrict__ cgp, long long pslot,
                                                int bx, int by, int bz, cplx<T>* lds) {
  constexpr int S = col_block_stride<T, N1, N2, W>();
  using SE = SpecElem<T, SL>;
  using E = typename SE::E;
  E* sp = reinterpret_cast<E*>(spec);
  const int t = threadIdx.x, w = t % W, q = t / W;
  const bool act = t < W * N2;
  const int f = bz;
  const int kind = ck.kind[f];
  const int nb = geo.nh - 1;
  const bool nyq = bx * W >= nb;  // uniform over the workgroup
  if (nyq && by * W >= geo.n1) {  // the whole workgroup: no pencil left
    if (cgp && threadIdx.x == 0) cgp[pslot] = 0.0;
    return;
  }
  const unsigned n01 = (unsigned)geo.n0 * (unsigned)geo.n1;
  const unsigned NB = n01 * (unsigned)nb;
  const int i1 = nyq ? by * W + w : by;
  const int c = nyq ? nb : bx * W + w;
  const bool ok = act && i1 < geo.n1;
  const unsigned off = nyq ? 3u * NB + (unsigned)f * n01 + (unsigned)i1
                           : (unsigned)f * NB + (unsigned)i1 * (unsigned)nb + (unsigned)c;
  const unsigned js = nyq ? (unsigned)geo.n1 : (unsigned)geo.n1 * (unsigned)nb;  // k0 stride
  E raw[N1];
#pragma unroll
  for (int n1 = 0; n1 < N1; ++n1) raw[n1] = ok ? at(sp, off + (unsigned)(N2 * n1 + q) * js) : E{};
  // WT multipliers from the table in the same split rows (wtk: the plan's
  // wtks): W contiguous entries of a row, or of the Nyquist plane, per k0
  T wm[N2];
  const bool wt = kind == OFDFT_CONV_WT && q < N1 && ok;
  const unsigned woff = nyq ? NB + (unsigned)i1 : (unsigned)i1 * (unsigned)nb + (unsigned)c;
  const unsigned wjs = nyq ? (unsigned)geo.n1 : (unsigned)geo.n1 * (unsigned)nb;
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) wm[k2] = wt ? at(wtk, woff + (unsigned)(q + N1 * k2) * wjs) : T(0);
  cplx<T> v[N2];
#pragma unroll
  for (int n1 = 0; n1 < N1; ++n1) v[n1] = SE::get(raw[n1]);
  if (act) {
    dftreg<T, N1>(v, false);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], tw[q * k1]);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) lds[w * S + k1 * (N2 + 1) + q] = v[k1];
  }
  __syncthreads();
  double pap = 0.0;
  if (act && q < N1) {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) v[n2] = lds[w * S + q * (N2 + 1) + n2];
    dftreg<T, N2>(v, false);
    pap = pencil_multiply<T, N1, N2>(v, geo, kind, ck.scale, wtk, wm, wt, ok, i1, c, q);
    dftreg<T, N2>(v, true);
#pragma unroll
    for (int n2 = 1; n2 < N2; ++n2) {
      cplx<T> wv = tw[n2 * q];
      wv.y = -wv.y;
      v[n2] = cmul(v[n2], wv);
    }
    // this thread re-reads nothing it overwrites: row (w, q) is private to it
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) lds[w * S + q * (N2 + 1) + n2] = v[n2];
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) v[k1] = lds[w * S + k1 * (N2 + 1) + q];
    dftreg<T, N1>(v, true);
    if (ok) {
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) at(sp, off + (unsigned)(N2 * n1 + q) * js) = SE::put(v[n1]);
    }
  }
  if (cgp) {
    // Hermitian weight 2 off the c = 0 and Nyquist planes
    const bool edge = nyq || c == 0;
    double acc[1] = {ok ? pap * ck.coef[f] * (edge ? 1.0 : 2.0) : 0.0};
    block_sums<1>(acc, cgp + pslot);
  }
}

// SL = 1, 2: split-row spectra (pencil_split_block).
template <typename T, int N1, int N2, int W, int SL = 0>
__global__ __launch_bounds__(W * N2) void k_pencil_reg(cplx<T>* __restrict__ spec, long long Nh, GridGeom geo,
                                                       const cplx<T>* __restrict__ tw, ConvKinds ck,
                                                       const T* __restrict__ wtk, const int* __restrict__ skip,
                                                       double* __restrict__ cgp) {
  if (skip && *skip) return;  // a finished device-side CG loop skips the rest of its queue
  __shared__ cplx<T> lds[W * col_block_stride<T, N1, N2, W>()];
  if constexpr (SL >= 1) {
    pencil_split_block<T, N1, N2, W, SL>(spec, geo, tw, ck, wtk, cgp,
                                         ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x,
                                         blockIdx.x, blockIdx.y, blockIdx.z, lds);
    return;
  }
  pencil_reg_block<T, N1, N2, W>(spec, Nh, geo, tw, ck, wtk, cgp,
                                 ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x, blockIdx.x,
                                 blockIdx.y, blockIdx.z, lds);
}


// WT kernel table K(|G|) (WTKernel kernel.py:325-334), evaluated in FP64.
template <typename T>
__global__ void k_wt_table(T* __restrict__ out, GridGeom geo, double tkf, double factor, double lbda) {
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nh;
       idx += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int i1 = (int)(r % geo.n1);
    const int i0 = (int)(r / geo.n1);
    double k = lindhard_dev(sqrt(gsq(geo, i0, i1, c)) / tkf, lbda, 1.0);
    if (geo.full) k = 0.5 * (k + lindhard_dev(sqrt(gsq(geo, i0, i1, c, 1)) / tkf, lbda, 1.0));
    out[idx] = T(k * factor);
  }
}

// The WT table in split rows (n2 even, nb = n2/2): entry (k0, k1, c < nb)
// at (k0*n1 + k1)*nb + c, the Nyquist column c = nb at n0*n1*nb + k0*n1 + k1;
// the layout of the split-row spectra, so that a pencil tile's multipliers
// are whole, aligned segments of a row or of the Nyquist plane.
template <typename T>
__global__ void k_wt_split(const T* __restrict__ wtk, T* __restrict__ out, GridGeom geo) {
  const long long n01 = (long long)geo.n0 * geo.n1, nb = geo.nh - 1;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n01 * geo.nh;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long r = idx / geo.nh;
    const long long c = idx - r * geo.nh;
    out[c < nb ? r * nb + c : n01 * nb + r] = wtk[idx];
  }
}

// This thread's share of sum_b partials[b * NS + s]: four independent
// accumulators keep four loads in flight (the partials were written by
// blocks on every XCD, so most loads miss this XCD's L2).
template <int NS>
__device__ __forceinline__ double thread_partials(const double* __restrict__ partials, long long nblk, int s) {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  const long long st = blockDim.x;
  long long b = threadIdx.x;
  for (; b + 3 * st < nblk; b += 4 * st) {
    a0 += partials[b * NS + s];
    a1 += partials[(b + st) * NS + s];
    a2 += partials[(b + 2 * st) * NS + s];
    a3 += partials[(b + 3 * st) * NS + s];
  }
  for (; b < nblk; b += st) a0 += partials[b * NS + s];
  return (a0 + a1) + (a2 + a3);
}

template <int NS>
__global__ void k_reduce_partials(const double* __restrict__ partials, long long nblk, double* __restrict__ out,
                                  double* __restrict__ hout, const int* __restrict__ skip) {
  if (skip && *skip) return;  // a finished device-side CG loop skips the rest of its queue
  const int s = blockIdx.x;
  double tmp[1] = {thread_partials<NS>(partials, nblk, s)};
  block_sums<1>(tmp, out + s);
  // thread 0 wrote out[s]; the mapped host copy saves a device-to-host copy
  if (hout && threadIdx.x == 0) hout[s] = out[s];
}

// --------------------------- vector kernels -------------------------------
template <typename T>
__global__ void k_dot3(const T* __restrict__ a, const T* __restrict__ b, long long N, double* __restrict__ partials) {
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
    const double x = a[i], y = b[i];
    acc[0] += x * y;
    acc[1] += x * x;
    acc[2] += y * y;
  }
  block_sums<3>(acc, partials + (long long)blockIdx.x * 3);
}

// alpha x + beta y with one fixed rounding order (beta y rounded, then one
// fused multiply-add), whichever kernel forms it
template <typename T>
__device__ __forceinline__ T lincomb_point(T alpha, T x, T beta, T y) {
  return fma(alpha, x, beta * y);
}

template <typename T>
__global__ void k_lincomb(T alpha, const T* x, T beta, const T* y, T* z, long long N) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
    z[i] = lincomb_point(alpha, x[i], beta, y[i]);
}

template <typename T>
__global__ void k_residual(const T* V, const T* phi, T mu, T* res, long long N,
                           double* __restrict__ partials) {
  double acc[1] = {0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
    const T r = (V[i] - mu) * phi[i];
    res[i] = r;
    acc[0] += (double)r * (double)r;
  }
  block_sums<1>(acc, partials + blockIdx.x);
}

template <typename T>
__global__ void k_cg_update(T a, const T* __restrict__ p, const T* __restrict__ Ap, T* __restrict__ d,
                            T* __restrict__ r, long long N, double* __restrict__ partials) {
  double acc[1] = {0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
    d[i] += a * p[i];
    const T rr = r[i] - a * Ap[i];
    r[i] = rr;
    acc[0] += (double)rr * (double)rr;
  }
  block_sums<1>(acc, partials + blockIdx.x);
}

// TN inner update with the step length computed on the device from the
// pAp just reduced into red[0]: if pAp > 0, d += a p and r -= a Ap with
// a = r0 / pAp; otherwise d and r are left untouched (the reference leaves
// the loop before updating).  Partial sums of r^2 either way.
template <typename T>
__global__ void k_cg_update_dev(const double* __restrict__ red, double r0, const T* __restrict__ p,
                                const T* __restrict__ Ap, T* __restrict__ d, T* __restrict__ r, long long N,
                                double* __restrict__ partials) {
  const double pAp = red[0];
  const bool upd = pAp > 0.0;
  const T a = upd ? T(r0 / pAp) : T(0);
  double acc[1] = {0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
    T rr = r[i];
    if (upd) {
      d[i] += a * p[i];
      rr -= a * Ap[i];
      r[i] = rr;
    }
    acc[0] += (double)rr * (double)rr;
  }
  block_sums<1>(acc, partials + blockIdx.x);
}

// ---------------------------------------------------------------------------
// Device-resident TN direction loop: get_direction_TN (optimization.py:92-148)
// with the exact HVP.  The host queues CG iterations without waiting; a
// one-block kernel reduces r^2, applies the reference's tests and sets
// `done`, after which every queued kernel returns at once.  The updates the
// tests decide are carried by the next iteration's own kernels: p = r +
// beta p by the forward row pass of the next product, Best = direction by
// the next step kernel; the terminal ones (restore Best, the negative-
// curvature scaling) by one kernel after the loop.
// ---------------------------------------------------------------------------
struct TnState {
  double r0, rconv, rmin, r1, pAp, beta, scale, pad;
  double wpp;  // sum W0 p^2 of the current direction p (the pointwise part of pAp)
  int done;    // loop finished
  int it;      // index of the CG iteration being run
  int niter;   // iterations taken (the reference's `number`)
  int status;  // OFDFT_TN_*
  int act;     // kActKeep / kActNext for the next iteration's kernels
  int maxit;
  int alias;   // `Best` still aliases `direction` (no copy taken yet)
  int fin;     // kActRestore / kActScale for k_tn_finish
};
constexpr int kActKeep = 1;     // Best = direction.copy()
constexpr int kActRestore = 2;  // direction = Best
constexpr int kActScale = 4;    // direction = r0 / pAp * p
constexpr int kActNext = 8;     // p = res + beta * p

// d = 0, r = -res, p = r; partial sums of r^2 and (with W0) of W0 r^2
template <typename T>
__global__ void k_tn_init(const T* __restrict__ res, const T* __restrict__ W0, T* __restrict__ d, T* __restrict__ r,
                          T* __restrict__ p, long long N, double* __restrict__ partials) {
  double acc[2] = {0.0, 0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
    const T v = -res[i];
    d[i] = T(0);
    r[i] = v;
    p[i] = v;
    const double v2 = (double)v * (double)v;
    acc[0] += v2;
    if (W0) acc[1] += (double)W0[i] * v2;
  }
  block_sums<2>(acc, partials + 2 * (long long)blockIdx.x);
}

__device__ __forceinline__ void tn_start_state(TnState* st, const double* __restrict__ rr, int maxit) {
  const double r0 = rr[0];
  st->r0 = r0;
  st->rconv = 0.1 * r0;
  st->rmin = 1e10;  // rLists = [1e10]
  st->r1 = r0;
  st->wpp = rr[1];
  st->pAp = 0.0;
  st->beta = 0.0;
  st->scale = 0.0;
  st->it = 0;
  st->niter = 0;
  st->status = OFDFT_TN_MAXFUN;
  st->act = 0;
  st->maxit = maxit;
  st->alias = 1;
  st->fin = 0;
  st->done = maxit > 0 ? 0 : 1;
}

// Copy of the final loop state into mapped host memory (hst, may be null),
// for a host that checks the loop at an event while later iterations still
// run: written once, by the control step that ends the loop (k_tn_start
// only clears `done`): every field but `done` first, then a system-scope
// fence, then `done`, so that a host that reads done == 1 (and fences)
// reads the fields of that same write.  Once done is set no later kernel of
// the loop writes again.
__device__ __forceinline__ void tn_publish(const TnState* st, TnState* hst) {
  if (!hst || !st->done) return;
  volatile TnState* h = hst;
  h->r0 = st->r0;
  h->rconv = st->rconv;
  h->rmin = st->rmin;
  h->r1 = st->r1;
  h->pAp = st->pAp;
  h->beta = st->beta;
  h->scale = st->scale;
  h->wpp = st->wpp;
  h->it = st->it;
  h->niter = st->niter;
  h->status = st->status;
  h->act = st->act;
  h->maxit = st->maxit;
  h->alias = st->alias;
  h->fin = st->fin;
  __threadfence_system();
  h->done = st->done;
  __threadfence_system();
}

__global__ void k_tn_start(TnState* st, const double* __restrict__ rr, int maxit, TnState* hst) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  tn_start_state(st, rr, maxit);
  if (hst) {
    reinterpret_cast<volatile TnState*>(hst)->done = st->done;
    __threadfence_system();
  }
}

// Best = direction when the previous iteration asked for a copy (direction
// has not changed since); then alpha = r0 / pAp; d += alpha p;
// r -= alpha Ap; partial sums of r^2.  Nothing happens when the loop is
// done or pAp < 0 (the reference leaves the loop before the update).
template <typename T>
__global__ void k_tn_step(const TnState* __restrict__ st, const double* __restrict__ red, const T* __restrict__ p,
                          const T* __restrict__ Ap, T* __restrict__ d, T* __restrict__ r, T* __restrict__ best,
                          long long N, double* __restrict__ partials) {
  if (st->done) return;
  const double pAp = red[0];
  if (pAp < 0.0) return;
  const T a = T(st->r0 / pAp);
  const bool keep = (st->act & kActKeep) != 0;
  double acc[1] = {0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
    const T di = d[i];
    if (keep) best[i] = di;
    d[i] = di + a * p[i];
    const T rr = r[i] - a * Ap[i];
    r[i] = rr;
    acc[0] += (double)rr * (double)rr;
  }
  block_sums<1>(acc, partials + blockIdx.x);
}

// The branch tests of optimization.py:116-145 on pAp and r1.
__device__ void tn_ctrl(TnState* st, double pAp, double r1) {
  st->pAp = pAp;
  int act = 0, fin = 0;
  if (pAp < 0.0) {
    if (st->it == 0) {
      st->scale = st->r0 / pAp;
      fin = kActScale;
      st->status = OFDFT_TN_WARN;
    } else {
      st->status = OFDFT_TN_FAILED;
    }
    st->done = 1;
  } else {
    const double r0 = st->r0, rmin = st->rmin;  // rmin = min(rLists[:-1])
    st->r1 = r1;
    if (r1 < rmin) {
      act |= kActKeep;
      st->rmin = r1;
    }
    if (r1 < st->rconv) {
      st->status = OFDFT_TN_CONV;
      st->done = 1;
    } else if (r1 > 1000.0 * rmin) {
      fin = kActRestore;
      st->status = OFDFT_TN_NOTREDUCE;
      st->done = 1;
    } else if (st->it > 10 && fabs(r0 - r1) < 0.1 * r0) {
      fin = kActRestore;
      st->status = OFDFT_TN_SMALL;
      st->done = 1;
    } else {
      st->beta = r1 / r0;
      st->r0 = r1;
      act |= kActNext;
    }
    // restoring Best while it still aliases direction leaves direction as
    // is; a copy taken in the same iteration equals direction too
    if (st->alias || (act & kActKeep)) fin &= ~kActRestore;
    if (act & kActKeep) st->alias = 0;
  }
  if (st->done) {
    st->niter = st->it + 1;
  } else if (st->it + 1 >= st->maxit) {
    st->done = 1;
    st->niter = st->maxit;
    st->status = OFDFT_TN_MAXFUN;
  }
  // once done, no later kernel reads `act`: the pending copy of Best is dropped
  st->it += 1;
  st->act = st->done ? 0 : act;
  st->fin = fin;
}

// r1 = sum of the step kernel's partials, then the tests; pAp = red[0].
__global__ void k_tn_reduce_ctrl(TnState* st, const double* __restrict__ partials, long long nblk,
                                 const double* __restrict__ red, double* __restrict__ rr, TnState* hst) {
  if (st->done) return;
  double tmp[1] = {thread_partials<1>(partials, nblk, 0)};
  block_sums<1>(tmp, rr);
  if (threadIdx.x == 0) {
    tn_ctrl(st, red[0], rr[0]);
    tn_publish(st, hst);
  }
}

// pAp = the pencil's spectral sums + sum W0 p^2 (kept in the state) into
// red[0], before the inverse row pass that takes the CG step.
__device__ __forceinline__ void tn_pap_block(const TnState* st, const double* __restrict__ partials, long long nblk,
                                             double* __restrict__ red) {
  double tmp[1] = {thread_partials<1>(partials, nblk, 0)};
  block_sums<1>(tmp, red);
  if (threadIdx.x == 0) red[0] += st->wpp;
}

__global__ void k_tn_pap(const TnState* st, const double* __restrict__ partials, long long nblk,
                         double* __restrict__ red) {
  if (st->done) return;
  tn_pap_block(st, partials, nblk, red);
}

// k_tn_pap's arguments, for the inverse axis pass that absorbs it.
struct PapArgs {
  const TnState* st;
  const double* part;
  long long nblk;
  double* red;
};

// The inverse axis-1 pass of the TN step with k_tn_pap folded in: the
// pencil's partial sums are complete when this kernel starts, so its first
// workgroup also completes p.Ap (one launch and one dispatch gap fewer per
// CG iteration; the inverse row pass that reads p.Ap runs after it).
template <typename T, int N1, int N2, int W, int SL = 0>
__global__ __launch_bounds__(W * N2) void k_axis_reg_pap(cplx<T>* __restrict__ spec, long long Nh, long long stride,
                                                         long long ostride, int ncol, const cplx<T>* __restrict__ tw,
                                                         int inv, const int* __restrict__ skip, PapArgs pa, int nouter) {
  if (skip && *skip) return;  // a finished device-side CG loop skips the rest of its queue
  __shared__ cplx<T> lds[W * col_block_stride<T, N1, N2, W>()];
  if constexpr (SL >= 1)
    axis_split_block<T, N1, N2, W, SL>(spec, Nh, stride, ostride, ncol, tw, inv, blockIdx.x, blockIdx.y, blockIdx.z,
                                       lds, nouter);
  else
    axis_reg_block<T, N1, N2, W>(spec, Nh, stride, ostride, ncol, tw, inv, blockIdx.x, blockIdx.y, blockIdx.z, lds);
  if ((blockIdx.x | blockIdx.y | blockIdx.z) == 0 && !pa.st->done) tn_pap_block(pa.st, pa.part, pa.nblk, pa.red);
}

// r1 and the three W0 sums of the inverse row pass, the tests of
// optimization.py:116-145, and sum W0 p^2 of the next direction
// p' = r + beta p from W0 r^2 + 2 beta W0 r p + beta^2 W0 p^2.
__device__ __forceinline__ void tn_ctrl4_block(TnState* st, const double* __restrict__ partials, long long nblk,
                                               const double* __restrict__ red, double* __restrict__ rr) {
  // the four sums of a block are adjacent: one pass over the blocks with
  // two independent accumulator sets keeps two 32-byte loads in flight
  double tmp[4] = {0.0, 0.0, 0.0, 0.0}, tm2[4] = {0.0, 0.0, 0.0, 0.0};
  const long long st2 = 2LL * blockDim.x;
  long long b = threadIdx.x;
  for (; b + blockDim.x < nblk; b += st2) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      tmp[s] += partials[b * 4 + s];
      tm2[s] += partials[(b + blockDim.x) * 4 + s];
    }
  }
  if (b < nblk) {
#pragma unroll
    for (int s = 0; s < 4; ++s) tmp[s] += partials[b * 4 + s];
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) tmp[s] += tm2[s];
  block_sums<4>(tmp, rr);
  __syncthreads();
  if (threadIdx.x == 0) {
    tn_ctrl(st, red[0], rr[0]);
    if (st->act & kActNext) {
      const double b = st->beta;
      st->wpp = rr[1] + 2.0 * b * rr[2] + b * b * rr[3];
    }
  }
}

__global__ void k_tn_reduce_ctrl4(TnState* st, const double* __restrict__ partials, long long nblk,
                                  const double* __restrict__ red, double* __restrict__ rr, TnState* hst) {
  if (st->done) return;
  tn_ctrl4_block(st, partials, nblk, red, rr);
  if (threadIdx.x == 0) tn_publish(st, hst);
}

// p = r + beta p when the control asked for it: used only where no forward
// row pass precedes the next product (no convolution fields).
template <typename T>
__global__ void k_tn_next(const TnState* __restrict__ st, const T* __restrict__ r, T* __restrict__ p, long long N) {
  if (st->done || !(st->act & kActNext)) return;
  const T beta = T(st->beta);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
    p[i] = r[i] + beta * p[i];
}

// The terminal update of the loop: direction = Best, or r0 / pAp * p.
template <typename T>
__global__ void k_tn_finish(const TnState* __restrict__ st, T* __restrict__ d, const T* __restrict__ p,
                            const T* __restrict__ best, long long N) {
  const int fin = st->fin;
  if (!fin) return;
  const T scale = T(st->scale);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
    d[i] = (fin & kActScale) ? scale * p[i] : best[i];
}

// The HVP's forward row functor with the CG update of the previous
// iteration folded in: p = r + beta p (optimization.py:144-145) when the
// control asked for it, written back for the step kernel and the next
// update; each point is loaded by exactly one thread of the row pass.
template <typename T, int LAY = 0>
struct PreHvpCG {
  static constexpr int kProf = PC_ROWS_FWD_HVP;
  PreHvp<T, LAY> h;  // its transform of (phi, p); its own pointers are not read
  const T* __restrict__ phi;
  T* __restrict__ p;  // the direction, updated in place
  const T* __restrict__ r;
  const TnState* __restrict__ st;
  __host__ __device__ int nin() const { return 4; }
  // raw loads only, so that the row kernel keeps every load of its points in
  // flight at once; the update and its store happen in apply
  struct In {
    T ph, pp, rr;
  };
  template <class I>
  __device__ __forceinline__ In load(I i) const { return {at(phi, i), at(p, i), at(r, i)}; }
  template <class I>
  __device__ __forceinline__ void apply(const In& x, T* out, I i) const {
    T pp = x.pp;
    if (st->act & kActNext) {
      pp = x.rr + T(st->beta) * pp;
      at(p, i) = pp;
    }
    h.apply(typename PreHvp<T, LAY>::In{x.ph, pp}, out);
  }
  __device__ __forceinline__ void operator()(long long i, T* out) const { apply(load(i), out, i); }
};

// The HVP's inverse row functor with the CG step folded in (the step of
// optimization.py:116-128 with pAp = red[0] already known from the pencil's
// spectral sums and the pointwise part): Best = direction when the previous
// iteration asked for it, d += a p, r -= a (H p) with a = r0 / pAp; H p is
// never stored.  acc: r^2 and, for the pointwise part of the next pAp,
// W0 r^2, W0 r p, W0 p^2.  Nothing changes when pAp < 0 (the reference
// leaves the loop before the update).
template <typename T, int LAY = 0>
struct PostHvpCG {
  static constexpr int NS = 4;
  static constexpr int kProf = PC_ROWS_INV_HVP;
  PostHvpFast<T, LAY> h;  // phi, p, W0 and the functional; its Ap is not written
  T* __restrict__ d;
  T* __restrict__ r;
  T* __restrict__ best;
  const TnState* __restrict__ st;
  const double* __restrict__ red;
  __host__ __device__ int nio() const { return 7; }
  struct In {
    typename PostHvpFast<T, LAY>::In x;
    T dd, rr;
  };
  template <class I>
  __device__ __forceinline__ In load(I i) const { return {h.load(i), at(d, i), at(r, i)}; }
  template <class I>
  __device__ __forceinline__ void apply(I i, const In& in, const T* c, double* acc) const {
    const double pAp = red[0];
    if (pAp < 0.0) return;
    const T ap = h.ap_of(in.x, c);
    const T a = T(st->r0 / pAp);
    if (st->act & kActKeep) at(best, i) = in.dd;
    at(d, i) = in.dd + a * in.x.pp;
    const T rn = in.rr - a * ap;
    at(r, i) = rn;
    const double rv = (double)rn, pv = (double)in.x.pp, w = (double)in.x.w0;
    acc[0] += rv * rv;
    acc[1] += w * rv * rv;
    acc[2] += w * rv * pv;
    acc[3] += w * pv * pv;
  }
  __device__ __forceinline__ void operator()(long long i, const T* c, double* acc) const { apply(i, load(i), c, acc); }
};

// The five sums one outer iteration needs after the direction is found:
// sum p phi, sum phi^2, sum p^2 (OrthogonalNormalization), sum V phi p and
// sum V phi phi (derivative at theta = 0 of the orthonormalised p).
template <typename T>
__global__ void k_orth_sums(const T* __restrict__ p, const T* __restrict__ phi, const T* __restrict__ V, long long N,
                            double* __restrict__ partials) {
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
    const double pp = (double)p[i], ph = (double)phi[i], vph = (double)V[i] * ph;
    acc[0] += pp * ph;
    acc[1] += ph * ph;
    acc[2] += pp * pp;
    acc[3] += vph * pp;
    acc[4] += vph * ph;
  }
  block_sums<5>(acc, partials + (long long)blockIdx.x * 5);
}

// OrthogonalNormalization's factor f = <p,phi>/<phi,phi> and scale
// s = sqrt(Ne / |p - f phi|^2) (optimization.py:205-226) from the five sums
// in red[0..4]: coef = {s, -s f}.  The volume element cancels in Ne / |.|^2
// and is left out, and every operation is rounded on its own (no
// contraction), so the host's formulas give the same bits.
__global__ void k_orth_coef(const double* __restrict__ red, double* __restrict__ coef) {
  if ((threadIdx.x | blockIdx.x) != 0) return;
  const double pphi = red[0], phiphi = red[1], pp = red[2];
  const double factor = __ddiv_rn(pphi, phiphi);
  const double pn = __dsub_rn(pp, __dmul_rn(factor, pphi));
  const double sc = __dsqrt_rn(__ddiv_rn(phiphi, pn));
  coef[0] = sc;
  coef[1] = __dmul_rn(-sc, factor);
}

// z = coef[0] x + coef[1] y with the coefficients read on the device
template <typename T>
__global__ void k_lincomb_coef(const double* __restrict__ coef, const T* x, const T* y, T* z, long long N) {
  const T alpha = T(coef[0]), beta = T(coef[1]);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
    z[i] = lincomb_point(alpha, x[i], beta, y[i]);
}

template <typename T>
__global__ void k_dot_vphip(const T* __restrict__ V, const T* __restrict__ phi, const T* __restrict__ p, long long N,
                            double* __restrict__ partials) {
  double acc[1] = {0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
    acc[0] += (double)V[i] * (double)phi[i] * (double)p[i];
  block_sums<1>(acc, partials + blockIdx.x);
}

template <typename Td, typename Ts>
__global__ void k_convert(Td* __restrict__ dst, const Ts* __restrict__ src, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = (Td)src[i];
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
constexpr int kColTile = kColTileC;

int vec_blocks(long long N) {
  long long b = (N + 255) / 256;
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (int)b;
}

FnParams make_fn(const ofdft_functional& f) {
  FnParams p{};
  p.tf = f.tf;
  p.vw = f.vw;
  p.wt = f.wt;
  p.lda = f.lda;
  p.hartree = f.hartree;
  p.pseudo = f.pseudo;
  p.x = f.x;
  p.y = f.y;
  p.alpha = f.alpha;
  p.beta = f.beta;
  p.beta56 = std::fabs(f.beta - 5.0 / 6.0) < 1e-15 ? 1 : 0;
  int F = 0;
  p.fH = f.hartree ? F++ : -1;
  p.fV = f.vw ? F++ : -1;
  p.fW = f.wt ? F++ : -1;
  p.F = F;
  return p;
}

GridGeom geom(const ofdft_plan_s* P) {
  GridGeom g;
  g.n0 = P->n[0];
  g.n1 = P->n[1];
  g.n2 = P->n[2];
  g.nh = P->nh;
  g.nhp = P->nhp;
  g.full = P->full;
  for (int i = 0; i < 9; ++i) {
    g.bg[i] = P->bg[i];
    g.bgf[i] = (float)P->bg[i];
  }
  return g;
}

// Compile-time FFT lengths; any other 2^a 3^b 5^c length takes the runtime path.
#define OFDFT_FIXED_SIZES(X) X(32) X(48) X(64) X(90) X(96) X(108) X(128) X(160) X(192)

// length n has compile-time row / column / pencil kernels
constexpr bool pencil_fixed(int n) {
#define OFDFT_IS(NN) || n == NN
  return false OFDFT_FIXED_SIZES(OFDFT_IS);
#undef OFDFT_IS
}

// the register pencil (and with it the spectral p.Ap sums of the fused TN
// step) runs on compile-time lengths of plans with 32-bit offsets
inline bool pencil_reg_ok(const ofdft_plan_s* P) { return P->idx32 && pencil_fixed(P->n[0]); }

// Split rows (n2/2 columns, the Nyquist column apart) are possible: not
// deselected, 32-bit offsets, compile-time lengths on all three axes, n2
// even.
inline bool split_rows_ok(const ofdft_plan_s* P) {
  return P->cg_spectra != OFDFT_CG_SPECTRA_PADDED && P->idx32 && pencil_fixed(P->n[0]) && pencil_fixed(P->n[1]) &&
         pencil_fixed(P->n[2]) && !(P->n[2] & 1);
}

// The spectrum layout of the passes of an evaluation or a Hessian product
// (ofdft_plan_set_cg_spectra) and its tile widths (*wa: axis pass, *wp:
// pencils): 2 split rows in the plan precision (the default) where
// split_rows_ok and n2/2 is a multiple of the plan's column tile widths
// (ofdft_plan_set_tiles), else 0 padded rows.
inline int pass_layout(const ofdft_plan_s* P, int* wa = nullptr, int* wp = nullptr) {
  const int nb = P->n[2] / 2;
  const bool split = split_rows_ok(P) && nb % P->tile_axis == 0 && nb % P->tile_pencil == 0;
  if (wa) *wa = split ? P->tile_axis : 0;
  if (wp) *wp = split ? P->tile_pencil : 0;
  return split ? 2 : 0;
}

// The layout of the TN inner loop's Hessian products: 1 (bf16 split rows,
// selected on an FP32 plan where split_rows_ok and n2/2 is a multiple of
// both bf16 tile widths; automatic: the widest of 32, 16, 8 that divides
// it), else that of pass_layout.
inline int cg_layout(const ofdft_plan_s* P, int* wa = nullptr, int* wp = nullptr) {
  const int nb = P->n[2] / 2;
  if (P->cg_spectra == OFDFT_CG_SPECTRA_BF16 && P->prec == OFDFT_F32 && split_rows_ok(P)) {
    int widest = 0;
    for (int w = 32; w >= 8 && !widest; w /= 2)
      if (nb % w == 0) widest = w;
    const int a = P->tile_bf_axis ? P->tile_bf_axis : widest;
    const int p = P->tile_bf_pencil ? P->tile_bf_pencil : widest;
    if (a && p && nb % a == 0 && nb % p == 0) {
      if (wa) *wa = a;
      if (wp) *wp = p;
      return 1;
    }
  }
  return pass_layout(P, wa, wp);
}

constexpr int ceil_div(int a, int b) { return (a + b - 1) / b; }
// columns per workgroup of the register column kernels
template <typename T>
struct RegTile {
  static constexpr int W = 16;
};
// threads of a rows_* workgroup for a compile-time length N: for N = 96
// and 192 (2 x 3 row pairs), 192 threads keep the load, radix-3 and
// point passes fully occupied where 256 would leave a quarter idle
template <int N>
constexpr int rows_th() {
  return (N == 96 || N == 192) ? 192 : 256;
}
constexpr int kColsTH = 256;
template <typename T, int N>
using RowsFFT = FFTFixed<N, kRowPB<T> * 3, rows_th<N>()>;
template <int N>
using ColsFFT = FFTFixed<N, kColTileC, kColsTH>;

// Spectrum layout of a pass: SL = 0 padded rows in the plan precision
// (field stride Nhp, row stride nhp); SL = 1 (bf16) and SL = 2 (plan
// precision) split rows (field stride n0*n1*n2/2 elements, row stride n2/2,
// the Nyquist column apart).
template <int SL>
long long spec_field_stride(const ofdft_plan_s* P) {
  return SL >= 1 ? (long long)P->n[0] * P->n[1] * (P->n[2] / 2) : P->Nhp;
}
template <int SL>
int spec_row_stride(const ofdft_plan_s* P) {
  return SL >= 1 ? P->n[2] / 2 : P->nhp;
}
// bytes of F spectra in a layout (read or written once)
template <int SL>
long long spec_bytes(const ofdft_plan_s* P, int F) {
  return (SL == 1 ? 4LL : 2LL * (long long)P->rs) * F * P->Nh;
}

template <typename T, int SL = 0, class Pre, class FFT>
void go_rows_fwd(ofdft_plan_s* P, const Pre& pre, const FFT& fft, int F, dim3 grid, int threads, size_t lds) {
  if (F == 3) {
    OFDFT_LAUNCH((k_rows_fwd<T, FFT, Pre, 3, SL>), grid, dim3(threads), lds, P, pre, fft,
                 reinterpret_cast<cplx<T>*>(P->spec), (long long)P->n[0] * P->n[1], spec_field_stride<SL>(P),
                 spec_row_stride<SL>(P), F, reinterpret_cast<const cplx<T>*>(P->tw[2]), P->skip);
    return;
  }
  OFDFT_LAUNCH((k_rows_fwd<T, FFT, Pre, 0, SL>), grid, dim3(threads), lds, P, pre, fft,
               reinterpret_cast<cplx<T>*>(P->spec), (long long)P->n[0] * P->n[1], spec_field_stride<SL>(P),
               spec_row_stride<SL>(P), F, reinterpret_cast<const cplx<T>*>(P->tw[2]), P->skip);
}

// SL = 1, 2 (split rows) run on compile-time lengths only (cg_layout)
template <typename T, class Pre, int SL = 0>
int launch_rows_fwd(ofdft_plan_s* P, const Pre& pre, int F) {
  const long long nrows = (long long)P->n[0] * P->n[1];
  const long long npairs = (nrows + 1) / 2;
  const int PB = kRowPB<T>;
  const size_t lds = (size_t)(PB * F + 1) * P->n[2] * sizeof(cplx<T>);
  dim3 grid((unsigned)((npairs + PB - 1) / PB));
  ProfScope ps(P, Pre::kProf, (long long)P->rs * pre.nin() * P->N + spec_bytes<SL>(P, F));
  // compile-time lengths address with 32-bit offsets: plans too large for
  // them take the runtime-schedule kernels
  switch (P->idx32 ? P->n[2] : 0) {
#define OFDFT_CASE(NN) \
    case NN: go_rows_fwd<T, SL>(P, pre, RowsFFT<T, NN>{}, F, grid, rows_th<NN>(), lds); break;
    OFDFT_FIXED_SIZES(OFDFT_CASE)
#undef OFDFT_CASE
    default: {
      if (SL >= 1) return fail(OFDFT_EINVAL, "split-row spectra need a compile-time row length");
      FFTLaunch L = fft_launch(P->pl[2], PB * F);
      // radix 11 / 13 schedules: one instantiation, up to four slots per thread
      if (has_large(P->pl[2])) go_rows_fwd<T>(P, pre, FFTRuntime<4, true>{P->pl[2]}, F, grid, L.threads, lds);
      else if (L.kb == 1) go_rows_fwd<T>(P, pre, FFTRuntime<1>{P->pl[2]}, F, grid, L.threads, lds);
      else if (L.kb == 2) go_rows_fwd<T>(P, pre, FFTRuntime<2>{P->pl[2]}, F, grid, L.threads, lds);
      else go_rows_fwd<T>(P, pre, FFTRuntime<4>{P->pl[2]}, F, grid, L.threads, lds);
    }
  }
  LAUNCHCHK();
  return 0;
}

template <typename T, int SL = 0, class Post, class FFT>
void go_rows_inv(ofdft_plan_s* P, const Post& post, const FFT& fft, int F, const cplx<T>* src, dim3 grid,
                 int threads, size_t lds, double* part) {
  if (F == 3) {
    OFDFT_LAUNCH((k_rows_inv<T, FFT, Post, 3, SL>), grid, dim3(threads), lds, P, post, fft, src,
                 (long long)P->n[0] * P->n[1], spec_field_stride<SL>(P), spec_row_stride<SL>(P), F,
                 reinterpret_cast<const cplx<T>*>(P->tw[2]), part, P->skip);
    return;
  }
  OFDFT_LAUNCH((k_rows_inv<T, FFT, Post, 0, SL>), grid, dim3(threads), lds, P, post, fft, src,
               (long long)P->n[0] * P->n[1], spec_field_stride<SL>(P), spec_row_stride<SL>(P), F,
               reinterpret_cast<const cplx<T>*>(P->tw[2]), part, P->skip);
}

// reduce: the post functor's sums end in P->red, and also in the mapped host
// buffer when to_host (for a fetch_sums that follows).
template <typename T, class Post, int SL = 0>
int launch_rows_inv(ofdft_plan_s* P, const Post& post, int F, const cplx<T>* src, bool reduce, bool to_host = true,
                    bool final_reduce = true) {
  const long long nrows = (long long)P->n[0] * P->n[1];
  const long long npairs = (nrows + 1) / 2;
  const int PB = kRowPB<T>;
  const size_t lds = (size_t)(PB * F + 1) * P->n[2] * sizeof(cplx<T>);
  const long long nblk = (npairs + PB - 1) / PB;
  if (reduce && nblk * Post::NS > P->npartials) return fail(OFDFT_ESTATE, "partials buffer too small");
  dim3 grid((unsigned)nblk);
  double* part = reduce ? P->partials : nullptr;
  {
    ProfScope ps(P, Post::kProf, (long long)P->rs * post.nio() * P->N + spec_bytes<SL>(P, F));
    switch (P->idx32 ? P->n[2] : 0) {
#define OFDFT_CASE(NN) \
      case NN: go_rows_inv<T, SL>(P, post, RowsFFT<T, NN>{}, F, src, grid, rows_th<NN>(), lds, part); break;
      OFDFT_FIXED_SIZES(OFDFT_CASE)
#undef OFDFT_CASE
      default: {
        if (SL >= 1) return fail(OFDFT_EINVAL, "split-row spectra need a compile-time row length");
        FFTLaunch L = fft_launch(P->pl[2], PB * F, 256);
        // radix 11 / 13 schedules: one instantiation, up to four slots per thread
        if (has_large(P->pl[2])) go_rows_inv<T>(P, post, FFTRuntime<4, true>{P->pl[2]}, F, src, grid, L.threads, lds, part);
        else if (L.kb == 1) go_rows_inv<T>(P, post, FFTRuntime<1>{P->pl[2]}, F, src, grid, L.threads, lds, part);
        else if (L.kb == 2) go_rows_inv<T>(P, post, FFTRuntime<2>{P->pl[2]}, F, src, grid, L.threads, lds, part);
        else go_rows_inv<T>(P, post, FFTRuntime<4>{P->pl[2]}, F, src, grid, L.threads, lds, part);
      }
    }
  }
  LAUNCHCHK();
  // final_reduce = false: the block partials stay for the caller's own reduction
  if (reduce && final_reduce) {
    hipLaunchKernelGGL((k_reduce_partials<Post::NS>), dim3(Post::NS), dim3(kRedThreads), 0, P->stream, P->partials, nblk, P->red,
                       to_host ? P->dhred : nullptr, P->skip);
    LAUNCHCHK();
  }
  return 0;
}

template <typename T, class FFT>
void go_axis(ofdft_plan_s* P, const FFT& fft, cplx<T>* spec, int axis, dim3 grid, int threads, size_t lds,
             long long stride, long long ostride, bool inv) {
  OFDFT_LAUNCH((k_axis_c2c<T, FFT>), grid, dim3(threads), lds, P, fft, spec, P->Nhp, stride, ostride,
                     P->nh, reinterpret_cast<const cplx<T>*>(P->tw[axis]), inv ? 1 : 0, P->skip);
}

template <typename T, int N1, int N2, int W, int SL>
void go_axis_reg(ofdft_plan_s* P, dim3 grid, cplx<T>* spec, long long Nh, long long stride, long long ostride,
                 int ncol, const cplx<T>* tw, bool inv, const PapArgs* pa) {
  if (pa) {
    OFDFT_LAUNCH((k_axis_reg_pap<T, N1, N2, W, SL>), grid, dim3(W * N2), 0, P, spec, Nh, stride, ostride, ncol, tw,
                 inv ? 1 : 0, P->skip, *pa, (int)grid.y);
    return;
  }
  OFDFT_LAUNCH((k_axis_reg<T, N1, N2, W, SL>), grid, dim3(W * N2), 0, P, spec, Nh, stride, ostride, ncol, tw,
               inv ? 1 : 0, P->skip, (int)grid.y);
}

template <typename T, int N1, int N2, int W>
void axis_reg(ofdft_plan_s* P, cplx<T>* spec, int axis, int F, long long stride, long long ostride, unsigned outer,
              bool inv, const PapArgs* pa) {
  go_axis_reg<T, N1, N2, W, 0>(P, dim3((unsigned)((P->nhp + W - 1) / W), outer, (unsigned)F), spec, P->Nhp, stride,
                               ostride, P->nh, reinterpret_cast<const cplx<T>*>(P->tw[axis]), inv, pa);
}

// pa: also complete p.Ap of the TN step (k_tn_pap), folded into the
// register kernel or launched after the runtime-schedule one.
template <typename T>
int launch_axis(ofdft_plan_s* P, cplx<T>* spec, int axis, int F, bool inv, const PapArgs* pa = nullptr) {
  const int W = kColTileC;
  const int n = P->n[axis];
  const size_t lds = (size_t)(W * (n + 1) + n) * sizeof(cplx<T>);
  long long stride, ostride;
  unsigned outer;
  if (axis == 1) {
    stride = P->nhp;
    ostride = (long long)P->n[1] * P->nhp;
    outer = P->n[0];
  } else {
    stride = (long long)P->n[1] * P->nhp;
    ostride = P->nhp;
    outer = P->n[1];
  }
  dim3 grid((unsigned)((P->nh + W - 1) / W), outer, (unsigned)F);
  ProfScope ps(P, PC_AXIS, (long long)P->rs * 4LL * F * P->Nh);
  // the register kernels address with 32-bit offsets (ofdft_plan_s::idx32)
  switch (P->idx32 ? n : 0) {
#define OFDFT_CASE(NN)                                                                                       \
    case NN:                                                                                                 \
      if (P->tile_axis == 8) axis_reg<T, Split<NN>::N1, Split<NN>::N2, 8>(P, spec, axis, F, stride, ostride, outer, inv, pa); \
      else if (sizeof(T) == 4 && P->tile_axis == 32)                                                           \
        axis_reg<T, Split<NN>::N1, Split<NN>::N2, (sizeof(T) == 4 ? 32 : 16)>(P, spec, axis, F, stride, ostride, outer, inv, pa); \
      else axis_reg<T, Split<NN>::N1, Split<NN>::N2, 16>(P, spec, axis, F, stride, ostride, outer, inv, pa);     \
      break;
    OFDFT_FIXED_SIZES(OFDFT_CASE)
#undef OFDFT_CASE
    default: {
      const Pass1D& pl = P->pl[axis];
      FFTLaunch L = fft_launch(pl, W);
      // radix 11 / 13 schedules: one instantiation, up to four slots per thread
      if (has_large(pl)) go_axis<T>(P, FFTRuntime<4, true>{pl}, spec, axis, grid, L.threads, lds, stride, ostride, inv);
      else if (L.kb == 1) go_axis<T>(P, FFTRuntime<1>{pl}, spec, axis, grid, L.threads, lds, stride, ostride, inv);
      else if (L.kb == 2) go_axis<T>(P, FFTRuntime<2>{pl}, spec, axis, grid, L.threads, lds, stride, ostride, inv);
      else go_axis<T>(P, FFTRuntime<4>{pl}, spec, axis, grid, L.threads, lds, stride, ostride, inv);
      if (pa) {
        LAUNCHCHK();
        hipLaunchKernelGGL(k_tn_pap, dim3(1), dim3(kRedThreads), 0, P->stream, pa->st, pa->part, pa->nblk, pa->red);
      }
    }
  }
  LAUNCHCHK();
  return 0;
}

template <typename T, class FFT>
void go_pencil(ofdft_plan_s* P, const FFT& fft, dim3 grid, int threads, size_t lds, const GridGeom& g,
               const ConvKinds& ck) {
  OFDFT_LAUNCH((k_pencil_conv<T, FFT>), grid, dim3(threads), lds, P, fft,
                     reinterpret_cast<cplx<T>*>(P->spec), P->Nhp, g, reinterpret_cast<const cplx<T>*>(P->tw[0]), ck,
                     reinterpret_cast<const T*>(P->wtk), P->skip);
}

template <typename T, int N1, int N2, int W, int SL>
void go_pencil_reg(ofdft_plan_s* P, dim3 grid, const GridGeom& g, const ConvKinds& ck, double* cgp, long long Nh) {
  OFDFT_LAUNCH((k_pencil_reg<T, N1, N2, W, SL>), grid, dim3(W * N2), 0, P, reinterpret_cast<cplx<T>*>(P->spec), Nh,
               g, reinterpret_cast<const cplx<T>*>(P->tw[0]), ck,
               reinterpret_cast<const T*>(SL >= 1 ? P->wtks : P->wtk), P->skip, cgp);
}

// workgroups of the register pencils of tile width W (bf: the bf16 layout's
// nb/W column tiles plus the Nyquist slot); one p.Ap partial sum each
inline dim3 pencil_grid(const ofdft_plan_s* P, int W, int F, bool bf) {
  const unsigned tx = bf ? (unsigned)(P->n[2] / 2 / W + 1) : (unsigned)((P->nhp + W - 1) / W);
  return dim3(tx, (unsigned)P->n[1], (unsigned)F);
}

template <typename T, int N1, int N2, int W>
void pencil_reg(ofdft_plan_s* P, int F, const GridGeom& g, const ConvKinds& ck, double* cgp) {
  go_pencil_reg<T, N1, N2, W, 0>(P, pencil_grid(P, W, F, false), g, ck, cgp, P->Nhp);
}

template <typename T>
int launch_pencil(ofdft_plan_s* P, int F, const int* kinds, double* cgp = nullptr, const double* coef = nullptr) {
  const int W = kColTileC;
  const size_t lds = (size_t)(W * (P->n[0] + 1) + P->n[0]) * sizeof(cplx<T>);
  dim3 grid((unsigned)((P->nh + W - 1) / W), (unsigned)P->n[1], (unsigned)F);
  ConvKinds ck;
  for (int f = 0; f < 3; ++f) ck.kind[f] = f < F ? kinds[f] : 0;
  ck.scale = 1.0 / (double)P->N;
  for (int f = 0; f < 3; ++f) ck.coef[f] = coef ? coef[f] : 0.0;
  if (cgp && !pencil_reg_ok(P)) return fail(OFDFT_EINVAL, "p.Ap sums need a compile-time pencil length");
  for (int f = 0; f < F; ++f)
    if (kinds[f] == OFDFT_CONV_WT && !(P->have_wtk && P->wtks)) return fail(OFDFT_ESTATE, "WT kernel table not built");
  GridGeom g = geom(P);
  long long nwt = 0;
  for (int f = 0; f < F; ++f) nwt += kinds[f] == OFDFT_CONV_WT ? 1 : 0;
  ProfScope ps(P, PC_PENCIL, (long long)P->rs * (4LL * F + nwt) * P->Nh);
  switch (P->idx32 ? P->n[0] : 0) {
#define OFDFT_CASE(NN)                                                                                   \
    case NN:                                                                                             \
      if (P->tile_pencil == 8) pencil_reg<T, Split<NN>::N1, Split<NN>::N2, 8>(P, F, g, ck, cgp);        \
      else if (sizeof(T) == 4 && P->tile_pencil == 32)                                                   \
        pencil_reg<T, Split<NN>::N1, Split<NN>::N2, (sizeof(T) == 4 ? 32 : 16)>(P, F, g, ck, cgp);      \
      else pencil_reg<T, Split<NN>::N1, Split<NN>::N2, 16>(P, F, g, ck, cgp);                           \
      break;
    OFDFT_FIXED_SIZES(OFDFT_CASE)
#undef OFDFT_CASE
    default: {
      FFTLaunch L = fft_launch(P->pl[0], W);
      // radix 11 / 13 schedules: one instantiation, up to four slots per thread
      if (has_large(P->pl[0])) go_pencil<T>(P, FFTRuntime<4, true>{P->pl[0]}, grid, L.threads, lds, g, ck);
      else if (L.kb == 1) go_pencil<T>(P, FFTRuntime<1>{P->pl[0]}, grid, L.threads, lds, g, ck);
      else if (L.kb == 2) go_pencil<T>(P, FFTRuntime<2>{P->pl[0]}, grid, L.threads, lds, g, ck);
      else go_pencil<T>(P, FFTRuntime<4>{P->pl[0]}, grid, L.threads, lds, g, ck);
    }
  }
  LAUNCHCHK();
  return 0;
}

// The axis-1 pass over split-row spectra (SL = 1 bf16, SL = 2 plan
// precision; axis_split_block): nb/W column tiles plus one Nyquist tile slot
// per plane, W columns per workgroup.
template <typename T, int N1, int N2, int W, int SL>
void axis_reg_split(ofdft_plan_s* P, int F, bool inv, const PapArgs* pa) {
  const int nb = P->n[2] / 2;
  go_axis_reg<T, N1, N2, W, SL>(P, dim3((unsigned)(nb / W + 1), (unsigned)P->n[0], (unsigned)F),
                                reinterpret_cast<cplx<T>*>(P->spec), spec_field_stride<SL>(P), nb,
                                (long long)P->n[1] * nb, nb, reinterpret_cast<const cplx<T>*>(P->tw[1]), inv, pa);
}

template <typename T, int SL>
int launch_axis_split(ofdft_plan_s* P, int F, bool inv, int W, const PapArgs* pa = nullptr) {
  static_assert(SL == 2 || (SL == 1 && sizeof(T) == 4), "bf16 spectra are for FP32 plans");
  ProfScope ps(P, PC_AXIS, 2 * spec_bytes<SL>(P, F));
  switch (P->n[1]) {
#define OFDFT_CASE(NN)                                                                                       \
    case NN:                                                                                                 \
      if (W == 8) axis_reg_split<T, Split<NN>::N1, Split<NN>::N2, 8, SL>(P, F, inv, pa);                    \
      else if (sizeof(T) == 4 && W == 32)                                                                    \
        axis_reg_split<T, Split<NN>::N1, Split<NN>::N2, (sizeof(T) == 4 ? 32 : 16), SL>(P, F, inv, pa);      \
      else axis_reg_split<T, Split<NN>::N1, Split<NN>::N2, 16, SL>(P, F, inv, pa);                          \
      break;
    OFDFT_FIXED_SIZES(OFDFT_CASE)
#undef OFDFT_CASE
    default:
      return fail(OFDFT_EINVAL, "split-row spectra need a compile-time axis-1 length");
  }
  LAUNCHCHK();
  return 0;
}

// The axis-0 pencils over split-row spectra (pencil_split_block); cgp
// receives one partial sum per workgroup, (nb/W + 1) * n1 * F of them.
template <typename T, int N1, int N2, int W, int SL>
void pencil_reg_split(ofdft_plan_s* P, int F, const GridGeom& g, const ConvKinds& ck, double* cgp) {
  go_pencil_reg<T, N1, N2, W, SL>(P, pencil_grid(P, W, F, true), g, ck, cgp, spec_field_stride<SL>(P));
}

template <typename T, int SL>
int launch_pencil_split(ofdft_plan_s* P, int F, const int* kinds, double* cgp, const double* coef, int W) {
  static_assert(SL == 2 || (SL == 1 && sizeof(T) == 4), "bf16 spectra are for FP32 plans");
  ConvKinds ck;
  for (int f = 0; f < 3; ++f) ck.kind[f] = f < F ? kinds[f] : 0;
  ck.scale = 1.0 / (double)P->N;
  for (int f = 0; f < 3; ++f) ck.coef[f] = coef ? coef[f] : 0.0;
  for (int f = 0; f < F; ++f)
    if (kinds[f] == OFDFT_CONV_WT && !(P->have_wtk && P->wtks)) return fail(OFDFT_ESTATE, "WT kernel table not built");
  GridGeom g = geom(P);
  long long nwt = 0;
  for (int f = 0; f < F; ++f) nwt += kinds[f] == OFDFT_CONV_WT ? 1 : 0;
  ProfScope ps(P, PC_PENCIL, 2 * spec_bytes<SL>(P, F) + (long long)P->rs * nwt * P->Nh);
  switch (P->n[0]) {
#define OFDFT_CASE(NN)                                                                                   \
    case NN:                                                                                             \
      if (W == 8) pencil_reg_split<T, Split<NN>::N1, Split<NN>::N2, 8, SL>(P, F, g, ck, cgp);           \
      else if (sizeof(T) == 4 && W == 32)                                                                \
        pencil_reg_split<T, Split<NN>::N1, Split<NN>::N2, (sizeof(T) == 4 ? 32 : 16), SL>(P, F, g, ck, cgp); \
      else pencil_reg_split<T, Split<NN>::N1, Split<NN>::N2, 16, SL>(P, F, g, ck, cgp);                 \
      break;
    OFDFT_FIXED_SIZES(OFDFT_CASE)
#undef OFDFT_CASE
    default:
      return fail(OFDFT_EINVAL, "split-row spectra need a compile-time pencil length");
  }
  LAUNCHCHK();
  return 0;
}

// the final reduction kernel also wrote its sums into the mapped host buffer
int fetch_sums(ofdft_plan_s* P, int ns, double* out) {
  HIPCHK(hipStreamSynchronize(P->stream));
  for (int i = 0; i < ns; ++i) out[i] = P->hred[i];
  return 0;
}

void finalize_energies(const ofdft_plan_s* P, double* s) {
  const ofdft_functional& f = P->fn;
  const double dV = P->dV;
  const double ctf = (3.0 / 10.0) * std::pow(3.0 * kPi * kPi, 2.0 / 3.0);
  s[OFDFT_E_TF] = f.tf ? s[OFDFT_E_TF] * ctf * dV * f.x : 0.0;
  s[OFDFT_E_VW] = f.vw ? s[OFDFT_E_VW] * dV * f.y : 0.0;
  s[OFDFT_E_NL] = f.wt ? s[OFDFT_E_NL] * dV : 0.0;
  s[OFDFT_E_XC] = f.lda ? s[OFDFT_E_XC] * dV : 0.0;
  s[OFDFT_E_H] = f.hartree ? s[OFDFT_E_H] * dV / 2.0 : 0.0;
  s[OFDFT_E_PP] = f.pseudo ? s[OFDFT_E_PP] * dV : 0.0;
  s[OFDFT_S_VRHO] *= dV;
  s[OFDFT_S_RHO] *= dV;
  s[OFDFT_S_DTHETA] *= 2.0 * dV;
}

// The five passes of an evaluation or a Hessian product: forward rows with
// `pre` (F fields), axis-1, the pencils with the multipliers `kinds`,
// inverse axis-1, inverse rows with `post` (reduce: its sums to P->red and,
// with to_host, the mapped host buffer).  The spectra take the split rows of
// pass_layout where the plan allows, else the padded rows.
template <typename T, class Pre, class Post>
int conv_passes(ofdft_plan_s* P, const Pre& pre, const Post& post, int F, const int* kinds, bool reduce,
                bool to_host = true) {
  auto* spec = reinterpret_cast<cplx<T>*>(P->spec);
  int rc, wa = 0, wp = 0;
  if (F > 0 && pass_layout(P, &wa, &wp) == 2) {
    if ((rc = launch_rows_fwd<T, Pre, 2>(P, pre, F))) return rc;
    if ((rc = launch_axis_split<T, 2>(P, F, false, wa))) return rc;
    if ((rc = launch_pencil_split<T, 2>(P, F, kinds, nullptr, nullptr, wp))) return rc;
    if ((rc = launch_axis_split<T, 2>(P, F, true, wa))) return rc;
    return launch_rows_inv<T, Post, 2>(P, post, F, spec, reduce, to_host);
  }
  if (F > 0) {
    if ((rc = launch_rows_fwd<T>(P, pre, F))) return rc;
    if ((rc = launch_axis<T>(P, spec, 1, F, false))) return rc;
    if ((rc = launch_pencil<T>(P, F, kinds))) return rc;
    if ((rc = launch_axis<T>(P, spec, 1, F, true))) return rc;
  }
  return launch_rows_inv<T>(P, post, F > 0 ? F : 1, spec, reduce, to_host);
}

// multipliers of the convolution fields of a functional
inline void conv_kinds(const FnParams& fp, int* kinds) {
  kinds[0] = kinds[1] = kinds[2] = 0;
  if (fp.fH >= 0) kinds[fp.fH] = OFDFT_CONV_HARTREE;
  if (fp.fV >= 0) kinds[fp.fV] = OFDFT_CONV_GG;
  if (fp.fW >= 0) kinds[fp.fW] = OFDFT_CONV_WT;
}

template <typename T>
int eval_impl(ofdft_plan_s* P, const T* rho, const T* phi, const T* p, double c, double s, T* V, T* aux,
              double* sums) {
  if (!P->have_fn) return fail(OFDFT_ESTATE, "functional not set");
  if (P->fn.pseudo && !P->have_vloc) return fail(OFDFT_ESTATE, "V_loc not set");
  FnParams fp = make_fn(P->fn);
  PreEval<T> pre{rho, phi, p, T(c), T(s), fp};
  int rc, kinds[3];
  conv_kinds(fp, kinds);
  PostEval<T> post{pre, reinterpret_cast<const T*>(P->vloc), V, aux, P->N};
  if ((rc = conv_passes<T>(P, pre, post, fp.F, kinds, true))) return rc;
  if (sums) {
    double tmp[OFDFT_NSUM];
    if ((rc = fetch_sums(P, OFDFT_NSUM, tmp))) return rc;
    finalize_energies(P, tmp);
    for (int i = 0; i < OFDFT_NSUM; ++i) sums[i] = tmp[i];
  }
  return 0;
}

template <typename T>
int hvp_impl(ofdft_plan_s* P, const T* phi, const T* p, const T* V, const T* aux, double mu, T* Ap, double* pAp,
             int raw = 0) {
  if (!P->have_fn) return fail(OFDFT_ESTATE, "functional not set");
  FnParams fp = make_fn(P->fn);
  int rc, kinds[3];
  conv_kinds(fp, kinds);
  PreHvp<T> pre{phi, p, fp};
  PostHvp<T> post{phi, p, V, aux, Ap, T(mu), P->N, fp, raw};
  if ((rc = conv_passes<T>(P, pre, post, fp.F, kinds, true))) return rc;
  if (pAp) {
    double t;
    if ((rc = fetch_sums(P, 1, &t))) return rc;
    *pAp = t;
  }
  return 0;
}

// The exact HVP; with `cg`, the forward row pass first applies the pending
// CG update p = r + beta p of the device-side TN loop (p is then updated in
// place; with no convolution fields the caller does that update itself).
template <typename T, int LAY>
int hvp_fast_lay(ofdft_plan_s* P, const FnParams& fp, const T* phi, const T* p, const T* W0, T* Ap, double* pAp,
                 const T* cg_r, const TnState* cg_st) {
  int rc, kinds[3];
  conv_kinds(fp, kinds);
  PreHvp<T, LAY> pre{phi, p, fp};
  PostHvpFast<T, LAY> post{phi, p, W0, Ap, fp};
  if (cg_st) {
    PreHvpCG<T, LAY> pcg{pre, phi, const_cast<T*>(p), cg_r, cg_st};
    rc = conv_passes<T>(P, pcg, post, fp.F, kinds, true, pAp != nullptr);
  } else {
    rc = conv_passes<T>(P, pre, post, fp.F, kinds, true, pAp != nullptr);
  }
  if (rc) return rc;
  if (pAp) {
    double t;
    if ((rc = fetch_sums(P, 1, &t))) return rc;
    *pAp = t;
  }
  return 0;
}

template <typename T>
int hvp_fast_impl(ofdft_plan_s* P, const T* phi, const T* p, const T* W0, T* Ap, double* pAp,
                  const T* cg_r = nullptr, const TnState* cg_st = nullptr) {
  if (!P->have_fn) return fail(OFDFT_ESTATE, "functional not set");
  const FnParams fp = make_fn(P->fn);
  return fixed_layout(fp) ? hvp_fast_lay<T, 1>(P, fp, phi, p, W0, Ap, pAp, cg_r, cg_st)
                          : hvp_fast_lay<T, 0>(P, fp, phi, p, W0, Ap, pAp, cg_r, cg_st);
}

// One inner CG step of the device TN loop: the HVP's transforms with the
// spectral part of pAp summed in the pencil, pAp completed with sum W0 p^2
// (k_tn_pap), and the step taken in the inverse row pass (PostHvpCG).
// *nblk_rows: the partial-sum blocks that pass leaves for k_tn_reduce_ctrl4.
template <typename T, int LAY>
int hvp_cg_lay(ofdft_plan_s* P, const FnParams& fp, const T* phi, T* p, const T* W0, T* d, T* r, T* best,
               const TnState* st, long long* nblk_rows) {
  const int F = fp.F;
  auto* spec = reinterpret_cast<cplx<T>*>(P->spec);
  int rc;
  long long nblk_pen = 0;
  // spectrum layout of the passes (cg_layout): split rows (the Nyquist
  // column apart) in the plan precision or, selected on FP32 plans, bf16
  int wa = 0, wp = 0;
  const int sl = F > 0 ? cg_layout(P, &wa, &wp) : 0;
  if (F > 0) {
    PreHvp<T, LAY> pre{phi, p, fp};
    PreHvpCG<T, LAY> pcg{pre, phi, p, r, st};
    int kinds[3];
    conv_kinds(fp, kinds);
    // p.Ap = sum W0 p^2 + sum_f coef_f g_f.(K_f * g_f) with g_f the field the
    // forward pass transforms: Hartree g = 2 phi p, vW g = sign(phi) p, WT
    // g = 2 beta rho^(beta-1) phi p (the weights of PostHvpFast over these)
    double coef[3] = {0.0, 0.0, 0.0};
    if (fp.fH >= 0) coef[fp.fH] = 0.5;
    if (fp.fV >= 0) coef[fp.fV] = 0.5 * fp.y;
    if (fp.fW >= 0) coef[fp.fW] = fp.alpha / fp.beta;
    {
      const dim3 pg = pencil_grid(P, sl ? wp : P->tile_pencil, F, sl != 0);
      nblk_pen = (long long)pg.x * pg.y * pg.z;
    }
    if (nblk_pen > P->npartials) return fail(OFDFT_ESTATE, "partials buffer too small");
    // p.Ap is completed by the inverse axis pass (k_tn_pap folded in)
    const PapArgs pa{st, (const double*)P->partials, nblk_pen, (double*)P->red};
    if (sl == 2) {
      if ((rc = launch_rows_fwd<T, PreHvpCG<T, LAY>, 2>(P, pcg, F))) return rc;
      if ((rc = launch_axis_split<T, 2>(P, F, false, wa))) return rc;
      if ((rc = launch_pencil_split<T, 2>(P, F, kinds, P->partials, coef, wp))) return rc;
      if ((rc = launch_axis_split<T, 2>(P, F, true, wa, &pa))) return rc;
    } else if (sl == 1) {
      if constexpr (sizeof(T) == 4) {
        if ((rc = launch_rows_fwd<T, PreHvpCG<T, LAY>, 1>(P, pcg, F))) return rc;
        if ((rc = launch_axis_split<T, 1>(P, F, false, wa))) return rc;
        if ((rc = launch_pencil_split<T, 1>(P, F, kinds, P->partials, coef, wp))) return rc;
        if ((rc = launch_axis_split<T, 1>(P, F, true, wa, &pa))) return rc;
      }
    } else {
      if ((rc = launch_rows_fwd<T>(P, pcg, F))) return rc;
      if ((rc = launch_axis<T>(P, spec, 1, F, false))) return rc;
      if ((rc = launch_pencil<T>(P, F, kinds, P->partials, coef))) return rc;
      if ((rc = launch_axis<T>(P, spec, 1, F, true, &pa))) return rc;
    }
  } else {
    hipLaunchKernelGGL(k_tn_pap, dim3(1), dim3(kRedThreads), 0, P->stream, st, (const double*)P->partials, nblk_pen,
                       P->red);
    LAUNCHCHK();
  }
  PostHvpFast<T, LAY> h{phi, p, W0, nullptr, fp};
  PostHvpCG<T, LAY> post{h, d, r, best, st, (const double*)P->red};
  if (sl == 2) {
    if ((rc = launch_rows_inv<T, PostHvpCG<T, LAY>, 2>(P, post, F, spec, true, false, false))) return rc;
  } else if (sl == 1) {
    if constexpr (sizeof(T) == 4) {
      if ((rc = launch_rows_inv<T, PostHvpCG<T, LAY>, 1>(P, post, F, spec, true, false, false))) return rc;
    }
  } else if ((rc = launch_rows_inv<T>(P, post, F > 0 ? F : 1, spec, true, false, false))) {
    return rc;
  }
  const long long npairs = ((long long)P->n[0] * P->n[1] + 1) / 2;
  *nblk_rows = (npairs + kRowPB<T> - 1) / kRowPB<T>;
  return 0;
}

template <typename T>
int hvp_cg_impl(ofdft_plan_s* P, const T* phi, T* p, const T* W0, T* d, T* r, T* best, const TnState* st,
                long long* nblk_rows) {
  const FnParams fp = make_fn(P->fn);
  return fixed_layout(fp) ? hvp_cg_lay<T, 1>(P, fp, phi, p, W0, d, r, best, st, nblk_rows)
                          : hvp_cg_lay<T, 0>(P, fp, phi, p, W0, d, r, best, st, nblk_rows);
}

template <typename T>
int r2c_impl(ofdft_plan_s* P, const T* in, cplx<T>* out) {
  // rows -> axis1 -> axis0 in the workspace, then copy the spectrum out.
  PreCopy<T> pre{in};
  auto* spec = reinterpret_cast<cplx<T>*>(P->spec);
  int rc;
  if ((rc = launch_rows_fwd<T>(P, pre, 1))) return rc;
  if ((rc = launch_axis<T>(P, spec, 1, 1, false))) return rc;
  if ((rc = launch_axis<T>(P, spec, 0, 1, false))) return rc;
  HIPCHK(hipMemcpy2DAsync(out, sizeof(cplx<T>) * P->nh, spec, sizeof(cplx<T>) * P->nhp, sizeof(cplx<T>) * P->nh,
                          (size_t)P->n[0] * P->n[1], hipMemcpyDeviceToDevice, P->stream));
  return 0;
}

template <typename T>
__global__ void k_scale_cplx(cplx<T>* z, long long n, T s) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    z[i].x *= s;
    z[i].y *= s;
  }
}

template <typename T>
int c2r_impl(ofdft_plan_s* P, const cplx<T>* in, T* out) {
  auto* spec = reinterpret_cast<cplx<T>*>(P->spec);
  int rc;
  HIPCHK(hipMemcpy2DAsync(spec, sizeof(cplx<T>) * P->nhp, in, sizeof(cplx<T>) * P->nh, sizeof(cplx<T>) * P->nh,
                          (size_t)P->n[0] * P->n[1], hipMemcpyDeviceToDevice, P->stream));
  hipLaunchKernelGGL((k_scale_cplx<T>), dim3(vec_blocks(P->Nhp)), dim3(256), 0, P->stream, spec, P->Nhp,
                     T(1.0 / (double)P->N));
  LAUNCHCHK();
  if ((rc = launch_axis<T>(P, spec, 0, 1, true))) return rc;
  if ((rc = launch_axis<T>(P, spec, 1, 1, true))) return rc;
  PostStore<T> post{out};
  return launch_rows_inv<T>(P, post, 1, spec, false);
}

template <typename T>
int conv_impl(ofdft_plan_s* P, int kind, const T* in, T* out) {
  PreCopy<T> pre{in};
  auto* spec = reinterpret_cast<cplx<T>*>(P->spec);
  int rc;
  if ((rc = launch_rows_fwd<T>(P, pre, 1))) return rc;
  if ((rc = launch_axis<T>(P, spec, 1, 1, false))) return rc;
  if ((rc = launch_pencil<T>(P, 1, &kind))) return rc;
  if ((rc = launch_axis<T>(P, spec, 1, 1, true))) return rc;
  PostStore<T> post{out};
  return launch_rows_inv<T>(P, post, 1, spec, false);
}

int reduce_blocks(ofdft_plan_s* P, int nblk, int ns, double* out) {
  int rc;
  switch (ns) {
    case 1: hipLaunchKernelGGL((k_reduce_partials<1>), dim3(1), dim3(kRedThreads), 0, P->stream, P->partials, (long long)nblk, P->red, P->dhred, P->skip); break;
    case 3: hipLaunchKernelGGL((k_reduce_partials<3>), dim3(3), dim3(kRedThreads), 0, P->stream, P->partials, (long long)nblk, P->red, P->dhred, P->skip); break;
    case 5: hipLaunchKernelGGL((k_reduce_partials<5>), dim3(5), dim3(kRedThreads), 0, P->stream, P->partials, (long long)nblk, P->red, P->dhred, P->skip); break;
    case 6: hipLaunchKernelGGL((k_reduce_partials<6>), dim3(6), dim3(kRedThreads), 0, P->stream, P->partials, (long long)nblk, P->red, P->dhred, P->skip); break;
    default: return fail(OFDFT_EINVAL, "unsupported number of sums");
  }
  LAUNCHCHK();
  if ((rc = fetch_sums(P, ns, out))) return rc;
  return 0;
}


// ===========================================================================
// ions: local pseudopotential (PME / exact structure factor), local forces,
// Ewald ion-ion energy.  FP64 plans only (set-up and forces are computed in
// FP64; an FP32 plan receives V_loc through ofdft_convert).
// ===========================================================================
struct SplineDev {
  const double* t;
  const double* c;
  int nt;
  double q1, dq, qmax, v0, v1, v2;
};

// splev of the cubic interpolating spline (FITPACK fpbspl recursion) on
// [q1, qmax); quartic interpolation below q1; zero from qmax on
// (calc_vlines, pseudo/__init__.py:398-423; math_utils.py:390-400).
__device__ double vline_eval(const SplineDev& sp, double x) {
  if (!(x < sp.qmax)) return 0.0;
  if (x < sp.q1) {
    const double f0 = sp.v2, f1 = sp.v1, f2 = sp.v0, f3 = sp.v1, f4 = sp.v2;
    const double t1 = 2.0 * f0 - 16.0 * f1 + 16.0 * f3 - 2.0 * f4;
    const double t2 = -f0 + 16.0 * f1 - 30.0 * f2 + 16.0 * f3 - f4;
    const double t3 = -2.0 * f0 + 4.0 * f1 - 4.0 * f3 + 2.0 * f4;
    const double t4 = f0 - 4.0 * f1 + 6.0 * f2 - 4.0 * f3 + f4;
    const double dx = x / sp.dq;
    return f2 + dx * (t1 + dx * (t2 + dx * (t3 + dx * t4))) / 24.0;
  }
  const int k = 3;
  const double* t = sp.t;
  const int hi = sp.nt - k - 2;
  int l;
  if (x >= t[hi + 1]) {
    l = hi;
  } else {
    int a = k, b = hi;
    while (a < b) {
      const int m = (a + b + 1) >> 1;
      if (t[m] <= x) a = m; else b = m - 1;
    }
    l = a;
  }
  double h[4] = {1.0, 0.0, 0.0, 0.0};
  double hh[3];
#pragma unroll
  for (int j = 1; j <= k; ++j) {
#pragma unroll
    for (int i = 0; i < j; ++i) hh[i] = h[i];
    h[0] = 0.0;
#pragma unroll
    for (int i = 1; i <= j; ++i) {
      const int li = l + i, lj = li - j;
      const double f = hh[i - 1] / (t[li] - t[lj]);
      h[i - 1] += f * (t[li] - x);
      h[i] = f * (x - t[lj]);
    }
  }
  double v = 0.0;
#pragma unroll
  for (int j = 0; j <= k; ++j) v += sp.c[l - k + j] * h[j];
  return v;
}

// Frequency vector at half-spectrum index (k0,k1,k2) (partner: index -k).
__device__ __forceinline__ void gvec(const GridGeom& g, int k0, int k1, int k2, int partner, double* G) {
  int m0 = fftfreq_i(k0, g.n0), m1 = fftfreq_i(k1, g.n1);
  int m2 = g.full ? fftfreq_i(k2, g.n2) : k2;
  if (partner) {
    m0 = fftfreq_i((g.n0 - k0) % g.n0, g.n0);
    m1 = fftfreq_i((g.n1 - k1) % g.n1, g.n1);
    m2 = fftfreq_i((g.n2 - k2) % g.n2, g.n2);
  }
  G[0] = m0 * g.bg[0] + m1 * g.bg[3] + m2 * g.bg[6];
  G[1] = m0 * g.bg[1] + m1 * g.bg[4] + m2 * g.bg[7];
  G[2] = m0 * g.bg[2] + m1 * g.bg[5] + m2 * g.bg[8];
}

// Half-space mask of ReciprocalGrid.mask (grid.py:658-695, serial offsets).
__device__ bool in_mask(const GridGeom& g, int k0, int k1, int k2) {
  const int D0 = g.n0 / 2, D1 = g.n1 / 2, D2 = g.n2 / 2;
  const bool e0 = (g.n0 % 2) == 0, e1 = (g.n1 % 2) == 0, e2 = (g.n2 % 2) == 0;
  if (g.full && k2 > D2) return false;
  if (k2 == 0) {
    if (k0 == 0 && k1 > D1) return false;
    if (k0 > D0) return false;
  }
  if (e2 && k2 == D2) {
    if (k0 == 0 && k1 == 0) return false;
    if (k0 == 0 && k1 > D1) return false;
    if (k0 > D0) return false;
    if (e1 && k0 == 0 && k1 == D1) return false;
    if (e0 && k0 == D0 && k1 == 0) return false;
    if (e0 && k0 == D0 && k1 > D1) return false;
  }
  if (e0 && k2 == 0 && k0 == D0) {
    if (k1 > D1) return false;
    if (e1 && k1 == D1) return false;
  }
  if (e1 && k2 == 0 && k0 == 0 && k1 == D1) return false;
  if (e0 && e1 && e2 && k0 == D0 && k1 == D1 && k2 == D2) return false;
  return true;
}

// Cardinal B-spline values M_n(x + j), stored at index 1..order of an
// (order_total+1) array (CBspline.calc_Mn, ewald.py:63-82).
__device__ void calc_Mn(double x, int order, double* Mn, int order_total) {
  for (int i = 0; i <= order_total; ++i) Mn[i] = 0.0;
  Mn[1] = x;
  Mn[2] = 1.0 - x;
  for (int i = 3; i <= order; ++i) {
    for (int j = 0; j < i; ++j) {
      const int n = i - j;
      Mn[n] = ((x + n - 1) * Mn[n] + (j + 1 - x) * Mn[n - 1]) / (i - 1);
    }
  }
}

constexpr int kMaxOrder = 12;

// PME Q-array spread of the atoms with species == sel (sel < 0: all atoms,
// weighted by w[i]) -- CBspline.get_PME_Qarray / _calc_PME_Qarray
// (ewald.py:137-194).  One workgroup per atom, one thread per stencil point.
__global__ void k_pme_spread(double* __restrict__ Q, const double* __restrict__ frac, const int* __restrict__ species,
                             int sel, const double* __restrict__ w, int n0, int n1, int n2, int order) {
  __shared__ double Mn[3][kMaxOrder + 1];
  __shared__ int base[3];
  const int i = blockIdx.x;
  const bool take = sel < 0 || species[i] == sel;
  if (!take) return;
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    const int n = d == 0 ? n0 : (d == 1 ? n1 : n2);
    const double up = frac[3 * i + d] * n;
    const double fl = floor(up);
    calc_Mn(up - fl, order, Mn[d], order);
    base[d] = (int)fl;
  }
  __syncthreads();
  const double wi = w ? w[i] : 1.0;
  const int o = order;
  for (int t = threadIdx.x; t < o * o * o; t += blockDim.x) {
    const int a = t / (o * o), b = (t / o) % o, c = t % o;
    int l0 = (1 + base[0] - a) % n0; if (l0 < 0) l0 += n0;
    int l1 = (1 + base[1] - b) % n1; if (l1 < 0) l1 += n1;
    int l2 = (1 + base[2] - c) % n2; if (l2 < 0) l2 += n2;
    const double v = wi * Mn[0][a + 1] * Mn[1][b + 1] * Mn[2][c + 1];
    atomicAdd(&Q[((long long)l0 * n1 + l1) * n2 + l2], v);
  }
}

__device__ __forceinline__ cplx<double> bmod(const double* bm, int n0, int n1, int k0, int k1, int k2) {
  const cplx<double>* b = reinterpret_cast<const cplx<double>*>(bm);
  return cmul(cmul(b[k0], b[n0 + k1]), b[n0 + n1 + k2]);
}

// acc += Qhat * [v(G) B(k) (+ v(G') conj B(-k) in full mode)] * scale
// (_PP_Reciprocal_PME, pseudo/__init__.py:437-452).
__global__ void k_vloc_accum_pme(cplx<double>* __restrict__ acc, const cplx<double>* __restrict__ qhat, GridGeom geo,
                                 const SplineDev* __restrict__ spl, int s, const double* __restrict__ bm,
                                 double scale) {
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  const SplineDev sp = spl[s];
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nh;
       idx += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int k1 = (int)(r % geo.n1), k0 = (int)(r / geo.n1);
    const long long ph = r * geo.nhp + c;
    cplx<double> m = cscale(bmod(bm, geo.n0, geo.n1, k0, k1, c), vline_eval(sp, sqrt(gsq(geo, k0, k1, c))));
    if (geo.full) {
      const int p0 = (geo.n0 - k0) % geo.n0, p1 = (geo.n1 - k1) % geo.n1, p2 = (geo.n2 - c) % geo.n2;
      cplx<double> bp = bmod(bm, geo.n0, geo.n1, p0, p1, p2);
      bp.y = -bp.y;
      const cplx<double> mp = cscale(bp, vline_eval(sp, sqrt(gsq(geo, k0, k1, c, 1))));
      m = cscale(cadd(m, mp), 0.5);
    }
    const cplx<double> z = cmul(qhat[ph], m);
    acc[ph].x += z.x * scale;
    acc[ph].y += z.y * scale;
  }
}

// acc = sum_s v_s(|G|) sum_{i in s} exp(-i G.R_i) * scale, Hermitian
// projection in full mode (_PP_Reciprocal, pseudo/__init__.py:425-435).
__global__ void k_vloc_exact(cplx<double>* __restrict__ acc, GridGeom geo, const SplineDev* __restrict__ spl,
                             const double* __restrict__ pos, const int* __restrict__ species, int nat,
                             double scale) {
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nh;
       idx += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int k1 = (int)(r % geo.n1), k0 = (int)(r / geo.n1);
    cplx<double> tot = {0.0, 0.0};
    const int npart = geo.full ? 2 : 1;
    for (int part = 0; part < npart; ++part) {
      double G[3];
      gvec(geo, k0, k1, c, part, G);
      const double q = sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
      cplx<double> v = {0.0, 0.0};
      int cur = -1;
      double vl = 0.0;
      for (int i = 0; i < nat; ++i) {
        if (species[i] != cur) {
          cur = species[i];
          vl = vline_eval(spl[cur], q);
        }
        double sn, cs;
        sincos(G[0] * pos[3 * i] + G[1] * pos[3 * i + 1] + G[2] * pos[3 * i + 2], &sn, &cs);
        v.x += vl * cs;
        v.y -= vl * sn;
      }
      if (part == 1) v.y = -v.y;  // conj(v(-k))
      tot = cadd(tot, v);
    }
    const double f = scale / npart;
    acc[r * geo.nhp + c] = cplx<double>{tot.x * f, tot.y * f};
  }
}

// Per-species radial table v_s(|G|) on the half spectrum (for forces).
__global__ void k_vline_table(double* __restrict__ out, GridGeom geo, const SplineDev* __restrict__ spl, int s) {
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  const SplineDev sp = spl[s];
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nh;
       idx += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int k1 = (int)(r % geo.n1), k0 = (int)(r / geo.n1);
    out[idx] = vline_eval(sp, sqrt(gsq(geo, k0, k1, c)));
  }
}

// Exact local-PP forces, one workgroup per atom (_Force, pseudo/__init__.py:505-521):
// F_i = (2/Omega) sum_{mask} G v_s(|G|) Im(rhoG exp(i G.R_i)), rhoG = fft(rho).
__global__ void k_force_exact(const cplx<double>* __restrict__ rhog, GridGeom geo, const SplineDev* __restrict__ spl,
                              const double* __restrict__ pos, const int* __restrict__ species, double dV,
                              double* __restrict__ partials) {
  const int i = blockIdx.x;
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  const SplineDev sp = spl[species[i]];
  const double R0 = pos[3 * i], R1 = pos[3 * i + 1], R2 = pos[3 * i + 2];
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long idx = threadIdx.x; idx < Nh; idx += blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int k1 = (int)(r % geo.n1), k0 = (int)(r / geo.n1);
    if (!in_mask(geo, k0, k1, c)) continue;
    double G[3];
    gvec(geo, k0, k1, c, 0, G);
    const double v = vline_eval(sp, sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]));
    double sn, cs;
    sincos(G[0] * R0 + G[1] * R1 + G[2] * R2, &sn, &cs);
    const cplx<double> z = rhog[r * geo.nhp + c];
    const double im = (z.y * cs + z.x * sn) * dV * v;
    acc[0] += G[0] * im;
    acc[1] += G[1] * im;
    acc[2] += G[2] * im;
  }
  block_sums<3>(acc, partials + 3 * i);
}

// spec = rhoG * conj(B) * v_s(|G|), G=0 removed, Hermitian projection in
// full mode (_ForcePME, pseudo/__init__.py:528-545).
__global__ void k_force_pme_mult(cplx<double>* __restrict__ out, const cplx<double>* __restrict__ rhog, GridGeom geo,
                                 const SplineDev* __restrict__ spl, int s, const double* __restrict__ bm) {
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  const SplineDev sp = spl[s];
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nh;
       idx += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int k1 = (int)(r % geo.n1), k0 = (int)(r / geo.n1);
    cplx<double> b = bmod(bm, geo.n0, geo.n1, k0, k1, c);
    b.y = -b.y;
    cplx<double> m = cscale(b, vline_eval(sp, sqrt(gsq(geo, k0, k1, c))));
    if (geo.full) {
      const int p0 = (geo.n0 - k0) % geo.n0, p1 = (geo.n1 - k1) % geo.n1, p2 = (geo.n2 - c) % geo.n2;
      const cplx<double> bp = bmod(bm, geo.n0, geo.n1, p0, p1, p2);
      m = cscale(cadd(m, cscale(bp, vline_eval(sp, sqrt(gsq(geo, k0, k1, c, 1))))), 0.5);
    }
    const long long ph = r * geo.nhp + c;
    out[ph] = idx == 0 ? cplx<double>{0.0, 0.0} : cmul(rhog[ph], m);
  }
}

// PME force gather for the atoms of species sel (every atom if sel < 0), one
// workgroup per atom:
// F_i = -sum_stencil (dQ^T inv(L)) rhoPB[l] (pseudo/__init__.py:546-570).
__global__ void k_force_pme_gather(const double* __restrict__ rhopb, const double* __restrict__ frac,
                                   const int* __restrict__ species, int sel, int n0, int n1, int n2, int order,
                                   const double* __restrict__ linv, double* __restrict__ partials) {
  __shared__ double Mn[3][kMaxOrder + 1];
  __shared__ double M2[3][kMaxOrder + 1];
  __shared__ int base[3];
  const int i = blockIdx.x;
  if (sel >= 0 && species[i] != sel) return;  // sel < 0: every atom
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    const int n = d == 0 ? n0 : (d == 1 ? n1 : n2);
    const double up = frac[3 * i + d] * n;
    const double fl = floor(up);
    calc_Mn(up - fl, order, Mn[d], order);
    calc_Mn(up - fl, order - 1, M2[d], order);
    base[d] = (int)fl;
  }
  __syncthreads();
  const int o = order;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int t = threadIdx.x; t < o * o * o; t += blockDim.x) {
    const int a = t / (o * o), b = (t / o) % o, c = t % o;
    int l0 = (1 + base[0] - a) % n0; if (l0 < 0) l0 += n0;
    int l1 = (1 + base[1] - b) % n1; if (l1 < 0) l1 += n1;
    int l2 = (1 + base[2] - c) % n2; if (l2 < 0) l2 += n2;
    const double dq0 = n0 * (M2[0][a + 1] - M2[0][a]) * Mn[1][b + 1] * Mn[2][c + 1];
    const double dq1 = n1 * Mn[0][a + 1] * (M2[1][b + 1] - M2[1][b]) * Mn[2][c + 1];
    const double dq2 = n2 * Mn[0][a + 1] * Mn[1][b + 1] * (M2[2][c + 1] - M2[2][c]);
    const double val = rhopb[((long long)l0 * n1 + l1) * n2 + l2];
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[d] -= (dq0 * linv[d] + dq1 * linv[3 + d] + dq2 * linv[6 + d]) * val;
  }
  block_sums<3>(acc, partials + 3 * i);
}

// Ewald real-space pair sum over images (ewald.py:280-322); one thread per
// ordered pair (i, j), looping over the image shifts.
__global__ void k_ewald_real(const double* __restrict__ pos, const double* __restrict__ q, int nat,
                             const double* __restrict__ shifts, int nshift, double rmax, double eta_sqrt,
                             double* __restrict__ partials) {
  double acc[1] = {0.0};
  const long long npair = (long long)nat * nat;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < npair;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t / nat), j = (int)(t % nat);
    const double dx0 = pos[3 * j] - pos[3 * i], dy0 = pos[3 * j + 1] - pos[3 * i + 1], dz0 = pos[3 * j + 2] - pos[3 * i + 2];
    double s = 0.0;
    for (int m = 0; m < nshift; ++m) {
      const double dx = dx0 - shifts[3 * m], dy = dy0 - shifts[3 * m + 1], dz = dz0 - shifts[3 * m + 2];
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      if (d < rmax && d > 0.001) s += erfc(eta_sqrt * d) / d;
    }
    acc[0] += q[i] * q[j] * s;
  }
  block_sums<1>(acc, partials + blockIdx.x);
}

// Ewald real-space stress sums (Stress_real, ewald.py:609-662): one thread
// per ordered pair (i, j) over the image shifts,
// q_i q_j [2 sqrt(eta/pi) exp(-eta d^2) + erfc(sqrt(eta) d)/d] d_a d_b / d^2.
__global__ void k_ewald_stress_real(const double* __restrict__ pos, const double* __restrict__ q, int nat,
                                    const double* __restrict__ shifts, int nshift, double rmax, double eta,
                                    double* __restrict__ partials) {
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const double es = sqrt(eta);
  const long long npair = (long long)nat * nat;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < npair;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t / nat), j = (int)(t % nat);
    const double qq = q[i] * q[j];
    for (int m = 0; m < nshift; ++m) {
      const double dx = pos[3 * i] - pos[3 * j] + shifts[3 * m];
      const double dy = pos[3 * i + 1] - pos[3 * j + 1] + shifts[3 * m + 1];
      const double dz = pos[3 * i + 2] - pos[3 * j + 2] + shifts[3 * m + 2];
      const double d2 = dx * dx + dy * dy + dz * dz;
      const double d = sqrt(d2);
      if (!(d < rmax && d > 0.001)) continue;
      const double w = qq * (1.1283791670955125739 * es * exp(-eta * d2) + erfc(es * d) / d) / d2;
      acc[0] += w * dx * dx;
      acc[1] += w * dx * dy;
      acc[2] += w * dx * dz;
      acc[3] += w * dy * dy;
      acc[4] += w * dy * dz;
      acc[5] += w * dz * dz;
    }
  }
  block_sums<6>(acc, partials + 6 * (long long)blockIdx.x);
}

// Ewald real-space forces (ewald.py:541-578), one workgroup per atom i:
// sum over atoms j and image shifts T of q_j (R_i - R_j + T) w(x),
// w = erfc(x)/x^3 + 2/sqrt(pi) exp(-x^2)/x^2, x = sqrt(eta) |R_i - R_j + T|;
// the caller multiplies by q_i eta^(3/2).
__global__ void k_ewald_force_real(const double* __restrict__ pos, const double* __restrict__ q, int nat,
                                   const double* __restrict__ shifts, int nshift, double rmax, double eta_sqrt,
                                   double* __restrict__ partials) {
  const int i = blockIdx.x;
  const double Ri0 = pos[3 * i], Ri1 = pos[3 * i + 1], Ri2 = pos[3 * i + 2];
  double acc[3] = {0.0, 0.0, 0.0};
  const long long n = (long long)nat * nshift;
  for (long long t = threadIdx.x; t < n; t += blockDim.x) {
    const int j = (int)(t % nat), m = (int)(t / nat);
    const double dx = Ri0 - pos[3 * j] + shifts[3 * m];
    const double dy = Ri1 - pos[3 * j + 1] + shifts[3 * m + 1];
    const double dz = Ri2 - pos[3 * j + 2] + shifts[3 * m + 2];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (d < rmax && d > 0.001) {
      const double x = eta_sqrt * d;
      const double w = q[j] * (erfc(x) / (x * x * x) + 1.1283791670955125739 * exp(-x * x) / (x * x));
      acc[0] += dx * w;
      acc[1] += dy * w;
      acc[2] += dz * w;
    }
  }
  block_sums<3>(acc, partials + 3 * i);
}

// Exact ionic structure factor S(G) = sum_j q_j exp(-i G.R_j) on the half
// spectrum (ions.py:121-124 summed with the charges, ewald.py:588-591).
__global__ void k_ewald_strf(cplx<double>* __restrict__ S, GridGeom geo, const double* __restrict__ pos,
                             const double* __restrict__ q, int nat) {
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nh;
       idx += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int k1 = (int)(r % geo.n1), k0 = (int)(r / geo.n1);
    double G[3];
    gvec(geo, k0, k1, c, 0, G);
    cplx<double> z = {0.0, 0.0};
    for (int j = 0; j < nat; ++j) {
      double sn, cs;
      sincos(G[0] * pos[3 * j] + G[1] * pos[3 * j + 1] + G[2] * pos[3 * j + 2], &sn, &cs);
      z.x += q[j] * cs;
      z.y -= q[j] * sn;
    }
    S[r * geo.nhp + c] = z;
  }
}

// Exact reciprocal Ewald forces (ewald.py:580-607), one workgroup per atom:
// sum over the mask of G (cos(G.R_i) Im S + sin(G.R_i) Re S) exp(-G^2/(4 eta))/G^2;
// the caller multiplies by q_i 8 pi / Omega.
__global__ void k_ewald_force_rec(const cplx<double>* __restrict__ S, GridGeom geo, const double* __restrict__ pos,
                                  double eta, double* __restrict__ partials) {
  const int i = blockIdx.x;
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  const double R0 = pos[3 * i], R1 = pos[3 * i + 1], R2 = pos[3 * i + 2];
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long idx = threadIdx.x; idx < Nh; idx += blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int k1 = (int)(r % geo.n1), k0 = (int)(r / geo.n1);
    if (!in_mask(geo, k0, k1, c)) continue;
    double G[3];
    gvec(geo, k0, k1, c, 0, G);
    const double g2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
    if (!(g2 > 0.0)) continue;
    double sn, cs;
    sincos(G[0] * R0 + G[1] * R1 + G[2] * R2, &sn, &cs);
    const cplx<double> z = S[r * geo.nhp + c];
    const double v = (cs * z.y + sn * z.x) * exp(-g2 / (4.0 * eta)) / g2;
    acc[0] += G[0] * v;
    acc[1] += G[1] * v;
    acc[2] += G[2] * v;
  }
  block_sums<3>(acc, partials + 3 * i);
}

// PME Ewald force multiplier (ewald.py:810-829): spec *= |B|^2 exp(-G^2/(4 eta))/G^2, G = 0 -> 0.
__global__ void k_ewald_pme_mult(cplx<double>* __restrict__ spec, GridGeom geo, const double* __restrict__ bm,
                                 double eta) {
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nh;
       idx += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int k1 = (int)(r % geo.n1), k0 = (int)(r / geo.n1);
    const double g2 = gsq(geo, k0, k1, c);
    const cplx<double> b = bmod(bm, geo.n0, geo.n1, k0, k1, c);
    const double m = g2 > 0.0 ? (b.x * b.x + b.y * b.y) * exp(-g2 / (4.0 * eta)) / g2 : 0.0;
    const long long ph = r * geo.nhp + c;
    spec[ph] = cscale(spec[ph], m);
  }
}

// Ewald reciprocal sum over the half-space mask: sum |S(G)|^2 exp(-G^2/(4 eta))/G^2
// with S = rfftn(Q) * B (PME, ewald.py:785-807) or the exact S (ewald.py:448-465).
__global__ void k_ewald_rec(const cplx<double>* __restrict__ qhat, const double* __restrict__ bm, GridGeom geo,
                            const double* __restrict__ pos, const double* __restrict__ q, int nat, double eta,
                            double* __restrict__ partials) {
  double acc[1] = {0.0};
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nh;
       idx += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int k1 = (int)(r % geo.n1), k0 = (int)(r / geo.n1);
    if (!in_mask(geo, k0, k1, c)) continue;
    double G[3];
    gvec(geo, k0, k1, c, 0, G);
    const double g2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
    if (!(g2 > 0.0)) continue;
    cplx<double> S;
    if (qhat) {
      S = cmul(qhat[r * geo.nhp + c], bmod(bm, geo.n0, geo.n1, k0, k1, c));
    } else {
      S = {0.0, 0.0};
      for (int i = 0; i < nat; ++i) {
        double sn, cs;
        sincos(G[0] * pos[3 * i] + G[1] * pos[3 * i + 1] + G[2] * pos[3 * i + 2], &sn, &cs);
        S.x += q[i] * cs;
        S.y -= q[i] * sn;
      }
    }
    acc[0] += (S.x * S.x + S.y * S.y) * exp(-g2 / (4.0 * eta)) / g2;
  }
  block_sums<1>(acc, partials + blockIdx.x);
}

// ---------------------------------------------------------------------------
// stress: reciprocal-space sums over the half-space mask and the vW
// gradient products (total_functional.py:176-209 and the per-term routines)
// ---------------------------------------------------------------------------
// d/dq of the cubic spline (splev der=1: FITPACK's derivative of the
// B-spline sum), first polynomial piece below the first knot, zero from
// qmax on (_PP_Derivative_One, pseudo/__init__.py:454-460).
__device__ double vline_deriv(const SplineDev& sp, double x) {
  if (!(x < sp.qmax)) return 0.0;
  const int k = 3;
  const double* t = sp.t;
  const int hi = sp.nt - k - 2;
  int l;
  if (x >= t[hi + 1]) {
    l = hi;
  } else {
    int a = k, b = hi;
    while (a < b) {
      const int m = (a + b + 1) >> 1;
      if (t[m] <= x) a = m; else b = m - 1;
    }
    l = a;
  }
  // degree-2 B-splines B_{l-2..l} at x
  double h[3] = {1.0, 0.0, 0.0};
  double hh[2];
#pragma unroll
  for (int j = 1; j <= k - 1; ++j) {
#pragma unroll
    for (int i = 0; i < j; ++i) hh[i] = h[i];
    h[0] = 0.0;
#pragma unroll
    for (int i = 1; i <= j; ++i) {
      const int li = l + i, lj = li - j;
      const double f = hh[i - 1] / (t[li] - t[lj]);
      h[i - 1] += f * (t[li] - x);
      h[i] = f * (x - t[lj]);
    }
  }
  // s'(x) = sum_j k (c_j - c_(j-1)) / (t_(j+k) - t_j) B_(j,k-1)(x)
  double d = 0.0;
#pragma unroll
  for (int i = 0; i < k; ++i) {
    const int j = l - 2 + i;
    d += k * (sp.c[j] - sp.c[j - 1]) / (t[j + k] - t[j]) * h[i];
  }
  return d;
}

// eta dF/deta of the inverse Lindhard function with the reference's three
// branches (LindhardDerivative, kedf/kernel.py:197-211).
__device__ double lindhard_deriv_eta(double eta, double mu) {
  const double atol = 1.0e-10;
  double d = 0.0;
  if (eta > atol && fabs(eta - 1.0) > atol) {
    const double ta = log(fabs((1.0 + eta) / (1.0 - eta)));
    const double den = 0.5 + 0.25 * (1.0 - eta * eta) / eta * ta;
    d = (0.5 / eta - 0.25 * (eta * eta + 1.0) / (eta * eta) * ta) / (den * den) + 6.0 * eta * mu;
  }
  if (eta < atol) d = -2.0 * eta * (1.0 / 3.0 - 3.0 * mu);
  if (fabs(eta - 1.0) < atol) d = -48.0;
  return d * eta;
}

enum StressMode { kStHartree = 0, kStWT = 1, kStPPExact = 2, kStPPPme = 3, kStEwald = 4 };

struct StressArgs {
  const cplx<double>* a;  // Hartree, PP: rfftn(rho); WT: rfftn(rho^alpha); Ewald PME: rfftn(Q)
  const cplx<double>* b;  // WT: rfftn(rho^beta); PP PME: rfftn(Q_s)
  const double* bm;       // PME b-moduli
  const SplineDev* spl;
  int nspecies, s;        // PP: tables; PP PME: the species of b
  const double* pos;
  const int* species;
  const double* q;        // Ewald charges
  int nat;
  double tkf, mu, eta;
};

// acc[k] = sum over the mask of the six pair weights (ab = xx xy xz yy yz
// zz) of one term:
//   Hartree   G_a G_b |a|^2 / G^4                      (hartree.py:68-90)
//   WT        (G_a G_b / G^2 - d_ab/3) eta F'(eta) Re(a conj b), eta = |G|/tkf
//                                                       (wt.py:75-106)
//   PP exact  G_a G_b Re(a sum_i v'_s(i)(|G|) e^(i G.R_i)) / |G|
//                                                       (pseudo/__init__.py:479-503)
//   PP PME    G_a G_b Re(conj(a) B v'_s(|G|) b) / |G|   (pseudo/__init__.py:573-602)
//   Ewald     |S|^2 e^(-G^2/(4 eta)) / G^2 (2 G_a G_b / G^2 (1 + G^2/(4 eta)) - d_ab),
//             S = a B (PME) or sum_i q_i e^(-i G.R_i)  (ewald.py:719-759, :915-958)
__global__ void k_stress_recip(StressArgs A, GridGeom geo, int mode, double* __restrict__ partials) {
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nh;
       idx += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int k1 = (int)(r % geo.n1), k0 = (int)(r / geo.n1);
    if (!in_mask(geo, k0, k1, c)) continue;
    double G[3];
    gvec(geo, k0, k1, c, 0, G);
    const double g2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
    const long long ph = r * geo.nhp + c;
    double w = 0.0, wd = 0.0, wg = 1.0;  // pair weight: wg * G_a G_b * w - d_ab * wd
    if (mode == kStHartree) {
      if (!(g2 > 0.0)) continue;
      const cplx<double> z = A.a[ph];
      w = (z.x * z.x + z.y * z.y) / (g2 * g2);
    } else if (mode == kStWT) {
      const cplx<double> x = A.a[ph], y = A.b[ph];
      const double ld = lindhard_deriv_eta(sqrt(g2) / A.tkf, A.mu) * (x.x * y.x + x.y * y.y);
      w = g2 > 0.0 ? ld / g2 : 0.0;
      wd = ld / 3.0;
    } else if (mode == kStPPExact || mode == kStPPPme) {
      if (!(g2 > 0.0)) continue;
      const double qv = sqrt(g2);
      const cplx<double> z = A.a[ph];
      if (mode == kStPPExact) {
        double vr = 0.0, vi = 0.0;
        for (int sp = 0; sp < A.nspecies; ++sp) {
          const double d = vline_deriv(A.spl[sp], qv);
          if (d == 0.0) continue;
          for (int i = 0; i < A.nat; ++i) {
            if (A.species[i] != sp) continue;
            double sn, cs;
            sincos(G[0] * A.pos[3 * i] + G[1] * A.pos[3 * i + 1] + G[2] * A.pos[3 * i + 2], &sn, &cs);
            vr += d * cs;
            vi += d * sn;
          }
        }
        w = (z.x * vr - z.y * vi) / qv;
      } else {
        const cplx<double> u = cmul(cmul(cplx<double>{z.x, -z.y}, bmod(A.bm, geo.n0, geo.n1, k0, k1, c)), A.b[ph]);
        w = u.x * vline_deriv(A.spl[A.s], qv) / qv;
      }
    } else {
      if (!(g2 > 0.0)) continue;
      cplx<double> S;
      if (A.a) {
        S = cmul(A.a[ph], bmod(A.bm, geo.n0, geo.n1, k0, k1, c));
      } else {
        S = {0.0, 0.0};
        for (int i = 0; i < A.nat; ++i) {
          double sn, cs;
          sincos(G[0] * A.pos[3 * i] + G[1] * A.pos[3 * i + 1] + G[2] * A.pos[3 * i + 2], &sn, &cs);
          S.x += A.q[i] * cs;
          S.y -= A.q[i] * sn;
        }
      }
      const double e = (S.x * S.x + S.y * S.y) * exp(-g2 / (4.0 * A.eta)) / g2;
      w = e;
      wg = 2.0 / g2 * (1.0 + g2 / (4.0 * A.eta));
      wd = e;
    }
    acc[0] += G[0] * G[0] * wg * w - wd;
    acc[1] += G[0] * G[1] * wg * w;
    acc[2] += G[0] * G[2] * wg * w;
    acc[3] += G[1] * G[1] * wg * w - wd;
    acc[4] += G[1] * G[2] * wg * w;
    acc[5] += G[2] * G[2] * wg * w - wd;
  }
  block_sums<6>(acc, partials + 6 * (long long)blockIdx.x);
}

// Gradient spectra for the vW stress: slot a = i G_a X / N, a = 0..2, from
// X in slot 0 (read before it is overwritten).  With full=1 frequencies
// G_a is replaced by its Hermitian part (G(k) - G(-k))/2, so that the
// half-spectrum inverse equals Re(ifftn) of the full array (vw.py:97-112
// with ReciprocalField.ifft(force_real=True)).
template <typename T>
__global__ void k_grad_spec(cplx<T>* __restrict__ spec, long long Nhp, GridGeom geo, double scale) {
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nh;
       idx += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int k1 = (int)(r % geo.n1), k0 = (int)(r / geo.n1);
    const long long ph = r * geo.nhp + c;
    const cplx<T> z = spec[ph];
    double G[3];
    gvec(geo, k0, k1, c, 0, G);
    if (geo.full) {
      double Gp[3];
      gvec(geo, k0, k1, c, 1, Gp);
      for (int a = 0; a < 3; ++a) G[a] = 0.5 * (G[a] - Gp[a]);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
      spec[a * Nhp + ph] = cplx<T>{T(-G[a] * (double)z.y * scale), T(G[a] * (double)z.x * scale)};
  }
}

// Divergence spectrum for the GGA potential: slot 0 = i (G . W) / N from the
// three spectra W_a in slots 0..2 (all read before slot 0 is written); the
// same Hermitian part of G as k_grad_spec with full=1 frequencies
// (kedf/gga.py:1284-1290).
template <typename T>
__global__ void k_div_spec(cplx<T>* __restrict__ spec, long long Nhp, GridGeom geo, double scale) {
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nh;
       idx += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int k1 = (int)(r % geo.n1), k0 = (int)(r / geo.n1);
    const long long ph = r * geo.nhp + c;
    double G[3];
    gvec(geo, k0, k1, c, 0, G);
    if (geo.full) {
      double Gp[3];
      gvec(geo, k0, k1, c, 1, Gp);
      for (int a = 0; a < 3; ++a) G[a] = 0.5 * (G[a] - Gp[a]);
    }
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const cplx<T> w = spec[a * Nhp + ph];
      re += G[a] * (double)w.x;
      im += G[a] * (double)w.y;
    }
    spec[ph] = cplx<T>{T(-im * scale), T(re * scale)};
  }
}

// sum V_xc rho of the LDA potential (the P of _LDAStress, xc/semilocal_xc.py:343-346)
__global__ void k_lda_vrho(const double* __restrict__ rho, long long N, double* __restrict__ partials) {
  double acc[1] = {0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
    const double r = rho[i];
    double ex, vxc;
    lda_point(r, cbrt(r), ex, vxc);
    acc[0] += vxc * r;
  }
  block_sums<1>(acc, partials + blockIdx.x);
}

// Python's float modulo x % 1.0, applied twice as ase's wrap does.
double pymod1(double x) {
  double r = std::fmod(x, 1.0);
  if (r != 0.0 && r < 0.0) r += 1.0;
  return r;
}

// b-moduli of the cardinal B-spline (CBspline._calc_bm, ewald.py:84-101).
std::vector<double> make_bm(const ofdft_plan_s* P, int order) {
  std::vector<double> M(order + 1, 0.0);
  M[1] = 1.0;
  M[2] = 0.0;
  for (int i = 3; i <= order; ++i)
    for (int j = 0; j < i; ++j) {
      const int n = i - j;
      M[n] = ((1.0 + n - 1) * M[n] + (j + 1 - 1.0) * M[n - 1]) / (i - 1);
    }
  const int nb2 = P->full ? P->n[2] : P->nh;
  std::vector<double> out;
  for (int ax = 0; ax < 3; ++ax) {
    const int cnt = ax < 2 ? P->n[ax] : nb2;
    for (int k = 0; k < cnt; ++k) {
      const double q = 2.0 * kPi * k / P->n[ax];
      double tr = std::cos(-(order - 1.0) * q), ti = std::sin(-(order - 1.0) * q);
      double fr = 0.0, fi = 0.0;
      for (int j = 1; j < order; ++j) {
        fr += M[j] * std::cos(-j * q);
        fi += M[j] * std::sin(-j * q);
      }
      const double den = fr * fr + fi * fi;
      out.push_back((tr * fr + ti * fi) / den);
      out.push_back((ti * fr - tr * fi) / den);
    }
  }
  return out;
}

// forward r2c of a real FP64 field into spectrum slot `slot` of the workspace
template <class Pre>
int r2c_pre_slot(ofdft_plan_s* P, const Pre& pre, int slot) {
  int rc;
  auto* spec = reinterpret_cast<cplx<double>*>(P->spec) + (long long)slot * P->Nhp;
  // rows kernel writes slot 0; move it if another slot is requested
  if ((rc = launch_rows_fwd<double>(P, pre, 1))) return rc;
  auto* s0 = reinterpret_cast<cplx<double>*>(P->spec);
  if ((rc = launch_axis<double>(P, s0, 1, 1, false))) return rc;
  if ((rc = launch_axis<double>(P, s0, 0, 1, false))) return rc;
  if (slot != 0) HIPCHK(hipMemcpyAsync(spec, s0, sizeof(cplx<double>) * P->Nhp, hipMemcpyDeviceToDevice, P->stream));
  return 0;
}

int r2c_slot(ofdft_plan_s* P, const double* in, int slot) { return r2c_pre_slot(P, PreCopy<double>{in}, slot); }

// The GGA kinetic functional at rho (or (c phi + s p)^2): forward transform
// of rho, three gradient fields, first pass (energy, stress sums, pot1 and
// the vector field), then -- with V -- the divergence of the vector field
// and V += pot1 - div (GGA, kedf/gga.py:1212-1311).
template <typename T>
int gga_impl(ofdft_plan_s* P, const GgaSpec& g, const T* rho, const T* phi, const T* p, double c, double s, T* V,
             double* out) {
  const long long N = P->N;
  if (V && !P->d_gga) HIPCHK(hipMalloc(&P->d_gga, 4 * (size_t)N * sizeof(T)));
  auto* spec = reinterpret_cast<cplx<T>*>(P->spec);
  PreRho<T> pr{rho, phi, p, T(c), T(s)};
  int rc;
  if ((rc = launch_rows_fwd<T>(P, pr, 1))) return rc;
  if ((rc = launch_axis<T>(P, spec, 1, 1, false))) return rc;
  if ((rc = launch_axis<T>(P, spec, 0, 1, false))) return rc;
  hipLaunchKernelGGL((k_grad_spec<T>), dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream, spec, P->Nhp, geom(P),
                     1.0 / (double)N);
  LAUNCHCHK();
  if ((rc = launch_axis<T>(P, spec, 0, 3, true))) return rc;
  if ((rc = launch_axis<T>(P, spec, 1, 3, true))) return rc;
  T* W = V ? reinterpret_cast<T*>(P->d_gga) : nullptr;
  PostGga1<T> post1{pr, g, W, N};
  if ((rc = launch_rows_inv<T>(P, post1, 3, spec, true))) return rc;
  double s1[8];
  if ((rc = fetch_sums(P, 8, s1))) return rc;
  const double dV = P->dV;
  for (int k = 0; k < OFDFT_GGA_NOUT; ++k) out[k] = 0.0;
  out[0] = s1[0] * dV;
  out[3] = s1[1] * dV;
  for (int k = 0; k < 6; ++k) out[4 + k] = s1[2 + k] * dV;
  if (!V) return 0;
  PreCopy3<T> p3{W + N, N};
  if ((rc = launch_rows_fwd<T>(P, p3, 3))) return rc;
  if ((rc = launch_axis<T>(P, spec, 1, 3, false))) return rc;
  if ((rc = launch_axis<T>(P, spec, 0, 3, false))) return rc;
  hipLaunchKernelGGL((k_div_spec<T>), dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream, spec, P->Nhp, geom(P),
                     1.0 / (double)N);
  LAUNCHCHK();
  if ((rc = launch_axis<T>(P, spec, 0, 1, true))) return rc;
  if ((rc = launch_axis<T>(P, spec, 1, 1, true))) return rc;
  PostGga2<T> post2{pr, W, V};
  if ((rc = launch_rows_inv<T>(P, post2, 1, spec, true))) return rc;
  double s2[2];
  if ((rc = fetch_sums(P, 2, s2))) return rc;
  out[2] = s2[0] * dV;
  out[1] = s2[1] * dV;
  return 0;
}

// ===========================================================================
// LWT / LMGP: nonlocal kinetic term whose kernel K(|G| / kf(r)) follows the
// local kf = 2 (3 pi^2 rho)^(1/3), interpolated between kernels at the kf
// nodes (LWTPotentialEnergy, kedf/lwt.py:83-341).  The reference loops over
// intervals with two nodes each; here every node j is one pass: its two
// convolutions p_j = IFFT[K_j B] and m_j = IFFT[K'_j B] (B = FFT(rho^beta))
// go straight into the interpolation sums pot1, pot3 of the points whose
// interval has j as an end, and its share K_j FFT(rho^alpha w_j) of the pot2
// spectrum is accumulated in reciprocal space, so that pot2 needs one
// inverse transform in all.
// ===========================================================================
struct LwtTabDev {
  const double* t;  // knots (kind 0)
  const double* c;  // coefficients (kind 0) or values (kind 1)
  int n, kind;
  double x0, dx, etamax, limit;
};

// Cubic B-spline value (FITPACK splev; the last interval beyond the end):
// the knot interval comes from the uniform data grid the spline was fitted
// on, corrected by a step or two at the not-a-knot ends.
__device__ double lwt_spline(const LwtTabDev& s, double x) {
  const int k = 3;
  const int hi = s.n - k - 2;
  const double* t = s.t;
  int l = (int)floor((x - s.x0) / s.dx) + 2;
  l = l < k ? k : (l > hi ? hi : l);
  while (l > k && t[l] > x) --l;
  while (l < hi && t[l + 1] <= x) ++l;
  double h[4] = {1.0, 0.0, 0.0, 0.0};
  double hh[3];
#pragma unroll
  for (int j = 1; j <= k; ++j) {
#pragma unroll
    for (int i = 0; i < j; ++i) hh[i] = h[i];
    h[0] = 0.0;
#pragma unroll
    for (int i = 1; i <= j; ++i) {
      const int li = l + i, lj = li - j;
      const double f = hh[i - 1] / (t[li] - t[lj]);
      h[i - 1] += f * (t[li] - x);
      h[i] = f * (x - t[lj]);
    }
  }
  double v = 0.0;
#pragma unroll
  for (int j = 0; j <= k; ++j) v += s.c[l - k + j] * h[j];
  return v;
}

// LWTKernelKf (kedf/kernel.py:379-405) at one eta: the table value below
// etamax, `limit` from there on; the nearest entry for a plain table.
__device__ __forceinline__ double lwt_kval(const LwtTabDev& s, double eta) {
  if (!(eta < s.etamax)) return s.limit;
  if (s.kind == 1) return s.c[(int)rint(eta / s.dx)];
  return lwt_spline(s, eta);
}

// MGPOmegaE (kernel.py:408-440, revke): 4 pi erf(c G^2)^2 a exp(-b G^2) / G^2
__device__ __forceinline__ double lwt_mgpe(double gg, double a, double b, double c) {
  const double e = erf(c * gg);
  return 4 * kPi * e * e * a * exp(-gg * b) / gg;
}

// local kf with the clamp at the upper bound (lwt.py:99, :138-139)
__device__ __forceinline__ double lwt_kf(double rho, double kfmax) {
  const double kf = 2.0 * cbrt(3.0 * (kPi * kPi) * rho);
  return kf > kfmax ? kfmax : kf;
}

// node j and its neighbours in the kf list
struct LwtNode {
  double km, k0, kp;  // kf_{j-1}, kf_j, kf_{j+1}
  int hasm, hasp, j;
};

// Weights of node j at one point: pot1 += ap p_j + am m_j, pot3 += bp p_j +
// bm m_j, and the densities rho^alpha wk, rho^alpha wd whose transforms are
// multiplied by K_j, K'_j for pot2 (lwt.py:213-315: the interval below the
// node makes it the upper end "1", the interval above the lower end "0";
// below the first node pot1 = p_0 (kf - kf_0)/kf_0, lwt.py:236-249).
struct LwtW {
  double ap, am, bp, bm, wk, wd;
};
__device__ __forceinline__ LwtW lwt_node_w(const LwtNode& nd, double kf, int herm) {
  LwtW w = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (nd.j == 0 && kf < nd.k0 + 1e-18) {
    const double t = (kf - nd.k0) / nd.k0;
    w.ap = t;
    w.wk = t;
    return w;
  }
  if (nd.hasm && kf > nd.km && kf < nd.k0 + 1e-18) {
    const double Dkf = nd.k0 - nd.km;
    const double t = (kf - nd.km) / Dkf;
    if (!herm) {
      w.ap = t;
      w.bm = t;
      w.wk = t;
    } else {
      const double t2 = t * t, t3 = t2 * t;
      const double h00 = 2.0 * t3 - 3.0 * t2 + 1.0;
      const double h01 = 1.0 - h00, h11 = t3 - t2;
      const double h00D = 6.0 * t2 - 6 * t, h11D = 3.0 * t2 - 2 * t;
      w.ap = h01;
      w.am = Dkf * h11;
      w.bp = -h00D / Dkf;
      w.bm = h11D;
      w.wk = h01;
      w.wd = Dkf * h11;
    }
  } else if (nd.hasp && kf > nd.k0 && kf < nd.kp + 1e-18) {
    const double Dkf = nd.kp - nd.k0;
    const double t = (kf - nd.k0) / Dkf;
    if (!herm) {
      const double t0 = 1 - t;
      w.ap = t0;
      w.bm = t0;
      w.wk = t0;
    } else {
      const double t2 = t * t, t3 = t2 * t;
      const double h00 = 2.0 * t3 - 3.0 * t2 + 1.0, h10 = t3 - 2.0 * t2 + t;
      const double h00D = 6.0 * t2 - 6 * t, h10D = 3.0 * t2 - 4 * t + 1.0;
      w.ap = h00;
      w.am = Dkf * h10;
      w.bp = h00D / Dkf;
      w.bm = h10D;
      w.wk = h00;
      w.wd = Dkf * h10;
    }
  }
  return w;
}

// rho^e with rho floored at 1e-30 where it is not positive (lwt.py:188-191)
template <typename T>
struct PreLwtPow {
  static constexpr int kProf = PC_ROWS_FWD;
  PreRho<T> r;
  double e;
  __host__ __device__ int nin() const { return r.nin(); }
  using In = typename PreRho<T>::In;
  __device__ __forceinline__ In load(long long i) const { return r.load(i); }
  __device__ __forceinline__ void apply(const In& x, T* out) const {
    const double rho = r.dens(x);
    out[0] = T(pow(rho > 0.0 ? rho : 1e-30, e));
  }
  __device__ __forceinline__ void apply(const In& x, T* out, long long) const { apply(x, out); }
  __device__ __forceinline__ void operator()(long long i, T* out) const { apply(load(i), out); }
};

// node j's weighted densities rho^alpha wk (and rho^alpha wd, Hermite)
template <typename T>
struct PreLwtW {
  static constexpr int kProf = PC_ROWS_FWD;
  PreRho<T> r;
  LwtNode nd;
  double kfmax, alpha;
  int herm;
  __host__ __device__ int nin() const { return r.nin(); }
  using In = typename PreRho<T>::In;
  __device__ __forceinline__ In load(long long i) const { return r.load(i); }
  __device__ __forceinline__ void apply(const In& x, T* out) const {
    const double rho = r.dens(x);
    const LwtW w = lwt_node_w(nd, lwt_kf(rho, kfmax), herm);
    const double ra = pow(rho > 0.0 ? rho : 1e-30, alpha);
    out[0] = T(ra * w.wk);
    if (herm) out[1] = T(ra * w.wd);
  }
  __device__ __forceinline__ void apply(const In& x, T* out, long long) const { apply(x, out); }
  __device__ __forceinline__ void operator()(long long i, T* out) const { apply(load(i), out); }
};

// node j's convolutions c[0] = p_j (c[1] = m_j) into pot1 (A1) and pot3
// (A3); the first node writes every point, the others only their points.
template <typename T>
struct PostLwtNode {
  static constexpr int NS = 1;
  static constexpr int kProf = PC_ROWS_INV;
  PreRho<T> r;
  LwtNode nd;
  double kfmax;
  int herm, kdd3, first;
  T* __restrict__ A1;
  T* __restrict__ A3;
  __host__ __device__ int nio() const { return r.nin() + (first ? 1 + kdd3 : 0); }
  using In = typename PreRho<T>::In;
  __device__ __forceinline__ In load(long long i) const { return r.load(i); }
  __device__ __forceinline__ void apply(long long i, const In& in, const T* c, double*) const {
    const LwtW w = lwt_node_w(nd, lwt_kf(r.dens(in), kfmax), herm);
    const double pv = c[0], mv = c[1];
    if (first) {
      A1[i] = T(w.ap * pv + w.am * mv);
      if (kdd3) A3[i] = T(w.bp * pv + w.bm * mv);
    } else if (w.ap != 0.0 || w.am != 0.0 || w.bp != 0.0 || w.bm != 0.0) {
      A1[i] = T((double)A1[i] + (w.ap * pv + w.am * mv));
      if (kdd3) A3[i] = T((double)A3[i] + (w.bp * pv + w.bm * mv));
    }
  }
  __device__ __forceinline__ void operator()(long long i, const T* c, double* acc) const { apply(i, load(i), c, acc); }
};

// The last pass (lwt.py:317-341): pot1 (+ mgpe), pot2 = c[0] (kdd > 1;
// pot1 otherwise), pot3, the low-density weight, E = sum rho^alpha pot1 and
// V += alpha rho^(alpha-1) pot1 + beta rho^(beta-1) pot2 + kf/3 rho^(alpha-1)
// pot3.  acc: [E, 2 sum V_LWT phi(t) dphi/dt, sum V_LWT rho].
template <typename T>
struct PostLwtFinal {
  static constexpr int NS = 3;
  static constexpr int kProf = PC_ROWS_INV;
  PreRho<T> r;
  double kfmax, alpha, beta, ld, rhomax_ld;
  int same, has2, mg;
  const T* __restrict__ A1;
  const T* __restrict__ A3;  // null: pot3 = 0
  T* __restrict__ V;
  __host__ __device__ int nio() const { return r.nin() + 1 + (A3 ? 1 : 0) + (V ? 2 : 0); }
  struct In {
    typename PreRho<T>::In x;
    T a1, a3, v;
  };
  __device__ __forceinline__ In load(long long i) const {
    return {r.load(i), A1[i], A3 ? A3[i] : T(0), V ? V[i] : T(0)};
  }
  __device__ __forceinline__ void apply(long long i, const In& in, const T* c, double* acc) const {
    const double rho = r.dens(in.x);
    const double kf = lwt_kf(rho, kfmax);
    const double rf = rho > 0.0 ? rho : 1e-30;
    const double ra = pow(rf, alpha), ra1 = ra / rf;
    const double rb1 = same ? ra1 : pow(rf, beta) / rf;
    double pot1 = in.a1;
    if (mg) pot1 = pot1 + (double)(has2 ? c[1] : c[0]);
    double pot2 = has2 ? (double)c[0] : pot1;
    double pot3 = in.a3;
    double f = 1.0;
    if (rho < 1e-6) f = pow(fabs(rho), ld) / rhomax_ld;
    if (rho < 0.0) f = 0.0;
    pot1 *= f;
    acc[0] += ra * pot1;
    if (V) {
      pot2 *= f;
      pot3 *= f;
      const double v = pot1 * (alpha * ra1) + (pot2 * (beta * rb1) + pot3 * ((kf / 3.0) * ra1));
      V[i] = T((double)in.v + v);
      acc[2] += v * rho;
      if (!r.rho && r.p) {
        const double ph = (double)r.c * in.x.a + (double)r.s * in.x.b;
        const double dph = (double)r.c * in.x.b - (double)r.s * in.x.a;
        acc[1] += 2.0 * v * ph * dph;
      }
    }
  }
  __device__ __forceinline__ void operator()(long long i, const T* c, double* acc) const { apply(i, load(i), c, acc); }
};

template <typename T>
__global__ void k_lwt_final_pw(PostLwtFinal<T> post, long long N, double* __restrict__ partials) {
  double acc[3] = {0.0, 0.0, 0.0};
  const T z[3] = {T(0), T(0), T(0)};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
    post.apply(i, post.load(i), z, acc);
  block_sums<3>(acc, partials + 3 * (long long)blockIdx.x);
}

// per-block min and max of the density into partials[2b], partials[2b+1]
template <typename T>
__global__ void k_lwt_range(PreRho<T> r, long long N, double* __restrict__ partials) {
  double lo = HUGE_VAL, hi = -HUGE_VAL;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
    const double v = r.dens(r.load(i));
    lo = fmin(lo, v);
    hi = fmax(hi, v);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, off, 64));
    hi = fmax(hi, __shfl_xor(hi, off, 64));
  }
  __shared__ double red[16][2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[wave][0] = lo;
    red[wave][1] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 1; w < nw; ++w) {
      lo = fmin(lo, red[w][0]);
      hi = fmax(hi, red[w][1]);
    }
    partials[2 * (long long)blockIdx.x] = lo;
    partials[2 * (long long)blockIdx.x + 1] = hi;
  }
}

// Node j in reciprocal space: Acc (+)= (K_j W_K + K'_j W_D) / N from the
// transforms of the node's weighted densities in slots 0 (and 1), then
// slot 0 = K_j B / N (and slot 1 = K'_j B / N) for the inverse transforms.
// Kernels are 0 at G = 0; with full=1 frequencies the Hermitian part
// (K(G) + K(-G))/2 gives Re(ifftn) of the full array.
template <typename T>
__global__ void k_lwt_node(cplx<T>* __restrict__ spec, long long Nhp, const cplx<T>* __restrict__ B,
                           cplx<T>* __restrict__ acc, GridGeom geo, LwtTabDev K, LwtTabDev D, double kf, int nw,
                           int mcalc, int first, double scale) {
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  const bool need_d = mcalc || nw == 2;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nh;
       idx += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int k1 = (int)(r % geo.n1), k0 = (int)(r / geo.n1);
    const long long ph = r * geo.nhp + c;
    double kv = 0.0, dv = 0.0;
    if (idx != 0) {
      const double eta = sqrt(gsq(geo, k0, k1, c)) / kf;
      kv = lwt_kval(K, eta);
      if (need_d) dv = lwt_kval(D, eta) / kf;
      if (geo.full) {
        const double etp = sqrt(gsq(geo, k0, k1, c, 1)) / kf;
        kv = 0.5 * (kv + lwt_kval(K, etp));
        if (need_d) dv = 0.5 * (dv + lwt_kval(D, etp) / kf);
      }
    }
    if (nw) {
      double ax = 0.0, ay = 0.0;
      if (!first) {
        const cplx<T> a = acc[ph];
        ax = a.x;
        ay = a.y;
      }
      const cplx<T> w0 = spec[ph];
      ax += kv * (double)w0.x * scale;
      ay += kv * (double)w0.y * scale;
      if (nw == 2) {
        const cplx<T> w1 = spec[Nhp + ph];
        ax += dv * (double)w1.x * scale;
        ay += dv * (double)w1.y * scale;
      }
      acc[ph] = cplx<T>{T(ax), T(ay)};
    }
    const cplx<T> b = B[ph];
    spec[ph] = cplx<T>{T(kv * (double)b.x * scale), T(kv * (double)b.y * scale)};
    if (mcalc) spec[Nhp + ph] = cplx<T>{T(dv * (double)b.x * scale), T(dv * (double)b.y * scale)};
  }
}

// The last spectra: slot 0 = Acc (+ M B / N) for pot2 (has2), then the MGP
// kinetic-electron convolution M B / N (mg) in the next slot.
template <typename T>
__global__ void k_lwt_final_spec(cplx<T>* __restrict__ spec, long long Nhp, const cplx<T>* __restrict__ B,
                                 const cplx<T>* __restrict__ acc, GridGeom geo, int has2, int mg, double ma,
                                 double mb, double mc, double scale) {
  const long long Nh = (long long)geo.n0 * geo.n1 * geo.nh;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nh;
       idx += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(idx % geo.nh);
    const long long r = idx / geo.nh;
    const int k1 = (int)(r % geo.n1), k0 = (int)(r / geo.n1);
    const long long ph = r * geo.nhp + c;
    double mx = 0.0, my = 0.0;
    if (mg) {
      double m = 0.0;
      if (idx != 0) {
        m = lwt_mgpe(gsq(geo, k0, k1, c), ma, mb, mc);
        if (geo.full) m = 0.5 * (m + lwt_mgpe(gsq(geo, k0, k1, c, 1), ma, mb, mc));
      }
      const cplx<T> b = B[ph];
      mx = m * (double)b.x * scale;
      my = m * (double)b.y * scale;
    }
    if (has2) {
      const cplx<T> a = acc[ph];
      spec[ph] = cplx<T>{T((double)a.x + mx), T((double)a.y + my)};
    }
    if (mg) spec[(has2 ? Nhp : 0) + ph] = cplx<T>{T(mx), T(my)};
  }
}

LwtTabDev lwt_tab_dev(const ofdft_plan_s* P, int i) {
  const ofdft_lwt_table& t = P->lwt_tab[i];
  return LwtTabDev{t.knots, t.coefs, t.n, t.kind, t.x0, t.dx, t.etamax, t.limit};
}

template <typename T>
int lwt_range_impl(ofdft_plan_s* P, const T* rho, const T* phi, const T* p, double c, double s, double* out) {
  const int nb = vec_blocks(P->N);
  PreRho<T> pr{rho, phi, p, T(c), T(s)};
  hipLaunchKernelGGL((k_lwt_range<T>), dim3(nb), dim3(256), 0, P->stream, pr, P->N, P->partials);
  LAUNCHCHK();
  std::vector<double> h(2 * (size_t)nb);
  HIPCHK(hipMemcpyAsync(h.data(), P->partials, sizeof(double) * h.size(), hipMemcpyDeviceToHost, P->stream));
  HIPCHK(hipStreamSynchronize(P->stream));
  double lo = h[0], hi = h[1];
  for (int b = 1; b < nb; ++b) {
    lo = std::min(lo, h[2 * b]);
    hi = std::max(hi, h[2 * b + 1]);
  }
  out[0] = lo;
  out[1] = hi;
  return 0;
}

template <typename T>
int lwt_impl(ofdft_plan_s* P, const ofdft_lwt_params& L, const double* kfl, int nsp, const T* rho, const T* phi,
             const T* p, double c, double s, T* V, double* out) {
  const long long N = P->N, Nhp = P->Nhp;
  if (!P->d_lwt) HIPCHK(hipMalloc(&P->d_lwt, 2 * (size_t)Nhp * sizeof(cplx<T>) + 2 * (size_t)N * sizeof(T)));
  auto* B = reinterpret_cast<cplx<T>*>(P->d_lwt);
  cplx<T>* Acc = B + Nhp;
  T* A1 = reinterpret_cast<T*>(Acc + Nhp);
  T* A3 = A1 + N;
  auto* spec = reinterpret_cast<cplx<T>*>(P->spec);
  PreRho<T> pr{rho, phi, p, T(c), T(s)};
  const int kdd = V ? L.kdd : 1;  // energy only: kdd = 1 (lwt.py:128-132)
  const int herm = L.hermite ? 1 : 0;
  const int mcalc = (kdd > 1 || herm) ? 1 : 0;
  const int nw = (V && kdd > 1) ? (herm ? 2 : 1) : 0;
  const int same = std::fabs(L.alpha - L.beta) < 1e-8 ? 1 : 0;
  const double beta = same ? L.alpha : L.beta;
  const double scale = 1.0 / (double)N;
  const LwtTabDev tK = lwt_tab_dev(P, 0), tD = lwt_tab_dev(P, 1);
  int rc;
  // B = FFT(rho^beta)
  if ((rc = launch_rows_fwd<T>(P, PreLwtPow<T>{pr, beta}, 1))) return rc;
  if ((rc = launch_axis<T>(P, spec, 1, 1, false))) return rc;
  if ((rc = launch_axis<T>(P, spec, 0, 1, false))) return rc;
  HIPCHK(hipMemcpyAsync(B, spec, sizeof(cplx<T>) * Nhp, hipMemcpyDeviceToDevice, P->stream));
  for (int j = 0; j < nsp; ++j) {
    LwtNode nd{j > 0 ? kfl[j - 1] : 0.0, kfl[j], j + 1 < nsp ? kfl[j + 1] : 0.0, j > 0 ? 1 : 0, j + 1 < nsp ? 1 : 0,
               j};
    if (nw) {
      if ((rc = launch_rows_fwd<T>(P, PreLwtW<T>{pr, nd, L.kfmax, L.alpha, herm}, nw))) return rc;
      if ((rc = launch_axis<T>(P, spec, 1, nw, false))) return rc;
      if ((rc = launch_axis<T>(P, spec, 0, nw, false))) return rc;
    }
    hipLaunchKernelGGL((k_lwt_node<T>), dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream, spec, Nhp,
                       (const cplx<T>*)B, Acc, geom(P), tK, tD, kfl[j], nw, mcalc, j == 0 ? 1 : 0, scale);
    LAUNCHCHK();
    const int F = 1 + mcalc;
    if ((rc = launch_axis<T>(P, spec, 0, F, true))) return rc;
    if ((rc = launch_axis<T>(P, spec, 1, F, true))) return rc;
    PostLwtNode<T> pn{pr, nd, L.kfmax, herm, kdd == 3 ? 1 : 0, j == 0 ? 1 : 0, A1, A3};
    if ((rc = launch_rows_inv<T>(P, pn, F, spec, false))) return rc;
  }
  const int has2 = nw ? 1 : 0, mg = L.mgpe ? 1 : 0;
  PostLwtFinal<T> pf{pr, L.kfmax, L.alpha, beta, L.ld, L.rhomax_ld, same, has2, mg, A1, kdd == 3 ? A3 : nullptr, V};
  double sm[3];
  if (has2 + mg) {
    hipLaunchKernelGGL((k_lwt_final_spec<T>), dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream, spec, Nhp,
                       (const cplx<T>*)B, (const cplx<T>*)Acc, geom(P), has2, mg, L.mgpe_a, L.mgpe_b, L.mgpe_c,
                       scale);
    LAUNCHCHK();
    const int F = has2 + mg;
    if ((rc = launch_axis<T>(P, spec, 0, F, true))) return rc;
    if ((rc = launch_axis<T>(P, spec, 1, F, true))) return rc;
    if ((rc = launch_rows_inv<T>(P, pf, F, spec, true))) return rc;
    if ((rc = fetch_sums(P, 3, sm))) return rc;
  } else {
    const int nb = vec_blocks(N);
    hipLaunchKernelGGL((k_lwt_final_pw<T>), dim3(nb), dim3(256), 0, P->stream, pf, N, P->partials);
    LAUNCHCHK();
    if ((rc = reduce_blocks(P, nb, 3, sm))) return rc;
  }
  const double dV = P->dV;
  out[0] = sm[0] * dV;
  out[1] = sm[1] * dV;
  out[2] = sm[2] * dV;
  return 0;
}

// six stress sums of k_stress_recip over the half spectrum into out[0..6)
int stress_recip(ofdft_plan_s* P, const StressArgs& A, int mode, double* out) {
  const int nb = vec_blocks(P->Nh);
  if ((long long)nb * 6 > P->npartials) return fail(OFDFT_ESTATE, "partials buffer too small");
  hipLaunchKernelGGL(k_stress_recip, dim3(nb), dim3(256), 0, P->stream, A, geom(P), mode, P->partials);
  if (hipGetLastError() != hipSuccess) return fail(OFDFT_EHIP, "stress kernel launch");
  return reduce_blocks(P, nb, 6, out);
}

// c2r of spectrum slot 0 (consumed) into a real FP64 field, 1/N normalised.
int c2r_slot0(ofdft_plan_s* P, double* out) {
  auto* s0 = reinterpret_cast<cplx<double>*>(P->spec);
  int rc;
  hipLaunchKernelGGL((k_scale_cplx<double>), dim3(vec_blocks(P->Nhp)), dim3(256), 0, P->stream, s0, P->Nhp,
                     1.0 / (double)P->N);
  LAUNCHCHK();
  if ((rc = launch_axis<double>(P, s0, 0, 1, true))) return rc;
  if ((rc = launch_axis<double>(P, s0, 1, 1, true))) return rc;
  PostStore<double> post{out};
  return launch_rows_inv<double>(P, post, 1, s0, false);
}

int upload_ions(ofdft_plan_s* P, int nat, const double* positions, const int* species) {
  std::vector<double> frac(3 * (size_t)nat);
  // frac = positions . inv(L) (rows of L are lattice vectors); inv(L) = bg^T/(2 pi)
  for (int i = 0; i < nat; ++i)
    for (int d = 0; d < 3; ++d) {
      double f = 0.0;
      for (int e = 0; e < 3; ++e) f += positions[3 * i + e] * P->bg[d * 3 + e] / (2.0 * kPi);
      frac[3 * i + d] = pymod1(pymod1(f));
    }
  if (P->d_frac) (void)hipFree(P->d_frac);
  if (P->d_pos) (void)hipFree(P->d_pos);
  if (P->d_species) (void)hipFree(P->d_species);
  P->d_frac = nullptr;
  P->d_pos = nullptr;
  P->d_species = nullptr;
  HIPCHK(hipMalloc((void**)&P->d_frac, sizeof(double) * 3 * nat));
  HIPCHK(hipMalloc((void**)&P->d_pos, sizeof(double) * 3 * nat));
  HIPCHK(hipMalloc((void**)&P->d_species, sizeof(int) * nat));
  HIPCHK(hipMemcpy(P->d_frac, frac.data(), sizeof(double) * 3 * nat, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P->d_pos, positions, sizeof(double) * 3 * nat, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P->d_species, species, sizeof(int) * nat, hipMemcpyHostToDevice));
  P->nat = nat;
  return 0;
}

int ensure_work(ofdft_plan_s* P) {
  if (!P->d_work) HIPCHK(hipMalloc((void**)&P->d_work, sizeof(double) * P->N));
  return 0;
}

int ensure_bm(ofdft_plan_s* P, int order) {
  std::vector<double> bm = make_bm(P, order);
  if (P->d_bm) (void)hipFree(P->d_bm);
  P->d_bm = nullptr;
  HIPCHK(hipMalloc((void**)&P->d_bm, sizeof(double) * bm.size()));
  HIPCHK(hipMemcpy(P->d_bm, bm.data(), sizeof(double) * bm.size(), hipMemcpyHostToDevice));
  return 0;
}

int reduce_to_host(ofdft_plan_s* P, long long nblk, int ns, double* out) {
  // partial layout [blk][ns]; reuse k_reduce_partials
  if (ns == 1)
    hipLaunchKernelGGL((k_reduce_partials<1>), dim3(1), dim3(kRedThreads), 0, P->stream, P->partials, nblk, P->red, P->dhred, P->skip);
  else
    hipLaunchKernelGGL((k_reduce_partials<3>), dim3(3), dim3(kRedThreads), 0, P->stream, P->partials, nblk, P->red, P->dhred, P->skip);
  LAUNCHCHK();
  return fetch_sums(P, ns, out);
}

bool valid(ofdft_plan P) { return P != nullptr; }

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
namespace {

// One fused CG iteration (the HVP with the CG step folded in, then the TN
// control) as a graph: captured on the plan stream at the first use with
// these buffers, then launched once per iteration.  The captured kernels
// read their control from the device TnState, so replays need no new
// arguments.
template <typename T>
int tn_graph_iteration(ofdft_plan_s* P, const T* phi, T* p, const T* W0, T* d, T* r, T* best, TnState* st,
                       double* rr) {
  const void* key[8] = {phi, p, W0, d, r, best, st, rr};
  if (!P->tn_exec || std::memcmp(key, P->tn_key, sizeof(key)) != 0) {
    if (P->tn_exec) {
      (void)hipGraphExecDestroy(P->tn_exec);
      P->tn_exec = nullptr;
    }
    if (!P->cap_stream) HIPCHK(hipStreamCreateWithFlags(&P->cap_stream, hipStreamNonBlocking));
    // the launchers enqueue on P->stream: point it at the capture stream
    hipStream_t run = P->stream;
    P->stream = P->cap_stream;
    const hipError_t eb = hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal);
    if (eb != hipSuccess) {
      P->stream = run;
      return fail(OFDFT_EHIP, std::string("graph capture begin: ") + hipGetErrorString(eb));
    }
    long long nbr = 0;
    int rc = hvp_cg_impl<T>(P, phi, p, W0, d, r, best, (const TnState*)st, &nbr);
    if (!rc) {
      hipLaunchKernelGGL(k_tn_reduce_ctrl4, dim3(1), dim3(kRedThreads), 0, P->stream, st, (const double*)P->partials,
                         nbr, (const double*)P->red, rr, reinterpret_cast<TnState*>(P->dhtnst));
      if (hipGetLastError() != hipSuccess) rc = fail(OFDFT_EHIP, "control launch during graph capture");
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(P->stream, &g);
    P->stream = run;
    if (!rc && e != hipSuccess) rc = fail(OFDFT_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    if (!rc) {
      const hipError_t ei = hipGraphInstantiate(&P->tn_exec, g, nullptr, nullptr, 0);
      if (ei != hipSuccess) {
        P->tn_exec = nullptr;
        rc = fail(OFDFT_EHIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
      }
    }
    if (g) (void)hipGraphDestroy(g);
    if (rc) return rc;
    std::memcpy(P->tn_key, key, sizeof(key));
  }
  HIPCHK(hipGraphLaunch(P->tn_exec, P->stream));
  return 0;
}

template <typename T>
int tn_direction_impl(ofdft_plan_s* P, const T* phi, const T* res, const T* W0, T* d, T* work, int maxfun, int chunk,
                      int* niter, int* status, double* pAp_last) {
  if (!P->have_fn) return fail(OFDFT_ESTATE, "functional not set");
  const long long N = P->N;
  T* r = work;
  T* p = work + N;
  T* Ap = work + 2 * N;
  T* best = work + 3 * N;
  auto* st = reinterpret_cast<TnState*>(P->tnst);
  const int nb = vec_blocks(N);
  double* rr = P->red + 8;  // r^2 lands in red[8]; pAp in red[0]
  P->skip = nullptr;
  {
    ProfScope ps(P, PC_VEC, (long long)P->rs * 4 * N);
    OFDFT_LAUNCH((k_tn_init<T>), dim3(nb), dim3(256), 0, P, res, W0, d, r, p, N, P->partials);
  }
  LAUNCHCHK();
  // rr[0] = r0, rr[1] = sum W0 p^2 of the first direction
  hipLaunchKernelGGL((k_reduce_partials<2>), dim3(2), dim3(kRedThreads), 0, P->stream, P->partials, (long long)nb, rr,
                     (double*)nullptr, (const int*)nullptr);
  LAUNCHCHK();
  auto* hst = reinterpret_cast<TnState*>(P->dhtnst);
  hipLaunchKernelGGL(k_tn_start, dim3(1), dim3(1), 0, P->stream, st, (const double*)rr, maxfun, hst);
  LAUNCHCHK();
  auto* h = reinterpret_cast<volatile TnState*>(P->htnst);
  const size_t prof_mark = P->prof_pending.size();
  P->skip = &st->done;
  const bool fused = make_fn(P->fn).F > 0;  // a forward row pass carries the update of p
  // the CG step rides on the inverse row pass when the pencil can sum p.Ap
  const bool cg_fused = !fused || pencil_reg_ok(P);
  // Queue k iterations, then 2 at a time.  The host checks the state that the
  // control kernels publish to mapped memory at an event after the
  // second-to-last queued iteration, so the device still runs the last one
  // while the host decides: no idle gap when more are needed, and a
  // finished loop leaves at most one iteration of launches that return at
  // entry.  The call returns without synchronising the stream.
  int rc = 0, queued = 0, k = std::max(1, chunk);
  bool done = false;
  while (queued < maxfun) {
    const int nq = std::min(k, maxfun - queued);
    const int mark = (nq >= 2 && queued + nq < maxfun) ? nq - 2 : nq - 1;
    for (int j = 0; j < nq; ++j, ++queued) {
      P->prof_tag = queued;
      if (!fused) {
        ProfScope ps(P, PC_VEC, (long long)P->rs * 3 * N);
        OFDFT_LAUNCH((k_tn_next<T>), dim3(nb), dim3(256), 0, P, (const TnState*)st, (const T*)r, p, N);
      }
      if (cg_fused && fused && P->use_graph && !P->prof_on) {
        if ((rc = tn_graph_iteration<T>(P, phi, p, W0, d, r, best, st, rr))) break;
      } else if (cg_fused) {
        long long nbr = 0;
        if ((rc = hvp_cg_impl<T>(P, phi, p, W0, d, r, best, (const TnState*)st, &nbr))) break;
        hipLaunchKernelGGL(k_tn_reduce_ctrl4, dim3(1), dim3(kRedThreads), 0, P->stream, st,
                           (const double*)P->partials, nbr, (const double*)P->red, rr, hst);
      } else {
        if ((rc = hvp_fast_impl<T>(P, phi, p, W0, Ap, nullptr, (const T*)r, (const TnState*)st))) break;
        {
          ProfScope ps(P, PC_VEC, (long long)P->rs * 6 * N);
          OFDFT_LAUNCH((k_tn_step<T>), dim3(nb), dim3(256), 0, P, st, (const double*)P->red,
                       (const T*)p, (const T*)Ap, d, r, best, N, P->partials);
        }
        hipLaunchKernelGGL(k_tn_reduce_ctrl, dim3(1), dim3(kRedThreads), 0, P->stream, st,
                           (const double*)P->partials, (long long)nb, (const double*)P->red, rr, hst);
      }
      if (hipGetLastError() != hipSuccess) {
        rc = fail(OFDFT_EHIP, "kernel launch failed in the TN loop");
        break;
      }
      if (j == mark && hipEventRecord(P->tn_ev, P->stream) != hipSuccess) {
        rc = fail(OFDFT_EHIP, "TN loop event");
        break;
      }
    }
    if (rc) break;
    P->prof_tag = -1;
    if (hipEventSynchronize(P->tn_ev) != hipSuccess) {
      rc = fail(OFDFT_EHIP, "TN loop event wait");
      break;
    }
    if (h->done) {
      done = true;
      break;
    }
    k = 2;
  }
  P->prof_tag = -1;
  P->skip = nullptr;
  if (rc) return rc;
  // every iteration was queued (maxfun) without the state showing done at a
  // check point: the last iteration's control ends the loop
  if (!done && hipStreamSynchronize(P->stream) != hipSuccess) return fail(OFDFT_EHIP, "TN state read-back");
  std::atomic_thread_fence(std::memory_order_acquire);
  if (!h->done) return fail(OFDFT_ESTATE, "TN loop did not finish");
  if (h->fin) {
    ProfScope ps(P, PC_VEC, (long long)P->rs * 2 * N);
    OFDFT_LAUNCH((k_tn_finish<T>), dim3(nb), dim3(256), 0, P, (const TnState*)st, d, (const T*)p,
                 (const T*)best, N);
    LAUNCHCHK();
  }
  // launches queued after the loop finished did no work: drop their timings
  P->prof_pending.erase(std::remove_if(P->prof_pending.begin() + prof_mark, P->prof_pending.end(),
                                       [&](const ofdft_plan_s::ProfRec& x) { return x.tag >= h->niter; }),
                        P->prof_pending.end());
  *niter = h->niter;
  *status = h->status;
  *pAp_last = h->pAp;
  return 0;
}

}  // namespace

namespace {

// Device copies of the ion data an Ewald sum needs, and the image shifts
// T = (ix, iy, iz) . L with |i| <= ceil(rmax / |a_i|) (ewald.py:287-300).
struct EwaldDev {
  double *pos = nullptr, *q = nullptr, *sh = nullptr, *frac = nullptr;
  int nshift = 0;
  ~EwaldDev() {
    if (pos) (void)hipFree(pos);
    if (q) (void)hipFree(q);
    if (sh) (void)hipFree(sh);
    if (frac) (void)hipFree(frac);
  }
};

// Host arrays of an Ewald sum: the image shifts T = (ix, iy, iz) . L with
// |i| <= ceil(rmax / |a_i|) (ewald.py:287-300) and the wrapped fractional
// coordinates of the ions.
void ewald_host_arrays(const ofdft_plan_s* P, int nat, const double* positions, double rmax, std::vector<double>& sh,
                       std::vector<double>& frac) {
  int Nimg[3];
  for (int d = 0; d < 3; ++d) {
    const double len = std::sqrt(P->lat[3 * d] * P->lat[3 * d] + P->lat[3 * d + 1] * P->lat[3 * d + 1] +
                                 P->lat[3 * d + 2] * P->lat[3 * d + 2]);
    Nimg[d] = (int)std::ceil(rmax / len);
  }
  sh.clear();
  for (int a = -Nimg[0]; a <= Nimg[0]; ++a)
    for (int b = -Nimg[1]; b <= Nimg[1]; ++b)
      for (int c = -Nimg[2]; c <= Nimg[2]; ++c)
        for (int d = 0; d < 3; ++d) sh.push_back(a * P->lat[d] + b * P->lat[3 + d] + c * P->lat[6 + d]);
  frac.assign(3 * (size_t)nat, 0.0);
  for (int i = 0; i < nat; ++i)
    for (int d = 0; d < 3; ++d) {
      double f = 0.0;
      for (int e = 0; e < 3; ++e) f += positions[3 * i + e] * P->bg[d * 3 + e] / (2.0 * kPi);
      frac[3 * i + d] = pymod1(pymod1(f));
    }
}

int ewald_upload(ofdft_plan_s* P, int nat, const double* positions, const double* charges, double rmax,
                 EwaldDev& E) {
  std::vector<double> sh, frac;
  ewald_host_arrays(P, nat, positions, rmax, sh, frac);
  E.nshift = (int)(sh.size() / 3);
  if (hipMalloc((void**)&E.pos, sizeof(double) * 3 * nat) != hipSuccess ||
      hipMalloc((void**)&E.q, sizeof(double) * nat) != hipSuccess ||
      hipMalloc((void**)&E.sh, sizeof(double) * sh.size()) != hipSuccess ||
      hipMalloc((void**)&E.frac, sizeof(double) * 3 * nat) != hipSuccess)
    return fail(OFDFT_ENOMEM, "Ewald buffers");
  if (hipMemcpy(E.pos, positions, sizeof(double) * 3 * nat, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(E.q, charges, sizeof(double) * nat, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(E.sh, sh.data(), sizeof(double) * sh.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(E.frac, frac.data(), sizeof(double) * 3 * nat, hipMemcpyHostToDevice) != hipSuccess)
    return fail(OFDFT_EHIP, "Ewald upload");
  return 0;
}

int ewald_check(ofdft_plan_s* P, int nat, double eta, double rmax, int pme, int order) {
  if (P->prec != OFDFT_F64) return fail(OFDFT_EINVAL, "Ewald sums run on an FP64 plan");
  if (nat < 1 || !(eta > 0) || !(rmax > 0)) return fail(OFDFT_EINVAL, "bad Ewald parameters");
  if (pme && (order < 3 || order > kMaxOrder)) return fail(OFDFT_EINVAL, "B-spline order must be in [3, 12]");
  return 0;
}

__global__ void k_zero(double* __restrict__ p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// Device buffers of one Ewald sum: ion data, image shifts and (PME) the
// charge grid (N doubles) and b-moduli.
struct EwaldBufs {
  const double *pos, *q, *sh, *frac;
  int nshift;
  double* work;
  const double* bm;
};

// The charge Q-array of the ions, spread with B-splines of the given order,
// transformed into spectrum slot 0 (ewald.py:137-194, :786-790).
int ewald_qhat_into(ofdft_plan_s* P, const EwaldBufs& B, int nat, int order) {
  hipLaunchKernelGGL(k_zero, dim3(vec_blocks(P->N)), dim3(256), 0, P->stream, B.work, P->N);
  if (hipGetLastError() != hipSuccess) return fail(OFDFT_EHIP, "zero launch");
  hipLaunchKernelGGL(k_pme_spread, dim3(nat), dim3(1024), 0, P->stream, B.work, B.frac, (const int*)nullptr, -1,
                     B.q, P->n[0], P->n[1], P->n[2], order);
  if (hipGetLastError() != hipSuccess) return fail(OFDFT_EHIP, "spread launch");
  return r2c_slot(P, B.work, 0);
}

int ewald_qhat(ofdft_plan_s* P, const EwaldDev& E, int nat, int order) {
  int rc;
  if ((rc = ensure_work(P)) || (rc = ensure_bm(P, order))) return rc;
  return ewald_qhat_into(P, EwaldBufs{E.pos, E.q, E.sh, E.frac, E.nshift, P->d_work, P->d_bm}, nat, order);
}

// Real-space and reciprocal Ewald sums on prepared buffers (the body of
// ofdft_ewald and ofdft_ewald_noalloc).
int ewald_energy_sums(ofdft_plan_s* P, const EwaldBufs& B, int nat, double eta, double rmax, int pme, int order,
                      double* e_real, double* e_rec) {
  int rc;
  const int nb = vec_blocks((long long)nat * nat);
  hipLaunchKernelGGL(k_ewald_real, dim3(nb), dim3(256), 0, P->stream, B.pos, B.q, nat, B.sh, B.nshift, rmax,
                     std::sqrt(eta), P->partials);
  if (hipGetLastError() != hipSuccess) return fail(OFDFT_EHIP, "Ewald real launch");
  double er;
  if ((rc = reduce_to_host(P, nb, 1, &er))) return rc;
  *e_real = er / 2.0;
  GridGeom g = geom(P);
  const cplx<double>* qhat = nullptr;
  if (pme) {
    if ((rc = ewald_qhat_into(P, B, nat, order))) return rc;
    qhat = reinterpret_cast<const cplx<double>*>(P->spec);
  }
  const int nb2 = vec_blocks(P->Nh);
  hipLaunchKernelGGL(k_ewald_rec, dim3(nb2), dim3(256), 0, P->stream, qhat, B.bm, g, B.pos, B.q, nat, eta,
                     P->partials);
  if (hipGetLastError() != hipSuccess) return fail(OFDFT_EHIP, "Ewald rec launch");
  double ek;
  if ((rc = reduce_to_host(P, nb2, 1, &ek))) return rc;
  *e_rec = 4.0 * kPi * ek / P->volume;
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Transform pipelines instantiated in their own translation units (one per
// family and precision, ofdft_tu_*.hip) so that the library compiles in
// parallel; the C ABI calls them through these wrappers.
// ---------------------------------------------------------------------------
namespace ofdft_tu {

template <typename T>
int eval(ofdft_plan_s* P, const T* rho, const T* phi, const T* p, double c, double s, T* V, T* aux, double* sums) {
  return eval_impl<T>(P, rho, phi, p, c, s, V, aux, sums);
}

template <typename T>
int hvp(ofdft_plan_s* P, const T* phi, const T* p, const T* V, const T* aux, double mu, T* Ap, double* pAp, int raw) {
  return hvp_impl<T>(P, phi, p, V, aux, mu, Ap, pAp, raw);
}

template <typename T>
int hvp_fast(ofdft_plan_s* P, const T* phi, const T* p, const T* W0, T* Ap, double* pAp) {
  return hvp_fast_impl<T>(P, phi, p, W0, Ap, pAp);
}

template <typename T>
int tn_direction(ofdft_plan_s* P, const T* phi, const T* res, const T* W0, T* d, T* work, int maxfun, int chunk,
                 int* niter, int* status, double* pAp_last) {
  return tn_direction_impl<T>(P, phi, res, W0, d, work, maxfun, chunk, niter, status, pAp_last);
}

template <typename T>
int gga(ofdft_plan_s* P, int functional, const double* params, int nparams, const T* rho, const T* phi, const T* p,
        double c, double s, T* V, double* out) {
  GgaSpec g{};
  g.id = functional;
  g.np = nparams;
  for (int k = 0; k < nparams; ++k) g.p[k] = params[k];
  return gga_impl<T>(P, g, rho, phi, p, c, s, V, out);
}

template <typename T>
int lwt(ofdft_plan_s* P, const ofdft_lwt_params& L, const double* kfl, int nsp, const T* rho, const T* phi,
        const T* p, double c, double s, T* V, double* out) {
  return lwt_impl<T>(P, L, kfl, nsp, rho, phi, p, c, s, V, out);
}

}  // namespace ofdft_tu

// EXT = extern: a declaration (the C ABI's translation unit); empty: the
// explicit instantiation (the unit that compiles it)
#define OFDFT_TU_EVAL(EXT, T)                                                                               \
  EXT template int ofdft_tu::eval<T>(ofdft_plan_s*, const T*, const T*, const T*, double, double, T*, T*, \
                                     double*);
#define OFDFT_TU_HVP(EXT, T)                                                                                 \
  EXT template int ofdft_tu::hvp<T>(ofdft_plan_s*, const T*, const T*, const T*, const T*, double, T*, double*, \
                                    int);                                                                    \
  EXT template int ofdft_tu::hvp_fast<T>(ofdft_plan_s*, const T*, const T*, const T*, T*, double*);           \
  EXT template int ofdft_tu::tn_direction<T>(ofdft_plan_s*, const T*, const T*, const T*, T*, T*, int, int,    \
                                             int*, int*, double*);
#define OFDFT_TU_GGA(EXT, T)                                                                                     \
  EXT template int ofdft_tu::gga<T>(ofdft_plan_s*, int, const double*, int, const T*, const T*, const T*, double, \
                                    double, T*, double*);
#define OFDFT_TU_LWT(EXT, T)                                                                                     \
  EXT template int ofdft_tu::lwt<T>(ofdft_plan_s*, const ofdft_lwt_params&, const double*, int, const T*,         \
                                    const T*, const T*, double, double, T*, double*);

namespace ofdft_tu {
// the TN inner loop of every plan of a batch in one persistent cooperative
// launch (defined in tn_batch.h, instantiated in ofdft_tu_batch_f*.hip)
template <typename T>
int tn_batch_run(ofdft_batch_s* B, const T* const* phi, const T* const* res, const T* const* W0, T* const* d,
                 T* const* work, int maxfun, int* niter, int* status, double* pAp_last);
}  // namespace ofdft_tu
