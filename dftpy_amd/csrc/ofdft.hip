// MI355X (gfx950) kernels and C ABI for the orbital-free DFT hot path.
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
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ofdft_amd.h"
#include "fft_lds.h"

using namespace ofdft;

namespace {

thread_local std::string g_err;

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
constexpr int kRedThreads = 256;

// ---------------------------------------------------------------------------
// host-side pass schedule and twiddles
// ---------------------------------------------------------------------------
bool make_pass(int n, Pass1D* pl) {
  pl->n = n;
  pl->npass = 0;
  int m = n;
  auto push = [&](int r) {
    if (pl->npass >= kMaxPass) return false;
    pl->radix[pl->npass++] = r;
    return true;
  };
  while (m % 4 == 0) { if (!push(4)) return false; m /= 4; }
  while (m % 2 == 0) { if (!push(2)) return false; m /= 2; }
  while (m % 3 == 0) { if (!push(3)) return false; m /= 3; }
  while (m % 5 == 0) { if (!push(5)) return false; m /= 5; }
  return m == 1;
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
  int nh;
  long long N, Nh;
  int prec;
  int full;   // reference full=True semantics (Hermitian-symmetrised multipliers)
  size_t rs;  // bytes per real
  double lat[9], bg[9], volume, dV;
  hipStream_t stream;
  Pass1D pl[3];
  void* tw[3];      // twiddle tables, plan precision complex
  void* spec;       // 3*Nh complex workspace
  void* wtk;        // WT kernel table, Nh reals (plan precision)
  void* vloc;       // N reals
  double* partials; // per-block partial sums
  long long npartials;
  double* red;      // device, kMaxSum doubles
  double* hred;     // pinned host mirror
  ofdft_functional fn;
  int have_fn, have_vloc, have_wtk;
};

namespace {

constexpr int kMaxSum = 16;

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T dev_cbrt(T x);
template <>
__device__ __forceinline__ double dev_cbrt<double>(double x) { return cbrt(x); }
template <>
__device__ __forceinline__ float dev_cbrt<float>(float x) { return cbrtf(x); }

template <typename T>
__device__ __forceinline__ T dev_pow(T x, T y);
template <>
__device__ __forceinline__ double dev_pow<double>(double x, double y) { return pow(x, y); }
template <>
__device__ __forceinline__ float dev_pow<float>(float x, float y) { return powf(x, y); }

template <typename T>
__device__ __forceinline__ T dev_log(T x);
template <>
__device__ __forceinline__ double dev_log<double>(double x) { return log(x); }
template <>
__device__ __forceinline__ float dev_log<float>(float x) { return logf(x); }

template <typename T>
__device__ __forceinline__ T dev_sqrt(T x);
template <>
__device__ __forceinline__ double dev_sqrt<double>(double x) { return sqrt(x); }
template <>
__device__ __forceinline__ float dev_sqrt<float>(float x) { return sqrtf(x); }

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
  double x, y, alpha, beta;
  int fH, fV, fW;  // spectrum slot of each convolution field (-1 if absent)
  int F;
};

// LDA exchange-correlation energy density per electron and potential:
// xc/semilocal_xc.py:287-320.
template <typename T>
__device__ __forceinline__ void lda_point(T rho, T& ex, T& vxc) {
  const T a = T(0.0311), b = T(-0.048), c = T(0.0020), d = T(-0.0116);
  const T g = T(-0.1423), b1 = T(1.0529), b2 = T(0.3334);
  T rc = dev_cbrt(rho);
  if (rc < T(1e-30)) rc = T(1e-30);
  const T rs = T(0.62035049089940001667) / rc;  // cbrt(3/(4 pi)) / cbrt(rho)
  ex = T(-0.73855876638202240588) * rc;         // -3/4 cbrt(3/pi)
  vxc = T(-0.98474502184269654118) * rc;        // cbrt(-3/pi)
  if (rs < T(1)) {
    const T lr = dev_log(rs);
    ex += a * lr + b + c * rs * lr + d * rs;
    vxc += lr * (a + T(2.0 / 3.0) * c * rs) + b - T(1.0 / 3.0) * a + T(1.0 / 3.0) * (T(2) * d - c) * rs;
  } else {
    const T sr = dev_sqrt(rs);
    const T den = T(1) + b1 * sr + b2 * rs;
    ex += g / den;
    vxc += (g + (T(7.0 / 6.0) * g * b1) * sr + (T(4.0 / 3.0) * g * b2 * rs)) / (den * den);
  }
}

// d V_xc / d rho: xc/semilocal_xc.py:321-333.
template <typename T>
__device__ __forceinline__ T lda_v2(T rho) {
  const T a = T(0.0311), c = T(0.0020), d = T(-0.0116);
  const T g = T(-0.1423), b1 = T(1.0529), b2 = T(0.3334);
  T rc0 = dev_cbrt(rho);
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
  int full;  // 1: frequencies of the full complex grid (grid.py:584-620, full=True)
  double bg[9];
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
  const T* __restrict__ in;
  __device__ __forceinline__ void operator()(long long i, T* out) const { out[0] = in[i]; }
};

// rho (mode 0) or phi(t) = c*phi + s*p (mode 1), then the convolution inputs
// rho (Hartree), sqrt(max(rho,ZERO)) (vW, vw.py:34-44), max(rho,ZERO)^beta
// (WT, wt.py:15-19).
template <typename T>
struct PreEval {
  const T* __restrict__ rho;
  const T* __restrict__ phi;
  const T* __restrict__ p;
  T c, s;
  FnParams fp;
  __device__ __forceinline__ T density(long long i) const {
    if (rho) return rho[i];
    T ph = phi[i];
    if (p) ph = c * ph + s * p[i];
    return ph * ph;
  }
  __device__ __forceinline__ void operator()(long long i, T* out) const {
    const T r = density(i);
    if (fp.fH >= 0) put3(out, fp.fH, r);
    if (fp.fV >= 0) put3(out, fp.fV, dev_sqrt(r > T(0) ? r : T(kZero)));
    if (fp.fW >= 0) put3(out, fp.fW, dev_pow(r < T(kZero) ? T(kZero) : r, T(fp.beta)));
  }
};

// HVP inputs for direction p at phi: drho = 2 phi p (Hartree), ds = drho/(2s)
// (vW), beta*rho^(beta-1)*drho (WT).
template <typename T>
struct PreHvp {
  const T* __restrict__ phi;
  const T* __restrict__ p;
  FnParams fp;
  __device__ __forceinline__ void operator()(long long i, T* out) const {
    const T ph = phi[i], pp = p[i];
    const T r = ph * ph;
    const T dr = T(2) * ph * pp;
    if (fp.fH >= 0) put3(out, fp.fH, dr);
    if (fp.fV >= 0) put3(out, fp.fV, r > T(0) ? dr / (T(2) * dev_sqrt(r)) : T(0));
    if (fp.fW >= 0)
      put3(out, fp.fW, r < T(kZero) ? T(0) : T(fp.beta) * dev_pow(r, T(fp.beta) - T(1)) * dr);
  }
};

// ---------------------------------------------------------------------------
// post-pointwise functors (consume rows_inv); c[f] are the normalised
// convolution results of the F fields at point i.
// ---------------------------------------------------------------------------
template <typename T>
struct PostStore {
  static constexpr int NS = 1;
  T* __restrict__ out;
  __device__ __forceinline__ void operator()(long long i, const T* c, double* acc) const { out[i] = c[0]; }
};

template <typename T>
struct PostEval {
  static constexpr int NS = OFDFT_NSUM;
  PreEval<T> pre;
  const T* __restrict__ vloc;
  T* __restrict__ V;
  T* __restrict__ aux;  // [0,N) vW raw potential, [N,2N) WT convolution
  long long N;
  __device__ __forceinline__ void operator()(long long i, const T* c, double* acc) const {
    const FnParams& fp = pre.fp;
    T ph = T(0), dph = T(0);
    T r;
    if (pre.rho) {
      r = pre.rho[i];
    } else {
      const T p0 = pre.phi[i];
      const T pp = pre.p ? pre.p[i] : T(0);
      ph = pre.c * p0 + pre.s * pp;
      dph = pre.c * pp - pre.s * p0;
      r = ph * ph;
    }
    T v = T(0);
    if (fp.tf) {
      // TF: energy density cbrt(rho^5), potential (5/3)C_TF cbrt(rho^2) (tf.py:13-35)
      const T r2 = r * r;
      acc[OFDFT_E_TF] += (double)dev_cbrt(r2 * r2 * r);
      v += T(4.7853900003136530265) * dev_cbrt(r2) * T(fp.x);
    }
    if (fp.fV >= 0) {
      // vW: 0.5 * IFFT[G^2 FFT(s)] / s, energy over rho > 0 (vw.py:29-53, :93)
      T s = dev_sqrt(r > T(0) ? r : T(kZero));
      if (s < T(kZero)) s = T(kZero);
      const T vraw = T(0.5) * pick3(c, fp.fV) / s;
      if (r > T(0)) acc[OFDFT_E_VW] += (double)(r * vraw);
      v += vraw * T(fp.y);
      if (aux) aux[i] = vraw;
    }
    if (fp.fW >= 0) {
      // WT: 2 alpha rho^beta / rho * conv, energy density V rho/(2 alpha) (wt.py:12-28, :125)
      const T rz = r < T(kZero) ? T(kZero) : r;
      const T vw = T(2 * fp.alpha) * (dev_pow(rz, T(fp.beta)) / rz) * pick3(c, fp.fW);
      acc[OFDFT_E_NL] += (double)(vw * r / T(2 * fp.alpha));
      v += vw;
      if (aux) aux[N + i] = pick3(c, fp.fW);
    }
    if (fp.lda) {
      T ex, vx;
      lda_point(r, ex, vx);
      acc[OFDFT_E_XC] += (double)(ex * r);
      v += vx;
    }
    if (fp.fH >= 0) {
      const T vh = pick3(c, fp.fH);
      acc[OFDFT_E_H] += (double)(vh * r);
      v += vh;
    }
    if (fp.pseudo) {
      const T vl = vloc[i];
      acc[OFDFT_E_PP] += (double)(vl * r);
      v += vl;
    }
    acc[OFDFT_S_VRHO] += (double)(v * r);
    acc[OFDFT_S_RHO] += (double)r;
    if (!pre.rho) acc[OFDFT_S_DTHETA] += (double)(v * ph * dph);
    if (V) V[i] = v;
  }
};

// Ap = (V - mu) p + phi * dV[drho]; acc[0] = sum p*Ap.
template <typename T>
struct PostHvp {
  static constexpr int NS = 1;
  const T* __restrict__ phi;
  const T* __restrict__ p;
  const T* __restrict__ V;
  const T* __restrict__ aux;
  T* __restrict__ Ap;
  T mu;
  long long N;
  FnParams fp;
  __device__ __forceinline__ void operator()(long long i, const T* c, double* acc) const {
    const T ph = phi[i], pp = p[i];
    const T r = ph * ph;
    const T dr = T(2) * ph * pp;
    T dv = T(0);
    if (fp.tf && r > T(0)) dv += T(fp.x) * T(3.1902600002091020177) * dr / dev_cbrt(r);  // (2/3)(5/3)C_TF
    if (fp.lda && r > T(0)) dv += lda_v2(r) * dr;
    if (fp.fH >= 0) dv += pick3(c, fp.fH);
    if (fp.fV >= 0 && r > T(0)) {
      const T s = dev_sqrt(r);
      const T ds = dr / (T(2) * s);
      dv += T(fp.y) * (T(0.5) * pick3(c, fp.fV) / s - aux[i] * ds / s);
    }
    if (fp.fW >= 0 && r >= T(kZero)) {
      const T rb1 = dev_pow(r, T(fp.beta) - T(1));
      dv += T(2 * fp.alpha) * ((T(fp.beta) - T(1)) * rb1 / r * dr * aux[N + i] + rb1 * pick3(c, fp.fW));
    }
    const T ap = (V[i] - mu) * pp + ph * dv;
    Ap[i] = ap;
    acc[0] += (double)(pp * ap);
  }
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// Axis-2 real-to-complex: block handles PB row pairs x F fields.  Rows r and
// r+1 of field f share one complex sequence z = a + i b; after the FFT,
// A[k] = (Z[k] + conj Z[-k])/2, B[k] = (Z[k] - conj Z[-k])/(2i).
template <typename T, int KB, class Pre>
__global__ void k_rows_fwd(Pre pre, cplx<T>* __restrict__ spec, long long nrows, int n2, int nh, long long Nh,
                           int F, int PB, Pass1D pl, const cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const long long pair0 = (long long)blockIdx.x * PB;
  const int ld = n2;
  for (int t = threadIdx.x; t < PB * n2; t += blockDim.x) {
    const int lp = t / n2, j = t - lp * n2;
    const long long r0 = 2 * (pair0 + lp), r1 = r0 + 1;
    T a[3] = {T(0), T(0), T(0)}, b[3] = {T(0), T(0), T(0)};
    if (r0 < nrows) pre(r0 * n2 + j, a);
    if (r1 < nrows) pre(r1 * n2 + j, b);
#pragma unroll
    for (int f = 0; f < 3; ++f)
      if (f < F) lds[(lp * F + f) * ld + j] = cplx<T>{a[f], b[f]};
  }
  __syncthreads();
  lds_fft<T, KB>(lds, ld, PB * F, pl, tw, false);
  for (int t = threadIdx.x; t < PB * F * nh; t += blockDim.x) {
    const int tr = t / nh, k = t - tr * nh;
    const int lp = tr / F, f = tr - lp * F;
    const long long r0 = 2 * (pair0 + lp), r1 = r0 + 1;
    const cplx<T> zk = lds[tr * ld + k];
    const cplx<T> zm = lds[tr * ld + (k == 0 ? 0 : n2 - k)];
    if (r0 < nrows) spec[f * Nh + r0 * nh + k] = cplx<T>{T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
    if (r1 < nrows) spec[f * Nh + r1 * nh + k] = cplx<T>{T(0.5) * (zk.y + zm.y), T(0.5) * (zm.x - zk.x)};
  }
}

// Axis-2 complex-to-real fused with the post functor.  Two half spectra A, B
// are combined into Z = A + iB on the full axis (Hermitian extension, DC and
// Nyquist imaginary parts dropped as irfft does); after the inverse FFT the
// real part is row r, the imaginary part row r+1.
template <typename T, int KB, class Post>
__global__ void k_rows_inv(Post post, const cplx<T>* __restrict__ spec, long long nrows, int n2, int nh,
                           long long Nh, int F, int PB, Pass1D pl, const cplx<T>* __restrict__ tw,
                           double* __restrict__ partials) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const long long pair0 = (long long)blockIdx.x * PB;
  const int ld = n2;
  const bool even = (n2 % 2) == 0;
  for (int t = threadIdx.x; t < PB * F * nh; t += blockDim.x) {
    const int tr = t / nh, k = t - tr * nh;
    const int lp = tr / F, f = tr - lp * F;
    const long long r0 = 2 * (pair0 + lp), r1 = r0 + 1;
    cplx<T> A = r0 < nrows ? spec[f * Nh + r0 * nh + k] : cplx<T>{T(0), T(0)};
    cplx<T> B = r1 < nrows ? spec[f * Nh + r1 * nh + k] : cplx<T>{T(0), T(0)};
    if (k == 0 || (even && k == n2 / 2)) {
      A.y = T(0);
      B.y = T(0);
    }
    lds[tr * ld + k] = cplx<T>{A.x - B.y, A.y + B.x};
    if (k != 0 && n2 - k != k) lds[tr * ld + (n2 - k)] = cplx<T>{A.x + B.y, B.x - A.y};
  }
  __syncthreads();
  lds_fft<T, KB>(lds, ld, PB * F, pl, tw, true);
  double acc[Post::NS];
#pragma unroll
  for (int s = 0; s < Post::NS; ++s) acc[s] = 0.0;
  for (int t = threadIdx.x; t < PB * n2; t += blockDim.x) {
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
  if (partials) block_sums<Post::NS>(acc, partials + (long long)blockIdx.x * Post::NS);
}

// Strided complex transform along one axis: blockIdx.y selects the outer
// index, blockIdx.x a tile of W contiguous columns, blockIdx.z the field.
template <typename T, int KB>
__global__ void k_axis_c2c(cplx<T>* __restrict__ spec, long long Nh, int n, long long stride, long long ostride,
                           int ncol, int W, Pass1D pl, const cplx<T>* __restrict__ tw, int inv) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* base = spec + (long long)blockIdx.z * Nh + (long long)blockIdx.y * ostride;
  const int c0 = blockIdx.x * W;
  const int ld = n + 1;
  for (int t = threadIdx.x; t < n * W; t += blockDim.x) {
    const int i = t / W, w = t - i * W;
    const int c = c0 + w;
    lds[w * ld + i] = c < ncol ? base[i * stride + c] : cplx<T>{T(0), T(0)};
  }
  __syncthreads();
  lds_fft<T, KB>(lds, ld, W, pl, tw, inv != 0);
  for (int t = threadIdx.x; t < n * W; t += blockDim.x) {
    const int i = t / W, w = t - i * W;
    const int c = c0 + w;
    if (c < ncol) base[i * stride + c] = lds[w * ld + i];
  }
}

struct ConvKinds {
  int kind[3];
  double scale;  // 1/N
};

// Axis-0 pencils: forward FFT, multiply by M_f(G)/N, inverse FFT.
template <typename T, int KB>
__global__ void k_pencil_conv(cplx<T>* __restrict__ spec, long long Nh, GridGeom geo, Pass1D pl,
                              const cplx<T>* __restrict__ tw, int W, ConvKinds ck, const T* __restrict__ wtk) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n = geo.n0;
  const int i1 = blockIdx.y;
  const int f = blockIdx.z;
  const long long stride = (long long)geo.n1 * geo.nh;
  cplx<T>* base = spec + (long long)f * Nh + (long long)i1 * geo.nh;
  const int c0 = blockIdx.x * W;
  const int ld = n + 1;
  for (int t = threadIdx.x; t < n * W; t += blockDim.x) {
    const int i = t / W, w = t - i * W;
    const int c = c0 + w;
    lds[w * ld + i] = c < geo.nh ? base[i * stride + c] : cplx<T>{T(0), T(0)};
  }
  __syncthreads();
  lds_fft<T, KB>(lds, ld, W, pl, tw, false);
  const int kind = ck.kind[f];
  for (int t = threadIdx.x; t < n * W; t += blockDim.x) {
    const int w = t / n, k0 = t - w * n;
    const int c = c0 + w;
    if (c >= geo.nh) continue;
    double m;
    if (kind == OFDFT_CONV_HARTREE) {
      const double g2 = gsq(geo, k0, i1, c);
      m = g2 > 0.0 ? 4.0 * kPi / g2 : 0.0;
      if (geo.full) {
        const double h2 = gsq(geo, k0, i1, c, 1);
        m = 0.5 * (m + (h2 > 0.0 ? 4.0 * kPi / h2 : 0.0));
      }
    } else if (kind == OFDFT_CONV_GG) {
      m = gsq(geo, k0, i1, c);
      if (geo.full) m = 0.5 * (m + gsq(geo, k0, i1, c, 1));
    } else if (kind == OFDFT_CONV_WT) {
      m = (double)wtk[((long long)k0 * geo.n1 + i1) * geo.nh + c];
    } else {
      m = 1.0;
    }
    const T mm = T(m * ck.scale);
    cplx<T> z = lds[w * ld + k0];
    lds[w * ld + k0] = cplx<T>{z.x * mm, z.y * mm};
  }
  __syncthreads();
  lds_fft<T, KB>(lds, ld, W, pl, tw, true);
  for (int t = threadIdx.x; t < n * W; t += blockDim.x) {
    const int i = t / W, w = t - i * W;
    const int c = c0 + w;
    if (c < geo.nh) base[i * stride + c] = lds[w * ld + i];
  }
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

template <int NS>
__global__ void k_reduce_partials(const double* __restrict__ partials, long long nblk, double* __restrict__ out) {
  const int s = blockIdx.x;
  double acc = 0.0;
  for (long long b = threadIdx.x; b < nblk; b += blockDim.x) acc += partials[b * NS + s];
  double tmp[1] = {acc};
  block_sums<1>(tmp, out + s);
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

template <typename T>
__global__ void k_lincomb(T alpha, const T* x, T beta, const T* y, T* z, long long N) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
    z[i] = alpha * x[i] + beta * y[i];
}

template <typename T>
__global__ void k_residual(const T* __restrict__ V, const T* __restrict__ phi, T mu, T* __restrict__ res, long long N,
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
constexpr int kRowPairs = 2;   // row pairs per rows_* block
constexpr int kColTile = 16;   // W for axis-1 / axis-0 tiles

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
  g.full = P->full;
  for (int i = 0; i < 9; ++i) g.bg[i] = P->bg[i];
  return g;
}

template <typename T, class Pre>
int launch_rows_fwd(ofdft_plan_s* P, const Pre& pre, int F) {
  const long long nrows = (long long)P->n[0] * P->n[1];
  const long long npairs = (nrows + 1) / 2;
  const int PB = kRowPairs;
  FFTLaunch L = fft_launch(P->pl[2], PB * F);
  const size_t lds = (size_t)PB * F * P->n[2] * sizeof(cplx<T>);
  dim3 grid((unsigned)((npairs + PB - 1) / PB));
  auto* spec = reinterpret_cast<cplx<T>*>(P->spec);
  auto* tw = reinterpret_cast<const cplx<T>*>(P->tw[2]);
  switch (L.kb) {
    case 1: hipLaunchKernelGGL((k_rows_fwd<T, 1, Pre>), grid, dim3(L.threads), lds, P->stream, pre, spec, nrows, P->n[2], P->nh, P->Nh, F, PB, P->pl[2], tw); break;
    case 2: hipLaunchKernelGGL((k_rows_fwd<T, 2, Pre>), grid, dim3(L.threads), lds, P->stream, pre, spec, nrows, P->n[2], P->nh, P->Nh, F, PB, P->pl[2], tw); break;
    default: hipLaunchKernelGGL((k_rows_fwd<T, 4, Pre>), grid, dim3(L.threads), lds, P->stream, pre, spec, nrows, P->n[2], P->nh, P->Nh, F, PB, P->pl[2], tw); break;
  }
  LAUNCHCHK();
  return 0;
}

template <typename T, class Post>
int launch_rows_inv(ofdft_plan_s* P, const Post& post, int F, const cplx<T>* src, bool reduce) {
  const long long nrows = (long long)P->n[0] * P->n[1];
  const long long npairs = (nrows + 1) / 2;
  const int PB = kRowPairs;
  FFTLaunch L = fft_launch(P->pl[2], PB * F, 256);
  const size_t lds = (size_t)PB * F * P->n[2] * sizeof(cplx<T>);
  const long long nblk = (npairs + PB - 1) / PB;
  if (reduce && nblk * Post::NS > P->npartials) return fail(OFDFT_ESTATE, "partials buffer too small");
  dim3 grid((unsigned)nblk);
  auto* tw = reinterpret_cast<const cplx<T>*>(P->tw[2]);
  double* part = reduce ? P->partials : nullptr;
  switch (L.kb) {
    case 1: hipLaunchKernelGGL((k_rows_inv<T, 1, Post>), grid, dim3(L.threads), lds, P->stream, post, src, nrows, P->n[2], P->nh, P->Nh, F, PB, P->pl[2], tw, part); break;
    case 2: hipLaunchKernelGGL((k_rows_inv<T, 2, Post>), grid, dim3(L.threads), lds, P->stream, post, src, nrows, P->n[2], P->nh, P->Nh, F, PB, P->pl[2], tw, part); break;
    default: hipLaunchKernelGGL((k_rows_inv<T, 4, Post>), grid, dim3(L.threads), lds, P->stream, post, src, nrows, P->n[2], P->nh, P->Nh, F, PB, P->pl[2], tw, part); break;
  }
  LAUNCHCHK();
  if (reduce) {
    hipLaunchKernelGGL((k_reduce_partials<Post::NS>), dim3(Post::NS), dim3(kRedThreads), 0, P->stream, P->partials, nblk, P->red);
    LAUNCHCHK();
  }
  return 0;
}

template <typename T>
int launch_axis(ofdft_plan_s* P, cplx<T>* spec, int axis, int F, bool inv) {
  const int W = kColTile;
  const int n = P->n[axis];
  const Pass1D& pl = P->pl[axis];
  FFTLaunch L = fft_launch(pl, W);
  const size_t lds = (size_t)W * (n + 1) * sizeof(cplx<T>);
  long long stride, ostride;
  unsigned outer;
  if (axis == 1) {
    stride = P->nh;
    ostride = (long long)P->n[1] * P->nh;
    outer = P->n[0];
  } else {
    stride = (long long)P->n[1] * P->nh;
    ostride = P->nh;
    outer = P->n[1];
  }
  dim3 grid((unsigned)((P->nh + W - 1) / W), outer, (unsigned)F);
  auto* tw = reinterpret_cast<const cplx<T>*>(P->tw[axis]);
  switch (L.kb) {
    case 1: hipLaunchKernelGGL((k_axis_c2c<T, 1>), grid, dim3(L.threads), lds, P->stream, spec, P->Nh, n, stride, ostride, P->nh, W, pl, tw, inv ? 1 : 0); break;
    case 2: hipLaunchKernelGGL((k_axis_c2c<T, 2>), grid, dim3(L.threads), lds, P->stream, spec, P->Nh, n, stride, ostride, P->nh, W, pl, tw, inv ? 1 : 0); break;
    default: hipLaunchKernelGGL((k_axis_c2c<T, 4>), grid, dim3(L.threads), lds, P->stream, spec, P->Nh, n, stride, ostride, P->nh, W, pl, tw, inv ? 1 : 0); break;
  }
  LAUNCHCHK();
  return 0;
}

template <typename T>
int launch_pencil(ofdft_plan_s* P, int F, const int* kinds) {
  const int W = kColTile;
  FFTLaunch L = fft_launch(P->pl[0], W);
  const size_t lds = (size_t)W * (P->n[0] + 1) * sizeof(cplx<T>);
  dim3 grid((unsigned)((P->nh + W - 1) / W), (unsigned)P->n[1], (unsigned)F);
  ConvKinds ck;
  for (int f = 0; f < 3; ++f) ck.kind[f] = f < F ? kinds[f] : 0;
  ck.scale = 1.0 / (double)P->N;
  for (int f = 0; f < F; ++f)
    if (kinds[f] == OFDFT_CONV_WT && !P->have_wtk) return fail(OFDFT_ESTATE, "WT kernel table not built");
  auto* spec = reinterpret_cast<cplx<T>*>(P->spec);
  auto* tw = reinterpret_cast<const cplx<T>*>(P->tw[0]);
  const T* wtk = reinterpret_cast<const T*>(P->wtk);
  GridGeom g = geom(P);
  switch (L.kb) {
    case 1: hipLaunchKernelGGL((k_pencil_conv<T, 1>), grid, dim3(L.threads), lds, P->stream, spec, P->Nh, g, P->pl[0], tw, W, ck, wtk); break;
    case 2: hipLaunchKernelGGL((k_pencil_conv<T, 2>), grid, dim3(L.threads), lds, P->stream, spec, P->Nh, g, P->pl[0], tw, W, ck, wtk); break;
    default: hipLaunchKernelGGL((k_pencil_conv<T, 4>), grid, dim3(L.threads), lds, P->stream, spec, P->Nh, g, P->pl[0], tw, W, ck, wtk); break;
  }
  LAUNCHCHK();
  return 0;
}

int fetch_sums(ofdft_plan_s* P, int ns, double* out) {
  HIPCHK(hipMemcpyAsync(P->hred, P->red, sizeof(double) * ns, hipMemcpyDeviceToHost, P->stream));
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

template <typename T>
int eval_impl(ofdft_plan_s* P, const T* rho, const T* phi, const T* p, double c, double s, T* V, T* aux,
              double* sums) {
  if (!P->have_fn) return fail(OFDFT_ESTATE, "functional not set");
  if (P->fn.pseudo && !P->have_vloc) return fail(OFDFT_ESTATE, "V_loc not set");
  FnParams fp = make_fn(P->fn);
  PreEval<T> pre{rho, phi, p, T(c), T(s), fp};
  const int F = fp.F;
  auto* spec = reinterpret_cast<cplx<T>*>(P->spec);
  int rc;
  if (F > 0) {
    if ((rc = launch_rows_fwd<T>(P, pre, F))) return rc;
    if ((rc = launch_axis<T>(P, spec, 1, F, false))) return rc;
    int kinds[3];
    if (fp.fH >= 0) kinds[fp.fH] = OFDFT_CONV_HARTREE;
    if (fp.fV >= 0) kinds[fp.fV] = OFDFT_CONV_GG;
    if (fp.fW >= 0) kinds[fp.fW] = OFDFT_CONV_WT;
    if ((rc = launch_pencil<T>(P, F, kinds))) return rc;
    if ((rc = launch_axis<T>(P, spec, 1, F, true))) return rc;
  }
  PostEval<T> post{pre, reinterpret_cast<const T*>(P->vloc), V, aux, P->N};
  if ((rc = launch_rows_inv<T>(P, post, F > 0 ? F : 1, spec, true))) return rc;
  if (sums) {
    double tmp[OFDFT_NSUM];
    if ((rc = fetch_sums(P, OFDFT_NSUM, tmp))) return rc;
    finalize_energies(P, tmp);
    for (int i = 0; i < OFDFT_NSUM; ++i) sums[i] = tmp[i];
  }
  return 0;
}

template <typename T>
int hvp_impl(ofdft_plan_s* P, const T* phi, const T* p, const T* V, const T* aux, double mu, T* Ap, double* pAp) {
  if (!P->have_fn) return fail(OFDFT_ESTATE, "functional not set");
  FnParams fp = make_fn(P->fn);
  const int F = fp.F;
  auto* spec = reinterpret_cast<cplx<T>*>(P->spec);
  int rc;
  if (F > 0) {
    PreHvp<T> pre{phi, p, fp};
    if ((rc = launch_rows_fwd<T>(P, pre, F))) return rc;
    if ((rc = launch_axis<T>(P, spec, 1, F, false))) return rc;
    int kinds[3];
    if (fp.fH >= 0) kinds[fp.fH] = OFDFT_CONV_HARTREE;
    if (fp.fV >= 0) kinds[fp.fV] = OFDFT_CONV_GG;
    if (fp.fW >= 0) kinds[fp.fW] = OFDFT_CONV_WT;
    if ((rc = launch_pencil<T>(P, F, kinds))) return rc;
    if ((rc = launch_axis<T>(P, spec, 1, F, true))) return rc;
  }
  PostHvp<T> post{phi, p, V, aux, Ap, T(mu), P->N, fp};
  if ((rc = launch_rows_inv<T>(P, post, F > 0 ? F : 1, spec, true))) return rc;
  if (pAp) {
    double t;
    if ((rc = fetch_sums(P, 1, &t))) return rc;
    *pAp = t;
  }
  return 0;
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
  HIPCHK(hipMemcpyAsync(out, spec, sizeof(cplx<T>) * P->Nh, hipMemcpyDeviceToDevice, P->stream));
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
  HIPCHK(hipMemcpyAsync(spec, in, sizeof(cplx<T>) * P->Nh, hipMemcpyDeviceToDevice, P->stream));
  hipLaunchKernelGGL((k_scale_cplx<T>), dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream, spec, P->Nh,
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
    case 1: hipLaunchKernelGGL((k_reduce_partials<1>), dim3(1), dim3(kRedThreads), 0, P->stream, P->partials, (long long)nblk, P->red); break;
    default: hipLaunchKernelGGL((k_reduce_partials<3>), dim3(3), dim3(kRedThreads), 0, P->stream, P->partials, (long long)nblk, P->red); break;
  }
  LAUNCHCHK();
  if ((rc = fetch_sums(P, ns, out))) return rc;
  return 0;
}

bool valid(ofdft_plan P) { return P != nullptr; }

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* ofdft_last_error(void) { return g_err.c_str(); }
int ofdft_version(void) { return 1; }

int ofdft_plan_create(ofdft_plan* out, int n0, int n1, int n2, const double* lattice9, int precision,
                      int full, void* hip_stream) {
  if (!out || !lattice9) return fail(OFDFT_EINVAL, "null argument");
  *out = nullptr;
  if (n0 < 2 || n1 < 2 || n2 < 2 || n0 > 4096 || n1 > 4096 || n2 > 4096)
    return fail(OFDFT_EINVAL, "grid dimensions must be in [2, 4096]");
  if (precision != OFDFT_F64 && precision != OFDFT_F32) return fail(OFDFT_EINVAL, "unknown precision");
  auto* P = new (std::nothrow) ofdft_plan_s();
  if (!P) return fail(OFDFT_ENOMEM, "host allocation failed");
  P->n[0] = n0;
  P->n[1] = n1;
  P->n[2] = n2;
  P->nh = n2 / 2 + 1;
  P->N = (long long)n0 * n1 * n2;
  P->Nh = (long long)n0 * n1 * P->nh;
  P->prec = precision;
  P->full = full ? 1 : 0;
  P->rs = precision == OFDFT_F64 ? sizeof(double) : sizeof(float);
  P->stream = reinterpret_cast<hipStream_t>(hip_stream);
  for (int i = 0; i < 3; ++i) {
    if (!make_pass(P->n[i], &P->pl[i])) {
      delete P;
      return fail(OFDFT_EINVAL, "grid sizes must be products of 2, 3 and 5 (at most 12 passes)");
    }
  }
  // LDS limits (64 KiB per workgroup)
  const size_t cs = 2 * P->rs;
  if ((size_t)kRowPairs * 3 * n2 * cs > 65536 || (size_t)kColTile * (n0 + 1) * cs > 65536 ||
      (size_t)kColTile * (n1 + 1) * cs > 65536) {
    delete P;
    return fail(OFDFT_EINVAL, "grid dimension too large for the LDS tile sizes");
  }
  std::memcpy(P->lat, lattice9, sizeof(double) * 9);
  // bg = 2 pi inv(L)^T (grid.py:409-410): rows of bg are reciprocal vectors
  const double* a = P->lat;
  const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                     a[2] * (a[3] * a[7] - a[4] * a[6]);
  if (std::fabs(det) < 1e-300) {
    delete P;
    return fail(OFDFT_EINVAL, "singular lattice");
  }
  double inv[9];
  inv[0] = (a[4] * a[8] - a[5] * a[7]) / det;
  inv[1] = (a[2] * a[7] - a[1] * a[8]) / det;
  inv[2] = (a[1] * a[5] - a[2] * a[4]) / det;
  inv[3] = (a[5] * a[6] - a[3] * a[8]) / det;
  inv[4] = (a[0] * a[8] - a[2] * a[6]) / det;
  inv[5] = (a[2] * a[3] - a[0] * a[5]) / det;
  inv[6] = (a[3] * a[7] - a[4] * a[6]) / det;
  inv[7] = (a[1] * a[6] - a[0] * a[7]) / det;
  inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) P->bg[i * 3 + j] = 2.0 * kPi * inv[j * 3 + i];
  P->volume = std::fabs(det);
  P->dV = P->volume / (double)P->N;

  auto cleanup = [&](int code, const std::string& msg) {
    ofdft_plan_destroy(P);
    return fail(code, msg);
  };
  for (int ax = 0; ax < 3; ++ax) {
    const int n = P->n[ax];
    std::vector<double> w(2 * n);
    for (int k = 0; k < n; ++k) {
      w[2 * k] = std::cos(2.0 * kPi * k / n);
      w[2 * k + 1] = -std::sin(2.0 * kPi * k / n);
    }
    if (hipMalloc(&P->tw[ax], cs * n) != hipSuccess) return cleanup(OFDFT_ENOMEM, "twiddle allocation");
    if (precision == OFDFT_F64) {
      if (hipMemcpy(P->tw[ax], w.data(), cs * n, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(OFDFT_EHIP, "twiddle upload");
    } else {
      std::vector<float> wf(2 * n);
      for (int i = 0; i < 2 * n; ++i) wf[i] = (float)w[i];
      if (hipMemcpy(P->tw[ax], wf.data(), cs * n, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(OFDFT_EHIP, "twiddle upload");
    }
  }
  if (hipMalloc(&P->spec, cs * 3 * P->Nh) != hipSuccess) return cleanup(OFDFT_ENOMEM, "spectrum workspace");
  const long long npairs = ((long long)n0 * n1 + 1) / 2;
  const long long nblk = (npairs + kRowPairs - 1) / kRowPairs;
  P->npartials = std::max<long long>(nblk * OFDFT_NSUM, 2048LL * 3);
  if (hipMalloc((void**)&P->partials, sizeof(double) * P->npartials) != hipSuccess)
    return cleanup(OFDFT_ENOMEM, "partials");
  if (hipMalloc((void**)&P->red, sizeof(double) * kMaxSum) != hipSuccess) return cleanup(OFDFT_ENOMEM, "red");
  if (hipHostMalloc((void**)&P->hred, sizeof(double) * kMaxSum, hipHostMallocDefault) != hipSuccess)
    return cleanup(OFDFT_ENOMEM, "pinned host buffer");
  *out = P;
  return 0;
}

int ofdft_plan_destroy(ofdft_plan P) {
  if (!P) return 0;
  for (int i = 0; i < 3; ++i)
    if (P->tw[i]) (void)hipFree(P->tw[i]);
  if (P->spec) (void)hipFree(P->spec);
  if (P->wtk) (void)hipFree(P->wtk);
  if (P->vloc) (void)hipFree(P->vloc);
  if (P->partials) (void)hipFree(P->partials);
  if (P->red) (void)hipFree(P->red);
  if (P->hred) (void)hipHostFree(P->hred);
  delete P;
  return 0;
}

int ofdft_plan_set_stream(ofdft_plan P, void* s) {
  if (!valid(P)) return fail(OFDFT_EINVAL, "null plan");
  P->stream = reinterpret_cast<hipStream_t>(s);
  return 0;
}

long long ofdft_plan_workspace_bytes(ofdft_plan P) {
  if (!valid(P)) return -1;
  long long b = 2 * (long long)P->rs * (3 * P->Nh + P->n[0] + P->n[1] + P->n[2]);
  if (P->wtk) b += (long long)P->rs * P->Nh;
  if (P->vloc) b += (long long)P->rs * P->N;
  b += sizeof(double) * (P->npartials + kMaxSum);
  return b;
}

int ofdft_set_functional(ofdft_plan P, const ofdft_functional* f) {
  if (!valid(P) || !f) return fail(OFDFT_EINVAL, "null argument");
  if (f->wt && std::fabs(f->alpha - f->beta) > 1e-9)
    return fail(OFDFT_EINVAL, "the WT path requires alpha == beta");
  if (f->wt && !(f->rho0 > 0.0)) return fail(OFDFT_EINVAL, "WT kernel needs rho0 > 0");
  P->fn = *f;
  P->have_fn = 1;
  if (f->wt) {
    if (!P->wtk && hipMalloc(&P->wtk, P->rs * P->Nh) != hipSuccess) return fail(OFDFT_ENOMEM, "WT table");
    const double rho0 = f->rho0;
    const double tkf = 2.0 * std::pow(3.0 * rho0 * kPi * kPi, 1.0 / 3.0);
    const double ctf = 0.3 * std::pow(3.0 * kPi * kPi, 2.0 / 3.0);
    double factor = 5.0 / (9.0 * f->alpha * f->beta * std::pow(rho0, f->alpha + f->beta - 5.0 / 3.0));
    factor *= ctf;
    GridGeom g = geom(P);
    if (P->prec == OFDFT_F64)
      hipLaunchKernelGGL((k_wt_table<double>), dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream,
                         reinterpret_cast<double*>(P->wtk), g, tkf, factor, f->x);
    else
      hipLaunchKernelGGL((k_wt_table<float>), dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream,
                         reinterpret_cast<float*>(P->wtk), g, tkf, factor, f->x);
    LAUNCHCHK();
    P->have_wtk = 1;
  }
  return 0;
}

int ofdft_set_vloc(ofdft_plan P, const void* v) {
  if (!valid(P) || !v) return fail(OFDFT_EINVAL, "null argument");
  if (!P->vloc && hipMalloc(&P->vloc, P->rs * P->N) != hipSuccess) return fail(OFDFT_ENOMEM, "V_loc");
  HIPCHK(hipMemcpyAsync(P->vloc, v, P->rs * P->N, hipMemcpyDeviceToDevice, P->stream));
  P->have_vloc = 1;
  return 0;
}

int ofdft_eval(ofdft_plan P, const void* rho, void* V, double* sums) {
  if (!valid(P) || !rho) return fail(OFDFT_EINVAL, "null argument");
  if (P->prec == OFDFT_F64)
    return eval_impl<double>(P, (const double*)rho, nullptr, nullptr, 1.0, 0.0, (double*)V, nullptr, sums);
  return eval_impl<float>(P, (const float*)rho, nullptr, nullptr, 1.0, 0.0, (float*)V, nullptr, sums);
}

int ofdft_eval_phi(ofdft_plan P, const void* phi, const void* p, double c, double s, void* V, void* aux,
                   double* sums) {
  if (!valid(P) || !phi) return fail(OFDFT_EINVAL, "null argument");
  if (P->prec == OFDFT_F64)
    return eval_impl<double>(P, nullptr, (const double*)phi, (const double*)p, c, s, (double*)V, (double*)aux, sums);
  return eval_impl<float>(P, nullptr, (const float*)phi, (const float*)p, c, s, (float*)V, (float*)aux, sums);
}

int ofdft_hvp(ofdft_plan P, const void* phi, const void* p, const void* V, const void* aux, double mu, void* Ap,
              double* pAp) {
  if (!valid(P) || !phi || !p || !V || !Ap) return fail(OFDFT_EINVAL, "null argument");
  if ((P->fn.vw || P->fn.wt) && !aux) return fail(OFDFT_EINVAL, "aux fields required for vW/WT");
  if (P->prec == OFDFT_F64)
    return hvp_impl<double>(P, (const double*)phi, (const double*)p, (const double*)V, (const double*)aux, mu,
                            (double*)Ap, pAp);
  return hvp_impl<float>(P, (const float*)phi, (const float*)p, (const float*)V, (const float*)aux, mu, (float*)Ap,
                         pAp);
}

int ofdft_fft_r2c(ofdft_plan P, const void* in, void* out) {
  if (!valid(P) || !in || !out) return fail(OFDFT_EINVAL, "null argument");
  if (P->prec == OFDFT_F64) return r2c_impl<double>(P, (const double*)in, (cplx<double>*)out);
  return r2c_impl<float>(P, (const float*)in, (cplx<float>*)out);
}

int ofdft_fft_c2r(ofdft_plan P, const void* in, void* out) {
  if (!valid(P) || !in || !out) return fail(OFDFT_EINVAL, "null argument");
  if (P->prec == OFDFT_F64) return c2r_impl<double>(P, (const cplx<double>*)in, (double*)out);
  return c2r_impl<float>(P, (const cplx<float>*)in, (float*)out);
}

int ofdft_convolve(ofdft_plan P, int kind, const void* in, void* out) {
  if (!valid(P) || !in || !out) return fail(OFDFT_EINVAL, "null argument");
  if (kind < OFDFT_CONV_NONE || kind > OFDFT_CONV_WT) return fail(OFDFT_EINVAL, "unknown convolution kind");
  if (P->prec == OFDFT_F64) return conv_impl<double>(P, kind, (const double*)in, (double*)out);
  return conv_impl<float>(P, kind, (const float*)in, (float*)out);
}

int ofdft_dot3(ofdft_plan P, const void* a, const void* b, double* out3) {
  if (!valid(P) || !a || !b || !out3) return fail(OFDFT_EINVAL, "null argument");
  const int nb = vec_blocks(P->N);
  if (P->prec == OFDFT_F64)
    hipLaunchKernelGGL((k_dot3<double>), dim3(nb), dim3(256), 0, P->stream, (const double*)a, (const double*)b, P->N, P->partials);
  else
    hipLaunchKernelGGL((k_dot3<float>), dim3(nb), dim3(256), 0, P->stream, (const float*)a, (const float*)b, P->N, P->partials);
  LAUNCHCHK();
  return reduce_blocks(P, nb, 3, out3);
}

int ofdft_axpby(ofdft_plan P, double alpha, const void* x, double beta, void* y) {
  return ofdft_lincomb(P, alpha, x, beta, y, y);
}

int ofdft_lincomb(ofdft_plan P, double alpha, const void* x, double beta, const void* y, void* z) {
  if (!valid(P) || !x || !y || !z) return fail(OFDFT_EINVAL, "null argument");
  const int nb = vec_blocks(P->N);
  if (P->prec == OFDFT_F64)
    hipLaunchKernelGGL((k_lincomb<double>), dim3(nb), dim3(256), 0, P->stream, alpha, (const double*)x, beta, (const double*)y, (double*)z, P->N);
  else
    hipLaunchKernelGGL((k_lincomb<float>), dim3(nb), dim3(256), 0, P->stream, (float)alpha, (const float*)x, (float)beta, (const float*)y, (float*)z, P->N);
  LAUNCHCHK();
  return 0;
}

int ofdft_residual(ofdft_plan P, const void* V, const void* phi, double mu, void* res, double* res2) {
  if (!valid(P) || !V || !phi || !res) return fail(OFDFT_EINVAL, "null argument");
  const int nb = vec_blocks(P->N);
  if (P->prec == OFDFT_F64)
    hipLaunchKernelGGL((k_residual<double>), dim3(nb), dim3(256), 0, P->stream, (const double*)V, (const double*)phi, mu, (double*)res, P->N, P->partials);
  else
    hipLaunchKernelGGL((k_residual<float>), dim3(nb), dim3(256), 0, P->stream, (const float*)V, (const float*)phi, (float)mu, (float*)res, P->N, P->partials);
  LAUNCHCHK();
  if (!res2) return 0;
  return reduce_blocks(P, nb, 1, res2);
}

int ofdft_cg_update(ofdft_plan P, double a, const void* p, const void* Ap, void* d, void* r, double* r2) {
  if (!valid(P) || !p || !Ap || !d || !r) return fail(OFDFT_EINVAL, "null argument");
  const int nb = vec_blocks(P->N);
  if (P->prec == OFDFT_F64)
    hipLaunchKernelGGL((k_cg_update<double>), dim3(nb), dim3(256), 0, P->stream, a, (const double*)p, (const double*)Ap, (double*)d, (double*)r, P->N, P->partials);
  else
    hipLaunchKernelGGL((k_cg_update<float>), dim3(nb), dim3(256), 0, P->stream, (float)a, (const float*)p, (const float*)Ap, (float*)d, (float*)r, P->N, P->partials);
  LAUNCHCHK();
  if (!r2) return 0;
  return reduce_blocks(P, nb, 1, r2);
}

int ofdft_dot_vphip(ofdft_plan P, const void* V, const void* phi, const void* p, double* out) {
  if (!valid(P) || !V || !phi || !p || !out) return fail(OFDFT_EINVAL, "null argument");
  const int nb = vec_blocks(P->N);
  if (P->prec == OFDFT_F64)
    hipLaunchKernelGGL((k_dot_vphip<double>), dim3(nb), dim3(256), 0, P->stream, (const double*)V, (const double*)phi, (const double*)p, P->N, P->partials);
  else
    hipLaunchKernelGGL((k_dot_vphip<float>), dim3(nb), dim3(256), 0, P->stream, (const float*)V, (const float*)phi, (const float*)p, P->N, P->partials);
  LAUNCHCHK();
  return reduce_blocks(P, nb, 1, out);
}

int ofdft_convert(ofdft_plan P, void* dst, int src_precision, const void* src, long long n) {
  if (!valid(P) || !dst || !src || n < 0) return fail(OFDFT_EINVAL, "bad argument");
  const int nb = vec_blocks(n);
  if (P->prec == OFDFT_F64 && src_precision == OFDFT_F32)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(nb), dim3(256), 0, P->stream, (double*)dst, (const float*)src, n);
  else if (P->prec == OFDFT_F32 && src_precision == OFDFT_F64)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(nb), dim3(256), 0, P->stream, (float*)dst, (const double*)src, n);
  else if (P->prec == OFDFT_F64)
    hipLaunchKernelGGL((k_convert<double, double>), dim3(nb), dim3(256), 0, P->stream, (double*)dst, (const double*)src, n);
  else
    hipLaunchKernelGGL((k_convert<float, float>), dim3(nb), dim3(256), 0, P->stream, (float*)dst, (const float*)src, n);
  LAUNCHCHK();
  return 0;
}

}  // extern "C"
