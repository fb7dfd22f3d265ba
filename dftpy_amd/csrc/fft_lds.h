// LDS-resident mixed-radix Stockham FFT building blocks for gfx950.
//
// A workgroup holds `ntr` complex sequences of length n in LDS (sequence b
// at lds[b*ld .. b*ld+n)).  Each pass of radix R maps one butterfly to one
// thread slot: every slot loads its R inputs into registers, the group
// synchronises, then every slot writes its R outputs (in-place Stockham
// autosort).  Twiddles come from a per-length table w[k] = exp(-2 pi i k/n)
// computed on the host in FP64 and stored in the transform precision.
//
// Forward = numpy convention exp(-2 pi i kn/N), unnormalised; inverse uses
// the conjugate table and is also unnormalised.
#pragma once
#include <hip/hip_runtime.h>

#include "roots.h"

namespace ofdft {

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by -i (forward) or +i (inverse)
template <typename T>
__device__ __forceinline__ cplx<T> rotq(cplx<T> a, bool inv) {
  return inv ? cplx<T>{-a.y, a.x} : cplx<T>{a.y, -a.x};
}

// Per-length pass schedule, passed by value to kernels.
constexpr int kMaxPass = 12;
struct Pass1D {
  int n;
  int npass;
  int radix[kMaxPass];
};

// ---------------------------------------------------------------------------
// small DFT kernels on registers (y_q = sum_r v_r w^{rq}, w = exp(-+2 pi i/R))
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void dft2(cplx<T>* v) {
  cplx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename T>
__device__ __forceinline__ void dft3(cplx<T>* v, bool inv) {
  const T h = T(0.86602540378443864676);  // sqrt(3)/2
  cplx<T> s = cadd(v[1], v[2]);
  cplx<T> d = csub(v[1], v[2]);
  cplx<T> m = {v[0].x - T(0.5) * s.x, v[0].y - T(0.5) * s.y};
  cplx<T> r = rotq(cscale(d, h), inv);  // -+ i*h*d
  v[0] = cadd(v[0], s);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <typename T>
__device__ __forceinline__ void dft4(cplx<T>* v, bool inv) {
  cplx<T> t0 = cadd(v[0], v[2]);
  cplx<T> t1 = csub(v[0], v[2]);
  cplx<T> t2 = cadd(v[1], v[3]);
  cplx<T> t3 = rotq(csub(v[1], v[3]), inv);
  v[0] = cadd(t0, t2);
  v[1] = cadd(t1, t3);
  v[2] = csub(t0, t2);
  v[3] = csub(t1, t3);
}

template <typename T>
__device__ __forceinline__ void dft5(cplx<T>* v, bool inv) {
  const T c1 = T(0.30901699437494742410);   // cos(2 pi/5)
  const T c2 = T(-0.80901699437494742410);  // cos(4 pi/5)
  const T s1 = T(0.95105651629515357212);   // sin(2 pi/5)
  const T s2 = T(0.58778525229247312917);   // sin(4 pi/5)
  cplx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  cplx<T> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  cplx<T> a0 = v[0];
  cplx<T> b1 = {a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y};
  cplx<T> b2 = {a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y};
  cplx<T> e1 = {s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y};
  cplx<T> e2 = {s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y};
  cplx<T> r1 = rotq(e1, inv);  // -+ i e1
  cplx<T> r2 = rotq(e2, inv);
  v[0] = {a0.x + t1.x + t2.x, a0.y + t1.y + t2.y};
  v[1] = cadd(b1, r1);
  v[4] = csub(b1, r1);
  v[2] = cadd(b2, r2);
  v[3] = csub(b2, r2);
}

// 8-point DFT as two 4-point DFTs (even / odd inputs) and a W8 twiddle stage.
template <typename T>
__device__ __forceinline__ void dft8(cplx<T>* v, bool inv) {
  const T h = T(0.70710678118654752440);  // 1/sqrt(2)
  cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4(e, inv);
  dft4(o, inv);
  // W8^1 = (1 -+ i)/sqrt2, W8^2 = -+ i, W8^3 = (-1 -+ i)/sqrt2
  const cplx<T> o1 = inv ? cplx<T>{h * (o[1].x - o[1].y), h * (o[1].x + o[1].y)}
                         : cplx<T>{h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)};
  const cplx<T> o2 = rotq(o[2], inv);
  const cplx<T> o3 = inv ? cplx<T>{-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y)}
                         : cplx<T>{h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y)};
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

// Odd prime R (7, 11, 13): inputs paired as v_j +- v_(R-j), then
// y_k = a_k -+ i b_k and y_(R-k) = a_k +- i b_k with a_k = v_0 + sum_j
// cos(2 pi jk/R)(v_j + v_(R-j)), b_k = sum_j sin(2 pi jk/R)(v_j - v_(R-j)).
template <typename T, int R>
__device__ __forceinline__ void dftp(cplx<T>* v, bool inv) {
  constexpr int H = (R - 1) / 2;
  cplx<T> t[H], d[H];
  cplx<T> y0 = v[0];
#pragma unroll
  for (int j = 1; j <= H; ++j) {
    t[j - 1] = cadd(v[j], v[R - j]);
    d[j - 1] = csub(v[j], v[R - j]);
    y0 = cadd(y0, t[j - 1]);
  }
  cplx<T> y[R];
  y[0] = y0;
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    cplx<T> a = v[0], b = {T(0), T(0)};
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      const T c = T(root_re<R>((j * k) % R));
      const T sn = T(-root_im<R>((j * k) % R));
      a = {a.x + c * t[j - 1].x, a.y + c * t[j - 1].y};
      b = {b.x + sn * d[j - 1].x, b.y + sn * d[j - 1].y};
    }
    const cplx<T> rb = rotq(b, inv);  // -+ i b
    y[k] = cadd(a, rb);
    y[R - k] = csub(a, rb);
  }
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = y[k];
}

template <typename T, int R>
__device__ __forceinline__ void bfly_load(const cplx<T>* __restrict__ x, int j, int nb, int k, int twstep,
                                          const cplx<T>* __restrict__ tw, bool inv, cplx<T>* v) {
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = x[j + r * nb];
  if (k != 0) {
#pragma unroll
    for (int r = 1; r < R; ++r) {
      cplx<T> w = tw[k * r * twstep];
      if (inv) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
  }
  if constexpr (R == 2) dft2(v);
  if constexpr (R == 3) dft3(v, inv);
  if constexpr (R == 4) dft4(v, inv);
  if constexpr (R == 5) dft5(v, inv);
  if constexpr (R == 7 || R == 11 || R == 13) dftp<T, R>(v, inv);
  if constexpr (R == 8) dft8(v, inv);
}

template <typename T, int R>
__device__ __forceinline__ void bfly_store(cplx<T>* __restrict__ x, int obase, int Ns, const cplx<T>* v) {
#pragma unroll
  for (int r = 0; r < R; ++r) x[obase + r * Ns] = v[r];
}

// In-place Stockham FFT of `ntr` LDS sequences.  Each thread handles up to KB
// butterfly slots per pass; the caller sizes blockDim.x so that
// KB * blockDim.x >= ntr * max_pass(n / R).  Must be called by the whole
// workgroup (contains barriers).
// One pass of radix 11 or 13 (register arrays of 13 points per slot, kept
// out of the common passes so that they do not size every runtime kernel).
template <typename T, int KB>
__device__ void lds_pass_large(cplx<T>* lds, int ld, int ntr, int n, int R, int Ns, const cplx<T>* __restrict__ tw,
                               bool inv) {
  const int nb = n / R;
  const int total = ntr * nb;
  const int twstep = n / (Ns * R);
  cplx<T> v[KB][13];
  int ob[KB];
#pragma unroll
  for (int kk = 0; kk < KB; ++kk) {
    const int t = threadIdx.x + kk * blockDim.x;
    ob[kk] = -1;
    if (t < total) {
      const int b = t / nb;
      const int j = t - b * nb;
      const int k = j % Ns;
      ob[kk] = b * ld + (j - k) * R + k;
      if (R == 11) bfly_load<T, 11>(lds + b * ld, j, nb, k, twstep, tw, inv, v[kk]);
      else bfly_load<T, 13>(lds + b * ld, j, nb, k, twstep, tw, inv, v[kk]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < KB; ++kk) {
    if (ob[kk] >= 0) {
      if (R == 11) bfly_store<T, 11>(lds, ob[kk], Ns, v[kk]);
      else bfly_store<T, 13>(lds, ob[kk], Ns, v[kk]);
    }
  }
  __syncthreads();
}

// LG: the schedule may hold radix 11 or 13 passes.
template <typename T, int KB, bool LG = false>
__device__ void lds_fft(cplx<T>* lds, int ld, int ntr, const Pass1D& pl, const cplx<T>* __restrict__ tw,
                        bool inv) {
  const int n = pl.n;
  int Ns = 1;
  for (int ip = 0; ip < pl.npass; ++ip) {
    const int R = pl.radix[ip];
    if constexpr (LG) {
      if (R > 8) {
        lds_pass_large<T, KB>(lds, ld, ntr, n, R, Ns, tw, inv);
        Ns *= R;
        continue;
      }
    }
    const int nb = n / R;
    const int total = ntr * nb;
    const int twstep = n / (Ns * R);
    cplx<T> v[KB][8];
    int ob[KB];
#pragma unroll
    for (int kk = 0; kk < KB; ++kk) {
      const int t = threadIdx.x + kk * blockDim.x;
      ob[kk] = -1;
      if (t < total) {
        const int b = t / nb;
        const int j = t - b * nb;
        const int k = j % Ns;
        const cplx<T>* xs = lds + b * ld;
        ob[kk] = b * ld + (j - k) * R + k;
        switch (R) {
          case 2: bfly_load<T, 2>(xs, j, nb, k, twstep, tw, inv, v[kk]); break;
          case 3: bfly_load<T, 3>(xs, j, nb, k, twstep, tw, inv, v[kk]); break;
          case 4: bfly_load<T, 4>(xs, j, nb, k, twstep, tw, inv, v[kk]); break;
          case 5: bfly_load<T, 5>(xs, j, nb, k, twstep, tw, inv, v[kk]); break;
          case 7: bfly_load<T, 7>(xs, j, nb, k, twstep, tw, inv, v[kk]); break;
          default: bfly_load<T, 8>(xs, j, nb, k, twstep, tw, inv, v[kk]); break;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KB; ++kk) {
      if (ob[kk] >= 0) {
        switch (R) {
          case 2: bfly_store<T, 2>(lds, ob[kk], Ns, v[kk]); break;
          case 3: bfly_store<T, 3>(lds, ob[kk], Ns, v[kk]); break;
          case 4: bfly_store<T, 4>(lds, ob[kk], Ns, v[kk]); break;
          case 5: bfly_store<T, 5>(lds, ob[kk], Ns, v[kk]); break;
          case 7: bfly_store<T, 7>(lds, ob[kk], Ns, v[kk]); break;
          default: bfly_store<T, 8>(lds, ob[kk], Ns, v[kk]); break;
        }
      }
    }
    __syncthreads();
    Ns *= R;
  }
}

// ---------------------------------------------------------------------------
// Compile-time schedules for the common lengths: radices, butterfly counts,
// strides and twiddle steps become constants and every pass is unrolled.
// ---------------------------------------------------------------------------
struct Sched {
  int np;
  int r[kMaxPass];
};

// Radix counts as 8s, then one 4 or 2, then 3s and 5s.  `asc`: the passes
// run in ascending radix order (the first pass stores at a stride of R
// points, and a small R there keeps its LDS bank conflicts low); otherwise
// 8s, 4, 2, 3s, 5s.  FP64 keeps the second order: its rounding, and so the
// forward-difference HVP trajectories built on it, stay as they are.
constexpr Sched make_sched(int n, bool asc) {
  Sched s{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int m = n, c8 = 0, c4 = 0, c2 = 0, c3 = 0, c5 = 0;
  while (m % 8 == 0) { ++c8; m /= 8; }
  while (m % 4 == 0) { ++c4; m /= 4; }
  while (m % 2 == 0) { ++c2; m /= 2; }
  while (m % 3 == 0) { ++c3; m /= 3; }
  while (m % 5 == 0) { ++c5; m /= 5; }
  if (asc) {
    for (int i = 0; i < c3; ++i) s.r[s.np++] = 3;
    for (int i = 0; i < c5; ++i) s.r[s.np++] = 5;
    for (int i = 0; i < c2; ++i) s.r[s.np++] = 2;
    for (int i = 0; i < c4; ++i) s.r[s.np++] = 4;
    for (int i = 0; i < c8; ++i) s.r[s.np++] = 8;
  } else {
    for (int i = 0; i < c8; ++i) s.r[s.np++] = 8;
    for (int i = 0; i < c4; ++i) s.r[s.np++] = 4;
    for (int i = 0; i < c2; ++i) s.r[s.np++] = 2;
    for (int i = 0; i < c3; ++i) s.r[s.np++] = 3;
    for (int i = 0; i < c5; ++i) s.r[s.np++] = 5;
  }
  return s;
}

constexpr int sched_maxnb(int n) {
  Sched s = make_sched(n, false);
  int mx = 0;
  for (int i = 0; i < s.np; ++i) mx = (n / s.r[i] > mx) ? n / s.r[i] : mx;
  return mx;
}

template <typename T, int N, int P, int NS, int NTR, int TH>
__device__ __forceinline__ void fixed_pass(cplx<T>* lds, int ld, int ntr, const cplx<T>* __restrict__ tw,
                                           bool inv) {
  constexpr Sched S = make_sched(N, sizeof(T) == 4);
  if constexpr (P < S.np) {
    constexpr int R = S.r[P];
    constexpr int nb = N / R;
    constexpr int twstep = N / (NS * R);
    constexpr int KB = (NTR * nb + TH - 1) / TH;  // butterfly slots per thread in this pass
    const int total = ntr * nb;
    cplx<T> v[KB][R];
    int ob[KB];
#pragma unroll
    for (int kk = 0; kk < KB; ++kk) {
      const int t = threadIdx.x + kk * TH;
      ob[kk] = -1;
      if (t < total) {
        const int b = t / nb;
        const int j = t - b * nb;
        const int k = j % NS;
        ob[kk] = b * ld + (j - k) * R + k;
        bfly_load<T, R>(lds + b * ld, j, nb, k, twstep, tw, inv, v[kk]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KB; ++kk)
      if (ob[kk] >= 0) bfly_store<T, R>(lds, ob[kk], NS, v[kk]);
    __syncthreads();
    fixed_pass<T, N, P + 1, NS * R, NTR, TH>(lds, ld, ntr, tw, inv);
  }
}

// FFT policy with a compile-time length N for up to NTR sequences per
// workgroup of TH threads.
template <int N, int NTR, int TH_>
struct FFTFixed {
  static constexpr int TH = TH_;
  static constexpr bool kFixed = true;
  __host__ __device__ static constexpr int n() { return N; }
  __device__ static int th() { return TH; }
  template <typename T>
  __device__ void run(cplx<T>* lds, int ld, int ntr, const cplx<T>* __restrict__ tw, bool inv) const {
    fixed_pass<T, N, 0, 1, NTR, TH>(lds, ld, ntr, tw, inv);
  }
};

// FFT policy for any 2^a 3^b 5^c length, schedule read at run time.
template <int KB_, bool LG = false>
struct FFTRuntime {
  static constexpr int KB = KB_;
  static constexpr int TH = 1024;
  static constexpr bool kFixed = false;
  Pass1D pl;
  __host__ __device__ int n() const { return pl.n; }
  __device__ static int th() { return blockDim.x; }
  template <typename T>
  __device__ void run(cplx<T>* lds, int ld, int ntr, const cplx<T>* __restrict__ tw, bool inv) const {
    lds_fft<T, KB, LG>(lds, ld, ntr, pl, tw, inv);
  }
};

}  // namespace ofdft
