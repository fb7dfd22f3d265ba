// Register-resident DFTs and the four-step column transform.
//
// A length N = N1*N2 transform is split as n = N2*n1 + n2, k = k1 + N1*k2:
//   X[k1 + N1 k2] = sum_n2 W_N2^(n2 k2) W_N^(n2 k1) sum_n1 W_N1^(n1 k1) x[N2 n1 + n2]
// Thread n2 loads its N1 strided inputs straight into registers, does the
// N1-point DFT and the W_N^(n2 k1) twiddle, one LDS transpose hands thread k1
// its N2 values, and the N2-point DFT finishes in registers.  The inverse runs
// the same steps in reverse order with conjugate roots.
#pragma once
#include "fft_lds.h"
#include "roots.h"

namespace ofdft {

template <typename T, int M>
__device__ __forceinline__ void dftreg(cplx<T>* v, bool inv);

// Cooley-Tukey split M = A*B inside registers: A-point DFTs over the strided
// subsequences, constant twiddles W_M^(b ka), B-point DFTs.
template <typename T, int A, int B>
__device__ __forceinline__ void dft_ct(cplx<T>* v, bool inv) {
  constexpr int M = A * B;
  cplx<T> y[M];
#pragma unroll
  for (int b = 0; b < B; ++b) {
    cplx<T> t[A];
#pragma unroll
    for (int a = 0; a < A; ++a) t[a] = v[B * a + b];
    dftreg<T, A>(t, inv);
#pragma unroll
    for (int ka = 0; ka < A; ++ka) {
      const int j = (b * ka) % M;
      if (j == 0) {
        y[ka * B + b] = t[ka];
      } else {
        const T wr = T(root_re<M>(j));
        const T wi = inv ? T(-root_im<M>(j)) : T(root_im<M>(j));
        y[ka * B + b] = cplx<T>{t[ka].x * wr - t[ka].y * wi, t[ka].x * wi + t[ka].y * wr};
      }
    }
  }
#pragma unroll
  for (int ka = 0; ka < A; ++ka) {
    cplx<T> t[B];
#pragma unroll
    for (int b = 0; b < B; ++b) t[b] = y[ka * B + b];
    dftreg<T, B>(t, inv);
#pragma unroll
    for (int kb = 0; kb < B; ++kb) v[ka + A * kb] = t[kb];
  }
}

template <typename T, int M>
__device__ __forceinline__ void dftreg(cplx<T>* v, bool inv) {
  if constexpr (M == 1) {
  } else if constexpr (M == 2) {
    dft2(v);
  } else if constexpr (M == 3) {
    dft3(v, inv);
  } else if constexpr (M == 4) {
    dft4(v, inv);
  } else if constexpr (M == 5) {
    dft5(v, inv);
  } else if constexpr (M == 8) {
    dft8(v, inv);
  } else if constexpr (M == 6) {
    dft_ct<T, 2, 3>(v, inv);
  } else if constexpr (M == 9) {
    dft_ct<T, 3, 3>(v, inv);
  } else if constexpr (M == 10) {
    dft_ct<T, 2, 5>(v, inv);
  } else if constexpr (M == 12) {
    dft_ct<T, 3, 4>(v, inv);
  } else if constexpr (M == 15) {
    dft_ct<T, 3, 5>(v, inv);
  } else if constexpr (M == 16) {
    dft_ct<T, 4, 4>(v, inv);
  } else {
    static_assert(M == 1, "unsupported register DFT size");
  }
}

// N = N1 * N2 split used by the register kernels (N1 <= N2 <= 16).
template <int N>
struct Split;
template <> struct Split<32> { static constexpr int N1 = 4, N2 = 8; };
template <> struct Split<48> { static constexpr int N1 = 4, N2 = 12; };
template <> struct Split<64> { static constexpr int N1 = 8, N2 = 8; };
template <> struct Split<90> { static constexpr int N1 = 9, N2 = 10; };
template <> struct Split<96> { static constexpr int N1 = 8, N2 = 12; };
template <> struct Split<108> { static constexpr int N1 = 9, N2 = 12; };
template <> struct Split<128> { static constexpr int N1 = 8, N2 = 16; };
template <> struct Split<160> { static constexpr int N1 = 10, N2 = 16; };
template <> struct Split<192> { static constexpr int N1 = 12, N2 = 16; };
template <> struct Split<256> { static constexpr int N1 = 16, N2 = 16; };

}  // namespace ofdft
