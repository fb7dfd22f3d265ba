// The truncated-Newton inner loop of many independent cells in one
// persistent cooperative launch per TN (outer) iteration.
//
// Each cell is an ordinary plan (its own grid length, lattice, functional,
// WT table, spectra, partial sums and TnState); the batch gives the kernel
// an array of per-cell descriptors.  The launch runs the same stages as the
// one-cell device loop (get_direction_TN, optimization.py:92-148, with the
// exact Hessian product and the CG step folded into the transforms):
//
//   init     d = 0, r = -res, p = r and the sums r0, W0 r^2      (per chunk)
//   start    the loop state of every cell                         (per cell)
//   repeat until every cell's loop is done (or maxfun):
//     rows_fwd  p = r + beta p, transform inputs, axis-2 R2C     (row pairs)
//     axis      axis-1 forward                                    (column tiles)
//     pencil    axis-0 forward, multiplier, inverse, p.Ap sums    (column tiles)
//     axis      axis-1 inverse                                    (column tiles)
//     pap       pAp of every cell                                 (per cell)
//     rows_inv  axis-2 C2R, H p, the CG step, r^2 and W0 sums     (row pairs)
//     ctrl      the branch tests of the reference                 (per cell)
//   finish   direction = Best or r0 / pAp * p                     (per chunk)
//
// Every stage is a work queue over (cell, tile) items, spread over the
// co-resident workgroups; a grid barrier separates the stages.  Cells of
// different lengths share a stage: an item dispatches on its cell's length
// to the compile-time transforms, run with one workgroup size (256).  Items
// of a cell whose loop has finished are skipped; the loop ends for all
// workgroups at the same iteration (they read the same states after the
// barrier).
#pragma once
#include <hip/hip_cooperative_groups.h>

#include "ofdft_core.h"

namespace {

constexpr int kBatchTH = 256;
constexpr int kBatchChunk = kBatchTH * 16;  // points per init / finish item
constexpr int kBatchTile = 16;              // columns per column-kernel item

template <typename T>
struct TnCell {
  int n, nh, nhp, F, tx;
  long long N, Nhp, nrows, rblocks, chunks;
  GridGeom geo;
  FnParams fp;
  ConvKinds ck;
  const cplx<T>* tw0;
  const cplx<T>* tw1;
  const cplx<T>* tw2;
  cplx<T>* spec;
  const T* wtk;
  const T* phi;
  const T* res;
  const T* W0;
  T* d;
  T* r;
  T* p;
  T* best;
  double* partials;
  double* red;
  TnState* st;
};

// stage item spaces: prefix sums over the cells, [kind][nb + 1]
enum { kPrefChunk = 0, kPrefRows = 1, kPrefAxis = 2, kPrefPen = 3, kPrefKinds = 4 };

__device__ __forceinline__ int batch_cell(const long long* __restrict__ pref, int nb, long long item) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pref[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T, int NN>
__device__ __noinline__ void batch_rows_fwd(const TnCell<T>& C, long long local, cplx<T>* lds) {
  PreHvp<T> pre{C.phi, C.p, C.fp};
  PreHvpCG<T> pcg{pre, C.phi, C.p, C.r, C.st};
  using FFT = FFTFixed<NN, kRowPB<T> * 3, kBatchTH>;
  if (C.F == 3)
    rows_fwd_block<T, FFT, PreHvpCG<T>, 3>(pcg, FFT{}, C.spec, C.nrows, C.Nhp, C.nhp, 3, C.tw2, local, lds);
  else
    rows_fwd_block<T, FFT, PreHvpCG<T>, 0>(pcg, FFT{}, C.spec, C.nrows, C.Nhp, C.nhp, C.F, C.tw2, local, lds);
}

template <typename T, int NN>
__device__ __noinline__ void batch_rows_inv(const TnCell<T>& C, long long local, cplx<T>* lds) {
  PostHvpFast<T> h{C.phi, C.p, C.W0, nullptr, C.fp};
  PostHvpCG<T> post{h, C.d, C.r, C.best, C.st, C.red};
  using FFT = FFTFixed<NN, kRowPB<T> * 3, kBatchTH>;
  if (C.F == 3)
    rows_inv_block<T, FFT, PostHvpCG<T>, 3>(post, FFT{}, C.spec, C.nrows, C.Nhp, C.nhp, 3, C.tw2, C.partials, local,
                                           lds);
  else
    rows_inv_block<T, FFT, PostHvpCG<T>, 0>(post, FFT{}, C.spec, C.nrows, C.Nhp, C.nhp, C.F, C.tw2, C.partials,
                                           local, lds);
}

template <typename T, int NN>
__device__ __noinline__ void batch_axis(const TnCell<T>& C, long long local, bool inv, cplx<T>* lds) {
  const int bx = (int)(local % C.tx);
  const long long rem = local / C.tx;
  const int by = (int)(rem % C.n), bz = (int)(rem / C.n);
  axis_reg_block<T, Split<NN>::N1, Split<NN>::N2, kBatchTile>(C.spec, C.Nhp, C.nhp, (long long)C.n * C.nhp, C.nhp,
                                                              C.tw1, inv ? 1 : 0, bx, by, bz, lds);
}

template <typename T, int NN>
__device__ __noinline__ void batch_pencil(const TnCell<T>& C, long long local, cplx<T>* lds) {
  const int bx = (int)(local % C.tx);
  const long long rem = local / C.tx;
  const int by = (int)(rem % C.n), bz = (int)(rem / C.n);
  pencil_reg_block<T, Split<NN>::N1, Split<NN>::N2, kBatchTile>(C.spec, C.Nhp, C.geo, C.tw0, C.ck, C.wtk, C.partials,
                                                                local, bx, by, bz, lds);
}

#define OFDFT_BATCH_SWITCH(NEXPR, CALL) \
  switch (NEXPR) {                      \
    case 32: CALL(32); break;           \
    case 48: CALL(48); break;           \
    case 64: CALL(64); break;           \
    case 90: CALL(90); break;           \
    case 96: CALL(96); break;           \
    case 108: CALL(108); break;         \
    case 128: CALL(128); break;         \
    case 160: CALL(160); break;         \
    case 192: CALL(192); break;         \
    default: break;                     \
  }

// Waves per SIMD the register allocation is held to: the stage functions are
// called (not inlined), so each keeps its own register set and the kernel
// saves what is live across a call; FP32 fits 4 waves (127 VGPRs), FP64 2.
template <typename T>
constexpr int kBatchWaves = sizeof(T) == 4 ? 4 : 2;

template <typename T>
__global__ __launch_bounds__(kBatchTH, kBatchWaves<T>) void k_tn_batch(const TnCell<T>* __restrict__ cells, int nb,
                                                       const long long* __restrict__ pref, int maxfun) {
  namespace cgr = cooperative_groups;
  cgr::grid_group grid = cgr::this_grid();
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const long long* pChunk = pref + kPrefChunk * (nb + 1);
  const long long* pRows = pref + kPrefRows * (nb + 1);
  const long long* pAxis = pref + kPrefAxis * (nb + 1);
  const long long* pPen = pref + kPrefPen * (nb + 1);
  // init: d = 0, r = -res, p = r; sums r^2 and W0 r^2 per chunk
  for (long long item = blockIdx.x; item < pChunk[nb]; item += gridDim.x) {
    const int c = batch_cell(pChunk, nb, item);
    const TnCell<T>& C = cells[c];
    const long long chunk = item - pChunk[c];
    double acc[2] = {0.0, 0.0};
    for (int k = 0; k < kBatchChunk / kBatchTH; ++k) {
      const long long i = chunk * kBatchChunk + (long long)k * kBatchTH + threadIdx.x;
      if (i < C.N) {
        const T v = -C.res[i];
        C.d[i] = T(0);
        C.r[i] = v;
        C.p[i] = v;
        const double v2 = (double)v * (double)v;
        acc[0] += v2;
        acc[1] += (double)C.W0[i] * v2;
      }
    }
    block_sums<2>(acc, C.partials + 2 * chunk);
  }
  grid.sync();
  for (int c = blockIdx.x; c < nb; c += gridDim.x) {
    const TnCell<T>& C = cells[c];
    double* rr = C.red + 8;
    double tmp[2] = {thread_partials<2>(C.partials, C.chunks, 0), thread_partials<2>(C.partials, C.chunks, 1)};
    block_sums<2>(tmp, rr);
    __syncthreads();
    if (threadIdx.x == 0) tn_start_state(C.st, rr, maxfun);
  }
  grid.sync();
  for (int it = 0; it < maxfun; ++it) {
    int live = 0;
    for (int c = 0; c < nb; ++c) live += cells[c].st->done ? 0 : 1;
    if (!live) break;  // the same decision in every workgroup: states are read after a barrier
    // forward row pass with the update of p
    for (long long item = blockIdx.x; item < pRows[nb]; item += gridDim.x) {
      const int c = batch_cell(pRows, nb, item);
      const TnCell<T>& C = cells[c];
      if (C.st->done) continue;
#define OFDFT_B_ROWSF(NN) batch_rows_fwd<T, NN>(C, item - pRows[c], lds)
      OFDFT_BATCH_SWITCH(C.n, OFDFT_B_ROWSF)
#undef OFDFT_B_ROWSF
    }
    grid.sync();
    for (long long item = blockIdx.x; item < pAxis[nb]; item += gridDim.x) {
      const int c = batch_cell(pAxis, nb, item);
      const TnCell<T>& C = cells[c];
      if (C.st->done) continue;
#define OFDFT_B_AXISF(NN) batch_axis<T, NN>(C, item - pAxis[c], false, lds)
      OFDFT_BATCH_SWITCH(C.n, OFDFT_B_AXISF)
#undef OFDFT_B_AXISF
    }
    grid.sync();
    for (long long item = blockIdx.x; item < pPen[nb]; item += gridDim.x) {
      const int c = batch_cell(pPen, nb, item);
      const TnCell<T>& C = cells[c];
      if (C.st->done) continue;
#define OFDFT_B_PEN(NN) batch_pencil<T, NN>(C, item - pPen[c], lds)
      OFDFT_BATCH_SWITCH(C.n, OFDFT_B_PEN)
#undef OFDFT_B_PEN
    }
    grid.sync();
    for (long long item = blockIdx.x; item < pAxis[nb]; item += gridDim.x) {
      const int c = batch_cell(pAxis, nb, item);
      const TnCell<T>& C = cells[c];
      if (C.st->done) continue;
#define OFDFT_B_AXISI(NN) batch_axis<T, NN>(C, item - pAxis[c], true, lds)
      OFDFT_BATCH_SWITCH(C.n, OFDFT_B_AXISI)
#undef OFDFT_B_AXISI
    }
    grid.sync();
    for (int c = blockIdx.x; c < nb; c += gridDim.x) {
      const TnCell<T>& C = cells[c];
      if (C.st->done) continue;
      tn_pap_block(C.st, C.partials, pPen[c + 1] - pPen[c], C.red);
    }
    grid.sync();
    for (long long item = blockIdx.x; item < pRows[nb]; item += gridDim.x) {
      const int c = batch_cell(pRows, nb, item);
      const TnCell<T>& C = cells[c];
      if (C.st->done) continue;
#define OFDFT_B_ROWSI(NN) batch_rows_inv<T, NN>(C, item - pRows[c], lds)
      OFDFT_BATCH_SWITCH(C.n, OFDFT_B_ROWSI)
#undef OFDFT_B_ROWSI
    }
    grid.sync();
    for (int c = blockIdx.x; c < nb; c += gridDim.x) {
      const TnCell<T>& C = cells[c];
      if (C.st->done) continue;
      tn_ctrl4_block(C.st, C.partials, C.rblocks, C.red, C.red + 8);
    }
    grid.sync();
  }
  // finish: direction = Best or r0 / pAp * p where the loop asked for it
  for (long long item = blockIdx.x; item < pChunk[nb]; item += gridDim.x) {
    const int c = batch_cell(pChunk, nb, item);
    const TnCell<T>& C = cells[c];
    const int fin = C.st->fin;
    if (!fin) continue;
    const T scale = T(C.st->scale);
    const long long chunk = item - pChunk[c];
    for (int k = 0; k < kBatchChunk / kBatchTH; ++k) {
      const long long i = chunk * kBatchChunk + (long long)k * kBatchTH + threadIdx.x;
      if (i < C.N) C.d[i] = (fin & kActScale) ? scale * C.p[i] : C.best[i];
    }
  }
}

#undef OFDFT_BATCH_SWITCH

constexpr bool batch_length_ok(int n) {
  return n == 32 || n == 48 || n == 64 || n == 90 || n == 96 || n == 108 || n == 128 || n == 160 || n == 192;
}

}  // namespace

namespace ofdft_tu {

template <typename T>
int tn_batch_run(ofdft_batch_s* B, const T* const* phi, const T* const* res, const T* const* W0, T* const* d,
                 T* const* work, int maxfun, int* niter, int* status, double* pAp_last) {
  const int nb = (int)B->plans.size();
  if (maxfun < 1) return fail(OFDFT_EINVAL, "maxfun must be positive");
  hipStream_t stream = B->plans[0]->stream;
  std::vector<TnCell<T>> cells(nb);
  std::vector<long long> pref((size_t)kPrefKinds * (nb + 1), 0);
  size_t lds = 0;
  for (int c = 0; c < nb; ++c) {
    ofdft_plan_s* P = B->plans[c];
    if (!phi[c] || !res[c] || !W0[c] || !d[c] || !work[c]) return fail(OFDFT_EINVAL, "null field in the batch");
    if (P->stream != stream) return fail(OFDFT_EINVAL, "the plans of a batch must share one stream");
    if (!P->have_fn) return fail(OFDFT_ESTATE, "functional not set on a plan of the batch");
    FnParams fp = make_fn(P->fn);
    if (fp.F < 1) return fail(OFDFT_EINVAL, "the batched TN loop needs a convolution term (Hartree, vW or WT)");
    if (P->fn.wt && !P->have_wtk) return fail(OFDFT_ESTATE, "WT kernel table not built");
    TnCell<T>& C = cells[c];
    C.n = P->n[0];
    C.nh = P->nh;
    C.nhp = P->nhp;
    C.F = fp.F;
    C.tx = (P->nhp + kBatchTile - 1) / kBatchTile;
    C.N = P->N;
    C.Nhp = P->Nhp;
    C.nrows = (long long)P->n[0] * P->n[1];
    const long long npairs = (C.nrows + 1) / 2;
    C.rblocks = (npairs + kRowPB<T> - 1) / kRowPB<T>;
    C.chunks = (C.N + kBatchChunk - 1) / kBatchChunk;
    C.geo = geom(P);
    C.fp = fp;
    int kinds[3] = {0, 0, 0};
    double coef[3] = {0.0, 0.0, 0.0};
    if (fp.fH >= 0) {
      kinds[fp.fH] = OFDFT_CONV_HARTREE;
      coef[fp.fH] = 0.5;
    }
    if (fp.fV >= 0) {
      kinds[fp.fV] = OFDFT_CONV_GG;
      coef[fp.fV] = 0.5 * fp.y;
    }
    if (fp.fW >= 0) {
      kinds[fp.fW] = OFDFT_CONV_WT;
      coef[fp.fW] = fp.alpha / fp.beta;
    }
    for (int f = 0; f < 3; ++f) {
      C.ck.kind[f] = f < fp.F ? kinds[f] : 0;
      C.ck.coef[f] = coef[f];
    }
    C.ck.scale = 1.0 / (double)P->N;
    C.tw0 = reinterpret_cast<const cplx<T>*>(P->tw[0]);
    C.tw1 = reinterpret_cast<const cplx<T>*>(P->tw[1]);
    C.tw2 = reinterpret_cast<const cplx<T>*>(P->tw[2]);
    C.spec = reinterpret_cast<cplx<T>*>(P->spec);
    C.wtk = reinterpret_cast<const T*>(P->wtk);
    C.phi = phi[c];
    C.res = res[c];
    C.W0 = W0[c];
    C.d = d[c];
    C.r = work[c];
    C.p = work[c] + C.N;
    C.best = work[c] + 3 * C.N;
    C.partials = P->partials;
    C.red = P->red;
    C.st = reinterpret_cast<TnState*>(P->tnst);
    const long long naxis = (long long)C.tx * C.n * C.F;
    if (std::max(std::max(C.rblocks * 4, naxis), C.chunks * 2) > P->npartials)
      return fail(OFDFT_ESTATE, "partials buffer too small for the batched TN loop");
    pref[kPrefChunk * (nb + 1) + c + 1] = pref[kPrefChunk * (nb + 1) + c] + C.chunks;
    pref[kPrefRows * (nb + 1) + c + 1] = pref[kPrefRows * (nb + 1) + c] + C.rblocks;
    pref[kPrefAxis * (nb + 1) + c + 1] = pref[kPrefAxis * (nb + 1) + c] + naxis;
    pref[kPrefPen * (nb + 1) + c + 1] = pref[kPrefPen * (nb + 1) + c] + naxis;
    const size_t rows_lds = (size_t)(kRowPB<T> * 3 + 1) * C.n * sizeof(cplx<T>);
    size_t col_lds = 0;
#define OFDFT_B_COL(NN) \
  col_lds = (size_t)kBatchTile * col_block_stride<T, Split<NN>::N1, Split<NN>::N2, kBatchTile>() * sizeof(cplx<T>)
    switch (C.n) {
      case 32: OFDFT_B_COL(32); break;
      case 48: OFDFT_B_COL(48); break;
      case 64: OFDFT_B_COL(64); break;
      case 90: OFDFT_B_COL(90); break;
      case 96: OFDFT_B_COL(96); break;
      case 108: OFDFT_B_COL(108); break;
      case 128: OFDFT_B_COL(128); break;
      case 160: OFDFT_B_COL(160); break;
      case 192: OFDFT_B_COL(192); break;
      default: break;
    }
#undef OFDFT_B_COL
    lds = std::max(lds, std::max(rows_lds, col_lds));
  }
  // device copies of the descriptors and prefix sums (grown, never shrunk)
  const size_t cbytes = sizeof(TnCell<T>) * nb, pbytes = sizeof(long long) * pref.size();
  if (cbytes > B->cells_cap) {
    if (B->d_cells) HIPCHK(hipFree(B->d_cells));
    B->d_cells = nullptr;
    HIPCHK(hipMalloc(&B->d_cells, cbytes));
    B->cells_cap = cbytes;
  }
  if (pbytes > B->pref_cap) {
    if (B->d_pref) HIPCHK(hipFree(B->d_pref));
    B->d_pref = nullptr;
    HIPCHK(hipMalloc((void**)&B->d_pref, pbytes));
    B->pref_cap = pbytes;
  }
  HIPCHK(hipMemcpyAsync(B->d_cells, cells.data(), cbytes, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(B->d_pref, pref.data(), pbytes, hipMemcpyHostToDevice, stream));
  // every workgroup of a cooperative launch must be resident at once
  int dev = 0, ncu = 0, per_cu = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tn_batch<T>, kBatchTH, lds));
  if (per_cu < 1) return fail(OFDFT_EHIP, "the batched TN kernel does not fit a compute unit");
  long long most = 0;
  for (int k = 0; k < kPrefKinds; ++k) most = std::max(most, pref[(size_t)k * (nb + 1) + nb]);
  const int grid = (int)std::max<long long>(1, std::min<long long>((long long)per_cu * ncu, most));
  const TnCell<T>* dc = reinterpret_cast<const TnCell<T>*>(B->d_cells);
  const long long* dp = B->d_pref;
  int nbv = nb, mf = maxfun;
  void* args[] = {(void*)&dc, (void*)&nbv, (void*)&dp, (void*)&mf};
  const hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_tn_batch<T>), dim3(grid),
                                                  dim3(kBatchTH), args, (unsigned)lds, stream);
  if (e != hipSuccess) return fail(OFDFT_EHIP, std::string("cooperative launch: ") + hipGetErrorString(e));
  for (int c = 0; c < nb; ++c) {
    ofdft_plan_s* P = B->plans[c];
    HIPCHK(hipMemcpyAsync(P->htnst, P->tnst, sizeof(TnState), hipMemcpyDeviceToHost, stream));
  }
  HIPCHK(hipStreamSynchronize(stream));
  for (int c = 0; c < nb; ++c) {
    const TnState* h = reinterpret_cast<const TnState*>(B->plans[c]->htnst);
    if (!h->done) return fail(OFDFT_ESTATE, "batched TN loop did not finish");
    niter[c] = h->niter;
    status[c] = h->status;
    pAp_last[c] = h->pAp;
  }
  return 0;
}

}  // namespace ofdft_tu

#define OFDFT_TU_BATCH(EXT, T)                                                                                   \
  EXT template int ofdft_tu::tn_batch_run<T>(ofdft_batch_s*, const T* const*, const T* const*, const T* const*, \
                                             T* const*, T* const*, int, int*, int*, double*);
