// The truncated-Newton inner loop of many independent cells per TN (outer)
// iteration, in one of two forms.
//
// Each cell is an ordinary plan (its own grid length, lattice, functional,
// WT table, spectra, partial sums and TnState); the batch gives the kernels
// an array of per-cell descriptors.  Both forms run the same stages as the
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
// Persistent form (k_tn_batch): ONE cooperative launch.  Every stage is a
// work queue over (cell, tile) items, spread over the co-resident
// workgroups; a grid barrier separates the stages.  Cells of different
// lengths share a stage: an item dispatches on its cell's length to the
// compile-time transforms, run with one workgroup size (256).  Items of a
// cell whose loop has finished are skipped; the loop ends for all
// workgroups at the same iteration (they read the same states after the
// barrier).
//
// Graph form (k_bt_*): the cells are grouped by length; each stage of one CG
// iteration is one launch per length class, with the one-cell kernels'
// workgroup shapes (row threads per length, the autotuned column tile
// widths) and a grid extended over the class's cells.  The launches of one
// CG iteration of all cells are captured once into a hipGraph and replayed
// per iteration; kernel boundaries replace the grid barriers.  The host
// checks a live-cell count every few iterations.
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
  long long npen;  // pencil partial-sum slots of one CG iteration
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
  axis_reg_block<T, Split<NN>::N1, Split<NN>::N2, kBatchTile>(C.spec, C.Nhp, C.nhp, (long long)C.n * C.nhp, C.nh,
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

// init of one chunk of a cell: d = 0, r = -res, p = r; sums r^2, W0 r^2
// (kBatchTH threads)
template <typename T>
__device__ __forceinline__ void bt_init_chunk(const TnCell<T>& C, long long chunk) {
  if (!C.phi) return;  // a cell left out of this call
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

// the loop state of one cell from its init sums (one workgroup)
template <typename T>
__device__ __forceinline__ void bt_start_cell(const TnCell<T>& C, int maxfun) {
  double* rr = C.red + 8;
  if (!C.phi) {  // a cell left out of this call: its loop is over before it starts
    if (threadIdx.x == 0) tn_start_state(C.st, rr, 0);
    return;
  }
  double tmp[2] = {thread_partials<2>(C.partials, C.chunks, 0), thread_partials<2>(C.partials, C.chunks, 1)};
  block_sums<2>(tmp, rr);
  __syncthreads();
  if (threadIdx.x == 0) tn_start_state(C.st, rr, maxfun);
}

// finish of one chunk: direction = Best or r0 / pAp * p where the loop asked
template <typename T>
__device__ __forceinline__ void bt_finish_chunk(const TnCell<T>& C, long long chunk) {
  const int fin = C.st->fin;
  if (!fin) return;
  const T scale = T(C.st->scale);
  for (int k = 0; k < kBatchChunk / kBatchTH; ++k) {
    const long long i = chunk * kBatchChunk + (long long)k * kBatchTH + threadIdx.x;
    if (i < C.N) C.d[i] = (fin & kActScale) ? scale * C.p[i] : C.best[i];
  }
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
    bt_init_chunk(cells[c], item - pChunk[c]);
  }
  grid.sync();
  for (int c = blockIdx.x; c < nb; c += gridDim.x) bt_start_cell(cells[c], maxfun);
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
    bt_finish_chunk(cells[c], item - pChunk[c]);
  }
}

// ---------------------------------------------------------------------------
// Graph form: one launch per (stage, length class); blockIdx.y (rows) or
// blockIdx.z / F (columns) selects the class's cell through idx.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBatchTH) void k_bt_init(const TnCell<T>* __restrict__ cells, int nb,
                                                      const long long* __restrict__ pchunk) {
  const int c = batch_cell(pchunk, nb, blockIdx.x);
  bt_init_chunk(cells[c], (long long)blockIdx.x - pchunk[c]);
}

template <typename T>
__global__ __launch_bounds__(kBatchTH) void k_bt_start(const TnCell<T>* __restrict__ cells, int maxfun) {
  bt_start_cell(cells[blockIdx.x], maxfun);
}

template <typename T>
__global__ __launch_bounds__(kBatchTH) void k_bt_finish(const TnCell<T>* __restrict__ cells, int nb,
                                                        const long long* __restrict__ pchunk) {
  const int c = batch_cell(pchunk, nb, blockIdx.x);
  bt_finish_chunk(cells[c], (long long)blockIdx.x - pchunk[c]);
}

template <typename T, int NN, int FC>
__global__ __launch_bounds__(rows_th<NN>()) void k_bt_rows_fwd(const TnCell<T>* __restrict__ cells,
                                                               const int* __restrict__ idx) {
  const TnCell<T>& C = cells[idx[blockIdx.y]];
  if (C.st->done) return;
  extern __shared__ unsigned char smem_raw[];
  PreHvp<T> pre{C.phi, C.p, C.fp};
  PreHvpCG<T> pcg{pre, C.phi, C.p, C.r, C.st};
  rows_fwd_block<T, RowsFFT<T, NN>, PreHvpCG<T>, FC>(pcg, RowsFFT<T, NN>{}, C.spec, C.nrows, C.Nhp, C.nhp, C.F,
                                                     C.tw2, blockIdx.x, reinterpret_cast<cplx<T>*>(smem_raw));
}

template <typename T, int NN, int FC>
__global__ __launch_bounds__(rows_th<NN>()) void k_bt_rows_inv(const TnCell<T>* __restrict__ cells,
                                                               const int* __restrict__ idx) {
  const TnCell<T>& C = cells[idx[blockIdx.y]];
  if (C.st->done) return;
  extern __shared__ unsigned char smem_raw[];
  PostHvpFast<T> h{C.phi, C.p, C.W0, nullptr, C.fp};
  PostHvpCG<T> post{h, C.d, C.r, C.best, C.st, C.red};
  rows_inv_block<T, RowsFFT<T, NN>, PostHvpCG<T>, FC>(post, RowsFFT<T, NN>{}, C.spec, C.nrows, C.Nhp, C.nhp, C.F,
                                                      C.tw2, C.partials, blockIdx.x,
                                                      reinterpret_cast<cplx<T>*>(smem_raw));
}

template <typename T, int N1, int N2, int W>
__global__ __launch_bounds__(W * N2) void k_bt_axis(const TnCell<T>* __restrict__ cells, const int* __restrict__ idx,
                                                    int F, int inv) {
  const int k = blockIdx.z / F, f = blockIdx.z - k * F;
  const TnCell<T>& C = cells[idx[k]];
  if (C.st->done || f >= C.F) return;
  __shared__ cplx<T> lds[W * col_block_stride<T, N1, N2, W>()];
  axis_reg_block<T, N1, N2, W>(C.spec, C.Nhp, C.nhp, (long long)C.n * C.nhp, C.nh, C.tw1, inv, blockIdx.x,
                               blockIdx.y, f, lds);
}

template <typename T, int N1, int N2, int W>
__global__ __launch_bounds__(W * N2) void k_bt_pencil(const TnCell<T>* __restrict__ cells,
                                                      const int* __restrict__ idx, int F) {
  const int k = blockIdx.z / F, f = blockIdx.z - k * F;
  const TnCell<T>& C = cells[idx[k]];
  if (C.st->done || f >= C.F) return;
  __shared__ cplx<T> lds[W * col_block_stride<T, N1, N2, W>()];
  // the cell's partial-sum slot of this tile, as the one-cell pencil numbers it
  const long long pslot = ((long long)f * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  pencil_reg_block<T, N1, N2, W>(C.spec, C.Nhp, C.geo, C.tw0, C.ck, C.wtk, C.partials, pslot, blockIdx.x,
                                 blockIdx.y, f, lds);
}

template <typename T>
__global__ __launch_bounds__(kBatchTH) void k_bt_pap(const TnCell<T>* __restrict__ cells) {
  const TnCell<T>& C = cells[blockIdx.x];
  if (C.st->done) return;
  tn_pap_block(C.st, C.partials, C.npen, C.red);
}

template <typename T>
__global__ __launch_bounds__(kBatchTH) void k_bt_ctrl(const TnCell<T>* __restrict__ cells) {
  const TnCell<T>& C = cells[blockIdx.x];
  if (C.st->done) return;
  tn_ctrl4_block(C.st, C.partials, C.rblocks, C.red, C.red + 8);
}

// number of cells whose loop is still running, into mapped host memory
template <typename T>
__global__ __launch_bounds__(kBatchTH) void k_bt_live(const TnCell<T>* __restrict__ cells, int nb,
                                                      int* __restrict__ live) {
  double acc[1] = {0.0};
  for (int c = threadIdx.x; c < nb; c += blockDim.x) acc[0] += cells[c].st->done ? 0.0 : 1.0;
  __shared__ double tot[1];
  block_sums<1>(acc, tot);
  __syncthreads();
  if (threadIdx.x == 0) *live = (int)(tot[0] + 0.5);
}

// the loop results of every cell, into mapped host memory (one copy for all)
template <typename T>
__global__ void k_bt_gather(const TnCell<T>* __restrict__ cells, int nb, TnState* __restrict__ out) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nb; c += gridDim.x * blockDim.x) out[c] = *cells[c].st;
}

#undef OFDFT_BATCH_SWITCH

#define OFDFT_BATCH_SWITCH_H(NEXPR, CALL) \
  switch (NEXPR) {                        \
    case 32: CALL(32); break;             \
    case 48: CALL(48); break;             \
    case 64: CALL(64); break;             \
    case 90: CALL(90); break;             \
    case 96: CALL(96); break;             \
    case 108: CALL(108); break;           \
    case 128: CALL(128); break;           \
    case 160: CALL(160); break;           \
    case 192: CALL(192); break;           \
    default: break;                       \
  }

constexpr bool batch_length_ok(int n) {
  return n == 32 || n == 48 || n == 64 || n == 90 || n == 96 || n == 108 || n == 128 || n == 160 || n == 192;
}

}  // namespace

namespace {

// One length class of the graph form: cells idx[off .. off + m) of length n.
struct BtClass {
  int n, m, off, F, fc3, wa, wp;
  long long rblocks;
};

constexpr int nhp_of(int n) { return (n / 2 + 1 + 15) / 16 * 16; }

// Descriptors, prefix sums and LDS size of a batch call; graph = 1 sizes the
// pencil partial sums for the plans' own tile widths, 0 for kBatchTile.
template <typename T>
int batch_setup(ofdft_batch_s* B, const T* const* phi, const T* const* res, const T* const* W0, T* const* d,
                T* const* work, bool graph, std::vector<TnCell<T>>& cells, std::vector<long long>& pref,
                size_t& lds) {
  const int nb = (int)B->plans.size();
  hipStream_t stream = B->plans[0]->stream;
  cells.assign(nb, TnCell<T>{});
  pref.assign((size_t)kPrefKinds * (nb + 1), 0);
  lds = 0;
  for (int c = 0; c < nb; ++c) {
    ofdft_plan_s* P = B->plans[c];
    // phi[c] null: cell c sits this call out (no fields are touched)
    if (phi[c] && (!res[c] || !W0[c] || !d[c] || !work[c])) return fail(OFDFT_EINVAL, "null field in the batch");
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
    const int wp = graph ? P->tile_pencil : kBatchTile;
    C.npen = (long long)((P->nhp + wp - 1) / wp) * C.n * C.F;
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
    C.p = work[c] ? work[c] + C.N : nullptr;
    C.best = work[c] ? work[c] + 3 * C.N : nullptr;
    C.partials = P->partials;
    C.red = P->red;
    C.st = reinterpret_cast<TnState*>(P->tnst);
    const long long naxis = (long long)C.tx * C.n * C.F;
    if (std::max(std::max(std::max(C.rblocks * 4, naxis), C.npen), C.chunks * 2) > P->npartials)
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
  const size_t sbytes = sizeof(TnState) * nb;
  if (sbytes > B->states_cap) {
    if (B->h_states) HIPCHK(hipHostFree(B->h_states));
    B->h_states = nullptr;
    HIPCHK(hipHostMalloc(&B->h_states, sbytes, hipHostMallocDefault));
    HIPCHK(hipHostGetDevicePointer(&B->d_states, B->h_states, 0));
    B->states_cap = sbytes;
  }
  HIPCHK(hipMemcpyAsync(B->d_cells, cells.data(), cbytes, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(B->d_pref, pref.data(), pbytes, hipMemcpyHostToDevice, stream));
  return 0;
}

// every cell's TnState into the mapped host array, then the results
template <typename T>
int batch_results(ofdft_batch_s* B, hipStream_t stream, int* niter, int* status, double* pAp_last) {
  const int nb = (int)B->plans.size();
  hipLaunchKernelGGL((k_bt_gather<T>), dim3((nb + 255) / 256), dim3(256), 0, stream,
                     reinterpret_cast<const TnCell<T>*>(B->d_cells), nb, reinterpret_cast<TnState*>(B->d_states));
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(stream));
  const TnState* h = reinterpret_cast<const TnState*>(B->h_states);
  int longest = 0;
  for (int c = 0; c < nb; ++c) {
    if (!h[c].done) return fail(OFDFT_ESTATE, "batched TN loop did not finish");
    niter[c] = h[c].niter;
    status[c] = h[c].status;
    pAp_last[c] = h[c].pAp;
    longest = std::max(longest, h[c].niter);
  }
  B->last_iters = longest;
  return 0;
}

// One stage of one length class (graph form).
template <typename T, int NN>
void bt_stage(int stage, const BtClass& k, const TnCell<T>* dc, const int* didx, hipStream_t s) {
  constexpr int N1 = Split<NN>::N1, N2 = Split<NN>::N2;
  constexpr int nhp = nhp_of(NN);
  const int* ix = didx + k.off;
  const unsigned fz = (unsigned)(k.F * k.m);
  if (stage == 0 || stage == 5) {
    const size_t lds = (size_t)(kRowPB<T> * k.F + 1) * NN * sizeof(cplx<T>);
    const dim3 g((unsigned)k.rblocks, (unsigned)k.m);
    if (stage == 0) {
      if (k.fc3) hipLaunchKernelGGL((k_bt_rows_fwd<T, NN, 3>), g, dim3(rows_th<NN>()), lds, s, dc, ix);
      else hipLaunchKernelGGL((k_bt_rows_fwd<T, NN, 0>), g, dim3(rows_th<NN>()), lds, s, dc, ix);
    } else {
      if (k.fc3) hipLaunchKernelGGL((k_bt_rows_inv<T, NN, 3>), g, dim3(rows_th<NN>()), lds, s, dc, ix);
      else hipLaunchKernelGGL((k_bt_rows_inv<T, NN, 0>), g, dim3(rows_th<NN>()), lds, s, dc, ix);
    }
    return;
  }
  if (stage == 1 || stage == 3) {
    const int inv = stage == 3 ? 1 : 0;
#define OFDFT_BT_AXIS(W) \
  hipLaunchKernelGGL((k_bt_axis<T, N1, N2, W>), dim3((nhp + W - 1) / W, NN, fz), dim3(W * N2), 0, s, dc, ix, k.F, inv)
    if (k.wa == 8) OFDFT_BT_AXIS(8);
    else if (sizeof(T) == 4 && k.wa == 32) OFDFT_BT_AXIS((sizeof(T) == 4 ? 32 : 16));
    else OFDFT_BT_AXIS(16);
#undef OFDFT_BT_AXIS
    return;
  }
#define OFDFT_BT_PEN(W) \
  hipLaunchKernelGGL((k_bt_pencil<T, N1, N2, W>), dim3((nhp + W - 1) / W, NN, fz), dim3(W * N2), 0, s, dc, ix, k.F)
  if (k.wp == 8) OFDFT_BT_PEN(8);
  else if (sizeof(T) == 4 && k.wp == 32) OFDFT_BT_PEN((sizeof(T) == 4 ? 32 : 16));
  else OFDFT_BT_PEN(16);
#undef OFDFT_BT_PEN
}

template <typename T>
void bt_stage_any(int stage, const BtClass& k, const TnCell<T>* dc, const int* didx, hipStream_t s) {
#define OFDFT_BT_CASE(NN) bt_stage<T, NN>(stage, k, dc, didx, s)
  OFDFT_BATCH_SWITCH_H(k.n, OFDFT_BT_CASE)
#undef OFDFT_BT_CASE
}

}  // namespace

namespace ofdft_tu {

// Persistent form: one cooperative launch runs every cell's loop.
template <typename T>
int tn_batch_persistent(ofdft_batch_s* B, const T* const* phi, const T* const* res, const T* const* W0,
                        T* const* d, T* const* work, int maxfun, int* niter, int* status, double* pAp_last) {
  const int nb = (int)B->plans.size();
  hipStream_t stream = B->plans[0]->stream;
  std::vector<TnCell<T>> cells;
  std::vector<long long> pref;
  size_t lds = 0;
  int rc = batch_setup<T>(B, phi, res, W0, d, work, false, cells, pref, lds);
  if (rc) return rc;
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
  return batch_results<T>(B, stream, niter, status, pAp_last);
}

// Graph form: per-length stage launches, one captured graph per CG
// iteration of all cells, replayed until every cell's loop has ended.
template <typename T>
int tn_batch_graph(ofdft_batch_s* B, const T* const* phi, const T* const* res, const T* const* W0, T* const* d,
                   T* const* work, int maxfun, int* niter, int* status, double* pAp_last) {
  const int nb = (int)B->plans.size();
  hipStream_t stream = B->plans[0]->stream;
  std::vector<TnCell<T>> cells;
  std::vector<long long> pref;
  size_t lds = 0;
  int rc = batch_setup<T>(B, phi, res, W0, d, work, true, cells, pref, lds);
  if (rc) return rc;
  // length classes (ascending length), cells in plan order within a class
  std::vector<int> order(nb);
  for (int c = 0; c < nb; ++c) order[c] = c;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells[a].n < cells[b].n; });
  std::vector<BtClass> cls;
  for (int i = 0; i < nb; ++i) {
    const TnCell<T>& C = cells[order[i]];
    const ofdft_plan_s* P = B->plans[order[i]];
    if (cls.empty() || cls.back().n != C.n) {
      cls.push_back(BtClass{C.n, 0, i, C.F, C.F == 3 ? 1 : 0, P->tile_axis, P->tile_pencil, C.rblocks});
    } else {
      BtClass& k = cls.back();
      if (P->tile_pencil != k.wp) return fail(OFDFT_EINVAL, "plans of one length in a batch need one pencil tile width");
      k.F = std::max(k.F, C.F);
      if (C.F != 3) k.fc3 = 0;
    }
    cls.back().m += 1;
  }
  const size_t ibytes = sizeof(int) * nb;
  if (ibytes > B->idx_cap) {
    if (B->d_idx) HIPCHK(hipFree(B->d_idx));
    B->d_idx = nullptr;
    HIPCHK(hipMalloc((void**)&B->d_idx, ibytes));
    B->idx_cap = ibytes;
  }
  if (!B->h_live) {
    HIPCHK(hipHostMalloc((void**)&B->h_live, sizeof(int), hipHostMallocDefault));
    HIPCHK(hipHostGetDevicePointer((void**)&B->d_live, B->h_live, 0));
  }
  HIPCHK(hipMemcpyAsync(B->d_idx, order.data(), ibytes, hipMemcpyHostToDevice, stream));
  const TnCell<T>* dc = reinterpret_cast<const TnCell<T>*>(B->d_cells);
  // the captured graph holds launch shapes and device addresses only
  std::vector<long long> key = {nb, (long long)(uintptr_t)B->d_cells, (long long)(uintptr_t)B->d_idx};
  for (const BtClass& k : cls) key.insert(key.end(), {k.n, k.m, k.off, k.F, k.fc3, k.wa, k.wp});
  if (!B->exec || key != B->key) {
    if (B->exec) {
      (void)hipGraphExecDestroy(B->exec);
      B->exec = nullptr;
    }
    if (!B->cap_stream) HIPCHK(hipStreamCreateWithFlags(&B->cap_stream, hipStreamNonBlocking));
    hipStream_t cs = B->cap_stream;
    HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    for (int stage = 0; stage < 7; ++stage) {
      if (stage == 4) hipLaunchKernelGGL((k_bt_pap<T>), dim3(nb), dim3(kBatchTH), 0, cs, dc);
      else if (stage == 6) hipLaunchKernelGGL((k_bt_ctrl<T>), dim3(nb), dim3(kBatchTH), 0, cs, dc);
      else
        for (const BtClass& k : cls) bt_stage_any<T>(stage, k, dc, B->d_idx, cs);
    }
    const hipError_t el = hipGetLastError();
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &g);
    if (el != hipSuccess || ec != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(OFDFT_EHIP, std::string("batch graph capture: ") + hipGetErrorString(el != hipSuccess ? el : ec));
    }
    const hipError_t ei = hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      B->exec = nullptr;
      return fail(OFDFT_EHIP, std::string("batch graph instantiate: ") + hipGetErrorString(ei));
    }
    B->key = key;
  }
  const long long nchunks = pref[kPrefChunk * (nb + 1) + nb];
  const long long* dchunk = B->d_pref + kPrefChunk * (nb + 1);
  hipLaunchKernelGGL((k_bt_init<T>), dim3((unsigned)nchunks), dim3(kBatchTH), 0, stream, dc, nb, dchunk);
  hipLaunchKernelGGL((k_bt_start<T>), dim3(nb), dim3(kBatchTH), 0, stream, dc, maxfun);
  LAUNCHCHK();
  // iterations queued between checks: first about the previous call's
  // longest loop, then two at a time (a queued iteration after every loop
  // has ended returns at once in each launch)
  int queued = 0, k = std::max(1, B->last_iters - 1);
  while (queued < maxfun) {
    for (int j = 0; j < k && queued < maxfun; ++j, ++queued) HIPCHK(hipGraphLaunch(B->exec, stream));
    hipLaunchKernelGGL((k_bt_live<T>), dim3(1), dim3(kBatchTH), 0, stream, dc, nb, B->d_live);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(stream));
    if (*B->h_live == 0) break;
    k = 2;
  }
  hipLaunchKernelGGL((k_bt_finish<T>), dim3((unsigned)nchunks), dim3(kBatchTH), 0, stream, dc, nb, dchunk);
  LAUNCHCHK();
  return batch_results<T>(B, stream, niter, status, pAp_last);
}

template <typename T>
int tn_batch_run(ofdft_batch_s* B, const T* const* phi, const T* const* res, const T* const* W0, T* const* d,
                 T* const* work, int maxfun, int* niter, int* status, double* pAp_last) {
  if (maxfun < 1) return fail(OFDFT_EINVAL, "maxfun must be positive");
  if (B->use_graph) return tn_batch_graph<T>(B, phi, res, W0, d, work, maxfun, niter, status, pAp_last);
  return tn_batch_persistent<T>(B, phi, res, W0, d, work, maxfun, niter, status, pAp_last);
}

}  // namespace ofdft_tu

#define OFDFT_TU_BATCH(EXT, T)                                                                                   \
  EXT template int ofdft_tu::tn_batch_run<T>(ofdft_batch_s*, const T* const*, const T* const*, const T* const*, \
                                             T* const*, T* const*, int, int*, int*, double*);
