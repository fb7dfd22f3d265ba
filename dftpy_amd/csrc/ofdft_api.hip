// C ABI of the orbital-free DFT hot path (declared in include/ofdft_amd.h).
// The transform pipelines of the evaluation, the Hessian products, the TN
// loop and the GGA / LWT terms are compiled in ofdft_tu_*.hip.
#include "ofdft_core.h"

namespace ofdft_shared {
thread_local std::string g_err;
}

OFDFT_TU_EVAL(extern, float)
OFDFT_TU_EVAL(extern, double)
OFDFT_TU_HVP(extern, float)
OFDFT_TU_HVP(extern, double)
OFDFT_TU_GGA(extern, float)
OFDFT_TU_GGA(extern, double)
OFDFT_TU_LWT(extern, float)
OFDFT_TU_LWT(extern, double)

#include <mutex>

// ---------------------------------------------------------------------------
// Allocation-free Ewald sums (ofdft_ewald_noalloc): the per-call staging of
// ofdft_ewald (ion positions, charges, fractional coordinates, image shifts)
// and its PME charge grid and b-moduli live in one static device array per
// GPU, held until the process ends, instead of buffers allocated per call.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kScratchBytes = size_t(8) << 20;  // 8 MiB on every GPU the library runs on
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of the scratch array on the GPU of `stream`: looked up once
// per device (hipGetSymbolAddress with that device current) and kept; in host
// code the array's name is not a device address.
int scratch_address(hipStream_t stream, unsigned char** out) {
  int dev = 0;
  HIPCHK(hipStreamGetDevice(stream, &dev));
  if (dev < 0 || dev >= kMaxScratchDevices) return fail(OFDFT_EINVAL, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return fail(OFDFT_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

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
  // rows padded to a multiple of 16 entries: every 16-column tile of the
  // column kernels starts on a 128-byte boundary and covers whole lines
  P->nhp = (P->nh + 15) / 16 * 16;
  P->N = (long long)n0 * n1 * n2;
  P->Nh = (long long)n0 * n1 * P->nh;
  P->Nhp = (long long)n0 * n1 * P->nhp;
  {
    const long long cs = precision == OFDFT_F64 ? 16 : 8;
    P->idx32 = (3 * P->Nhp * cs < (1LL << 31) && P->N * cs < (1LL << 31)) ? 1 : 0;
  }
  P->prec = precision;
  P->full = full ? 1 : 0;
  P->rs = precision == OFDFT_F64 ? sizeof(double) : sizeof(float);
  P->stream = reinterpret_cast<hipStream_t>(hip_stream);
  for (int i = 0; i < 3; ++i) {
    if (!make_pass(P->n[i], &P->pl[i], P->prec == OFDFT_F32)) {
      delete P;
      return fail(OFDFT_EINVAL, "grid sizes must be products of 2, 3, 5, 7, 11 and 13 (at most 12 passes)");
    }
  }
  // LDS of the runtime-schedule kernels: a row workgroup holds (3 PB + 1)
  // rows of n2 points, a column workgroup 16 columns of n0 (or n1) points;
  // both must fit gfx950's 160 KiB per workgroup.  FP64: n0, n1 <= 601 and
  // n2 <= 1462; FP32: n0, n1 <= 1203 and n2 <= 2925.
  const size_t cs = 2 * P->rs;
  const size_t rows_lds = (size_t)((precision == OFDFT_F64 ? kRowPB<double> : kRowPB<float>) * 3 + 1) * n2 * cs;
  const size_t col0_lds = (size_t)(kColTile * (n0 + 1) + n0) * cs, col1_lds = (size_t)(kColTile * (n1 + 1) + n1) * cs;
  if (rows_lds > kMaxLdsBytes || col0_lds > kMaxLdsBytes || col1_lds > kMaxLdsBytes) {
    delete P;
    return fail(OFDFT_EINVAL, std::string("grid dimension too large for the LDS tile sizes (") +
                                  (precision == OFDFT_F64 ? "FP64: n0, n1 <= 601, n2 <= 1462" :
                                                            "FP32: n0, n1 <= 1203, n2 <= 2925") + ")");
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
  if (hipMalloc(&P->spec, cs * 3 * P->Nhp) != hipSuccess) return cleanup(OFDFT_ENOMEM, "spectrum workspace");
  if (hipMemset(P->spec, 0, cs * 3 * P->Nhp) != hipSuccess) return cleanup(OFDFT_EHIP, "spectrum workspace clear");
  const long long npairs = ((long long)n0 * n1 + 1) / 2;
  const long long nblk = (npairs + kRowPBMin - 1) / kRowPBMin;
  P->npartials = std::max<long long>(nblk * OFDFT_NSUM, 2048LL * 3);
  if (hipMalloc((void**)&P->partials, sizeof(double) * P->npartials) != hipSuccess)
    return cleanup(OFDFT_ENOMEM, "partials");
  if (hipMalloc((void**)&P->red, sizeof(double) * kMaxSum) != hipSuccess) return cleanup(OFDFT_ENOMEM, "red");
  if (hipHostMalloc((void**)&P->hred, sizeof(double) * kMaxSum, hipHostMallocDefault) != hipSuccess)
    return cleanup(OFDFT_ENOMEM, "pinned host buffer");
  if (hipHostGetDevicePointer((void**)&P->dhred, P->hred, 0) != hipSuccess)
    return cleanup(OFDFT_EHIP, "pinned host buffer mapping");
  if (hipMalloc(&P->tnst, sizeof(TnState)) != hipSuccess) return cleanup(OFDFT_ENOMEM, "TN state");
  if (hipMemset(P->tnst, 0, sizeof(TnState)) != hipSuccess) return cleanup(OFDFT_EHIP, "TN state clear");
  if (hipHostMalloc(&P->htnst, sizeof(TnState), hipHostMallocCoherent) != hipSuccess)
    return cleanup(OFDFT_ENOMEM, "pinned TN state");
  std::memset(P->htnst, 0, sizeof(TnState));
  if (hipHostGetDevicePointer(&P->dhtnst, P->htnst, 0) != hipSuccess)
    return cleanup(OFDFT_EHIP, "pinned TN state mapping");
  if (hipEventCreateWithFlags(&P->tn_ev, hipEventDisableTiming) != hipSuccess)
    return cleanup(OFDFT_EHIP, "TN loop event");
  *out = P;
  return 0;
}

int ofdft_plan_destroy(ofdft_plan P) {
  if (!P) return 0;
  for (int i = 0; i < 3; ++i)
    if (P->tw[i]) (void)hipFree(P->tw[i]);
  if (P->spec) (void)hipFree(P->spec);
  if (P->wtk) (void)hipFree(P->wtk);
  if (P->wtks) (void)hipFree(P->wtks);
  if (P->vloc) (void)hipFree(P->vloc);
  if (P->partials) (void)hipFree(P->partials);
  if (P->d_knots) (void)hipFree(P->d_knots);
  if (P->d_coefs) (void)hipFree(P->d_coefs);
  if (P->d_splines) (void)hipFree(P->d_splines);
  if (P->d_frac) (void)hipFree(P->d_frac);
  if (P->d_pos) (void)hipFree(P->d_pos);
  if (P->d_species) (void)hipFree(P->d_species);
  if (P->d_bm) (void)hipFree(P->d_bm);
  if (P->d_work) (void)hipFree(P->d_work);
  if (P->d_gga) (void)hipFree(P->d_gga);
  if (P->d_lwttab) (void)hipFree(P->d_lwttab);
  if (P->d_lwt) (void)hipFree(P->d_lwt);
  if (P->red) (void)hipFree(P->red);
  if (P->hred) (void)hipHostFree(P->hred);
  if (P->tnst) (void)hipFree(P->tnst);
  if (P->htnst) (void)hipHostFree(P->htnst);
  if (P->tn_exec) (void)hipGraphExecDestroy(P->tn_exec);
  if (P->tn_ev) (void)hipEventDestroy(P->tn_ev);
  if (P->cap_stream) (void)hipStreamDestroy(P->cap_stream);
  for (hipEvent_t e : P->ev_pool) (void)hipEventDestroy(e);
  delete P;
  return 0;
}

int ofdft_batch_create(ofdft_batch* out, const ofdft_plan* plans, int nb) {
  if (!out || !plans) return fail(OFDFT_EINVAL, "null argument");
  *out = nullptr;
  if (nb < 1) return fail(OFDFT_EINVAL, "a batch needs at least one plan");
  for (int c = 0; c < nb; ++c) {
    const ofdft_plan_s* P = plans[c];
    if (!P) return fail(OFDFT_EINVAL, "null plan in the batch");
    if (P->prec != plans[0]->prec) return fail(OFDFT_EINVAL, "the plans of a batch must share one precision");
    const int n = P->n[0];
    if (P->n[1] != n || P->n[2] != n || !(n == 32 || n == 48 || n == 64 || n == 90 || n == 96 || n == 108 ||
                                           n == 128 || n == 160 || n == 192))
      return fail(OFDFT_EINVAL, "batched plans must be cubic with a compile-time length "
                                "(32, 48, 64, 90, 96, 108, 128, 160, 192)");
  }
  auto* B = new (std::nothrow) ofdft_batch_s();
  if (!B) return fail(OFDFT_ENOMEM, "host allocation failed");
  B->plans.assign(plans, plans + nb);
  B->prec = plans[0]->prec;
  *out = B;
  return 0;
}

int ofdft_batch_destroy(ofdft_batch B) {
  if (!B) return 0;
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->cap_stream) (void)hipStreamDestroy(B->cap_stream);
  if (B->d_cells) (void)hipFree(B->d_cells);
  if (B->d_pref) (void)hipFree(B->d_pref);
  if (B->d_idx) (void)hipFree(B->d_idx);
  if (B->h_states) (void)hipHostFree(B->h_states);
  if (B->h_live) (void)hipHostFree(B->h_live);
  delete B;
  return 0;
}

int ofdft_batch_set_graph(ofdft_batch B, int on) {
  if (!B) return fail(OFDFT_EINVAL, "null batch");
  B->use_graph = on ? 1 : 0;
  return 0;
}

int ofdft_batch_tn_direction(ofdft_batch B, const void* const* phi, const void* const* res, const void* const* W0,
                             void* const* d, void* const* work, int maxfun, int* niter, int* status,
                             double* pAp_last) {
  if (!B || !phi || !res || !W0 || !d || !work || !niter || !status || !pAp_last)
    return fail(OFDFT_EINVAL, "null argument");
  if (B->prec == OFDFT_F64)
    return ofdft_tu::tn_batch_run<double>(B, reinterpret_cast<const double* const*>(phi),
                                          reinterpret_cast<const double* const*>(res),
                                          reinterpret_cast<const double* const*>(W0),
                                          reinterpret_cast<double* const*>(d), reinterpret_cast<double* const*>(work),
                                          maxfun, niter, status, pAp_last);
  return ofdft_tu::tn_batch_run<float>(B, reinterpret_cast<const float* const*>(phi),
                                       reinterpret_cast<const float* const*>(res),
                                       reinterpret_cast<const float* const*>(W0), reinterpret_cast<float* const*>(d),
                                       reinterpret_cast<float* const*>(work), maxfun, niter, status, pAp_last);
}

int ofdft_plan_set_tiles(ofdft_plan P, int tile_axis, int tile_pencil) {
  if (!valid(P)) return fail(OFDFT_EINVAL, "null plan");
  auto ok = [&](int w) { return w == 8 || w == 16 || (w == 32 && P->prec == OFDFT_F32); };
  if (!ok(tile_axis) || !ok(tile_pencil))
    return fail(OFDFT_EINVAL, "tile widths must be 8 or 16 (FP64), 8, 16 or 32 (FP32)");
  P->tile_axis = tile_axis;
  P->tile_pencil = tile_pencil;
  if (P->tn_exec) {  // a captured CG iteration holds the old launch shapes
    (void)hipGraphExecDestroy(P->tn_exec);
    P->tn_exec = nullptr;
  }
  return 0;
}

int ofdft_plan_set_cg_spectra(ofdft_plan P, int mode, int tile_axis, int tile_pencil) {
  if (!valid(P)) return fail(OFDFT_EINVAL, "null plan");
  if (mode != OFDFT_CG_SPECTRA_PLAN && mode != OFDFT_CG_SPECTRA_BF16 && mode != OFDFT_CG_SPECTRA_PADDED)
    return fail(OFDFT_EINVAL, "CG spectra mode must be OFDFT_CG_SPECTRA_PLAN, _BF16 or _PADDED");
  auto ok = [](int w) { return w == 0 || w == 8 || w == 16 || w == 32; };
  if (!ok(tile_axis) || !ok(tile_pencil))
    return fail(OFDFT_EINVAL, "bf16 tile widths must be 0 (automatic), 8, 16 or 32");
  const bool change = mode != P->cg_spectra || tile_axis != P->tile_bf_axis || tile_pencil != P->tile_bf_pencil;
  P->cg_spectra = mode;
  P->tile_bf_axis = tile_axis;
  P->tile_bf_pencil = tile_pencil;
  if (change && P->tn_exec) {  // a captured CG iteration holds the old kernels
    (void)hipGraphExecDestroy(P->tn_exec);
    P->tn_exec = nullptr;
  }
  return 0;
}

int ofdft_plan_cg_spectra(ofdft_plan P, int* tile_axis, int* tile_pencil) {
  if (!valid(P)) return fail(OFDFT_EINVAL, "null plan");
  int a = 0, p = 0;
  const int sl = cg_layout(P, &a, &p);
  if (tile_axis) *tile_axis = a;
  if (tile_pencil) *tile_pencil = p;
  return sl == 1 ? OFDFT_CG_SPECTRA_BF16 : (sl == 2 ? OFDFT_CG_SPECTRA_PLAN : OFDFT_CG_SPECTRA_PADDED);
}

int ofdft_plan_set_graph(ofdft_plan P, int on) {
  if (!valid(P)) return fail(OFDFT_EINVAL, "null plan");
  P->use_graph = on ? 1 : 0;
  return 0;
}

int ofdft_plan_set_stream(ofdft_plan P, void* s) {
  if (!valid(P)) return fail(OFDFT_EINVAL, "null plan");
  P->stream = reinterpret_cast<hipStream_t>(s);
  return 0;
}

long long ofdft_plan_workspace_bytes(ofdft_plan P) {
  if (!valid(P)) return -1;
  long long b = 2 * (long long)P->rs * (3 * P->Nhp + P->n[0] + P->n[1] + P->n[2]);
  if (P->wtk) b += (long long)P->rs * P->Nh;
  if (P->wtks) b += (long long)P->rs * P->Nh;
  if (P->vloc) b += (long long)P->rs * P->N;
  b += sizeof(double) * (P->npartials + kMaxSum);
  return b;
}

int ofdft_plan_set_profiling(ofdft_plan P, int on) {
  if (!valid(P)) return fail(OFDFT_EINVAL, "null plan");
  P->prof_on = on & ((1 << PC_NCLS) - 1);
  P->prof_stride = 1 + ((on >> 8) & 0xff);
  for (int i = 0; i < 8; ++i) {
    P->prof_seq[i] = 0;
    P->prof_ms[i] = 0.0;
    P->prof_count[i] = 0;
    P->prof_bytes[i] = 0.0;
  }
  return 0;
}

int ofdft_plan_get_profile(ofdft_plan P, double* ms, long long* count, double* bytes) {
  if (!valid(P) || !ms || !count || !bytes) return fail(OFDFT_EINVAL, "null argument");
  if (!P->prof_pending.empty()) {
    HIPCHK(hipStreamSynchronize(P->stream));
    for (const auto& r : P->prof_pending) {
      float t = 0.f;
      // a scope whose launch did not happen leaves its events unrecorded
      if (hipEventElapsedTime(&t, P->ev_pool[r.e0], P->ev_pool[r.e1]) != hipSuccess) {
        (void)hipGetLastError();
        continue;
      }
      P->prof_ms[r.cls] += t;
      P->prof_count[r.cls] += 1;
      P->prof_bytes[r.cls] += (double)r.bytes;
    }
    P->prof_pending.clear();
    P->ev_used = 0;
  }
  for (int i = 0; i < PC_NCLS; ++i) {
    ms[i] = P->prof_ms[i];
    count[i] = P->prof_count[i];
    bytes[i] = P->prof_bytes[i];
  }
  return 0;
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
    // the split-row copy for the pencils of split-row spectra (n2 even)
    const bool split = P->n[2] % 2 == 0;
    if (split && !P->wtks && hipMalloc(&P->wtks, P->rs * P->Nh) != hipSuccess)
      return fail(OFDFT_ENOMEM, "WT table (split rows)");
    if (P->prec == OFDFT_F64) {
      hipLaunchKernelGGL((k_wt_table<double>), dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream,
                         reinterpret_cast<double*>(P->wtk), g, tkf, factor, f->x);
      if (split)
        hipLaunchKernelGGL((k_wt_split<double>), dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream,
                           reinterpret_cast<const double*>(P->wtk), reinterpret_cast<double*>(P->wtks), g);
    } else {
      hipLaunchKernelGGL((k_wt_table<float>), dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream,
                         reinterpret_cast<float*>(P->wtk), g, tkf, factor, f->x);
      if (split)
        hipLaunchKernelGGL((k_wt_split<float>), dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream,
                           reinterpret_cast<const float*>(P->wtk), reinterpret_cast<float*>(P->wtks), g);
    }
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
    return ofdft_tu::eval<double>(P, (const double*)rho, nullptr, nullptr, 1.0, 0.0, (double*)V, nullptr, sums);
  return ofdft_tu::eval<float>(P, (const float*)rho, nullptr, nullptr, 1.0, 0.0, (float*)V, nullptr, sums);
}

int ofdft_eval_phi(ofdft_plan P, const void* phi, const void* p, double c, double s, void* V, void* aux,
                   double* sums) {
  if (!valid(P) || !phi) return fail(OFDFT_EINVAL, "null argument");
  if (P->prec == OFDFT_F64)
    return ofdft_tu::eval<double>(P, nullptr, (const double*)phi, (const double*)p, c, s, (double*)V, (double*)aux, sums);
  return ofdft_tu::eval<float>(P, nullptr, (const float*)phi, (const float*)p, c, s, (float*)V, (float*)aux, sums);
}

int ofdft_hvp(ofdft_plan P, const void* phi, const void* p, const void* V, const void* aux, double mu, void* Ap,
              double* pAp) {
  if (!valid(P) || !phi || !p || !V || !Ap) return fail(OFDFT_EINVAL, "null argument");
  if ((P->fn.vw || P->fn.wt) && !aux) return fail(OFDFT_EINVAL, "aux fields required for vW/WT");
  if (P->prec == OFDFT_F64)
    return ofdft_tu::hvp<double>(P, (const double*)phi, (const double*)p, (const double*)V, (const double*)aux, mu,
                                 (double*)Ap, pAp, 0);
  return ofdft_tu::hvp<float>(P, (const float*)phi, (const float*)p, (const float*)V, (const float*)aux, mu, (float*)Ap,
                              pAp, 0);
}

int ofdft_potential_jvp(ofdft_plan P, const void* phi, const void* p, const void* V, const void* aux, void* dV) {
  if (!valid(P) || !phi || !p || !V || !dV) return fail(OFDFT_EINVAL, "null argument");
  if ((P->fn.vw || P->fn.wt) && !aux) return fail(OFDFT_EINVAL, "aux fields required for vW/WT");
  if (P->prec == OFDFT_F64)
    return ofdft_tu::hvp<double>(P, (const double*)phi, (const double*)p, (const double*)V, (const double*)aux, 0.0,
                                 (double*)dV, nullptr, 1);
  return ofdft_tu::hvp<float>(P, (const float*)phi, (const float*)p, (const float*)V, (const float*)aux, 0.0,
                              (float*)dV, nullptr, 1);
}

int ofdft_hvp_prepare(ofdft_plan P, const void* phi, const void* V, const void* aux, double mu, void* res,
                      void* W0, double* res2) {
  if (!valid(P) || !phi || !V || !res || !W0) return fail(OFDFT_EINVAL, "null argument");
  if (!P->have_fn) return fail(OFDFT_ESTATE, "functional not set");
  if ((P->fn.vw || P->fn.wt) && !aux) return fail(OFDFT_EINVAL, "aux fields required for vW/WT");
  FnParams fp = make_fn(P->fn);
  const int nb = vec_blocks(P->N);
  {
    ProfScope ps(P, PC_VEC, (long long)P->rs * 5 * P->N);
    if (P->prec == OFDFT_F64)
      OFDFT_LAUNCH((k_hvp_prepare<double>), dim3(nb), dim3(256), 0, P, (const double*)phi,
                         (const double*)V, (const double*)aux, mu, (double*)res, (double*)W0, P->N, fp, P->partials);
    else
      OFDFT_LAUNCH((k_hvp_prepare<float>), dim3(nb), dim3(256), 0, P, (const float*)phi,
                         (const float*)V, (const float*)aux, (float)mu, (float*)res, (float*)W0, P->N, fp,
                         P->partials);
  }
  LAUNCHCHK();
  if (!res2) return 0;
  return reduce_blocks(P, nb, 1, res2);
}

// mapped host slots of the deferred sums: hred[kDeferRes] the sum res^2 of
// ofdft_hvp_prepare_async, hred[kDeferOrth ..] the five sums of
// ofdft_orth_normalize (the other reductions use hred[0 .. 9])
constexpr int kDeferRes = 10;
constexpr int kDeferOrth = 11;
static_assert(kDeferOrth + 5 <= kMaxSum, "deferred sums exceed the mapped slots");

int ofdft_hvp_prepare_async(ofdft_plan P, const void* phi, const void* V, const void* aux, double mu, void* res,
                            void* W0) {
  int rc = ofdft_hvp_prepare(P, phi, V, aux, mu, res, W0, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL((k_reduce_partials<1>), dim3(1), dim3(kRedThreads), 0, P->stream, P->partials,
                     (long long)vec_blocks(P->N), P->red + kDeferRes, P->dhred + kDeferRes, (const int*)nullptr);
  LAUNCHCHK();
  return 0;
}

int ofdft_orth_normalize(ofdft_plan P, void* p, const void* phi, const void* V) {
  if (!valid(P) || !p || !phi || !V) return fail(OFDFT_EINVAL, "null argument");
  const int nb = vec_blocks(P->N);
  if ((long long)nb * 5 > P->npartials) return fail(OFDFT_ESTATE, "partials buffer too small");
  double* coef = P->red + 12;  // after the five sums in red[0..4]
  if (P->prec == OFDFT_F64)
    hipLaunchKernelGGL((k_orth_sums<double>), dim3(nb), dim3(256), 0, P->stream, (const double*)p, (const double*)phi,
                       (const double*)V, P->N, P->partials);
  else
    hipLaunchKernelGGL((k_orth_sums<float>), dim3(nb), dim3(256), 0, P->stream, (const float*)p, (const float*)phi,
                       (const float*)V, P->N, P->partials);
  LAUNCHCHK();
  hipLaunchKernelGGL((k_reduce_partials<5>), dim3(5), dim3(kRedThreads), 0, P->stream, P->partials, (long long)nb,
                     P->red, P->dhred + kDeferOrth, (const int*)nullptr);
  LAUNCHCHK();
  hipLaunchKernelGGL(k_orth_coef, dim3(1), dim3(1), 0, P->stream, (const double*)P->red, coef);
  LAUNCHCHK();
  if (P->prec == OFDFT_F64)
    hipLaunchKernelGGL((k_lincomb_coef<double>), dim3(nb), dim3(256), 0, P->stream, (const double*)coef,
                       (const double*)p, (const double*)phi, (double*)p, P->N);
  else
    hipLaunchKernelGGL((k_lincomb_coef<float>), dim3(nb), dim3(256), 0, P->stream, (const double*)coef,
                       (const float*)p, (const float*)phi, (float*)p, P->N);
  LAUNCHCHK();
  return 0;
}

int ofdft_plan_deferred_sums(ofdft_plan P, double* out6) {
  if (!valid(P) || !out6) return fail(OFDFT_EINVAL, "null argument");
  HIPCHK(hipStreamSynchronize(P->stream));
  out6[0] = P->hred[kDeferRes];
  for (int i = 0; i < 5; ++i) out6[1 + i] = P->hred[kDeferOrth + i];
  return 0;
}

int ofdft_hvp_fast(ofdft_plan P, const void* phi, const void* p, const void* W0, void* Ap, double* pAp) {
  if (!valid(P) || !phi || !p || !W0 || !Ap) return fail(OFDFT_EINVAL, "null argument");
  if (P->prec == OFDFT_F64)
    return ofdft_tu::hvp_fast<double>(P, (const double*)phi, (const double*)p, (const double*)W0, (double*)Ap, pAp);
  return ofdft_tu::hvp_fast<float>(P, (const float*)phi, (const float*)p, (const float*)W0, (float*)Ap, pAp);
}

int ofdft_tn_inner(ofdft_plan P, const void* phi, const void* p, const void* W0, void* Ap, void* d, void* r,
                   double r0, double* out2) {
  if (!valid(P) || !phi || !p || !W0 || !Ap || !d || !r || !out2) return fail(OFDFT_EINVAL, "null argument");
  int rc;
  const int nb = vec_blocks(P->N);
  double* red2 = P->red + 8;  // red[8]: r1 (red[0] holds pAp)
  if (P->prec == OFDFT_F64) {
    if ((rc = ofdft_tu::hvp_fast<double>(P, (const double*)phi, (const double*)p, (const double*)W0, (double*)Ap, nullptr)))
      return rc;
    ProfScope ps(P, PC_VEC, (long long)P->rs * 6 * P->N);
    OFDFT_LAUNCH((k_cg_update_dev<double>), dim3(nb), dim3(256), 0, P, P->red, r0, (const double*)p,
                       (const double*)Ap, (double*)d, (double*)r, P->N, P->partials);
  } else {
    if ((rc = ofdft_tu::hvp_fast<float>(P, (const float*)phi, (const float*)p, (const float*)W0, (float*)Ap, nullptr)))
      return rc;
    ProfScope ps(P, PC_VEC, (long long)P->rs * 6 * P->N);
    OFDFT_LAUNCH((k_cg_update_dev<float>), dim3(nb), dim3(256), 0, P, P->red, r0, (const float*)p,
                       (const float*)Ap, (float*)d, (float*)r, P->N, P->partials);
  }
  LAUNCHCHK();
  hipLaunchKernelGGL((k_reduce_partials<1>), dim3(1), dim3(kRedThreads), 0, P->stream, P->partials, (long long)nb, red2, (double*)nullptr, P->skip);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(P->hred, P->red, sizeof(double) * 9, hipMemcpyDeviceToHost, P->stream));
  HIPCHK(hipStreamSynchronize(P->stream));
  out2[0] = P->hred[0];
  out2[1] = P->hred[8];
  return 0;
}

int ofdft_tn_direction(ofdft_plan P, const void* phi, const void* res, const void* W0, void* d, void* work,
                       int maxfun, int chunk, int* niter, int* status, double* pAp_last) {
  if (!valid(P) || !phi || !res || !W0 || !d || !work || !niter || !status || !pAp_last)
    return fail(OFDFT_EINVAL, "null argument");
  if (maxfun < 1) return fail(OFDFT_EINVAL, "maxfun must be positive");
  if (P->prec == OFDFT_F64)
    return ofdft_tu::tn_direction<double>(P, (const double*)phi, (const double*)res, (const double*)W0, (double*)d,
                                     (double*)work, maxfun, chunk, niter, status, pAp_last);
  return ofdft_tu::tn_direction<float>(P, (const float*)phi, (const float*)res, (const float*)W0, (float*)d, (float*)work,
                                  maxfun, chunk, niter, status, pAp_last);
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


int ofdft_vloc_build(ofdft_plan P, int nspecies, const ofdft_vline* tables, int nat, const double* positions,
                     const int* species, int pme, int order, void* vloc_out) {
  if (!valid(P) || !tables || !positions || !species || !vloc_out) return fail(OFDFT_EINVAL, "null argument");
  if (P->prec != OFDFT_F64) return fail(OFDFT_EINVAL, "V_loc is built on an FP64 plan");
  if (nspecies < 1 || nat < 1) return fail(OFDFT_EINVAL, "need at least one species and one atom");
  if (pme && (order < 3 || order > kMaxOrder)) return fail(OFDFT_EINVAL, "B-spline order must be in [3, 12]");
  for (int i = 0; i < nat; ++i)
    if (species[i] < 0 || species[i] >= nspecies) return fail(OFDFT_EINVAL, "species index out of range");
  for (int s = 0; s < nspecies; ++s)
    if (tables[s].nknots < 8 || !tables[s].knots || !tables[s].coefs) return fail(OFDFT_EINVAL, "bad spline table");
  int rc;
  // spline tables
  long long tot = 0;
  for (int s = 0; s < nspecies; ++s) tot += tables[s].nknots;
  if (P->d_knots) (void)hipFree(P->d_knots);
  if (P->d_coefs) (void)hipFree(P->d_coefs);
  if (P->d_splines) (void)hipFree(P->d_splines);
  P->d_knots = P->d_coefs = nullptr;
  P->d_splines = nullptr;
  HIPCHK(hipMalloc((void**)&P->d_knots, sizeof(double) * tot));
  HIPCHK(hipMalloc((void**)&P->d_coefs, sizeof(double) * tot));
  std::vector<SplineDev> sd(nspecies);
  long long off = 0;
  for (int s = 0; s < nspecies; ++s) {
    const ofdft_vline& t = tables[s];
    HIPCHK(hipMemcpy(P->d_knots + off, t.knots, sizeof(double) * t.nknots, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P->d_coefs + off, t.coefs, sizeof(double) * t.nknots, hipMemcpyHostToDevice));
    sd[s] = SplineDev{P->d_knots + off, P->d_coefs + off, t.nknots, t.q1, t.dq, t.qmax, t.v0, t.v1, t.v2};
    off += t.nknots;
  }
  HIPCHK(hipMalloc(&P->d_splines, sizeof(SplineDev) * nspecies));
  HIPCHK(hipMemcpy(P->d_splines, sd.data(), sizeof(SplineDev) * nspecies, hipMemcpyHostToDevice));
  P->nspecies = nspecies;
  P->pme = pme ? 1 : 0;
  P->order = order;
  if ((rc = upload_ions(P, nat, positions, species))) return rc;
  GridGeom g = geom(P);
  auto* acc = reinterpret_cast<cplx<double>*>(P->spec) + P->Nhp;  // slot 1
  const auto* spl = reinterpret_cast<const SplineDev*>(P->d_splines);
  if (P->pme) {
    if ((rc = ensure_work(P))) return rc;
    if ((rc = ensure_bm(P, order))) return rc;
    HIPCHK(hipMemsetAsync(acc, 0, sizeof(cplx<double>) * P->Nhp, P->stream));
    for (int s = 0; s < nspecies; ++s) {
      HIPCHK(hipMemsetAsync(P->d_work, 0, sizeof(double) * P->N, P->stream));
      hipLaunchKernelGGL(k_pme_spread, dim3(nat), dim3(1024), 0, P->stream, P->d_work, P->d_frac, P->d_species, s,
                         (const double*)nullptr, P->n[0], P->n[1], P->n[2], order);
      LAUNCHCHK();
      if ((rc = r2c_slot(P, P->d_work, 0))) return rc;
      // v += vline * fft(Q) ; v *= B * N/Omega ; V = ifft(v): irfftn(rfftn(Q) vline B N/Omega)
      hipLaunchKernelGGL(k_vloc_accum_pme, dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream, acc,
                         reinterpret_cast<const cplx<double>*>(P->spec), g, spl, s, P->d_bm,
                         (double)P->N / P->volume);
      LAUNCHCHK();
    }
  } else {
    // V = ifft(v) = irfftn(v)/dV
    hipLaunchKernelGGL(k_vloc_exact, dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream, acc, g, spl, P->d_pos,
                       P->d_species, nat, 1.0 / P->dV);
    LAUNCHCHK();
  }
  HIPCHK(hipMemcpyAsync(P->spec, acc, sizeof(cplx<double>) * P->Nhp, hipMemcpyDeviceToDevice, P->stream));
  if ((rc = c2r_slot0(P, (double*)vloc_out))) return rc;
  return 0;
}

int ofdft_local_forces(ofdft_plan P, const void* rho, double* forces_out) {
  if (!valid(P) || !rho || !forces_out) return fail(OFDFT_EINVAL, "null argument");
  if (P->prec != OFDFT_F64) return fail(OFDFT_EINVAL, "forces are computed on an FP64 plan");
  if (!P->d_splines || P->nat < 1) return fail(OFDFT_ESTATE, "call ofdft_vloc_build first");
  if ((long long)3 * P->nat > P->npartials) return fail(OFDFT_EINVAL, "too many atoms for the partials buffer");
  int rc;
  GridGeom g = geom(P);
  const auto* spl = reinterpret_cast<const SplineDev*>(P->d_splines);
  auto* rhog = reinterpret_cast<cplx<double>*>(P->spec) + P->Nhp;  // slot 1 = rfftn(rho)
  if ((rc = r2c_slot(P, (const double*)rho, 1))) return rc;
  std::vector<double> F(3 * (size_t)P->nat, 0.0);
  if (!P->pme) {
    hipLaunchKernelGGL(k_force_exact, dim3(P->nat), dim3(256), 0, P->stream, rhog, g, spl, P->d_pos, P->d_species,
                       P->dV, P->partials);
    LAUNCHCHK();
    std::vector<double> h(3 * (size_t)P->nat);
    HIPCHK(hipMemcpyAsync(h.data(), P->partials, sizeof(double) * h.size(), hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    for (size_t i = 0; i < h.size(); ++i) F[i] = h[i] * 2.0 / P->volume;
  } else {
    if ((rc = ensure_work(P))) return rc;
    if (!P->d_bm) return fail(OFDFT_ESTATE, "PME tables missing");
    double linv[9];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) linv[a * 3 + b] = P->bg[b * 3 + a] / (2.0 * kPi);  // inv(L)
    double* d_linv = P->red;  // 9 doubles of scratch (kMaxSum >= 9)
    HIPCHK(hipMemcpyAsync(d_linv, linv, sizeof(linv), hipMemcpyHostToDevice, P->stream));
    std::vector<double> h(3 * (size_t)P->nat);
    for (int s = 0; s < P->nspecies; ++s) {
      // rhoPB = ifft(fft(rho) conj(B) v_s) = irfftn(rfftn(rho) conj(B) v_s)
      hipLaunchKernelGGL(k_force_pme_mult, dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream,
                         reinterpret_cast<cplx<double>*>(P->spec), rhog, g, spl, s, P->d_bm);
      LAUNCHCHK();
      if ((rc = c2r_slot0(P, P->d_work))) return rc;
      HIPCHK(hipMemsetAsync(P->partials, 0, sizeof(double) * 3 * P->nat, P->stream));
      hipLaunchKernelGGL(k_force_pme_gather, dim3(P->nat), dim3(1024), 0, P->stream, P->d_work, P->d_frac,
                         P->d_species, s, P->n[0], P->n[1], P->n[2], P->order, d_linv, P->partials);
      LAUNCHCHK();
      HIPCHK(hipMemcpyAsync(h.data(), P->partials, sizeof(double) * h.size(), hipMemcpyDeviceToHost, P->stream));
      HIPCHK(hipStreamSynchronize(P->stream));
      for (int i = 0; i < P->nat; ++i) {
        // partials of other species' atoms were zeroed and skipped
        F[3 * i] += h[3 * i];
        F[3 * i + 1] += h[3 * i + 1];
        F[3 * i + 2] += h[3 * i + 2];
      }
    }
  }
  for (size_t i = 0; i < F.size(); ++i) forces_out[i] = F[i];
  return 0;
}

int ofdft_ewald(ofdft_plan P, int nat, const double* positions, const double* charges, double eta, double rmax,
                int pme, int order, double* e_real, double* e_rec) {
  if (!valid(P) || !positions || !charges || !e_real || !e_rec) return fail(OFDFT_EINVAL, "null argument");
  int rc;
  if ((rc = ewald_check(P, nat, eta, rmax, pme, order))) return rc;
  EwaldDev E;
  if ((rc = ewald_upload(P, nat, positions, charges, rmax, E))) return rc;
  if (pme && ((rc = ensure_work(P)) || (rc = ensure_bm(P, order)))) return rc;
  return ewald_energy_sums(P, EwaldBufs{E.pos, E.q, E.sh, E.frac, E.nshift, P->d_work, P->d_bm}, nat, eta, rmax, pme,
                           order, e_real, e_rec);
}

int ofdft_ewald_noalloc(ofdft_plan P, int nat, const double* positions, const double* charges, double eta,
                        double rmax, int pme, int order, double* e_real, double* e_rec) {
  if (!valid(P) || !positions || !charges || !e_real || !e_rec) return fail(OFDFT_EINVAL, "null argument");
  int rc;
  if ((rc = ewald_check(P, nat, eta, rmax, pme, order))) return rc;
  std::vector<double> sh, frac, bm;
  ewald_host_arrays(P, nat, positions, rmax, sh, frac);
  if (pme) bm = make_bm(P, order);
  // layout: pos | q | frac | shifts | [charge grid | b-moduli], 256-byte aligned
  const size_t o_pos = 0;
  const size_t o_q = o_pos + align256(sizeof(double) * 3 * nat);
  const size_t o_frac = o_q + align256(sizeof(double) * nat);
  const size_t o_sh = o_frac + align256(sizeof(double) * 3 * nat);
  const size_t o_work = o_sh + align256(sizeof(double) * sh.size());
  const size_t o_bm = o_work + (pme ? align256(sizeof(double) * P->N) : 0);
  const size_t total = o_bm + align256(sizeof(double) * bm.size());
  if (total > kScratchBytes)
    return fail(OFDFT_EINVAL, "ofdft_ewald_noalloc needs " + std::to_string(total) + " bytes of scratch (" +
                                  std::to_string(nat) + " atoms, " + std::to_string(sh.size() / 3) +
                                  " image shifts" + (pme ? ", PME grid of " + std::to_string(P->N) + " points" : "") +
                                  "); its static array holds " + std::to_string(kScratchBytes) + " bytes");
  unsigned char* base = nullptr;
  if ((rc = scratch_address(P->stream, &base))) return rc;
  auto* pos = reinterpret_cast<double*>(base + o_pos);
  auto* q = reinterpret_cast<double*>(base + o_q);
  auto* fr = reinterpret_cast<double*>(base + o_frac);
  auto* shd = reinterpret_cast<double*>(base + o_sh);
  double* work = pme ? reinterpret_cast<double*>(base + o_work) : nullptr;
  double* bmd = pme ? reinterpret_cast<double*>(base + o_bm) : nullptr;
  HIPCHK(hipMemcpyAsync(pos, positions, sizeof(double) * 3 * nat, hipMemcpyHostToDevice, P->stream));
  HIPCHK(hipMemcpyAsync(q, charges, sizeof(double) * nat, hipMemcpyHostToDevice, P->stream));
  HIPCHK(hipMemcpyAsync(fr, frac.data(), sizeof(double) * frac.size(), hipMemcpyHostToDevice, P->stream));
  HIPCHK(hipMemcpyAsync(shd, sh.data(), sizeof(double) * sh.size(), hipMemcpyHostToDevice, P->stream));
  if (pme) HIPCHK(hipMemcpyAsync(bmd, bm.data(), sizeof(double) * bm.size(), hipMemcpyHostToDevice, P->stream));
  // the sums read their results back on the plan stream; on an error the
  // stream is drained here, so the host arrays above outlive every copy
  rc = ewald_energy_sums(P, EwaldBufs{pos, q, shd, fr, (int)(sh.size() / 3), work, bmd}, nat, eta, rmax, pme, order,
                         e_real, e_rec);
  if (rc) (void)hipStreamSynchronize(P->stream);
  return rc;
}

int ofdft_ewald_forces(ofdft_plan P, int nat, const double* positions, const double* charges, double eta,
                       double rmax, int pme, int order, double* forces_out) {
  if (!valid(P) || !positions || !charges || !forces_out) return fail(OFDFT_EINVAL, "null argument");
  int rc;
  if ((rc = ewald_check(P, nat, eta, rmax, pme, order))) return rc;
  if ((long long)3 * nat > P->npartials) return fail(OFDFT_EINVAL, "too many atoms for the partials buffer");
  EwaldDev E;
  if ((rc = ewald_upload(P, nat, positions, charges, rmax, E))) return rc;
  std::vector<double> h(3 * (size_t)nat);
  // real space
  hipLaunchKernelGGL(k_ewald_force_real, dim3(nat), dim3(256), 0, P->stream, E.pos, E.q, nat, E.sh, E.nshift, rmax,
                     std::sqrt(eta), P->partials);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(h.data(), P->partials, sizeof(double) * h.size(), hipMemcpyDeviceToHost, P->stream));
  HIPCHK(hipStreamSynchronize(P->stream));
  const double es3 = eta * std::sqrt(eta);
  for (int i = 0; i < nat; ++i)
    for (int d = 0; d < 3; ++d) forces_out[3 * i + d] = h[3 * i + d] * charges[i] * es3;
  // reciprocal space
  GridGeom g = geom(P);
  auto* spec0 = reinterpret_cast<cplx<double>*>(P->spec);
  if (!pme) {
    hipLaunchKernelGGL(k_ewald_strf, dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream, spec0, g, E.pos, E.q, nat);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_ewald_force_rec, dim3(nat), dim3(256), 0, P->stream, spec0, g, E.pos, eta, P->partials);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(h.data(), P->partials, sizeof(double) * h.size(), hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    for (int i = 0; i < nat; ++i)
      for (int d = 0; d < 3; ++d) forces_out[3 * i + d] += h[3 * i + d] * charges[i] * 8.0 * kPi / P->volume;
  } else {
    if ((rc = ewald_qhat(P, E, nat, order))) return rc;
    hipLaunchKernelGGL(k_ewald_pme_mult, dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream, spec0, g, P->d_bm, eta);
    LAUNCHCHK();
    if ((rc = c2r_slot0(P, P->d_work))) return rc;
    double linv[9];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) linv[a * 3 + b] = P->bg[b * 3 + a] / (2.0 * kPi);  // inv(L)
    double* d_linv = P->red;  // 9 doubles of scratch (kMaxSum >= 9)
    HIPCHK(hipMemcpyAsync(d_linv, linv, sizeof(linv), hipMemcpyHostToDevice, P->stream));
    hipLaunchKernelGGL(k_force_pme_gather, dim3(nat), dim3(1024), 0, P->stream, P->d_work, E.frac, (const int*)nullptr,
                       -1, P->n[0], P->n[1], P->n[2], order, d_linv, P->partials);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(h.data(), P->partials, sizeof(double) * h.size(), hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    for (int i = 0; i < nat; ++i)
      for (int d = 0; d < 3; ++d) forces_out[3 * i + d] += h[3 * i + d] * charges[i] * 4.0 * kPi / P->dV;
  }
  return 0;
}

int ofdft_functional_stress(ofdft_plan P, const void* rho, double tkf, double* out) {
  if (!valid(P) || !rho || !out) return fail(OFDFT_EINVAL, "null argument");
  if (P->prec != OFDFT_F64) return fail(OFDFT_EINVAL, "stress is computed on an FP64 plan");
  if (!P->have_fn) return fail(OFDFT_ESTATE, "functional not set");
  const ofdft_functional& f = P->fn;
  if (f.wt && !(tkf > 0.0)) return fail(OFDFT_EINVAL, "tkf must be positive");
  for (int i = 0; i < OFDFT_NSTRESS; ++i) out[i] = 0.0;
  const double* r = (const double*)rho;
  auto* spec = reinterpret_cast<cplx<double>*>(P->spec);
  StressArgs A{};
  int rc;
  if (f.hartree) {
    if ((rc = r2c_slot(P, r, 0))) return rc;
    A.a = spec;
    if ((rc = stress_recip(P, A, kStHartree, out))) return rc;
  }
  if (f.vw) {
    if ((rc = r2c_slot(P, r, 0))) return rc;
    hipLaunchKernelGGL((k_grad_spec<double>), dim3(vec_blocks(P->Nh)), dim3(256), 0, P->stream, spec, P->Nhp,
                       geom(P), 1.0 / (double)P->N);
    LAUNCHCHK();
    if ((rc = launch_axis<double>(P, spec, 0, 3, true))) return rc;
    if ((rc = launch_axis<double>(P, spec, 1, 3, true))) return rc;
    PostStressVW<double> post{r};
    if ((rc = launch_rows_inv<double>(P, post, 3, spec, true))) return rc;
    if ((rc = fetch_sums(P, 6, out + 6))) return rc;
  }
  if (f.wt) {
    if ((rc = r2c_pre_slot(P, PrePow<double>{r, f.alpha}, 0))) return rc;
    const bool same = f.beta == f.alpha;
    if (!same && (rc = r2c_pre_slot(P, PrePow<double>{r, f.beta}, 1))) return rc;
    A.a = spec;
    A.b = same ? spec : spec + P->Nhp;
    A.tkf = tkf;
    A.mu = f.y;
    if ((rc = stress_recip(P, A, kStWT, out + 12))) return rc;
  }
  if (f.lda) {
    const int nb = vec_blocks(P->N);
    hipLaunchKernelGGL(k_lda_vrho, dim3(nb), dim3(256), 0, P->stream, r, P->N, P->partials);
    LAUNCHCHK();
    if ((rc = reduce_blocks(P, nb, 1, out + 18))) return rc;
  }
  return 0;
}

int ofdft_local_stress(ofdft_plan P, const void* rho, double* out6) {
  if (!valid(P) || !rho || !out6) return fail(OFDFT_EINVAL, "null argument");
  if (P->prec != OFDFT_F64) return fail(OFDFT_EINVAL, "stress is computed on an FP64 plan");
  if (!P->d_splines || P->nat < 1) return fail(OFDFT_ESTATE, "call ofdft_vloc_build first");
  int rc;
  auto* spec = reinterpret_cast<cplx<double>*>(P->spec);
  auto* rhog = spec + P->Nhp;  // slot 1 = rfftn(rho)
  if ((rc = r2c_slot(P, (const double*)rho, 1))) return rc;
  StressArgs A{};
  A.a = rhog;
  A.spl = reinterpret_cast<const SplineDev*>(P->d_splines);
  A.nspecies = P->nspecies;
  for (int k = 0; k < 6; ++k) out6[k] = 0.0;
  if (!P->pme) {
    A.pos = P->d_pos;
    A.species = P->d_species;
    A.nat = P->nat;
    return stress_recip(P, A, kStPPExact, out6);
  }
  if ((rc = ensure_work(P))) return rc;
  if (!P->d_bm) return fail(OFDFT_ESTATE, "PME tables missing");
  A.bm = P->d_bm;
  for (int s = 0; s < P->nspecies; ++s) {
    // Q_s: the unit-weight B-spline spread of the atoms of species s
    HIPCHK(hipMemsetAsync(P->d_work, 0, sizeof(double) * P->N, P->stream));
    hipLaunchKernelGGL(k_pme_spread, dim3(P->nat), dim3(1024), 0, P->stream, P->d_work, P->d_frac, P->d_species, s,
                       (const double*)nullptr, P->n[0], P->n[1], P->n[2], P->order);
    LAUNCHCHK();
    if ((rc = r2c_slot(P, P->d_work, 0))) return rc;
    A.b = spec;
    A.s = s;
    double t[6];
    if ((rc = stress_recip(P, A, kStPPPme, t))) return rc;
    for (int k = 0; k < 6; ++k) out6[k] += t[k];
  }
  return 0;
}

int ofdft_ewald_stress(ofdft_plan P, int nat, const double* positions, const double* charges, double eta,
                       double rmax, int pme, int order, double* real6, double* rec6) {
  if (!valid(P) || !positions || !charges || !real6 || !rec6) return fail(OFDFT_EINVAL, "null argument");
  int rc;
  if ((rc = ewald_check(P, nat, eta, rmax, pme, order))) return rc;
  EwaldDev E;
  if ((rc = ewald_upload(P, nat, positions, charges, rmax, E))) return rc;
  const int nb = vec_blocks((long long)nat * nat);
  if ((long long)nb * 6 > P->npartials) return fail(OFDFT_ESTATE, "partials buffer too small");
  hipLaunchKernelGGL(k_ewald_stress_real, dim3(nb), dim3(256), 0, P->stream, E.pos, E.q, nat, E.sh, E.nshift, rmax,
                     eta, P->partials);
  if (hipGetLastError() != hipSuccess) return fail(OFDFT_EHIP, "Ewald stress launch");
  if ((rc = reduce_blocks(P, nb, 6, real6))) return rc;
  StressArgs A{};
  A.eta = eta;
  if (pme) {
    if ((rc = ewald_qhat(P, E, nat, order))) return rc;
    A.a = reinterpret_cast<const cplx<double>*>(P->spec);
    A.bm = P->d_bm;
  } else {
    A.pos = E.pos;
    A.q = E.q;
    A.nat = nat;
  }
  return stress_recip(P, A, kStEwald, rec6);
}

int ofdft_gga(ofdft_plan P, int functional, const double* params, int nparams, const void* rho, const void* phi,
              const void* p, double c, double s, void* V_inout, double* out) {
  if (!valid(P) || !out || (!rho && !phi)) return fail(OFDFT_EINVAL, "null argument");
  if (functional < 0 || functional >= OFDFT_GGA_COUNT) return fail(OFDFT_EINVAL, "unknown GGA functional");
  if (nparams < 0 || nparams > 8 || (nparams > 0 && !params)) return fail(OFDFT_EINVAL, "bad GGA parameters");
  if (P->prec == OFDFT_F64)
    return ofdft_tu::gga<double>(P, functional, params, nparams, (const double*)rho, (const double*)phi,
                                 (const double*)p, c, s, (double*)V_inout, out);
  return ofdft_tu::gga<float>(P, functional, params, nparams, (const float*)rho, (const float*)phi,
                              (const float*)p, c, s, (float*)V_inout, out);
}


int ofdft_lwt_set_kernel(ofdft_plan P, const ofdft_lwt_table* K, const ofdft_lwt_table* D) {
  if (!valid(P) || !K || !D) return fail(OFDFT_EINVAL, "null argument");
  const ofdft_lwt_table* in[2] = {K, D};
  size_t tot = 0;
  for (int i = 0; i < 2; ++i) {
    const ofdft_lwt_table& t = *in[i];
    if (t.kind == 0) {
      if (t.n < 8 || !t.knots || !t.coefs || !(t.dx > 0.0)) return fail(OFDFT_EINVAL, "bad LWT spline table");
      tot += 2 * (size_t)t.n;
    } else if (t.kind == 1) {
      if (t.n < 2 || !t.coefs || !(t.dx > 0.0)) return fail(OFDFT_EINVAL, "bad LWT table");
      // eta < etamax reads entry rint(eta / dx): it must stay below n
      if (t.etamax / t.dx > (double)t.n - 0.5) return fail(OFDFT_EINVAL, "LWT table shorter than etamax / dx");
      tot += (size_t)t.n;
    } else {
      return fail(OFDFT_EINVAL, "unknown LWT table kind");
    }
    if (!(t.etamax > 0.0)) return fail(OFDFT_EINVAL, "bad LWT etamax");
  }
  if (P->d_lwttab) {
    HIPCHK(hipStreamSynchronize(P->stream));
    (void)hipFree(P->d_lwttab);
    P->d_lwttab = nullptr;
  }
  HIPCHK(hipMalloc(&P->d_lwttab, tot * sizeof(double)));
  double* cur = P->d_lwttab;
  for (int i = 0; i < 2; ++i) {
    ofdft_lwt_table t = *in[i];
    if (t.kind == 0) {
      HIPCHK(hipMemcpy(cur, t.knots, sizeof(double) * t.n, hipMemcpyHostToDevice));
      t.knots = cur;
      cur += t.n;
    } else {
      t.knots = nullptr;
    }
    HIPCHK(hipMemcpy(cur, t.coefs, sizeof(double) * t.n, hipMemcpyHostToDevice));
    t.coefs = cur;
    cur += t.n;
    P->lwt_tab[i] = t;
  }
  return 0;
}

int ofdft_lwt_range(ofdft_plan P, const void* rho, const void* phi, const void* p, double c, double s,
                    double* out2) {
  if (!valid(P) || !out2 || (!rho && !phi)) return fail(OFDFT_EINVAL, "null argument");
  if (P->prec == OFDFT_F64)
    return lwt_range_impl<double>(P, (const double*)rho, (const double*)phi, (const double*)p, c, s, out2);
  return lwt_range_impl<float>(P, (const float*)rho, (const float*)phi, (const float*)p, c, s, out2);
}

int ofdft_lwt(ofdft_plan P, const ofdft_lwt_params* prm, const double* kflist, int nsp, const void* rho,
              const void* phi, const void* p, double c, double s, void* V_inout, double* out3) {
  if (!valid(P) || !prm || !kflist || !out3 || (!rho && !phi)) return fail(OFDFT_EINVAL, "null argument");
  if (!P->d_lwttab) return fail(OFDFT_ESTATE, "LWT kernel tables not set");
  if (nsp < 2) return fail(OFDFT_EINVAL, "LWT needs at least two kf nodes");
  for (int j = 0; j < nsp; ++j)
    if (!(kflist[j] > 0.0) || (j > 0 && !(kflist[j] > kflist[j - 1])))
      return fail(OFDFT_EINVAL, "LWT kf nodes must be positive and increasing");
  if (prm->kdd < 1 || prm->kdd > 3) return fail(OFDFT_EINVAL, "LWT kdd must be 1, 2 or 3");
  if (!(prm->rhomax_ld > 0.0)) return fail(OFDFT_EINVAL, "LWT rhomax_ld must be positive");
  if (P->prec == OFDFT_F64)
    return ofdft_tu::lwt<double>(P, *prm, kflist, nsp, (const double*)rho, (const double*)phi, (const double*)p, c, s,
                            (double*)V_inout, out3);
  return ofdft_tu::lwt<float>(P, *prm, kflist, nsp, (const float*)rho, (const float*)phi, (const float*)p, c, s,
                         (float*)V_inout, out3);
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

int ofdft_orth_sums(ofdft_plan P, const void* p, const void* phi, const void* V, double* out5) {
  if (!valid(P) || !p || !phi || !V || !out5) return fail(OFDFT_EINVAL, "null argument");
  const int nb = vec_blocks(P->N);
  if ((long long)nb * 5 > P->npartials) return fail(OFDFT_ESTATE, "partials buffer too small");
  if (P->prec == OFDFT_F64)
    hipLaunchKernelGGL((k_orth_sums<double>), dim3(nb), dim3(256), 0, P->stream, (const double*)p, (const double*)phi,
                       (const double*)V, P->N, P->partials);
  else
    hipLaunchKernelGGL((k_orth_sums<float>), dim3(nb), dim3(256), 0, P->stream, (const float*)p, (const float*)phi,
                       (const float*)V, P->N, P->partials);
  LAUNCHCHK();
  return reduce_blocks(P, nb, 5, out5);
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

