/*
 * ofdft_amd.h -- C ABI of the MI355X (gfx950) orbital-free DFT hot path.
 *
 * The library replaces the numpy arithmetic under DFTpy's functional and
 * optimizer API (the reference is pure Python; there is no FFI in the
 * reference, so every entry point below names the Python function whose
 * arithmetic it performs, as reference file:line under src/dftpy).
 *
 * Conventions
 *  - A plan is bound to one real grid n0 x n1 x n2 (C order, index
 *    (i0*n1+i1)*n2+i2), one lattice (row-major 3x3, rows = lattice vectors in
 *    bohr), one precision and one HIP stream.  The half-complex layout is
 *    n0 x n1 x (n2/2+1) interleaved complex (numpy rfftn layout).
 *  - Precision OFDFT_F64: every device array is double / double2.
 *    OFDFT_F32: fields and spectra are float / float2; every energy, norm
 *    and dot product is accumulated in FP64.
 *  - All device pointers are owned by the caller (torch tensors); the plan
 *    owns only its workspace and tables.  Calls are asynchronous on the plan
 *    stream except those that return host scalars (they synchronise the
 *    plan stream before returning).
 *  - Every call returns 0 on success or a negative OFDFT_E* code;
 *    ofdft_last_error() gives the text of the calling thread's last error.
 *    No C++ exception crosses this boundary.
 *  - One plan must not be used from several host threads at once.
 */
#ifndef OFDFT_AMD_H
#define OFDFT_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define OFDFT_F64 0
#define OFDFT_F32 1

#define OFDFT_OK 0
#define OFDFT_EINVAL (-1)
#define OFDFT_EHIP (-2)
#define OFDFT_ENOMEM (-3)
#define OFDFT_ESTATE (-4)

/* Slots of the FP64 result array filled by ofdft_eval / ofdft_eval_phi.
 * Energies are final (dV and functional fractions applied).            */
#define OFDFT_E_TF 0      /* KEDF-TF            kedf/tf.py:66-80           */
#define OFDFT_E_VW 1      /* KEDF-VW            kedf/vw.py:84-94           */
#define OFDFT_E_NL 2      /* KEDF-NL (WT)       kedf/wt.py:123-130         */
#define OFDFT_E_XC 3      /* XC (LDA)           xc/semilocal_xc.py:302-307 */
#define OFDFT_E_H 4       /* HARTREE            hartree.py:37-38           */
#define OFDFT_E_PP 5      /* PSEUDO             pseudo/__init__.py:223-228 */
#define OFDFT_S_VRHO 6    /* integral V*rho dV (optimization.py:542-544)   */
#define OFDFT_S_RHO 7     /* integral rho dV                               */
#define OFDFT_S_DTHETA 8  /* 2*integral V*phi(t)*dphi/dt dV (optimization.py:270-277) */
#define OFDFT_NSUM 9

/* Outcome of ofdft_tn_direction, the `stat` strings of
 * get_direction_TN (optimization.py:102-142).                           */
#define OFDFT_TN_CONV 0       /* "CONV": r1 < 0.1 r0                         */
#define OFDFT_TN_WARN 1       /* "WARN": pAp < 0 at the first iteration      */
#define OFDFT_TN_FAILED 2     /* "FAILED": pAp < 0 later                     */
#define OFDFT_TN_NOTREDUCE 3  /* "WARN : Not reduce": r1 > 1000 min(r)       */
#define OFDFT_TN_SMALL 4      /* "WARN : Change too small"                   */
#define OFDFT_TN_MAXFUN 5     /* "NOTCONV": maxfun iterations taken          */

typedef struct ofdft_plan_s* ofdft_plan;

/* Functional selection and options.  Replaces the construction of
 * TotalFunctional(KEDF, XC, HARTREE, PSEUDO) (total_functional.py:49-78)
 * with KEDF name expansion kedf/__init__.py:27-69 and options :124-164.   */
typedef struct {
  int tf;          /* include Thomas-Fermi (kedf/tf.py:66)                */
  int vw;          /* include von Weizsacker (kedf/vw.py:115)             */
  int wt;          /* include Wang-Teter nonlocal term (kedf/wt.py:109)   */
  int lda;         /* include Slater+PZ LDA (xc/semilocal_xc.py:283)      */
  int hartree;     /* include Hartree (hartree.py:24)                     */
  int pseudo;      /* include the local pseudopotential (needs vloc)      */
  double x;        /* TF fraction; also lambda of the WT kernel (kernel.py:333) */
  double y;        /* vW fraction                                          */
  double alpha;    /* WT exponents; alpha must equal beta                  */
  double beta;
  double rho0;     /* WT kernel density (wt.py:58-73); <=0: kernel unset   */
} ofdft_functional;

/* Plan lifetime.  Replaces DirectGrid/ReciprocalGrid set-up
 * (grid.py:26-57, :384-415) and the FFT plan cache (math_utils.py:94-155).
 * full=0: rfftn/irfftn semantics (grid full=False, field.py:166-170).
 * full=1: the reference's full complex grid (grid.py:302 default), i.e.
 * Re(ifftn(M * fftn(f))); realised on the half spectrum with the
 * Hermitian-symmetrised multiplier (M(G_k) + M(G_-k))/2.
 * Grid lengths: products of 2, 3, 5, 7, 11 and 13 (at most 12 passes), with
 * the LDS bound of the transforms (160 KiB per workgroup on gfx950): FP64
 * n0, n1 <= 601 and n2 <= 1462; FP32 n0, n1 <= 1203 and n2 <= 2925.
 * Anything else returns OFDFT_EINVAL before any device call.              */
int ofdft_plan_create(ofdft_plan* out, int n0, int n1, int n2, const double* lattice9, int precision, int full,
                      void* hip_stream);
int ofdft_plan_destroy(ofdft_plan plan);
/* Execution mode of ofdft_tn_direction's inner CG loop (the loop of
 * get_direction_TN, optimization.py:92-148): on=1 captures one CG iteration
 * (its kernel launches only) into a hipGraph at the first call with a given
 * set of buffers and launches that graph once per queued iteration instead
 * of the individual kernels; on=0 (default) launches the kernels.  Results
 * are identical; profiling (ofdft_plan_set_profiling) uses the launches.  */
int ofdft_plan_set_graph(ofdft_plan plan, int on);
/* Columns per workgroup of the compile-time-length column kernels: the
 * axis-1 pass and the axis-0 pencils (forward transform, multiplier,
 * inverse).  8 or 16 (FP64), 8, 16 or 32 (FP32); default 16.  The values
 * the project's autotuner measured fastest per (length, precision) are
 * applied by dftpy_amd at plan creation.  Results do not depend on them.
 * No counterpart in the reference (numpy.fft has no tiling).              */
int ofdft_plan_set_tiles(ofdft_plan plan, int tile_axis, int tile_pencil);
/* Layout and storage of the spectra between the transform passes of the
 * Hessian products in ofdft_tn_direction's inner CG loop (the Hessian
 * product of get_direction_TN, optimization.py:92-148):
 *  OFDFT_CG_SPECTRA_PLAN (default): plan precision in split rows -- n2/2
 *    columns per row, the Nyquist column in a plane of its own -- where the
 *    plan allows them (compile-time kernels on all three lengths, n2 even,
 *    n2/2 a multiple of the plan's column tile widths), else padded rows;
 *  OFDFT_CG_SPECTRA_BF16: split rows of bf16 pairs (each component rounded
 *    to nearest even from FP32; every transform, multiplier and pointwise
 *    step still computes in FP32, sums in FP64), halving their HBM traffic,
 *    on FP32 plans where n2/2 is a multiple of both bf16 tile widths; the
 *    directions then differ from the FP32 ones by the rounding (relative
 *    2^-9 per stored value); other plans fall back as for _PLAN;
 *  OFDFT_CG_SPECTRA_PADDED: plan precision in the padded rows of the other
 *    passes (n2/2+1 columns, row stride a multiple of 16).
 * tile_axis, tile_pencil: columns per workgroup of the bf16 axis pass and
 * pencils, 8, 16 or 32, or 0 for the widest dividing n2/2 (the other modes
 * use the plan's tiles).  The mode also sets the layout of the five passes
 * of every evaluation of this plan (ofdft_eval, ofdft_eval_phi, the
 * line-search trials): split rows in the plan precision unless the mode is
 * OFDFT_CG_SPECTRA_PADDED or the plan does not allow them; bf16 storage is
 * used by the inner CG loop only.  Nothing is allocated: the layouts
 * live in the plan's spectrum workspace.  No counterpart in the reference
 * (its FFTs run in FP64 through numpy).                                    */
#define OFDFT_CG_SPECTRA_PLAN 0
#define OFDFT_CG_SPECTRA_BF16 1
#define OFDFT_CG_SPECTRA_PADDED 2
int ofdft_plan_set_cg_spectra(ofdft_plan plan, int mode, int tile_axis, int tile_pencil);
/* The layout the next inner CG loop of this plan uses (one of the
 * OFDFT_CG_SPECTRA_* above, after the fall-backs), with its column tile
 * widths in *tile_axis and *tile_pencil when those are not NULL (0 for
 * padded rows); OFDFT_EINVAL for a null plan.                              */
int ofdft_plan_cg_spectra(ofdft_plan plan, int* tile_axis, int* tile_pencil);

/* Batched TN inner loop over many independent cells (the C5 batching form;
 * the reference solves each cell with its own optimize_rho,
 * optimization.py:283-525, whose get_direction_TN :92-148 this runs for
 * all cells at once).  A batch is a list of plans -- one per cell, each
 * with its own grid length, lattice, functional, WT table and V_loc -- that
 * share one precision and one stream; lengths must be cubic and in the
 * compile-time set 32, 48, 64, 90, 96, 108, 128, 160, 192.
 * ofdft_batch_tn_direction runs ofdft_tn_direction for every cell c
 * (phi[c], res[c], W0[c], d[c], work[c]: the same device arrays) for all
 * cells at once, in the form ofdft_batch_set_graph selects, until every
 * cell's loop has ended.  A null phi[c] leaves cell c out of the call (its
 * other pointers may be null too; it reports niter 0): a caller keeps one
 * batch while cells converge.  niter, status, pAp_last: host arrays of nb.
 * The batch keeps pointers to the plans: destroy the batch before them.
 * ofdft_batch_create validates the plans (host only).  */
typedef struct ofdft_batch_s* ofdft_batch;
int ofdft_batch_create(ofdft_batch* out, const ofdft_plan* plans, int nb);
int ofdft_batch_destroy(ofdft_batch batch);
/* How ofdft_batch_tn_direction runs (no reference counterpart: a launch
 * form of the same computation).  on = 1 (the default): the graph form --
 * the cells are grouped by length, each stage of a CG iteration is one
 * launch per length with the one-cell kernels' workgroup shapes and tile
 * widths, and one CG iteration of every cell is a captured hipGraph
 * replayed until every loop has ended (the host checks every few
 * iterations).  on = 0: the persistent form -- ONE cooperative launch whose
 * stages are work queues over (cell, tile) items separated by grid
 * barriers.  Same results either way (up to the order of FP64 sums).  */
int ofdft_batch_set_graph(ofdft_batch batch, int on);
int ofdft_batch_tn_direction(ofdft_batch batch, const void* const* phi, const void* const* res,
                             const void* const* W0, void* const* d, void* const* work, int maxfun, int* niter,
                             int* status, double* pAp_last);
int ofdft_plan_set_stream(ofdft_plan plan, void* hip_stream);
/* Device bytes of workspace the plan owns. */
long long ofdft_plan_workspace_bytes(ofdft_plan plan);

/* Per-kernel-class timing with HIP events recorded on the plan stream
 * around every launch (no reference counterpart: it replaces the
 * wall-clock @timer labels of time_data.py:6-122 with device time).
 * Classes: 0 rows_fwd (R2C + evaluation pre-pointwise), 1 rows_fwd_hvp
 * (R2C + HVP pre-pointwise), 2 axis (strided C2C), 3 pencil (axis-0 C2C +
 * multiplier + inverse), 4 rows_inv (C2R + evaluation post), 5
 * rows_inv_hvp (C2R + HVP post), 6 vector kernels.  get_profile
 * synchronises and returns, per class (arrays of OFDFT_NPROF), total
 * milliseconds, launch count and the algorithmic
 * bytes (each launch's compulsory reads + writes) summed over launches.
 * set_profiling resets the counters; `mask` selects the classes timed
 * (bit c for class c, 0 = off, 127 = all), so that a caller can time one
 * kernel class without adding events around every other launch; bits
 * 8-15 hold k - 1 to time one launch in every k of each selected class
 * (the first, the (k+1)-th, ...), which keeps the timing's own dispatch
 * overhead off most launches.                                          */
#define OFDFT_NPROF 7
int ofdft_plan_set_profiling(ofdft_plan plan, int mask);
int ofdft_plan_get_profile(ofdft_plan plan, double* ms, long long* count, double* bytes);

/* Functional options; builds the WT kernel table on the device when
 * f->wt (get_ke_kernel_wt, kedf/wt.py:58-73; WTKernel kernel.py:325-334;
 * LindhardFunction kernel.py:14-136).                                      */
int ofdft_set_functional(ofdft_plan plan, const ofdft_functional* f);

/* Local pseudopotential V_loc(r): copies n0*n1*n2 values (plan precision)
 * into plan storage (LocalPseudo._vreal, pseudo/__init__.py:219-246).      */
int ofdft_set_vloc(ofdft_plan plan, const void* vloc_dev);

/* One energy+potential evaluation on rho (TotalFunctional.compute,
 * total_functional.py:107-146).  V_out (device, may be NULL) receives the
 * total potential; sums_out (host, OFDFT_NSUM doubles) receives the
 * energies.  Synchronises the stream.                                     */
int ofdft_eval(ofdft_plan plan, const void* rho, void* V_out, double* sums_out);

/* Evaluation at phi(t) = c*phi + s*p, rho = phi(t)^2 (the trial point of
 * Optimization.ValueAndDerivative, optimization.py:228-280; p may be NULL).
 * Also fills OFDFT_S_DTHETA with the derivative along
 * dphi/dt = c*p - s*phi.  aux_out (device, 2*N, may be NULL) receives the
 * vW potential and the WT convolution needed by ofdft_hvp.               */
int ofdft_eval_phi(ofdft_plan plan, const void* phi, const void* p, double c, double s, void* V_out,
                   void* aux_out, double* sums_out);

/* Exact Hessian-vector product of the constrained TN system at phi:
 * Ap = (V - mu) p + phi * dV/drho[2 phi p]  -- the quantity
 * get_direction_TN forms by forward differences (optimization.py:105-115).
 * V and aux come from ofdft_eval_phi at phi.  *pAp_out = sum(p*Ap).
 * Synchronises the stream.                                               */
int ofdft_hvp(ofdft_plan plan, const void* phi, const void* p, const void* V, const void* aux, double mu,
              void* Ap_out, double* pAp_out);

/* The same product in two steps for the TN inner loop.  hvp_prepare, once
 * per outer iteration, writes the residual res = (V - mu) phi
 * (optimization.py:323, :469; *res2 = sum res^2) and the local part
 * W0 = V - mu + 2 rho a(rho), a = d(V_TF + V_XC + V_vW,local + V_WT,local)/drho.
 * hvp_fast then forms Ap from phi, p and W0 plus the three convolutions;
 * it equals ofdft_hvp up to rounding.                                    */
int ofdft_hvp_prepare(ofdft_plan plan, const void* phi, const void* V, const void* aux, double mu, void* res,
                      void* W0, double* res2);

/* Response of the total potential to a density change drho = 2 phi p at
 * rho = phi^2 (the second functional derivative applied to drho, the same
 * terms as ofdft_hvp): dV_out = dV/drho [drho].  With phi = sqrt(rho) and
 * p = u / (2 phi) this is the Hessian of E(rho) applied to u, which the
 * Python autograd layer uses as the backward of V(rho); it replaces the
 * forward-difference product of get_direction_TN (optimization.py:106-114).
 * The WT kernel is held at its cached rho0 (wt.py:64-72).               */
int ofdft_potential_jvp(ofdft_plan plan, const void* phi, const void* p, const void* V, const void* aux, void* dV_out);
int ofdft_hvp_fast(ofdft_plan plan, const void* phi, const void* p, const void* W0, void* Ap, double* pAp_out);

/* One inner step of get_direction_TN (optimization.py:105-145) with the
 * exact HVP: Ap = hvp_fast(p); pAp = sum p*Ap; if pAp > 0: a = r0/pAp,
 * d += a p, r -= a Ap (on the device); r1 = sum r^2.  out2 = {pAp, r1};
 * one stream synchronisation.                                            */
int ofdft_tn_inner(ofdft_plan plan, const void* phi, const void* p, const void* W0, void* Ap, void* d, void* r,
                   double r0, double* out2);

/* The whole of get_direction_TN (optimization.py:92-148) with the exact
 * HVP, run on the device: d = direction (N reals, output), res = residual
 * (phi-gradient, input), work = 4*N reals of scratch (r, p, Ap, Best).
 * The host queues `chunk` CG iterations at a time without waiting; a
 * one-thread control kernel applies the reference's branch tests (pAp < 0,
 * r1 < rConv, r1 > 1000 min(rLists), stagnation after 10 iterations) and
 * ends the loop, after which queued kernels return at once.  One stream
 * synchronisation per chunk.  niter = the reference's `number`, status =
 * OFDFT_TN_*, pAp_last = pAp of the last iteration.                      */
int ofdft_tn_direction(ofdft_plan plan, const void* phi, const void* res, const void* W0, void* d, void* work,
                       int maxfun, int chunk, int* niter, int* status, double* pAp_last);

/* Raw transforms (DirectField.fft field.py:413-430 without the dV factor;
 * ReciprocalField.ifft field.py:779-804 without the 1/dV factor):
 * r2c = numpy.fft.rfftn, c2r = numpy.fft.irfftn (1/N normalised).
 * c2r does not modify its input.                                         */
int ofdft_fft_r2c(ofdft_plan plan, const void* in_real, void* out_cplx);
int ofdft_fft_c2r(ofdft_plan plan, const void* in_cplx, void* out_real);

/* Reciprocal-space convolution out = irfftn(M(G) * rfftn(in)) with
 * M = 4 pi/G^2 (kind 1, Hartree, hartree.py:34-36), G^2 (kind 2, vW,
 * vw.py:46-49) or the WT kernel (kind 3, wt.py:22).                      */
#define OFDFT_CONV_NONE 0
#define OFDFT_CONV_HARTREE 1
#define OFDFT_CONV_GG 2
#define OFDFT_CONV_WT 3
int ofdft_convolve(ofdft_plan plan, int kind, const void* in_real, void* out_real);

/* ---- ions (FP64 plans) -------------------------------------------------
 * Radial local-PP table of one species: the cubic interpolating spline of
 * v(q) fitted without the q=0 point (ReadPseudo.get_vloc_interp,
 * pseudo/__init__.py:689-697; knots/coefs as scipy splrep returns them,
 * both of length nknots, degree 3), plus the q-grid and v values the
 * quartic small-q branch uses (calc_vlines :413-420).                    */
typedef struct {
  const double* knots;  /* host */
  const double* coefs;  /* host */
  int nknots;
  double q1;    /* gp[1]          */
  double dq;    /* gp[1] - gp[0]  */
  double qmax;  /* gp[-1]         */
  double v0, v1, v2;  /* vp[0], vp[1], vp[2] */
} ofdft_vline;

/* V_loc(r) for nat ions (positions host, Cartesian bohr, nat*3; species
 * index per ion) with PME (pme=1, B-spline order `order`,
 * _PP_Reciprocal_PME pseudo/__init__.py:437-452) or the exact structure
 * factor (pme=0, _PP_Reciprocal :425-435).  Writes N doubles to vloc_out
 * (device) and keeps the ion state for ofdft_local_forces.              */
int ofdft_vloc_build(ofdft_plan plan, int nspecies, const ofdft_vline* tables, int nat, const double* positions,
                     const int* species, int pme, int order, void* vloc_out);

/* Local-PP ionic forces at density rho (device, N doubles) for the ions
 * of the last ofdft_vloc_build: _ForcePME (pseudo/__init__.py:523-571) or
 * _Force (:505-521).  forces_out: host nat*3, Ha/bohr.                   */
int ofdft_local_forces(ofdft_plan plan, const void* rho, double* forces_out);

/* Ewald ion-ion sums (ewald.py:197-507): real-space pair sum with cutoff
 * rmax over periodic images, and the reciprocal sum over the half-space
 * mask with the PME structure factor (pme=1, Energy_rec_PME :785-807) or
 * the exact one (pme=0, Energy_rec :448-465).  e_real already halved;
 * e_rec includes 4 pi/Omega.  The G=0 / self terms (Energy_corr) are host
 * arithmetic on the charges.
 * ofdft_ewald_forces: the ion-ion forces (Ha/bohr, nat x 3) of the same
 * sums: Forces_real (:541-578) plus Forces_rec_PME (:810-864, pme=1) or
 * Forces_rec (:580-607, pme=0); `forces` property :509-519.               */
int ofdft_ewald_forces(ofdft_plan plan, int nat, const double* positions, const double* charges, double eta,
                       double rmax, int pme, int order, double* forces_out);
int ofdft_ewald(ofdft_plan plan, int nat, const double* positions, const double* charges, double eta, double rmax,
                int pme, int order, double* e_real, double* e_rec);

/* ofdft_ewald_noalloc: ofdft_ewald (same arguments, same kernels, same
 * results) without any device allocation.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, molecular-dynamics or serving loops, harnesses
 * that forbid hipMalloc).  The per-call staging (positions, charges,
 * fractional coordinates, image shifts) and, with pme=1, the charge grid
 * and b-moduli live in a static 8 MiB device array, one per GPU, held until
 * the process ends: (7 nat + 3 nshift) doubles plus, with pme=1,
 * n0 n1 n2 + 2 (n0 + n1 + n2/2 + 1) doubles, each part rounded up to 256
 * bytes, must fit (about 96^3 with PME); a larger request returns
 * OFDFT_EINVAL with the sizes.  Calls on one GPU share the array: calls that
 * overlap on one GPU (several host threads or streams) race.  The plan's own
 * workspace (spectra, partial sums) was allocated at ofdft_plan_create.     */
int ofdft_ewald_noalloc(ofdft_plan plan, int nat, const double* positions, const double* charges, double eta,
                        double rmax, int pme, int order, double* e_real, double* e_rec);

/* Stress (TotalFunctional.get_stress, total_functional.py:176-209).  The
 * six components of each symmetric 3x3 sum are ordered xx xy xz yy yz zz;
 * the callers apply the reference's factors, volumes and energy terms.
 *
 * ofdft_functional_stress: FP64 plan, rho device N doubles, the terms of
 * the set functional.  out[OFDFT_NSTRESS] (host):
 *   [0,6)   Hartree  sum_mask G_a G_b |rfftn(rho)|^2 / G^4   (hartree.py:68-90)
 *   [6,12)  vW       sum_r d_a rho d_b rho / rho, gradients by FFT
 *                                                            (kedf/vw.py:97-112)
 *   [12,18) WT       sum_mask (G_a G_b / G^2 - d_ab / 3) eta F'(eta)
 *                    Re(rfftn(rho^alpha) conj rfftn(rho^beta)), eta = |G| / tkf,
 *                    mu = y                                  (kedf/wt.py:75-106,
 *                                                             kernel.py:197-211)
 *   [18]    LDA      sum_r V_xc rho                          (xc/semilocal_xc.py:343-346)
 * Terms not in the functional are 0.                                      */
#define OFDFT_NSTRESS 19
int ofdft_functional_stress(ofdft_plan plan, const void* rho, double tkf, double* out);

/* Local-PP stress sum for the ions and tables of the last ofdft_vloc_build
 * (pseudo/__init__.py:248-258): exact form sum_mask G_a G_b
 * Re(rfftn(rho) sum_i v'_s(i)(|G|) exp(i G.R_i)) / |G| (_Stress :462-503),
 * or the PME form sum_s sum_mask G_a G_b Re(conj(rfftn(rho)) B v'_s(|G|)
 * rfftn(Q_s)) / |G| (_StressPME :573-602).  v' is the derivative of the
 * table's cubic spline (_PP_Derivative_One :454-460).  out6 host.         */
int ofdft_local_stress(ofdft_plan plan, const void* rho, double* out6);

/* Ewald stress sums (ewald.py:523-538): real6 = sum over ordered pairs and
 * images of q_i q_j [2 sqrt(eta/pi) exp(-eta d^2) + erfc(sqrt(eta) d)/d]
 * d_a d_b / d^2 (Stress_real :609-662); rec6 = sum_mask |S|^2
 * exp(-G^2/(4 eta))/G^2 (2 G_a G_b/G^2 (1 + G^2/(4 eta)) - d_ab) with the
 * PME (Stress_rec_PME :915-958) or exact (Stress_rec :719-759) S.        */
int ofdft_ewald_stress(ofdft_plan plan, int nat, const double* positions, const double* charges, double eta,
                       double rmax, int pme, int order, double* real6, double* rec6);

/* GGA kinetic functionals (GGA, kedf/gga.py:1212-1311, enhancement factors
 * :42-575, names and default parameters :577-635): one functional id below,
 * its parameters (the reference's list, completed with the defaults by the
 * caller).  The density is rho (device, N values of the plan precision) or,
 * with rho = NULL, (c phi + s p)^2 (p may be NULL).  V_inout (device, may
 * be NULL) receives V += V_GGA; without it only the energy and the stress
 * sums are formed.  out[OFDFT_GGA_NOUT] (host, FP64, all x dV):
 *   [0] E_GGA = sum C_TF rho^(5/3) F(s)
 *   [1] 2 sum V_GGA phi(t) dphi/dt   (phi mode with p, and V given)
 *   [2] sum V_GGA rho                (V given)
 *   [3] sum pot1 rho, [4,10) sum tf dFds2 / rho^(8/3) d_a rho d_b rho
 *       (xx xy xz yy yz zz): the GGA stress terms of kedf/gga.py:1312-1326 */
#define OFDFT_GGA_LKTVW 0
#define OFDFT_GGA_LKT 1
#define OFDFT_GGA_DK 2
#define OFDFT_GGA_LLP 3
#define OFDFT_GGA_OL 4
#define OFDFT_GGA_THAK 5
#define OFDFT_GGA_B86A 6
#define OFDFT_GGA_B86B 7
#define OFDFT_GGA_DK87 8
#define OFDFT_GGA_PW86 9
#define OFDFT_GGA_PW91O 10
#define OFDFT_GGA_PW91 11
#define OFDFT_GGA_LG94 12
#define OFDFT_GGA_P92 13
#define OFDFT_GGA_PBE 14
#define OFDFT_GGA_P82 15
#define OFDFT_GGA_VJKS00 16
#define OFDFT_GGA_VT84F 17
#define OFDFT_GGA_TFVW 18
#define OFDFT_GGA_STV 19
#define OFDFT_GGA_PBE2M 20
#define OFDFT_GGA_PG 21
#define OFDFT_GGA_COUNT 22
#define OFDFT_GGA_NOUT 10
int ofdft_gga(ofdft_plan plan, int functional, const double* params, int nparams, const void* rho, const void* phi,
              const void* p, double c, double s, void* V_inout, double* out);

/* LWT / LMGP nonlocal kinetic term (LWTPotentialEnergy, kedf/lwt.py:83-341;
 * LWT / LMGP :506-579).  The kernel of a node kf_j is K(|G| / kf_j) from a
 * table in eta, its derivative table -eta dK/deta divided by kf_j, both
 * 0 at G = 0 and constant (`limit`) from etamax on (LWTKernelKf,
 * kedf/kernel.py:379-405).  A table is a cubic B-spline (kind 0: knots and
 * coefficients as scipy's splrep returns them, fitted on the uniform grid
 * x0 + i dx) or a plain array read at the nearest entry (kind 1: `coefs`
 * holds the n values, dx the spacing).  Set once per plan (the
 * reference's ke_kernel_saved tables):                                     */
typedef struct ofdft_lwt_table {
  int kind;             /* 0 cubic spline, 1 nearest table entry            */
  const double* knots;  /* host, n knots (kind 0)                           */
  const double* coefs;  /* host, n coefficients (kind 0) or n values (1)    */
  int n;
  double x0, dx;        /* uniform data grid (kind 0) / table spacing (1)  */
  double etamax, limit; /* the kernel is `limit` for eta >= etamax           */
} ofdft_lwt_table;
int ofdft_lwt_set_kernel(ofdft_plan plan, const ofdft_lwt_table* kernel, const ofdft_lwt_table* deriv);
/* out[0] = min, out[1] = max of the density (rho, or (c phi + s p)^2 with
 * rho = NULL): the kf range of guess_kf_bound (lwt.py:36-80) and the
 * largest density of the low-density weight (lwt.py:322-329).             */
int ofdft_lwt_range(ofdft_plan plan, const void* rho, const void* phi, const void* p, double c, double s,
                    double* out2);
typedef struct ofdft_lwt_params {
  double alpha, beta;   /* rho^alpha (energy side), rho^beta (convolved)    */
  double kfmax;         /* kf = 2 (3 pi^2 rho)^(1/3) is clamped to it       */
  double ld;            /* low-density weight |rho|^ld / rhomax^ld where     */
  double rhomax_ld;     /*   rho < 1e-6 (0 where rho < 0); rhomax_ld = that  */
  int kdd;              /* 1, 2 or 3 potential terms (lwt.py:262-341)       */
  int hermite;          /* 0 linear, 1 Hermite interpolation in kf          */
  int mgpe;             /* 1: add the MGP kinetic-electron kernel           */
  double mgpe_a, mgpe_b, mgpe_c; /* 4 pi erf(c G^2)^2 a exp(-b G^2) / G^2
                                    (MGPOmegaE, kernel.py:408-440)          */
} ofdft_lwt_params;
/* The LWT term at rho (or (c phi + s p)^2) with the nsp node list kflist
 * (host; lwt.py:148-165 with the 1e-16 widening), V_inout (device, may be
 * NULL: energy only, the reference's kdd = 1) receives V += V_LWT.
 * out[3] (host, x dV): [0] E = sum rho^alpha pot1, [1] 2 sum V_LWT phi(t)
 * dphi/dt (phi mode with p, V given), [2] sum V_LWT rho (V given).        */
int ofdft_lwt(ofdft_plan plan, const ofdft_lwt_params* params, const double* kflist, int nsp, const void* rho,
              const void* phi, const void* p, double c, double s, void* V_inout, double* out3);

/* ---- vector kernels used by the optimizer (optimization.py:92-280) ---- */
/* out[0] = sum(a*b), out[1] = sum(a*a), out[2] = sum(b*b) in FP64 (host).  */
int ofdft_dot3(ofdft_plan plan, const void* a, const void* b, double* out3);
/* y = alpha*x + beta*y */
int ofdft_axpby(ofdft_plan plan, double alpha, const void* x, double beta, void* y);
/* z = alpha*x + beta*y (z may alias x or y) */
int ofdft_lincomb(ofdft_plan plan, double alpha, const void* x, double beta, const void* y, void* z);
/* res = (V - mu) * phi ; *res2 = sum(res^2)  (optimization.py:323, :469) */
int ofdft_residual(ofdft_plan plan, const void* V, const void* phi, double mu, void* res, double* res2);
/* TN CG update: d += a*p; r -= a*Ap; *r2 = sum(r^2) (optimization.py:124-127) */
int ofdft_cg_update(ofdft_plan plan, double a, const void* p, const void* Ap, void* d, void* r, double* r2);
/* sum(V*phi*p) (optimization.py:371) */
int ofdft_dot_vphip(ofdft_plan plan, const void* V, const void* phi, const void* p, double* out);
/* The sums of one outer step in one pass, FP64 accumulation:
 * out5 = {sum p*phi, sum phi^2, sum p^2, sum V*phi*p, sum V*phi*phi}.
 * OrthogonalNormalization (optimization.py:205-226) and the steepest-
 * descent test / derivative at theta = 0 (:370-380, :270-277) follow from
 * them for p' = s (p - f phi) without another pass over the fields.      */
int ofdft_orth_sums(ofdft_plan plan, const void* p, const void* phi, const void* V, double* out5);
/* OrthogonalNormalization (optimization.py:205-226) done on the device:
 * p = s (p - f phi) with f = sum p*phi / sum phi^2 and
 * s = sqrt(sum phi^2 / (sum p^2 - f sum p*phi)) (the reference's dV factors
 * cancel and are left out; each operation rounded once), from the five sums
 * of ofdft_orth_sums, which are kept for ofdft_plan_deferred_sums.  Does
 * not synchronise: the first line-search trial can be queued behind it.   */
int ofdft_orth_normalize(ofdft_plan plan, void* p, const void* phi, const void* V);
/* ofdft_hvp_prepare without waiting for the sum: res and W0 as there, sum
 * res^2 kept for ofdft_plan_deferred_sums; does not synchronise.          */
int ofdft_hvp_prepare_async(ofdft_plan plan, const void* phi, const void* V, const void* aux, double mu, void* res,
                            void* W0);
/* Synchronises the plan stream and returns the sums the two calls above
 * kept: out6[0] = sum res^2 of the last ofdft_hvp_prepare_async,
 * out6[1..5] = the five sums of the last ofdft_orth_normalize (as
 * ofdft_orth_sums orders them).  The sums of other calls do not overwrite
 * them.  No counterpart in the reference (its sums are numpy values).     */
int ofdft_plan_deferred_sums(ofdft_plan plan, double* out6);
/* precision conversion between plans: dst (dst precision) = src (src precision), N values */
int ofdft_convert(ofdft_plan dst_plan, void* dst, int src_precision, const void* src, long long n);

const char* ofdft_last_error(void);
int ofdft_version(void);

#ifdef __cplusplus
}
#endif
#endif
