/* mdft_hip.h -- C ABI of the MI355X (gfx950) excess-functional hot path.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success and
 * a nonzero status on failure; mdft_last_error() then returns the message of
 * the calling thread's last failure (the reference throws std::runtime_error
 * from Impl::Throw_If, common/MDFT_Asserts.hpp:15-28; the C++ facade in
 * include/mdft/ rethrows it).  No exception crosses this boundary.
 *
 * Device pointers are HIP device memory; `stream` is a hipStream_t (NULL =
 * the default stream).  Calls are stream ordered and do not synchronise,
 * except where a host scalar is returned (ff).  A handle is not thread safe:
 * one handle per stream.
 *
 * Layouts are the reference's LayoutRight views:
 *   orientation fields o, xi, vexc, df : (nx*ny*nz, ntheta, nphi, npsi) real
 *   projections p                     : (np, nx, ny, nz) complex (re, im)
 *
 * Precision forms (MDFT_FP64 / MDFT_FP32 / MDFT_MIXED*) fix the element types:
 *   form        o / xi / df   p               arithmetic
 *   MDFT_FP64   double        complex double  FP64 everywhere
 *   MDFT_FP32   float         complex float   FP32 (reference ScalarType=float)
 *   MDFT_MIXED  double        complex float   projections: BF16x3 MFMA with
 *                                             FP32 accumulation; FFT, rotation,
 *                                             OZ in FP32; R(q), interpolation
 *                                             weights, ff in FP64
 *   MDFT_MIXED_X6  as MDFT_MIXED, but angl2proj splits data and weights in
 *                  three BF16 parts (BF16x6 MFMA: data exact to FP32,
 *                  weights to ~2^-26) and proj2angl runs in FP32 on the
 *                  vector ALUs; |dF| ~1e-3 kJ/mol at 64^3, mmax = 5
 *   MDFT_MIXED_X6F the arithmetic of MDFT_MIXED_X6 on float fields: o, xi
 *                  and df are float (FP32 copies of the solvent fields,
 *                  half their HBM bytes); delta_rho, the get_delta_f products
 *                  and every reduction stay FP64.  The functional and the
 *                  server keep a double iterate and evaluate the excess on
 *                  its FP32 copy
 */
#ifndef MDFT_HIP_H
#define MDFT_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDFT_FP64 0
#define MDFT_FP32 1
#define MDFT_MIXED 2
#define MDFT_MIXED_X6 3
#define MDFT_MIXED_X6F 4

/* element type codes of the free functions */
#define MDFT_DOUBLE 0
#define MDFT_FLOAT 1

typedef struct mdft_opt_s* mdft_opt_t;
typedef struct mdft_conv_s* mdft_conv_t;
typedef struct mdft_solver_s* mdft_solver_t;

const char* mdft_last_error(void);
int mdft_version(void);

/* ---- init-time tables (host, FP64) ------------------------------------ */

/* common/MDFT_Math_Utils.hpp:24-73 (gauss_legendre): nodes descending,
 * weights summing to 1. */
int mdft_gauss_legendre(int n, double* x, double* w);

/* common/MDFT_Wigner.hpp:62-130 (wigner_small_d): d^m_{mu,mup}(theta). */
int mdft_wigner_small_d(int m, int mu, int mup, double theta, double* out);

/* common/MDFT_Grid.hpp:129-152 (AngularGrid sizes): sizes[0..5] = ntheta,
 * nphi, npsi, no, np, mmax/mrso + 1. */
int mdft_angular_sizes(int mmax, int mrso, int* sizes);

/* common/MDFT_Grid.hpp:129-263 (AngularGrid) and
 * common/MDFT_OrientationProjectionTransform.hpp:55-147
 * (OrientationProjectionMap): theta and wtheta (ntheta), w (no), p map
 * ((mmax+1) x (2mmax+1) x (mmax/mrso+1)), m, mup, mu, mu2 (np), and the
 * Wigner table (ntheta x np).  Any output pointer may be NULL. */
int mdft_angular_tables(int mmax, int mrso, double* theta, double* wtheta, double* w, int* pmap,
                        int* m, int* mup, int* mu, int* mu2, double* wigner_d);

/* common/MDFT_Grid.hpp:87-103 (SpatialGrid::kproj): k_i for a box side L. */
int mdft_kproj(int n, double length, double* k);

/* common/MDFT_Rotation.hpp:103-283
 * (rotation_matrix_between_complex_spherical_harmonics_lu): R has
 * (mmax+1) x (2mmax+1) x (2mmax+1) complex entries (re, im interleaved). */
int mdft_rotation_matrix(int mmax, const double* q, double* R);

/* common/MDFT_Convolution.hpp:126-176 (the (q, -q) list, gamma_p_map):
 * the count, and the list in the reference's order (ix, iy, iz triples). */
int mdft_qpair_count(int nx, int ny, int nz);
int mdft_qpairs(int nx, int ny, int nz, int* ixyz);

/* common/IO/MDFT_ReadCLuc.hpp:40-259 (LucData): reads np_luc tuples and
 * nq = int(qmaxwanted/dq + 0.5) + 2 rows of a Luc c(q) file.  tuples:
 * np_luc x 5 ints (m, n, mu, nu, khi), c: np_luc x nq complex (re, im); either
 * may be NULL (call once with NULLs to get nq). */
int mdft_luc_read(const char* path, int np_luc, double qmaxwanted, int* nq, double* dq, int* tuples,
                  double* c);

/* ---- OrientationProjectionTransform --------------------------------------
 * common/MDFT_OrientationProjectionTransform.hpp:161-485: the constructor
 * (:217-257), angl2proj (:267-380) and proj2angl (:388-484). */
int mdft_opt_create(int nx, int ny, int nz, int mmax, int mrso, int precision, mdft_opt_t* out);
int mdft_opt_destroy(mdft_opt_t h);
int mdft_angl2proj(mdft_opt_t h, const void* o, void* p, void* stream);
int mdft_proj2angl(mdft_opt_t h, const void* p, void* o, void* stream);

/* ---- Convolution -----------------------------------------------------------
 * common/MDFT_Convolution.hpp:20-549: the constructor (:106-313) reading the
 * Luc file `luc_path` with np_luc tuples, and execute (:324-548), in place on
 * p.  mdft_conv_create_from_table takes the tuples and values that
 * mdft_luc_read returns instead of a path (nq_table >= the rows the grid
 * needs). */
int mdft_conv_create(const char* luc_path, int nx, int ny, int nz, double lx, double ly, double lz,
                     int mmax, int mrso, int np_luc, int precision, mdft_conv_t* out);
int mdft_conv_create_from_table(int nx, int ny, int nz, double lx, double ly, double lz, int mmax,
                                int mrso, int np_luc, const int* tuples, const double* c, int nq_table,
                                double dq, int precision, mdft_conv_t* out);
/* accessors dq() and the sizes of gamma_p_map() / mnmunukhi_q() (:316-318) */
int mdft_conv_info(mdft_conv_t h, int* nq, int* np_new, int* npairs, double* dq);
/* mnmunukhi_q (:316-318): np_new x nq complex (re, im), signs folded in */
int mdft_conv_table(mdft_conv_t h, double* c);
int mdft_conv_execute(mdft_conv_t h, void* p, void* stream);
int mdft_conv_destroy(mdft_conv_t h);

/* Batched unnormalised 3D c2c FFT, in place over the last three axes of a
 * (batch, nx, ny, nz) complex array (the KokkosFFT plans of
 * common/MDFT_Convolution.hpp:198-208); sign -1 forward, +1 backward;
 * dtype MDFT_DOUBLE or MDFT_FLOAT. */
int mdft_fft3d(void* data, int batch, int nx, int ny, int nz, int sign, int dtype, void* stream);

/* The same transform in a given configuration (no counterpart in the
 * reference): plane != 0 runs z and y as one pass over each (y, z) plane
 * staged in LDS (plane_threads 256 or 512; radices 2-5; the plane must fit in
 * 160 KB of LDS), else one pass per axis with wz / wy / wx lines per tile
 * (0: the default).  Every configuration runs the same butterflies on the
 * same lines: results are bitwise those of mdft_fft3d.  A Convolution tunes
 * its configuration on first use by timing these candidates;
 * mdft_conv_fft_config reports it (tuned = 0 before the first execute). */
int mdft_fft3d_config(void* data, int batch, int nx, int ny, int nz, int sign, int dtype, int plane,
                      int plane_threads, int wz, int wy, int wx, void* stream);
int mdft_conv_fft_config(mdft_conv_t h, int* tuned, int* plane, int* plane_threads, int* wz, int* wy, int* wx);

/* ---- Solvent free functions ---------------------------------------------- */

/* common/MDFT_Solvent.hpp:575-596 (get_delta_rho): delta_rho = rho0 (xi^2 - 1)
 * over n elements. */
int mdft_delta_rho(const void* xi, void* delta_rho, size_t n, double rho0, int dtype, void* stream);

/* common/MDFT_Solvent.hpp:613-660 (get_delta_f): delta_f = 2 rho0 xi vexc and
 * *ff = sum rho0 prefactor vexc w (xi^2 - 1) (FP64 accumulation, fixed
 * order: deterministic).  w is a device array of no doubles; ff is a host
 * pointer and the call synchronises the stream to fill it. */
int mdft_delta_f(const void* xi, const void* vexc, const double* w, void* delta_f, size_t nxyz, int no,
                 double rho0, double prefactor, double* ff, int dtype, void* stream);

/* No counterpart in the reference.  mdft_delta_f for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same kernels and the same result bit
 * for bit, with the partial sums and the reduced ff in a fixed device array
 * of the library (1 MiB per GPU, held until the process ends) instead of
 * memory from hipMallocAsync.  nxyz above MDFT_DELTA_F_NOALLOC_MAX_POINTS is
 * rejected.  The stream must belong to the current device.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other streams,
 * other host threads) race and must be serialised by the caller.  ff is a
 * host pointer and the call synchronises the stream to fill it. */
#define MDFT_DELTA_F_NOALLOC_MAX_POINTS 1048568
int mdft_delta_f_noalloc(const void* xi, const void* vexc, const double* w, void* delta_f, size_t nxyz, int no,
                         double rho0, double prefactor, double* ff, int dtype, void* stream);

/* ---- Solver: one c-proj functional + gradient evaluation -------------------
 * src/mini-mdft.hpp:137-193 (Solver::energy_cproj_mrso): delta_rho ->
 * angl2proj -> Convolution::execute -> proj2angl -> get_delta_f, with
 * delta_rho fused into the projection load and get_delta_f into the
 * proj2angl epilogue.  xi and df are (N, no) doubles for MDFT_FP64 and
 * MDFT_MIXED, floats for MDFT_FP32.  The solver owns its work buffers
 * (mdft_solver_workspace_bytes). */
int mdft_solver_create(mdft_conv_t conv, double rho0, double prefactor, int precision,
                       mdft_solver_t* out);
int mdft_solver_workspace_bytes(mdft_solver_t h, size_t* bytes);
int mdft_solver_destroy(mdft_solver_t h);
/* ff: host pointer, the call synchronises the stream */
int mdft_energy_cproj(mdft_solver_t h, const void* xi, void* df, double* ff, void* stream);
/* ff_device: one device double written on the stream, no synchronisation */
int mdft_energy_cproj_async(mdft_solver_t h, const void* xi, void* df, double* ff_device, void* stream);

/* Same-shape batches (no counterpart in the reference, which evaluates one
 * field per Solver): a solver created with capacity B evaluates up to B
 * fields in one pass of every stage.  xi and df hold nbatch fields back to
 * back (nbatch x N x no elements); ff (host, synchronising) or ff_device
 * receives nbatch values.  Each field's result is bitwise the one it gets
 * alone (same per-point arithmetic, same partial sums in the same order).
 * Needs N = nx ny nz to be a multiple of 64 when capacity > 1;
 * mdft_solver_create is capacity 1. */
int mdft_solver_create_batch(mdft_conv_t conv, double rho0, double prefactor, int precision, int capacity,
                             mdft_solver_t* out);
int mdft_energy_cproj_batch(mdft_solver_t h, int nbatch, const void* xi, void* df, double* ff, void* stream);
int mdft_energy_cproj_batch_async(mdft_solver_t h, int nbatch, const void* xi, void* df, double* ff_device,
                                  void* stream);

/* Stage timers (the Solver's per-stage Kokkos::Timer printout,
 * src/mini-mdft.hpp:140-192), as hipEvents recorded on the work stream
 * without synchronising between evaluations.  set_timing(1) starts (and
 * resets) them, set_timing(2) times only the OZ pass, whose launch records its
 * own start and end (hipExtLaunchKernel; the other stages read 0: each event
 * recorded between kernels leaves the GPU idle for a few microseconds),
 * set_timing(0) stops them.  stage_ms waits
 * for the recorded evaluations and returns the accumulated milliseconds of
 * ms[0..4] = angl2proj (delta_rho fused), forward 3D FFT, rotation/OZ pass,
 * backward 3D FFT, proj2angl + delta_f, and the number of evaluations in
 * *count. */
int mdft_solver_set_timing(mdft_solver_t h, int enable);

/* No counterpart in the reference (its path is FP64 throughout).  The r-mean
 * (q = 0) mode of ff evaluated in FP64: proj2angl's epilogue also sums
 * xi^2 - 1 and vexc over r for every orientation, the q = 0 value is carried
 * through angl2proj, the OZ product at q = 0 and proj2angl in FP64, and it
 * replaces the mixed path's r-mean contribution to ff (df is unchanged).  With
 * xi near 1 this mode carries almost all of ff, and an FP32 path's ~1e-7
 * relative error on it would reach ff whole.  On by default for MDFT_MIXED_X6;
 * available (off by default) for MDFT_FP64 and MDFT_FP32, where it changes ff
 * only by rounding; not available for MDFT_MIXED (proj2angl on the matrix
 * cores) or angular grids without a compile-time proj2angl kernel. */
int mdft_solver_set_mean_fp64(mdft_solver_t h, int on);
int mdft_solver_stage_ms(mdft_solver_t h, double* ms, long* count);

/* No counterpart in the reference (its path stores every projection).  The
 * compact channel order of the fused evaluation: delta_rho is real, so its
 * mu = 0 projection with mup < 0 equals (-1)^mup conj of its (m, -mup, 0)
 * partner, before the FFT and (at -q) after it; and proj2angl reads the
 * mu = 0 column through its real part, to which the pair contributes only as
 * g_+ + (-1)^mup conj(g_-).  In this order every stage holds the other
 * projections (73 of 88 at mmax 5, mrso 2): angl2proj writes them, the FFTs
 * transform them, the OZ pass reads an implied input from its partner at the
 * mirrored point and stores the pair's combination in the partner's channel,
 * and proj2angl takes the implied sums as zero.  Nothing is assumed of c(q).
 * On by default in every form where the grid has the compile-time projection
 * (or GEMM) and OZ kernels; results change by rounding only.
 * mdft_solver_compact reports whether it is on. */
int mdft_solver_set_compact(mdft_solver_t h, int on);
int mdft_solver_compact(mdft_solver_t h, int* on);

/* ---- utilities --------------------------------------------------------------
 * Counter-based uniforms: x_i = lo + (hi - lo) u_i with u_i the splitmix64
 * stream of (seed, offset + i) (same definition as mini_mdft_amd/rng.py). */
int mdft_fill_uniform(void* x, size_t n, unsigned long long seed, unsigned long long offset, double lo,
                      double hi, int dtype, void* stream);

/* ---- external potential -------------------------------------------------------
 * AngularGrid's rotation tables (common/MDFT_Grid.hpp:265-333, m_rotxx ..
 * m_rotzz): rot[io * 9 + 3 a + b] = R_ab of orientation io, Euler ZYZ with
 * theta from the GL nodes, phi_j = j 2 pi / nphi, psi_k = k 2 pi / mrso / npsi
 * (host output). */
int mdft_rotation_matrices(int mmax, int mrso, double* rot);

/* V_ext(r, o) in kJ/mol, (N, no) doubles on the device: the sum over the sites
 * of the solvent molecule in orientation o at grid node r = (i lx/nx, j ly/ny,
 * k lz/nz) and over the solute sites of Lennard-Jones (Lorentz-Berthelot
 * mixing, between sites with sigma > 0) and Constants::qfact q q / d
 * (common/MDFT_Constants.hpp:46), d the minimum-image distance floored at
 * 1e-6 A, capped at vmax.  solute (nsolute x 6) and solvent (nsolvent x 6) are
 * host rows (charge e, sigma A, epsilon kJ/mol, x, y, z A), the solvent's in
 * the molecular frame (Solute sites: common/MDFT_Solute.hpp:126-191; solvent
 * sites: common/MDFT_Solvent.hpp:417-532).  The reference builds these tables
 * but never evaluates V_ext (common/MDFT_Solvent.hpp:160).  Synchronises. */
int mdft_vext(int nx, int ny, int nz, double lx, double ly, double lz, int mmax, int mrso, int nsolute,
              const double* solute, int nsolvent, const double* solvent, double vmax, double* vext,
              void* stream);

/* ---- full functional and minimiser (no counterpart in the reference: its
 * Solver::run makes one excess evaluation, src/mini-mdft.hpp:129-133, and the
 * iteration settings are unused, common/MDFT_System.hpp:188-189) --------------
 *   F_id  = kT rho0 dv sum w (xi^2 ln xi^2 - xi^2 + 1)
 *   F_ext = rho0 dv sum w xi^2 V
 *   F_exc = (dv / N) ff              (ff of mdft_energy_cproj)
 *   g     = dv w (2 rho0 xi (kT ln xi^2 + V) + (2 prefactor / N) df)
 * dv = lx ly lz / N.  The functional takes its grid, rho0, prefactor and
 * precision form from the solver (any form but MDFT_FP32; xi and g are
 * doubles, and MDFT_MIXED_X6F evaluates the excess on an FP32 copy of xi and
 * reads its df in FP32); vext (device, N x no) is borrowed and must outlive
 * the handle.  F (host) receives F, F_id, F_ext, F_exc.  Both calls synchronise. */
/* xi = exp(-V / (2 kT)) for n device values (V / kT below -600 taken as -600):
 * the minimum of F_id + F_ext alone, the density guess from V_ext of the
 * original MDFT, as a start point for mdft_lbfgs_minimize.  The reference
 * fills xi with 1 because it has no V_ext (common/MDFT_Solvent.hpp:287-296).
 * Synchronises. */
int mdft_xi_from_vext(const double* vext, long long n, double kT, double* xi, void* stream);

typedef struct mdft_functional_s* mdft_functional_t;
int mdft_functional_create(mdft_solver_t solver, const double* vext, double kT, mdft_functional_t* out);
int mdft_functional_destroy(mdft_functional_t h);
int mdft_functional_eval(mdft_functional_t h, const double* xi, double* g, double* F, void* stream);

/* L-BFGS (FP64 iterate, gradient, dot products and line search) from xi, in
 * place: m stored pairs (1..8; FP32 storage for the mixed forms), diagonal
 * initial inverse Hessian gamma h0 with h0 = 1 / (2 rho0 dv w kT) and
 * gamma = s'y / y'(h0 y) of the newest pair, backtracking from t = 1, or
 * from t = min(1, 1 / max |h0 g|) along -h0 g (no stored pair: no xi moves
 * by more than 1) (at most 30 trials; after a rejected trial t the next is the minimiser of the
 * quadratic through F(0), F'(0) and F(t), kept within [0.1 t, 0.5 t], or t/2
 * where F(t) is not finite) until Armijo (c1 = 1e-4) or the approximate Wolfe
 * test of Hager and Zhang holds (g(t).d <= (2 c1 - 1) g.d with F(t) <= F +
 * 1e-6 (|F_id| + |F_ext| + |F_exc|): Armijo read off the slopes, which stays
 * decisive where F's rounding hides the decrease), a pair kept when
 * s'y > 1e-12 |s| |y| (a full memory whose new pair fails this restarts), a
 * restart from -h0 g on a non-descent direction;
 * converged when max |h0 g| < gtol.  Outputs (host; history may be
 * NULL, else max_iter + 1 doubles receiving F at the start and after every
 * iteration): F[4] at the returned point, the iterations made, the
 * functional evaluations, status (0 converged, 1 max_iter reached, 2 no
 * acceptable point within 30 trials, 4 F or g not finite at the start point:
 * NaN or Inf in xi or V_ext) and max |h0 g|.  The handle keeps its
 * work space between calls: N no doubles (x_t) and 2 (m + 2) N no values of
 * the pair type (g, g_t and the pairs), besides the evaluation's df (N no
 * doubles, or for MDFT_MIXED_X6F N no floats and the FP32 copy of xi). */
int mdft_lbfgs_minimize(mdft_functional_t h, double* xi, int m, int max_iter, double gtol, double* F,
                        int* iterations, int* evaluations, int* status, double* residual, double* history,
                        void* stream);

/* ---- batched and served minimisation (no counterpart in the reference) ----
 * A server minimises many solutes (each: V_ext from its sites, then L-BFGS
 * from xi = 1 as mdft_lbfgs_minimize does) under one HBM budget.  Solutes of
 * one shape (grid, box, mmax) share an engine of `capacity` slots that
 * advances all of them together: every tick gives each slot one trial point
 * and evaluates the functional of all slots in one batched pass of every
 * stage, the tick's kernels replayed as one HIP graph.  Requests wait in a
 * queue and are admitted first-fit in arrival order while their footprint
 * (the solute's fields, plus an engine's tables and slabs for a new shape)
 * fits the budget; a request that could never fit ends with status 3.  Each
 * solute's result is bitwise the one it gets alone (its arithmetic does not
 * depend on the other slots).
 *   create: the c(q) table (tuples, values as mdft_luc_read returns them) for
 *     every mmax that will be requested, the solvent's mrso and n0 (A^-3),
 *     the temperature (K), the precision form (MDFT_FP64 or a mixed form),
 *     the slots per engine, the L-BFGS m, max_iter, gtol and the budget (B);
 *   submit: one solute (host arrays as mdft_vext takes them, copied); out_xi
 *     (device, N no doubles, may be NULL) receives the final xi; returns its id;
 *     fails at once when the c(q) table does not cover the request (fewer q
 *     rows than the grid's qmax needs, a tuple of mmax missing);
 *   step: admits what fits, then one tick of every engine (engines of
 *     different shapes run concurrently on their own streams); synchronises;
 *   poll: one finished solute per call (*got = 0 when none): F[4], the
 *     iterations and evaluations, status (0 converged, 1 max_iter, 2 no
 *     acceptable step, 3 rejected by the budget, 4 F or g not finite at the
 *     start point, 5 its set-up failed on admission: engine or V_ext, the
 *     text from mdft_server_poll_error), max |h0 g|, and the seconds
 *     from submit to admission and to the end;
 *   use_graphs: graph replay on (default) or off, before the first step. */
typedef struct mdft_server_s* mdft_server_t;
int mdft_server_create(int np_luc, const int* tuples, const double* c, int nq_table, double dq, int mrso, double n0,
                       double temperature, int precision, int capacity, int m, int max_iter, double gtol,
                       double hbm_budget_bytes, mdft_server_t* out);
int mdft_server_destroy(mdft_server_t h);
int mdft_server_use_graphs(mdft_server_t h, int on);
/* Start points of the served solves (before the first step): on = 0, xi = 1
 * (the default, Solvent::init_density, common/MDFT_Solvent.hpp:287-296);
 * on = 1, xi = exp(-V / (2 kT)) from each solute's V_ext (mdft_xi_from_vext).
 * No counterpart in the reference, which has no V_ext to start from. */
int mdft_server_start_from_vext(mdft_server_t h, int on);
int mdft_server_submit(mdft_server_t h, int nx, int ny, int nz, double lx, double ly, double lz, int mmax,
                       int nsolute, const double* solute, int nsolvent, const double* solvent, double vmax,
                       double* out_xi, long long* id);
int mdft_server_step(mdft_server_t h, int* active, int* pending);
int mdft_server_poll(mdft_server_t h, long long* id, double* F, int* iterations, int* evaluations, int* status,
                     double* residual, double* queued_s, double* total_s, int* got);
int mdft_server_info(mdft_server_t h, double* bytes_in_use, int* engines, int* active, int* pending);
/* the error text of the result mdft_server_poll returned last: why its
 * admission failed (status 5), else the empty string; truncated to buflen - 1
 * characters */
int mdft_server_poll_error(mdft_server_t h, char* buf, int buflen);

#ifdef __cplusplus
}
#endif

#endif /* MDFT_HIP_H */
