/* C ABI of the MI355X-native RCWA solve path (librcwa_amd.so).
 *
 * Drop-in boundary for rcwa3d (LukhaGian/rcwa-cuda).  The reference has no
 * plugin layer: its interface is the header rcwa3d/include/rcwa.hpp, whose
 * stage functions are called in sequence by rcwa3d/src/main.cpp:126-191.
 * Each entry point below names the reference interface it replaces.
 *
 * Conventions shared with the reference:
 *   - complex numbers are (re, im) double pairs, the layout of
 *     std::complex<double> (rcwa.hpp:12);
 *   - matrices are column-major (rcwa.hpp:13);
 *   - permittivity/permeability grids are [layer][iy][ix] (rcwa.hpp:56);
 *   - harmonic index m = ix*Q + iy, x slow, y fast (rcwa.cpp:63-64);
 *   - R and T are Q x P, column-major, rows indexed by the y order
 *     (rcwa.cpp:576, 582);
 *   - angles are in radians, lengths in any consistent unit.
 * Differences by design: nothing aborts or prints (rcwa.cpp:300 aborts);
 * every call returns an rcwa_status and rcwa_last_error() describes the
 * last failure of the calling thread.  No memory allocated by the library
 * crosses the ABI: host outputs are caller-owned, device memory is owned by
 * the context or server.
 */
#ifndef RCWA_AMD_H
#define RCWA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } rcwa_c128;

typedef enum {
  RCWA_OK = 0,
  RCWA_E_ARG = 1,         /* invalid argument or shape */
  RCWA_E_NOMEM = 2,       /* request exceeds the context's HBM budget */
  RCWA_E_EIG_NOCONV = 3,  /* QR iteration did not converge (rcwa.cpp:300 aborts here) */
  RCWA_E_SINGULAR = 4,    /* an exactly singular matrix was inverted, or k_z = 0 in a
                             half-space or the gap medium (the divisions of
                             rcwa.cpp:225, 329, 541, 560 return inf/NaN there) */
  RCWA_E_NONFINITE = 5,   /* non-finite values in inputs (grids, thicknesses, sources,
                             media, sweep materials) or results */
  RCWA_E_HIP = 6,         /* HIP runtime error */
  RCWA_E_NOTREADY = 7,    /* server ticket not finished yet */
  RCWA_E_CANCELLED = 8    /* server ticket cancelled */
} rcwa_status;

typedef enum { RCWA_FP64 = 0, RCWA_SPLIT = 1 } rcwa_precision;

/* rcwa.hpp:48-78 (struct Device).  er/ur point to num_layers*Ny*Nx complex
 * values, [layer][iy][ix]; t to num_layers thicknesses.  When
 * grids_on_device is non-zero, er and ur are device pointers (HBM-resident
 * inputs); otherwise host pointers. */
typedef struct {
  int Nx, Ny, num_layers;
  double Lx, Ly;
  const rcwa_c128* er;
  const rcwa_c128* ur;
  const double* t;
  int grids_on_device;
} rcwa_device;

/* rcwa.hpp:81-95 (struct Source); k0 = 2 pi / lambda0 is derived. */
typedef struct {
  double lambda0, theta, phi, pte, ptm;
} rcwa_source;

/* rcwa.hpp:98-112 (struct RCWAParams). */
typedef struct {
  int Nx_harmonics, Ny_harmonics;
  rcwa_c128 er_ref, ur_ref, er_trn, ur_trn;
} rcwa_params;

/* rcwa.hpp:116-128 (struct Results).  R and T are caller-allocated arrays of
 * (2*Ny_harmonics+1)*(2*Nx_harmonics+1) = PQ doubles (Q x P, column-major).
 * r and t (optional, NULL to skip) receive the per-order complex amplitudes
 * of the reflected and transmitted fields, the 3PQ-vectors of
 * ComputeReflectedField / ComputeTransmittedField (rcwa.cpp:526-561): the x
 * components of every order (harmonic index m = ix*Q + iy), then y, then z.
 * The reference keeps them inside main.cpp:184-190; exposing them is a
 * widening of the interface for callers that need phases (Jones matrices). */
typedef struct {
  double* R;
  double* T;
  double R_tot, T_tot;
  rcwa_c128* r;
  rcwa_c128* t;
  /* Output: conditioning of the solve's mode matrices, the largest over its
   * layers of sqrt(2PQ) ||W^-1||_F for the unit-column eigenvector matrix W,
   * an upper bound on the 2-norm condition number of W (1 for a layer whose
   * eigenvectors are orthogonal; 0 when the device has no layers).  The
   * reference has no such output (rcwa.cpp:298-306 takes W as it comes);
   * a large value flags nearly defective layer matrices, whose R and T carry
   * correspondingly less accuracy. */
  double cond;
} rcwa_results;

/* One independent solve: a device, its media and nsrc sources.  Sources
 * with equal (lambda0, theta, phi) share one S-matrix (polarisation enters
 * only at rcwa.cpp:495); each distinct incidence gets its own. */
typedef struct {
  const rcwa_device* device;
  const rcwa_params* params;
  const rcwa_source* src;
  int nsrc;
} rcwa_problem;

/* Per-call statistics (stage times from HIP events). */
typedef struct {
  double ms_total, ms_convmat, ms_layers, ms_eig, ms_fold, ms_out;
  double flops_algorithmic; /* 2384 N^3 per layer + 1621 N^3 per S-matrix */
  int n_smatrices, n_layers;
  /* eigensolver phases, filled only when rcwa_ctx_set_timing(ctx, 1): the
   * stream is synchronised after each layer batch to read the events.  A
   * small layer batch run as two staggered halves on two streams counts
   * each half, so these sums can exceed the wall time they cover. */
  double ms_eig_hess, ms_eig_schur, ms_eig_vec;
  int n_eig_launches, n_eig_matrices;
} rcwa_stats;

typedef struct rcwa_ctx rcwa_ctx;

const char* rcwa_last_error(void);
const char* rcwa_version(void);

/* Context: one per host thread; owns a HIP stream, a device memory pool of at
 * most hbm_budget bytes (0 = 80% of free HBM) and the eigensolver limits. */
rcwa_status rcwa_ctx_create(int device, size_t hbm_budget, rcwa_ctx** out);
rcwa_status rcwa_ctx_destroy(rcwa_ctx* ctx);
/* QR sweep limit per eigenvalue (default 30, as Eigen's ComplexSchur). */
rcwa_status rcwa_ctx_set_eig_maxit(rcwa_ctx* ctx, int maxit_per_row);
/* Rotation-log capacity of the QR iteration, in rotations per matrix
 * (0 = default, n^2).  The QR sweeps update only the active block and log
 * their rotations; the log is applied to the Schur vectors and the rest of
 * the Schur form after the iteration, or earlier when it fills up.  A tuning
 * and test knob with no counterpart in the reference (Eigen's ComplexSchur,
 * rcwa.cpp:298-306, has no such parameter). */
rcwa_status rcwa_ctx_set_qr_log(rcwa_ctx* ctx, long long rotations_per_matrix);
rcwa_status rcwa_ctx_get_stats(const rcwa_ctx* ctx, rcwa_stats* out);
rcwa_status rcwa_ctx_reset_stats(rcwa_ctx* ctx);
/* Upper bound on the number of S-matrices batched into one chunk (0 = as
 * many as the HBM budget allows). */
rcwa_status rcwa_ctx_set_chunk(rcwa_ctx* ctx, int max_chunk);
/* Iterative refinement of the solves with the explicit inverse in the
 * Redheffer fold, (I - S11 P22)^-1 (rcwa.cpp:602-606 solves with LU): one
 * step Y += F (B - M Y).  Level 0 none, 1 the source columns the
 * efficiencies use, 2 (default) also the n x n solve that carries the fold
 * forward.
 * Stacks whose layer S-matrices have large entries make M ill-conditioned;
 * there the refined solves keep R + T = 1 for lossless stacks. */
rcwa_status rcwa_ctx_set_refine(rcwa_ctx* ctx, int level);
/* Tile-shape autotuning of the complex GEMMs (no counterpart in the
 * reference, whose products are Eigen's): while on, the first product of a
 * shape class (M, N, K, batch by its power of two) that a solve, sweep or
 * server worker of this context meets is timed in every tile shape on
 * scratch operands (a device allocation and a stream synchronisation, once
 * per class and process) and the fastest is kept for the process.  Every
 * tile shape gives the same bits, so results do not depend on the choice.
 * RCWA_TUNE_FILE names a file the table is read from at first use and
 * appended to.  Off by default. */
rcwa_status rcwa_ctx_set_autotune(rcwa_ctx* ctx, int on);
/* The tuned table: six ints per entry (M, N, K, batch class, device, tile
 * shape) into out, up to max_entries; returns the number of entries. */
int rcwa_autotune_entries(int* out, int max_entries);
/* Forces one tile shape (0 .. 4) for every complex GEMM of the process, or
 * -1 for the tuned ones: for tests and comparisons. */
rcwa_status rcwa_zgemm_force_shape(int shape);
/* Per-phase eigensolver timing with HIP events (adds a synchronisation). */
rcwa_status rcwa_ctx_set_timing(rcwa_ctx* ctx, int on);

/* Full solve, replacing the main.cpp:126-191 sequence (ConvMat per layer ->
 * ComputeWaveVectors -> GapMedium -> SMatrixInit -> per-layer SMatrixLayer
 * and RedhefferProduct -> SMatrixReflection/SMatrixTransmission -> two
 * RedhefferProducts -> ComputeSourceModeCoeff -> Compute{Reflected,
 * Transmitted}Field -> ComputeDiffractionEfficiencies).  out has nsrc
 * entries. */
rcwa_status rcwa_solve(rcwa_ctx* ctx, const rcwa_device* dev, const rcwa_params* params,
                       const rcwa_source* src, int nsrc, rcwa_precision precision,
                       rcwa_results* out);

/* Batched form of rcwa_solve over independent problems (e.g. a spectral
 * sweep).  out is laid out problem by problem, nsrc entries each. */
rcwa_status rcwa_solve_batch(rcwa_ctx* ctx, const rcwa_problem* probs, int nprob,
                             rcwa_precision precision, rcwa_results* out);

/* Optional: global S-matrix blocks S11, S21 columns for the two source
 * entries (indices PQ/2 and PQ/2+PQ) of the last solve_batch, for tests.
 * cols: nprob_smatrices x 4 x (2N) complex. */
rcwa_status rcwa_last_source_columns(rcwa_ctx* ctx, rcwa_c128* cols, size_t count);

/* ----- stage entry points (host buffers), one per reference function ----- */

/* rcwa.hpp:172 ConvMat: out is PQ x PQ column-major for layer `layer`. */
rcwa_status rcwa_convmat(rcwa_ctx* ctx, const rcwa_c128* field, int layer, int Nx, int Ny,
                         int Nx_harmonics, int Ny_harmonics, rcwa_c128* out);

/* rcwa.hpp:175 ComputeWaveVectors: k_inc[3], Kx, Ky, Kz_ref, Kz_trn [PQ]. */
rcwa_status rcwa_wave_vectors(double Lx, double Ly, const rcwa_source* src,
                              const rcwa_params* params, rcwa_c128* k_inc, rcwa_c128* Kx,
                              rcwa_c128* Ky, rcwa_c128* Kz_ref, rcwa_c128* Kz_trn);

/* rcwa.hpp:178 GapMedium: W0, V0 are 2PQ x 2PQ column-major. */
rcwa_status rcwa_gap_medium(int PQ, const rcwa_c128* Kx, const rcwa_c128* Ky, rcwa_c128* W0,
                            rcwa_c128* V0);

/* rcwa.hpp:184 SMatrixLayer for `batch` layers that share Kx, Ky (one
 * incidence): erc/urc are batch x (PQ x PQ), k0t[b] = k0 * t_b; outputs S11
 * and S12 batch x (2PQ x 2PQ) (S22 = S11, S21 = S12, rcwa.cpp:351). */
rcwa_status rcwa_layer_smatrix(rcwa_ctx* ctx, int PQ, int batch, const rcwa_c128* erc,
                               const rcwa_c128* urc, const rcwa_c128* Kx, const rcwa_c128* Ky,
                               const double* k0t, rcwa_c128* S11, rcwa_c128* S12);

/* rcwa.hpp:190/193 SMatrixReflection (which = 0) / SMatrixTransmission
 * (which = 1): four 2PQ x 2PQ blocks S11, S12, S21, S22; W_end = I. */
rcwa_status rcwa_end_smatrix(int which, int PQ, const rcwa_params* params, const rcwa_c128* Kx,
                             const rcwa_c128* Ky, const rcwa_c128* Kz, rcwa_c128* S11,
                             rcwa_c128* S12, rcwa_c128* S21, rcwa_c128* S22);

/* rcwa.hpp:214 RedhefferProduct on `batch` pairs: A[4], B[4], S[4] each
 * point to batch x (n x n) blocks in the order S11, S12, S21, S22. */
rcwa_status rcwa_star_product(rcwa_ctx* ctx, int n, int batch, const rcwa_c128* const* A,
                              const rcwa_c128* const* B, rcwa_c128* const* S);

/* rcwa.hpp:233 extraction(ComplexEigenSolver) on `batch` n x n matrices:
 * eigenvalues w (batch x n) and unit-norm eigenvectors V (batch x n x n). */
rcwa_status rcwa_eig(rcwa_ctx* ctx, int n, int batch, const rcwa_c128* A, rcwa_c128* w,
                     rcwa_c128* V);

/* Batched inverse (the LU solves of the reference, rcwa.cpp:267-281). */
rcwa_status rcwa_inverse(rcwa_ctx* ctx, int n, int batch, const rcwa_c128* A, rcwa_c128* Ainv);

/* C = A * B for batch column-major matrices (the Eigen products). */
rcwa_status rcwa_zgemm(rcwa_ctx* ctx, int m, int n, int k, int batch, const rcwa_c128* A,
                       const rcwa_c128* B, rcwa_c128* C);
/* The same product in the split-precision form (FP16 slices on the matrix
 * cores, FP32 accumulation); no counterpart in the reference. */
rcwa_status rcwa_zgemm_split(rcwa_ctx* ctx, int m, int n, int k, int batch, const rcwa_c128* A,
                             const rcwa_c128* B, rcwa_c128* C);

/* rcwa.hpp:200-209 ComputeSourceModeCoeff + ComputeReflectedField +
 * ComputeTransmittedField + ComputeDiffractionEfficiencies from the
 * S11/S21 columns at the two source entries (W_ref = W_trn = I):
 * s11c, s21c are 2 x 2PQ complex (column PQ/2 then PQ/2+PQ). */
rcwa_status rcwa_efficiencies(const rcwa_source* src, const rcwa_params* params,
                              const rcwa_c128* k_inc, const rcwa_c128* Kx, const rcwa_c128* Ky,
                              const rcwa_c128* Kz_ref, const rcwa_c128* Kz_trn,
                              const rcwa_c128* s11c, const rcwa_c128* s21c, rcwa_results* out);

/* ----- device-pointer batched stage entry points (HBM-resident) ----- */

/* Convolution matrices for `batch` grids [b][iy][ix] (device) -> C (device,
 * batch x PQ x PQ).  Enqueued on the context stream. */
rcwa_status rcwa_convmat_batched(rcwa_ctx* ctx, const rcwa_c128* grids_dev, int Nx, int Ny,
                                 int Nx_harmonics, int Ny_harmonics, int batch,
                                 rcwa_c128* C_dev);

/* rcwa.hpp:184 SMatrixLayer for `batch` layers, HBM-resident: erc_dev and
 * urc_dev are device pointers to batch x (PQ x PQ) convolution matrices,
 * S11_dev and S12_dev device pointers to batch x (2PQ x 2PQ) outputs
 * (S22 = S11, S21 = S12, rcwa.cpp:351).  Each layer has its own incidence:
 * Kx, Ky are host arrays of batch x PQ (from rcwa_wave_vectors), k0t a host
 * array of k0 * t per layer.  Work buffers come from the context's arena.
 * Enqueued on the context stream; returns when S11 and S12 are complete. */
rcwa_status rcwa_layer_smatrix_batched(rcwa_ctx* ctx, int PQ, int batch, const rcwa_c128* erc_dev,
                                       const rcwa_c128* urc_dev, const rcwa_c128* Kx,
                                       const rcwa_c128* Ky, const double* k0t,
                                       rcwa_c128* S11_dev, rcwa_c128* S12_dev);

/* rcwa.hpp:214 RedhefferProduct on `batch` pairs, HBM-resident: A_dev[4],
 * B_dev[4] and S_dev[4] (host arrays of four device pointers, blocks S11,
 * S12, S21, S22) each point to batch x (n x n) column-major blocks.  S may
 * alias A (in-place update of A), not B.  Work buffers come from the
 * context's arena.  Returns when S is complete. */
rcwa_status rcwa_star_product_batched(rcwa_ctx* ctx, int n, int batch,
                                      const rcwa_c128* const* A_dev,
                                      const rcwa_c128* const* B_dev, rcwa_c128* const* S_dev);

/* Allocation-free form of rcwa_star_product_batched: same arguments, same
 * kernels and results, but no device memory is allocated -- its work slots,
 * Gauss-Jordan workspace and flags live in a fixed device array of the
 * library (52 MiB per GPU, held until the process ends).  No counterpart in
 * the reference; for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, hosts that forbid hipMalloc).  Limits:
 * n <= 1024, batch <= 64, n * n * batch <= 524288; larger calls return
 * RCWA_E_ARG.  All calls on one GPU share the array, so calls that overlap on
 * one GPU race: issue them one at a time per device. */
rcwa_status rcwa_star_product_batched_noalloc(rcwa_ctx* ctx, int n, int batch,
                                              const rcwa_c128* const* A_dev,
                                              const rcwa_c128* const* B_dev,
                                              rcwa_c128* const* S_dev);

/* ----- structure rasterisation on the device ----- */

/* The reference rasterises its structure on the host before ConvMat
 * (main.cpp:48-91: a circle sampled at cell centres into er[layer][iy][ix],
 * rcwa.hpp:56).  These shapes are sampled the same way on the device:
 * x = (ix + 0.5) Lx / Nx, y = (iy + 0.5) Ly / Ny, and a sample is inside
 *   RCWA_SHAPE_ELLIPSE:  ((x-cx)/ax)^2 + ((y-cy)/ay)^2 <= 1
 *   RCWA_SHAPE_RECT:     |x-cx| <= ax and |y-cy| <= ay  (ay = inf: a line)
 *   RCWA_SHAPE_CIRCLE:   (x-cx)^2 + (y-cy)^2 <= ax^2
 * evaluated in FP64 without fused multiply-adds, so a grid matches a host
 * rasterisation written with the same expressions bit for bit.  No
 * periodic wrap: a shape across the cell edge is given once per image. */
typedef enum { RCWA_SHAPE_ELLIPSE = 0, RCWA_SHAPE_RECT = 1, RCWA_SHAPE_CIRCLE = 2 } rcwa_shape_kind;

typedef struct {
  int layer, kind;
  double cx, cy, ax, ay;
  rcwa_c128 value; /* permittivity (or permeability) inside the shape */
} rcwa_shape;

/* Fills grid_dev (device, num_layers x Ny x Nx, [layer][iy][ix]) with
 * background[layer] (host array) and then the shapes (host array) of each
 * layer in their order, later shapes painting over earlier ones.  Returns
 * when the grid is complete; grid_dev can then be passed to rcwa_solve* with
 * grids_on_device = 1 (the step before ConvMat, without the host-to-device
 * copy of the grids). */
rcwa_status rcwa_rasterize(rcwa_ctx* ctx, int Nx, int Ny, double Lx, double Ly, int num_layers,
                           const rcwa_c128* background, const rcwa_shape* shapes, int nshapes,
                           rcwa_c128* grid_dev);

/* ----- sweep planner ----- */

/* main.cpp:126-141 recomputes every stage for each call, ConvMat of the
 * grids included.  A sweep plan holds a structure whose layers are made of
 * materials: material[layer][iy][ix] (host, values 0 .. num_materials-1)
 * with the permittivity (and permeability) of each material given per sweep
 * point.  ConvMat is linear in the grid, so the plan computes the
 * convolution matrix of every (layer, material) mask once, at creation, and
 * each point's [[eps]] and [[mu]] are the sums eps_m [[mask_m]]: no grid is
 * built or transformed per point.  Points whose eps (and mu) rows are equal
 * (an angle sweep over non-dispersive media) share one [[eps]]^-1 (and
 * [[mu]]^-1); with mu = NULL every material has mu = 1, [[mu]] = I and its
 * inverse is not computed.  rcwa_sweep_run splits the points into chunks of
 * equal size that fit the context's HBM budget (and its chunk limit) and
 * runs each chunk as one batched solve. */
typedef struct rcwa_sweep rcwa_sweep;

typedef struct {
  int Nx, Ny, num_layers, num_materials;
  double Lx, Ly;
  const unsigned char* material; /* host, num_layers * Ny * Nx */
  const double* t;               /* num_layers thicknesses */
} rcwa_sweep_structure;

/* The plan keeps a pointer to ctx (which must outlive it) and device memory
 * for the mask convolution matrices: (layer, material) pairs x PQ^2
 * complex128. */
rcwa_status rcwa_sweep_create(rcwa_ctx* ctx, const rcwa_sweep_structure* structure,
                              const rcwa_params* params, rcwa_sweep** out);
/* npoints sweep points.  Point p has the nsrc sources src[p*nsrc ..
 * p*nsrc+nsrc-1] (sources of equal incidence share one S-matrix, as in
 * rcwa_solve), the permittivities eps[p*num_materials + m] and, unless mu is
 * NULL, the permeabilities mu[p*num_materials + m].  out[p*nsrc + s] as in
 * rcwa_solve.  Same results as rcwa_solve_batch on the rasterised grids,
 * to the rounding of the convolution matrices. */
rcwa_status rcwa_sweep_run(rcwa_sweep* sweep, int npoints, const rcwa_source* src, int nsrc,
                           const rcwa_c128* eps, const rcwa_c128* mu, rcwa_precision precision,
                           rcwa_results* out);
rcwa_status rcwa_sweep_destroy(rcwa_sweep* sweep);

/* ----- serving form (continuous batching) ----- */

typedef struct rcwa_server rcwa_server;

typedef struct {
  size_t hbm_budget;     /* bytes of HBM the server may hold (0 = 60% of free) */
  int max_batch;         /* largest batched launch (problems), 0 = auto */
  int n_workers;         /* concurrent GPU streams, 0 = 2 */
  rcwa_precision precision;
  int autotune;          /* 1: ZGEMM tile shapes tuned on first use (rcwa_ctx_set_autotune) */
} rcwa_server_config;

typedef struct {
  double t_submit, t_start, t_finish; /* seconds on the server clock */
  int batch_size;                     /* size of the launch it ran in */
} rcwa_ticket_info;

rcwa_status rcwa_server_create(int device, const rcwa_server_config* cfg, rcwa_server** out);
/* Thread-safe.  Copies the problem's host data; the caller may reuse its
 * buffers as soon as this returns. */
rcwa_status rcwa_server_submit(rcwa_server* srv, const rcwa_problem* prob, uint64_t* ticket);
/* Non-blocking: RCWA_E_NOTREADY until done; then fills out (nsrc entries). */
rcwa_status rcwa_server_poll(rcwa_server* srv, uint64_t ticket, rcwa_results* out,
                             rcwa_ticket_info* info);
rcwa_status rcwa_server_wait(rcwa_server* srv, uint64_t ticket, rcwa_results* out,
                             rcwa_ticket_info* info);
rcwa_status rcwa_server_cancel(rcwa_server* srv, uint64_t ticket);
double rcwa_server_clock(rcwa_server* srv);
rcwa_status rcwa_server_destroy(rcwa_server* srv);

#ifdef __cplusplus
}
#endif

#endif /* RCWA_AMD_H */
