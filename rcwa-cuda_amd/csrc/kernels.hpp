// Launchers for the gfx950 kernels of the RCWA solve path.  Every launcher
// takes the stream it enqueues on and returns 0 or a hipError_t value.
#pragma once
#include "common.hpp"
#include "rcwa_amd.h"

// zgemm.hip ------------------------------------------------------------------
int zgemm_batched(hipStream_t s, int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta,
                  ZCMat D, ZMat C, int batch);
// Structure flags of zgemm_batched_tri: ZGEMM_TRI_C forms only the 64 x 64
// tiles of C that meet its upper triangle (the tiles strictly below are not
// written); ZGEMM_TRI_B takes B as upper triangular (its entries below the
// diagonal are not read); ZGEMM_TRI_A likewise takes A as upper triangular.
constexpr int ZGEMM_TRI_C = 1, ZGEMM_TRI_B = 2, ZGEMM_TRI_A = 4;
// Tile-shape autotuner of zgemm_batched*: while a ZgemmTuneScope(true) is
// alive on a thread, a product of a shape class (M, N, K, batch by its power
// of two, device) met for the first time is timed on scratch operands in
// every tile shape (a hipMalloc and a synchronisation of the stream) and the
// fastest is kept for the process; elsewhere known classes use their tuned
// shape and unknown ones the default.  Every shape gives the same bits.
struct ZgemmTuneScope {
  explicit ZgemmTuneScope(bool on);
  ~ZgemmTuneScope();
  bool prev;
};
int zgemm_shape_count();
// Forces one tile shape for every product (tests and comparisons); -1 undoes.
void zgemm_force_shape(int shape);
// The tuned table: 6 ints per entry (M, N, K, batch class, device, shape)
// into out (up to max_entries); returns the number of entries.
int zgemm_tuned_entries(int* out, int max_entries);
int zgemm_batched_tri(hipStream_t s, int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta,
                      ZCMat D, ZMat C, int batch, int tri);
// zgemm_batched for the products b with sel[b] != 0 (device flags); C of the
// others is left as it was.
int zgemm_batched_sel(hipStream_t s, int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta,
                      ZCMat D, ZMat C, int batch, const int* sel);

// linalg.hip -----------------------------------------------------------------
// In-place inverse of a batch of n x n matrices by blocked Gauss-Jordan
// elimination with partial (row) pivoting.  ws must hold
// gj_workspace_bytes(n, batch) bytes.  info[b] is set to 1 when a pivot is
// exactly zero (the matrix is singular); it is not cleared.
size_t gj_workspace_bytes(int n, int batch);
int gj_inverse_batched(hipStream_t s, int n, ZMat M, int batch, void* ws, int* info);
// In-place inverse for the fold's I - X11 P22: for n > 512 in batches of at
// most 64 (where Gauss-Jordan's n/16 .. n/4 panel steps on one workgroup per
// matrix leave most of the chip idle) by 2x2-block recursion on the Schur
// complement, the blocks of at most 512 rows inverted by Gauss-Jordan with
// partial pivoting inside the block and the rest as half-size ZGEMMs;
// otherwise gj_inverse_batched.  No pivoting across blocks: for matrices whose leading
// blocks are well conditioned, as I - K with small K.  ws must hold
// block_inverse_ws_bytes(n, cap) bytes for a batch <= cap.
size_t block_inverse_ws_bytes(int n, int cap);
int block_inverse_batched(hipStream_t s, int n, ZMat M, int batch, void* ws, int* info);

// flag[b] = 1 when ||M_b||_1 ||F_b||_1 > kappa (column sums of |re| + |im|,
// within a factor 2 of the complex 1-norms), else 0: with F = M^-1 an
// estimate of M's 1-norm condition number.  scratch: device memory of
// cond1_scratch_bytes(n, batch) bytes.
size_t cond1_scratch_bytes(int n, int batch);
int cond1_flags(hipStream_t s, int n, ZCMat M, ZCMat F, double kappa, int* flag, int batch,
                double* scratch);

// Elementwise helpers.
int zset_identity(hipStream_t s, int n, ZMat M, int batch, zc diag);
// dst (cols x rows, leading dimension cols) = src^H (rows x cols).
int zconj_transpose_batched(hipStream_t s, int rows, int cols, ZCMat src, ZMat dst, int batch);
int zcopy_batched(hipStream_t s, int rows, int cols, ZCMat src, ZMat dst, int batch);
// M(i,j) = r[b*rstride + i] * M(i,j) * c[b*cstride + j]; r or c may be null.
int zscale_rc(hipStream_t s, int rows, int cols, ZMat M, const zc* r, long long rstride,
              const zc* c, long long cstride, int batch);
// dst = src * diag(c) (column scaling) into a separate buffer.
int zscale_cols_to(hipStream_t s, int rows, int cols, ZCMat src, const zc* c, long long cstride,
                   ZMat dst, int batch);
// Normalise every column to unit 2-norm.
int znormalize_cols(hipStream_t s, int n, ZMat M, int batch);
// out[b] = scale * ||M_b||_F (Frobenius norm) for a batch of rows x cols matrices.
int zfrob_norms(hipStream_t s, int rows, int cols, ZCMat M, double scale, double* out,
                int batch);
// Count non-finite entries of `count` complex values (adds to *flag).
int zcount_nonfinite(hipStream_t s, const zc* p, long long count, int* flag);

// eig.hip --------------------------------------------------------------------
// Eigen-decomposition of a batch of general complex matrices:
// A (overwritten by its Schur factor) -> eigenvalues w[b*n + i] and unit
// 2-norm eigenvectors W.  ws must hold eig_workspace_bytes(n, batch) bytes.
// info[b] = 1 if the QR iteration did not converge within maxit_per_row*n
// sweeps.
size_t eig_workspace_bytes(int n, int batch);
// marks (optional): four events recorded before the Hessenberg reduction,
// before the QR sweeps, before the eigenvector phase and at the end.
// log_cap > 0 caps the QR rotation log per matrix (entries; default 2n^2):
// when it fills, the sweep kernel applies it itself.  A tuning and test knob.
// inverse: W receives the inverse of the unit-norm eigenvector matrix
// instead (formed from the Schur vectors and the triangular eigenvectors of
// the Schur form, without a general inverse).
// phase: 0 the whole decomposition; 1 its first part, which returns once the
// QR iteration's first round is queued (nothing waits on the stream); 2 the
// rest, called with the same arguments after phase 1 (it first waits for that
// round).  Phases 1 and 2 let a caller queue work on other streams while the
// QR iteration runs.
int eig_batched(hipStream_t s, int n, ZMat A, zc* w, ZMat W, int batch, void* ws, int* info,
                int maxit_per_row, hipEvent_t* marks = nullptr, long long log_cap = 0,
                bool inverse = false, int phase = 0);
// The three phases, exposed for tests.
int hessenberg_batched(hipStream_t s, int n, ZMat A, ZMat Q, int batch);
// Blocked (compact-WY panels + ZGEMM updates); ws holds hess_blocked_ws_bytes(n, batch).
size_t hess_blocked_ws_bytes(int n, int batch);
int hessenberg_blocked_batched(hipStream_t s, int n, ZMat A, ZMat Q, int batch, void* ws);
// ws (schur_ws_bytes, required) holds the rotation and block logs: the sweeps
// run on the active block and log their rotations, which are then applied to
// the Schur vectors Z.  phase as for eig_batched: 1 queues the first round
// only, 2 waits for it and runs any further rounds, 0 both.
size_t schur_ws_bytes(int n, int batch);
int schur_batched(hipStream_t s, int n, ZMat T, ZMat Z, int batch, int* info, int maxit_per_row,
                  void* ws, long long log_cap = 0, int phase = 0);
// unit: T has a unit diagonal; X = T^-1 (no shift).
int trevc_batched(hipStream_t s, int n, ZCMat T, ZMat X, int batch, bool unit = false);

// rcwa_kernels.hip -----------------------------------------------------------
// Cell-centre rasterisation of layer-sorted shapes (shapes first[l] ..
// first[l+1]-1 belong to layer l) over background bg[l]: grid [l][iy][ix].
int rasterize(hipStream_t s, int Nx, int Ny, double Lx, double Ly, int L, const zc* bg,
              const rcwa_shape* shapes, const int* first, zc* grid);
// Truncated 2-D DFT of grids [b][iy][ix] at frequencies |fx| <= 2hx,
// |fy| <= 2hy, normalised by 1/(Nx Ny): F[b][fy+2hy][fx+2hx].
int convmat_dft(hipStream_t s, const zc* grids, int Nx, int Ny, int hx, int hy, zc* F,
                int batch, zc* rowbuf);
size_t convmat_rowbuf_elems(int Nx, int Ny, int hx, int batch);
// Toeplitz gather C[b](m, n) = F[b](dq, dp) (rcwa.cpp:56-78), zero outside
// the grid's frequency range.
int convmat_gather(hipStream_t s, const zc* F, int Nx, int Ny, int hx, int hy, ZMat C,
                   int batch, const int* map = nullptr);  // dest slot map[b] (device) if given
// P and Q of rcwa.cpp:264-285 from the inverses of [[eps]] and [[mu]].
// Wave vectors Kx, Ky are per batch entry (stride N).
int build_pq(hipStream_t s, int N, ZCMat erc, ZCMat urc, ZCMat erci, ZCMat urci, const zc* Kx,
             const zc* Ky, ZMat P, ZMat Q, int batch);
// lam = usqrt(lam2), ilam = 1/lam, X = exp(-k0 t lam) (rcwa.cpp:317, 339).
int layer_modes(hipStream_t s, int n, const zc* lam2, zc* lam, zc* ilam, zc* X,
                const double* k0t, int batch);
// out = M V0 for the gap medium's V0 given by its four diagonals
// v0[b][4][N] (rcwa.cpp:225).
int mul_v0(hipStream_t s, int N, ZCMat M, const zc* v0, ZMat out, int batch);
// A, B of rcwa.cpp:335-338 from W^-1 and Y = W^-1 P V0: with V = Q W L^-1
// and P Q = W L^2 W^-1, V^-1 V0 = L^-1 Y, so A = W^-1 + L^-1 Y and
// B = W^-1 - L^-1 Y (L = diag(lam), ilam = 1/lam).  A overwrites Y.
int build_ab(hipStream_t s, int n, ZCMat Wi, ZMat Y, const zc* ilam, ZMat B, int batch);
// Dense n x n matrix from a 2x2 block structure of diagonals d + b*dstride, [4][N].
int expand_block_diag(hipStream_t s, int N, const zc* d, long long dstride, ZMat M, int batch);
// Gather columns cols[0..ncol) of M into out[b][c][n].
int gather_cols(hipStream_t s, int n, ZCMat M, const int* cols, int ncol, zc* out, int batch);
// Sweep plans: mask grids out[k] = (material[layer[k]] == mat[k]) of npairs
// (layer, material) pairs; out[map[g L + l]] = sum_k coef[g M + mat[k]] C[k]
// over layer l's pairs first[l] .. first[l+1]-1 (N x N matrices, G
// structures); dst[dmap[k]] = src[smap[k]] for matrices of `elems` entries.
int material_masks(hipStream_t s, int Nx, int Ny, const unsigned char* material,
                   const int* layer, const int* mat, int npairs, zc* out);
int mix_convmat(hipStream_t s, int N, int L, int M, const int* first, const int* mat,
                const zc* C, const zc* coef, const int* map, zc* out, int G);
int gather_mats(hipStream_t s, long long elems, const zc* src, const int* smap, zc* dst,
                const int* dmap, int count);
