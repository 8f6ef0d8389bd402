/*
 * qmgrid: C-ABI of the MI355X (gfx950) grid hot path of QMTools.
 *
 * Every entry point replaces one reference mechanism: a pycuda SourceModule
 * JIT-compiled kernel launched with prepared_call, plus the host logic around
 * it in qmtools/qmtools.py.  The comment of each entry point cites the
 * reference kernel and the Python call site it replaces.
 *
 * Conventions
 *  - Plain C types only; device buffers are raw device pointers (the Python
 *    host passes torch tensors' data_ptr()), `stream` is a hipStream_t passed
 *    as void* (NULL = the null stream).
 *  - Grids are stored x-fastest: point (i,j,k) at i + nx*(j + ny*k), as the
 *    reference kernels index them (kernel_density.cu:151-154).
 *  - Lengths in bohr unless stated; the FFT solver works in Angstrom/volt
 *    units like the reference's kernel_fft.cu.
 *  - Return value: QM_OK (0) or an error code; qm_last_error(ctx) describes
 *    the last failure.  Nothing throws across this boundary.
 *  - Asynchronous on `stream`, except where a result is returned to a host
 *    scalar (those synchronise the stream before returning).
 *  - A context is used by one host thread at a time.
 */
#ifndef QMGRID_H
#define QMGRID_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qm_ctx qm_ctx;   /* per-device context: workspaces, plans, tables */
typedef struct qm_mol qm_mol;   /* device-resident molecule: nuclei, basis tables, density matrix */

enum {
  QM_OK = 0,
  QM_EINVAL = 1,     /* bad argument */
  QM_ENOMEM = 2,     /* device allocation failed */
  QM_EHIP = 3,       /* HIP runtime error */
  QM_ENOCONV = 4,    /* iteration cap reached before the tolerance */
  QM_ELIMIT = 5      /* size beyond a documented limit */
};

/* element type of a grid buffer */
enum { QM_F64 = 0, QM_F32 = 1 };

/* arithmetic of the density contraction */
enum {
  QM_PREC_FP64 = 0,       /* FP64 basis values, FP64 MFMA contraction */
  QM_PREC_F16X3 = 1,      /* FP32 basis values, FP16 hi/lo split, 3 MFMA products, FP32 accumulate */
  QM_PREC_BF16X3 = 2,     /* FP32 basis values, BF16 hi/lo split, 3 MFMA products, FP32 accumulate */
  QM_PREC_FP32 = 3        /* FP32 arithmetic (Jacobi / FFT mixed path) */
};

/* Jacobi boundary handling */
enum { QM_BC_ZERO = 0, QM_BC_MULTIPOLE = 1 };

/* FFT Poisson mode */
enum { QM_FFT_PERIODIC = 0, QM_FFT_FREESPACE = 1 };

/* reductions */
enum { QM_RED_SUM = 0, QM_RED_ABSDIFF = 1, QM_RED_RELDIFF = 2 };

typedef struct {
  int nx, ny, nz;        /* points per axis (any positive size; the Python layer enforces multiples of 8) */
  double step;           /* grid step, bohr */
  double origin[3];      /* grid origin, bohr */
} qm_geom;

/* ---------------------------------------------------------------- context */

/* Creates the context on `device` (replaces `import pycuda.autoinit`, qmtools.py:9). */
int qm_ctx_create(int device, qm_ctx** out);
int qm_ctx_destroy(qm_ctx* ctx);
const char* qm_last_error(const qm_ctx* ctx);

/* Device bytes a context holds between calls: its grow-only workspaces plus the FFT twiddle
 * tables and cached Green's-function transforms (at most 4 GiB of those per context, least
 * recently used released first).  No reference counterpart; lets a request loop charge each
 * lane's resident memory against its device-memory budget. */
int qm_ctx_memory(const qm_ctx* ctx, int64_t* bytes);
/* ABI version: major*10000 + minor*100 + patch */
int qm_version(void);

/* ---------------------------------------------------------------- molecule */

/* Uploads nuclei and the per-shell basis tables of a molecule.
 * Replaces the device copies made by BasisSet.__init__ (qmtools.py:102-106) and
 * Molecule.__init__ (qmtools.py:174-178).  Shells are listed in basis-function
 * order (atom, then shell); shell s has L = shell_L[s] <= 2, its 2L+1 functions
 * in m order (0, +1, -1, +2, -2), and primitives alpha/coeff[shell_poff[s] ..
 * shell_poff[s] + shell_nprim[s]). */
int qm_mol_create(qm_ctx* ctx, int natoms, const int* Z, const double* xyz_bohr,
                  int nshell, const int* shell_atom, const int* shell_L,
                  const int* shell_nprim, const int* shell_poff,
                  int nprim, const double* alpha, const double* coeff, qm_mol** out);

/* Sets the alpha-spin density matrix (nbf x nbf, row-major, host memory).  The
 * kernels use its lower triangle, mirrored, times 2 for the closed shell, as
 * the reference kernels do (kernel_density.cu:74-147, kernel_hartree.cu:836-838).
 * The upload is ordered on `stream` (work queued earlier on it still sees the old
 * matrix; work on other streams that uses the molecule is ordered by the caller) and
 * waits for that stream only, never for the whole device.
 * Replaces the d_dm upload of Molecule.__init__ (qmtools.py:176). */
int qm_mol_set_dm(qm_ctx* ctx, qm_mol* mol, void* stream, const double* dm);

/* New nuclear positions (bohr), ordered on `stream` like qm_mol_set_dm.  The basis
 * tables and the density matrix stay resident; the GTO potential's distribution
 * table is rebuilt on its next use.
 * Replaces Molecule.Translate's re-upload (qmtools.py:182-188). */
int qm_mol_set_coords(qm_ctx* ctx, qm_mol* mol, void* stream, const double* xyz_bohr);
int qm_mol_nbf(const qm_mol* mol);

/* Resident-data counters of a molecule handle: counts[0] = density-matrix uploads,
 * counts[1] = coordinate updates, counts[2] = builds of the GTO distribution table.
 * No reference counterpart (the reference re-uploads per call); lets a caller check
 * that a moved molecule reuses its resident basis tables and density matrix. */
int qm_mol_stats(const qm_mol* mol, int64_t* counts);
int qm_mol_destroy(qm_ctx* ctx, qm_mol* mol);

/* ---------------------------------------------------------------- density */

/* Electrons per voxel on the cell-centred grid:
 *   Q(v) = 2 h^3 n^-3 sum_{s in n^3} sum_pq P_pq phi_p(r_s) phi_q(r_s),
 *   r_s = origin + (i + (s+1/2)/n) h.
 * If `normalize`, Q is rescaled so that sum Q = sum Z (Compute_density's
 * qtot/qscale step).  `raw_total` (host, may be NULL) receives sum Q before
 * rescaling; this synchronises `stream`.
 * Replaces gpu_densityqube_shmem_subgrid (kernel_density.cu:34-155) and
 * QMTools.Compute_density (qmtools.py:2381-2430).
 * precision: QM_PREC_FP64, QM_PREC_F16X3 or QM_PREC_BF16X3; out_dtype QM_F64 or QM_F32. */
int qm_density(qm_ctx* ctx, qm_mol* mol, void* stream, const qm_geom* g, int subgrid,
               int precision, int out_dtype, void* out, int normalize, double* raw_total);

/* Basis functions on arbitrary points: phi[f * npts + t] (function-major).
 * FP64 only; used by tests and by the density path's first stage. */
int qm_basis_points(qm_ctx* ctx, qm_mol* mol, void* stream, const double* pts, int64_t npts,
                    double* phi);

/* ---------------------------------------------------------------- reductions */

/* op QM_RED_SUM: sum a; QM_RED_ABSDIFF: sum |a-b|; QM_RED_RELDIFF: sum |a-b|/b.
 * FP64 accumulation in a fixed order (bitwise reproducible).  Result to the host.
 * Replaces gpu_total / gpu_qdiff / gpu_qdiff_rel (kernel_utils.cu:68-168) and
 * Compute_qtot / Compute_qdiff (qmtools.py:2796-2831). */
int qm_reduce(qm_ctx* ctx, void* stream, int op, int dtype, const void* a, const void* b,
              int64_t n, double* result);

/* qm_reduce that allocates nothing: same operations, same kernels, the same bits in *result.
 * No counterpart in the reference; for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid allocation).  The block partials and the device
 * result live in a fixed file-scope device array of 16.6 KB per GPU (2048 partials + result),
 * which holds what any n needs; n at or beyond 2^34 (the library's grid limit) is refused with
 * QM_ELIMIT.  Every call on a GPU shares that GPU's array, so calls that overlap on one GPU
 * (other streams, other host threads) race; nothing in it needs to start at zero.  Synchronises
 * `stream` like qm_reduce. */
int qm_reduce_noalloc(qm_ctx* ctx, void* stream, int op, int dtype, const void* a, const void* b,
                      int64_t n, double* result);

/* a *= factor.  Replaces gpu_rescale (kernel_utils.cu:54-65), Compute_qscale (qmtools.py:2807-2812). */
int qm_scale(qm_ctx* ctx, void* stream, int dtype, void* a, int64_t n, double factor);

/* Nuclear seed: writes Z * prod_d |a_d - i_d h|/h to the 8 nodes around each
 * nucleus (node positions origin + i h); other points are left untouched.
 * Nodes outside the grid are skipped.  A node shared by several nuclei gets
 * the value of the nucleus with the highest index (the reference lets such
 * writes race).  Replaces gpu_density_seed
 * (kernel_utils.cu:5-50) and Compute_qseed (qmtools.py:2645-2672). */
int qm_seed(qm_ctx* ctx, qm_mol* mol, void* stream, const qm_geom* g, int dtype, void* out);

/* ---------------------------------------------------------------- Jacobi */

/* Multipole (q, dipole, traceless quadrupole) of the node charges rho about the
 * box centre; writes V_b = -phi_mp/(4 pi) to the outer voxel layer of `v`
 * (interior untouched).  No reference counterpart (multipole Dirichlet option). */
int qm_multipole_boundary(qm_ctx* ctx, void* stream, const qm_geom* g, int dtype,
                          const void* rho, void* v, double moments_out[10]);

/* Jacobi relaxation of lap_h V = rho/h^3 (V = -phi/(4 pi)):
 *   V'(x) = (sum_6 V_nb - rho(x)/h)/6 on interior points; the outer layer of
 *   `v` is held (0 reproduces the reference, multipole values from
 *   qm_multipole_boundary give open boundaries).
 * precision QM_PREC_FP64: FP64 sweeps in pairs v->tmp->v; every `check_every`
 *   sweeps (even, > 0) computes delta = sum|v - tmp| like the reference loop
 *   (which checks every 100 iterations = 200 sweeps) and stops when
 *   delta <= tolerance.
 * precision QM_PREC_FP32: rho and v are FP64 buffers, tmp is unused; every
 *   `check_every` sweeps (> 0, e.g. 50) the FP64 residual charge is formed
 *   and the block's sweeps run in FP32 on the correction (defect correction:
 *   the same iterate as plain Jacobi in exact arithmetic); delta is the last
 *   sweep's sum|e_k - e_(k-1)|.
 * Stops after max_sweeps sweeps at the latest; then returns QM_ENOCONV
 * unless tolerance < 0, which means: run exactly max_sweeps sweeps, no checks.
 * `sweeps_done` and `last_delta` (host, may be NULL) report the outcome.
 * Replaces gpu_hartree_iteration (kernel_hartree.cu:85-171) and the loop of
 * Compute_hartree_iter (qmtools.py:2547-2561). */
int qm_jacobi(qm_ctx* ctx, void* stream, const qm_geom* g, int precision, const void* rho,
              void* v, void* tmp, int64_t max_sweeps, double tolerance, int check_every,
              int64_t* sweeps_done, double* last_delta);

/* ---------------------------------------------------------------- FFT Poisson */

/* Total-charge right-hand side and Poisson solve by a hand-written 3-D FFT.
 *   rhs = -( -Q/h^3 + sum_A Z_A N_sigma(r - R_A) [r <= cutoff*sigma] ) / eps0   (Angstrom, volts)
 * mode QM_FFT_PERIODIC: V^ = rhs^ * (-1/(4 pi^2 |k|^2)), V^(0) = 0 (reference semantics);
 * mode QM_FFT_FREESPACE: convolution with 1/(4 pi eps0 r) on a 2x zero-padded grid.
 * q: electrons per voxel, x-fastest (dtype); out: V in volts, C order [i][j][k] (dtype).
 * precision QM_PREC_FP64 (complex128) or QM_PREC_FP32 (complex64).
 * Replaces gpu_add_nuclear_charge + gpu_poisson_frequency_solve (kernel_fft.cu:23-106),
 * the cuFFT plans and QMTools.Compute_hartree_fft (qmtools.py:2578-2639). */
int qm_hartree_fft(qm_ctx* ctx, qm_mol* mol, void* stream, const qm_geom* g, int precision,
                   int dtype, const void* q, double sigma, double cutoff, int mode, void* out);

/* ---------------------------------------------------------------- GTO V_H */

/* Total potential (Hartree a.u.) of nuclei + GTO density at arbitrary points:
 *   V(C) = sum_A Z_A/|C - R_A| (|C-R_A| > 1e-8) - sum_pq 2 P_pq int phi_p phi_q / |r - C|
 * FP64, McMurchie-Davidson/Boys.  pts: device double[npts*3]; out: device double[npts].
 * Replaces gpu_hartree_GTO (kernel_hartree.cu:288-869) for point lists. */
int qm_hartree_gto_points(qm_ctx* ctx, qm_mol* mol, void* stream, const double* pts,
                          int64_t npts, double* out);

/* Same on the grid nodes origin + (i,j,k) h, written x-fastest (dtype).
 * Replaces gpu_hartree_GTO + QMTools.Compute_hartree (qmtools.py:2436-2489). */
int qm_hartree_gto_grid(qm_ctx* ctx, qm_mol* mol, void* stream, const qm_geom* g, int dtype,
                        void* out);

/* ---------------------------------------------------------------- V_ne */

/* V(v) = [sum_{s in ndiv^3} sum_A Z_A/|r_s - R_A|] / (cnt h^3), r_s = origin + i h + (s+1/2) h/ndiv,
 * cnt = number of (s, A) with r_s != R_A.  FP64.
 * Replaces gpu_vne_classic (kernel_vne_classic.cu:13-84), Compute_VNe_classic (qmtools.py:2750-2790). */
int qm_vne_classic(qm_ctx* ctx, qm_mol* mol, void* stream, const qm_geom* g, int ndiv, int dtype,
                   void* out);

/* Diffusion / adsorption pseudo-nuclear field: reference QMTools.Compute_VNe
 * (qmtools.py:2677-2745) and gpu_vne (kernel_vne.cu:13-128).  v (result) and
 * tmp are caller buffers of npts elements of `dtype`; v is cleared first.
 * Runs pairs of sweeps v -> tmp -> v until sum|v - tmp| / npts <= tolerance
 * (checked after each pair) or until the pair count exceeds maxiters (finite;
 * the reference's default is 10 npts^(1/3)), then *converged = 0.  A
 * tolerance >= 1 runs no sweep, as in the reference.  Synchronises `stream`. */
int qm_vne_diffusion(qm_ctx* ctx, qm_mol* mol, void* stream, const qm_geom* g, int dtype, double adsorb,
                     double diff, double tolerance, double maxiters, void* v, void* tmp, int64_t* pairs,
                     double* delta, int* converged);

/* ---------------------------------------------------------------- batches */

/* Density of nmol molecules on their own grids (sizes and steps may differ),
 * subgrid 1, in fused launches: the grids are cut into chunks packed into
 * rounds of the basis-value workspace; each round is one basis-value launch
 * plus one contraction launch per column-tile class.  outs[i] receives
 * molecule i's density (out_dtype), normalised to its sum Z if `normalize`.
 * The per-point arithmetic is the single-molecule path's (qm_density).
 * No reference counterpart (the reference runs one molecule per call). */
int qm_density_batch(qm_ctx* ctx, void* stream, int nmol, qm_mol* const* mols, const qm_geom* geoms,
                     void* const* outs, int precision, int out_dtype, int normalize);

/* nsweeps Jacobi sweeps on n grids at once (each its own size/step), one
 * launch per sweep for the whole batch.  With use_graph, each block of
 * `block` sweeps (even) is captured once into a hipGraph and replayed.
 * precision QM_PREC_FP64: rho[i], v[i], tmp[i] FP64; the outer layer of v[i]
 *   is held (zero or multipole values).
 * precision QM_PREC_FP32: rho[i], v[i] FP64, tmp unused; each block forms the
 *   FP64 residual and runs its sweeps in FP32 on the correction (as qm_jacobi).
 * Waits for the caller's stream before enqueueing (the item table is
 * uploaded from host memory), then returns with the work enqueued and the
 * caller's stream ordered after it.  No tolerance checks (fixed sweep
 * count).  No reference counterpart. */
int qm_jacobi_batch(qm_ctx* ctx, void* stream, int n, const qm_geom* geoms, const void* const* rho,
                    void* const* v, void* const* tmp, int precision, int64_t nsweeps, int block, int use_graph);

/* ---------------------------------------------------------------- tuning */

/* Tile-shape knobs of the density contraction and the Jacobi sweep (no reference counterpart). */
enum {
  QM_TUNE_SPLIT_VARIANT = 0, /* 16-bit split contraction, 128 x 128 tiles with per-block partial sums:
                               0 / 2 = FP32 basis values split while staged (8 / 4 waves), 1 = basis values
                               written pre-split, staged through registers, 3 = pre-split, fetched by LDS DMA
                               (default); 4 = basis values evaluated inside the contraction kernel, 128 points
                               x 320 columns per workgroup (no basis-value buffer; slower on the 60-atom
                               workload); 5 = pre-split, DMA, 256 points x 320 columns per workgroup and
                               128 x 80 per wave (slower); 64 or 128 = points per workgroup of the
                               all-columns kernel */
  QM_TUNE_F64_VARIANT = 1,  /* FP64 contraction: 0 = all columns per workgroup (wave-cyclic tiles), 1 = per-wave
                               loads of the dense P, no barriers, 2 = 128 x 128 tiles with per-block partial sums,
                               3 = 128 x 64 tiles with diagonal skipping, 4 = 128 x 128 tiles fed by LDS DMA with
                               diagonal skipping (default) */
  QM_TUNE_JACOBI_ZS = 2,    /* z planes per Jacobi sweep workgroup, 0 = chosen per grid (the two-sweep passes
                               of qm_jacobi_batch: one length for the whole batch), else 1..4096 */
  QM_TUNE_JACOBI_PAIRS = 3, /* FP32 two-sweep Jacobi passes: 1 = two columns per lane on grids with an even nx
                               (default), 0 = one column per lane; the results are bitwise the same */
  QM_TUNE_TILE_SKIP = 4,    /* FP64 contraction with LDS DMA: 1 = skip the products of basis-function groups
                               that are zero on a whole row tile (default), 0 = form every product; they add
                               exact zeros, so the results are the same */
  QM_TUNE_COUNT = 5
};
int qm_tune_set(qm_ctx* ctx, int knob, int value);
int qm_tune_get(qm_ctx* ctx, int knob, int* value);

/* Autotuner: times every candidate value of the knobs that apply to
 * `precision` on a slab of grid `g` (at most ~1.6M points, written to a
 * library workspace, not to any caller buffer), keeps the fastest in the
 * context and returns its time (ms, host) in best_ms.  Synchronises `stream`. */
int qm_autotune_density(qm_ctx* ctx, qm_mol* mol, void* stream, const qm_geom* g, int precision,
                        double* best_ms);

/* ---------------------------------------------------------------- instrumentation */

/* Kernel ids for qm_profile_read. */
enum {
  QM_K_DENSITY_CONTRACT = 0,   /* FP64 MFMA contraction */
  QM_K_DENSITY_SPLIT = 1,      /* FP16/BF16 hi/lo MFMA contraction */
  QM_K_PHI = 2,                /* basis values, FP64 */
  QM_K_PHI_SPLIT = 3,          /* basis values, FP32 + 16-bit split */
  QM_K_FFT_AXIS = 4,           /* one FFT pass along one axis */
  QM_K_GTO = 5,                /* Hermite/Boys potential */
  QM_K_JACOBI = 6,             /* one Jacobi sweep (FP64 or FP32) */
  QM_K_VNE = 7,                /* one pair of diffusion V_ne sweeps with its convergence test */
  QM_K_COUNT = 8
};

/* No reference counterpart (the reference times nothing on the device).
 * enable != 0 starts recording hipEvents around every launch of the kernels
 * above, on the stream they are launched on, and clears earlier records;
 * enable == 0 stops recording.  qm_profile_read synchronises on the recorded
 * events and returns the summed duration (ms) and the launch count of one
 * kernel id. */
int qm_profile(qm_ctx* ctx, int enable);
int qm_profile_read(qm_ctx* ctx, int kernel, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif /* QMGRID_H */
