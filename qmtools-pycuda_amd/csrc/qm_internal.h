// Internal declarations shared by the qmgrid translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "qmgrid.h"

// ------------------------------------------------------------------ device-side molecule view
// Basis values skip a primitive where alpha r^2 > kPhiExpCut (below 2e-22 of its
// coefficient) and a whole shell where r^2 > rcut2 = kPhiExpCut / min alpha.
constexpr double kPhiExpCut = 50.0;

struct DevShell {
  int atom;     // index into atoms
  int L;        // 0..2
  int nprim;    // primitives
  int poff;     // first primitive
  int bf;       // first basis function
  float rcut2;  // squared distance (bohr^2) beyond which every primitive is cut
  int pad_[2];
};

struct MolView {
  const double4* atoms;   // x, y, z (bohr), Z
  const DevShell* shells;
  const double2* prims;   // alpha, coeff
  int natoms, nshell, nbf, nprim;
};

struct GeomView {
  int nx, ny, nz;
  double h;
  double ox, oy, oz;
};

inline GeomView geom_view(const qm_geom* g) {
  GeomView v;
  v.nx = g->nx; v.ny = g->ny; v.nz = g->nz;
  v.h = g->step;
  v.ox = g->origin[0]; v.oy = g->origin[1]; v.oz = g->origin[2];
  return v;
}

// ------------------------------------------------------------------ host-side objects
struct Workspace {
  void* ptr = nullptr;
  size_t bytes = 0;
};

enum WsSlot {
  WS_PHI = 0,      // basis values of a chunk of points
  WS_PHI2,         // second plane (lo parts / scales)
  WS_RHO,          // per-sub-point density
  WS_RED,          // reduction partials
  WS_SCALAR,       // small device scalars
  WS_FFT_A,        // FFT work buffers
  WS_FFT_B,
  WS_JAC_A,        // Jacobi correction buffers
  WS_JAC_B,
  WS_JAC_C,
  WS_GTO,          // Hermite distributions
  WS_MISC,
  WS_BATCH,        // batched-launch tables
  WS_BATCH2,
  WS_VNE,          // diffusion V_ne partials and loop state
  WS_PART,         // per-column-block partial densities of the tiled contraction
  WS_MASK,         // per-row-tile activity bits of the basis-function groups
  WS_COUNT
};

struct FftPlanCache;

// kernel timing records (qm_profile / qm_profile_read)
struct ProfRec {
  int id;
  hipEvent_t a, b;
  int weight;   // launches covered by the interval (a graph replay covers several)
};

// an instantiated graph kept alive until the event recorded after its last launch fires
struct LiveGraph {
  hipGraph_t g;
  hipGraphExec_t x;
  hipEvent_t done;
};

struct qm_ctx {
  int device = 0;
  std::string err;
  Workspace ws[WS_COUNT];
  double* host_scalars = nullptr;   // pinned host buffer for results
  FftPlanCache* fft = nullptr;
  bool prof = false;
  hipStream_t graph_stream = nullptr;   // non-blocking stream for graph capture (default priority)
  std::vector<LiveGraph> graphs;        // graphs whose replays may still run
  // the last batched Jacobi block graph, reused while a call's item table (and so every kernel
  // argument and grid) is the same: a server lane replays one set of molecules round after round
  std::vector<char> jgraph_key;
  hipGraph_t jgraph = nullptr;
  hipGraphExec_t jgraph_exec = nullptr;
  // pinned staging of batched item tables: the upload is stream-ordered, no host wait for the stream
  void* table_pinned = nullptr;
  size_t table_pinned_bytes = 0;
  hipEvent_t table_uploaded = nullptr;  // recorded after the last upload from table_pinned
  void* gto_staged = nullptr;           // pinned staging of GTO distribution tables (grown as needed;
  size_t gto_staged_bytes = 0;          // every upload from it completes before the build returns)
  int split_tm = 3;                     // split contraction variant (0..3 tiled, 4 fused, 64 / 128 all-columns)
  int f64_variant = 4;                  // FP64 contraction: 0 wave-cyclic, 1 dense per-wave, 2 tiled, 3 rows, 4 tiled + DMA
  int jacobi_zs = 0;                    // z planes per Jacobi sweep workgroup (0 = chosen per grid)
  int jacobi_pairs = 1;                 // FP32 two-sweep passes with two columns per lane (even nx)
  int tile_skip = 1;                    // contractions skip basis-function groups that are zero on a row tile
  int ncu = 256;                        // compute units of the device
  std::vector<ProfRec> recs;
};

struct qm_mol {
  int natoms = 0, nshell = 0, nprim = 0, nbf = 0, nbf_pad = 0;
  double4* atoms = nullptr;
  DevShell* shells = nullptr;
  double2* prims = nullptr;
  float2* prims_l2e = nullptr;      // (alpha log2 e, coeff) in FP32 for the FP32 basis kernel
  int2* kgroups = nullptr;          // per 8 functions (nbf_pad / 8): shell of the first function, its m index
  // host copies (used to build per-request tables)
  std::vector<double> xyz, Z;
  std::vector<DevShell> h_shells;
  std::vector<double> h_alpha, h_coeff;
  // density matrix: total P_eff = 2 * (tril(dm) + tril(dm,-1)^T), nbf_pad^2, zero padded
  bool has_dm = false;
  std::vector<double> h_P;          // nbf x nbf total P_eff (host)
  double* P64 = nullptr;            // [nbf_pad][nbf_pad]
  double* M64 = nullptr;            // upper triangle of P, off-diagonal doubled (same shape)
  void* P16 = nullptr;              // fp16 hi/lo, K8-interleaved, 2 planes
  void* Pb16 = nullptr;             // bf16 hi/lo, K8-interleaved, 2 planes
  void* gto = nullptr;              // Hermite distributions of the GTO potential (gto.hip)
  bool gto_dirty = true;            // set when P or the coordinates change
  double pfac = 1.0;                // power-of-two scale applied to P before the 16-bit split
  double qtot = 0.0;
  int64_t n_p_uploads = 0;          // density-matrix uploads (qm_mol_set_dm)
  int64_t n_coord_updates = 0;      // coordinate updates (qm_mol_set_coords)
  int64_t n_gto_builds = 0;         // builds of the Hermite distribution tables
  MolView view() const {
    MolView v;
    v.atoms = atoms; v.shells = shells; v.prims = prims;
    v.natoms = natoms; v.nshell = nshell; v.nbf = nbf; v.nprim = nprim;
    return v;
  }
};

// ------------------------------------------------------------------ error helpers
int qm_fail(qm_ctx* ctx, int code, const char* fmt, ...);

#define QM_CHECK_HIP(ctx, call)                                                   \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return qm_fail((ctx), QM_EHIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define QM_CHECK_LAUNCH(ctx) QM_CHECK_HIP(ctx, hipGetLastError())

#define QM_REQUIRE(ctx, cond, code, ...)                   \
  do {                                                     \
    if (!(cond)) return qm_fail((ctx), (code), __VA_ARGS__); \
  } while (0)

// Grow-only device workspace from the stream-ordered allocator.  Growth frees the old
// buffer and allocates the new one in the order of stream s, so work queued earlier on s
// (or ordered before s by the caller) still reads the old buffer; no device-wide wait.
int qm_ws(qm_ctx* ctx, hipStream_t s, int slot, size_t bytes, void** out);
int qm_enter(qm_ctx* ctx);   // hipSetDevice + clear error

inline int64_t npoints(const qm_geom* g) { return (int64_t)g->nx * g->ny * g->nz; }
int qm_check_geom(qm_ctx* ctx, const qm_geom* g);

// ------------------------------------------------------------------ cross-unit internals
// deterministic FP64 reductions (reduce.hip)
int qm_reduce_dev(qm_ctx* ctx, hipStream_t s, int op, int dtype, const void* a, const void* b,
                  int64_t n, double* dev_result);
int qm_scale_dev(qm_ctx* ctx, hipStream_t s, int dtype, void* a, int64_t n, const double* dev_num,
                 double num_const, const double* dev_den);
// density-matrix device forms (density.hip)
int qm_density_upload_P(qm_ctx* ctx, qm_mol* mol, hipStream_t s);
void qm_fft_cache_free(FftPlanCache* c);
size_t qm_fft_cache_bytes(const FftPlanCache* c);
void qm_gto_cache_free(void* c);
void qm_reclaim_graphs(qm_ctx* ctx, bool wait);
// device address of the allocation-free reduction's scratch, looked up once per GPU (reduce.hip)
int qm_noalloc_setup(int device);

// Brackets one kernel launch with events on its stream when profiling is on.
struct ProfScope {
  qm_ctx* ctx;
  hipStream_t s;
  int id, weight;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(qm_ctx* c, hipStream_t st, int kid, int w = 1, bool on = true) : ctx(c), s(st), id(kid), weight(w) {
    if (on && ctx->prof && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
      (void)hipEventRecord(a, s);
  }
  ~ProfScope() {
    if (a && b) {
      (void)hipEventRecord(b, s);
      ctx->recs.push_back({id, a, b, weight});
    }
  }
};
