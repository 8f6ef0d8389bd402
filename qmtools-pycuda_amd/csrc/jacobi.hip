// Jacobi Poisson relaxation and multipole Dirichlet boundaries.
//
// Sweep (reference kernel_hartree.cu:85-171):
//   V'(x) = (sum_6 V_nb - rho(x)/h) / 6   on interior points,
//   V'    = V                             on the outer voxel layer.
// The reference writes 0 on the outer layer; holding whatever the layer
// contains reproduces that when the layer starts at 0 (its V0 = 0,
// qmtools.py:2535-2536) and gives open boundaries when the layer holds
// multipole values.
//
// FP32 path (defect correction): V stays in FP64.  Every block of sweeps
// computes the FP64 residual charge r = rho - h (sum_6 V - 6 V), runs the
// block's sweeps in FP32 on the correction e (e0 = 0, e = 0 on the outer
// layer, source r), and adds e to V.  In exact arithmetic V + e after k
// sweeps equals k plain sweeps from V, so the iterate is the same as the
// FP64 path's; FP32 rounding enters only the correction.
#include <algorithm>
#include <cmath>

#include "qm_internal.h"

namespace {

constexpr int kBX = 64, kBY = 4;

template <typename T>
__global__ __launch_bounds__(kBX* kBY) void k_sweep(const T* __restrict__ rho, const T* __restrict__ v,
                                                    T* __restrict__ vout, int nx, int ny, int nz, T h) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= nx || j >= ny) return;
  const int64_t sx = 1, sy = nx, sz = (int64_t)nx * ny;
  const int64_t idx = i + sy * j + sz * k;
  if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) {
    vout[idx] = v[idx];
    return;
  }
  T s = v[idx - sx] + v[idx + sx];
  s += v[idx + sy];
  s += v[idx - sy];
  s += v[idx + sz];
  s += v[idx - sz];
  s -= rho[idx] / h;
  vout[idx] = s / (T)6;
}

// r = rho - h (sum_6 V - 6 V) on interior points (FP64), 0 on the outer layer, stored FP32
__global__ __launch_bounds__(kBX* kBY) void k_residual(const double* __restrict__ rho,
                                                       const double* __restrict__ v, float* __restrict__ r,
                                                       int nx, int ny, int nz, double h) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= nx || j >= ny) return;
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  const int64_t idx = i + sy * j + sz * k;
  if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) {
    r[idx] = 0.f;
    return;
  }
  const double c = v[idx];
  double s = v[idx - 1] + v[idx + 1];
  s += v[idx + sy];
  s += v[idx - sy];
  s += v[idx + sz];
  s += v[idx - sz];
  r[idx] = (float)(rho[idx] - h * (s - 6.0 * c));
}

__global__ void k_add_correction(double* __restrict__ v, const float* __restrict__ e, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) v[i] += (double)e[i];
}

// ---------------------------------------------------------------- multipoles
constexpr int kMomThreads = 256;
constexpr int kMomBlocks = 1024;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <typename T>
__global__ __launch_bounds__(kMomThreads) void k_moments_pass1(const T* __restrict__ rho, GeomView g,
                                                               int64_t seg, double* __restrict__ part) {
  __shared__ double sm[kMomThreads / 64][10];
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  const int64_t lo = (int64_t)blockIdx.x * seg;
  const int64_t hi = min(n, lo + seg);
  const double cx = 0.5 * g.h * (g.nx - 1), cy = 0.5 * g.h * (g.ny - 1), cz = 0.5 * g.h * (g.nz - 1);
  double m[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t p = lo + threadIdx.x; p < hi; p += kMomThreads) {
    const double q = (double)rho[p];
    if (q == 0.0) continue;
    const int64_t i = p % g.nx, j = (p / g.nx) % g.ny, k = p / ((int64_t)g.nx * g.ny);
    const double dx = i * g.h - cx, dy = j * g.h - cy, dz = k * g.h - cz;
    const double d2 = dx * dx + dy * dy + dz * dz;
    m[0] += q;
    m[1] += q * dx;
    m[2] += q * dy;
    m[3] += q * dz;
    m[4] += q * (3 * dx * dx - d2);
    m[5] += q * (3 * dy * dy - d2);
    m[6] += q * (3 * dz * dz - d2);
    m[7] += q * 3 * dx * dy;
    m[8] += q * 3 * dx * dz;
    m[9] += q * 3 * dy * dz;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 10; ++c) {
    const double s = wave_sum(m[c]);
    if (lane == 0) sm[wid][c] = s;
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    double s = 0;
    for (int w = 0; w < kMomThreads / 64; ++w) s += sm[w][threadIdx.x];
    part[(int64_t)blockIdx.x * 10 + threadIdx.x] = s;
  }
}

__global__ void k_moments_pass2(const double* __restrict__ part, int nb, double* __restrict__ out) {
  if (threadIdx.x >= 10) return;
  double s = 0;
  for (int b = 0; b < nb; ++b) s += part[(int64_t)b * 10 + threadIdx.x];
  out[threadIdx.x] = s;
}

// V_b = -phi_mp / (4 pi) on the six faces; one thread per face point (edges written twice, same value)
template <typename T>
__global__ void k_boundary(const double* __restrict__ mom, GeomView g, T* __restrict__ v) {
  const int64_t fx = (int64_t)g.ny * g.nz, fy = (int64_t)g.nx * g.nz, fz = (int64_t)g.nx * g.ny;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t i, j, k;
  if (t < 2 * fz) {
    const int64_t side = t / fz, r = t % fz;
    i = r % g.nx; j = r / g.nx; k = side ? g.nz - 1 : 0;
  } else if ((t -= 2 * fz) < 2 * fy) {
    const int64_t side = t / fy, r = t % fy;
    i = r % g.nx; k = r / g.nx; j = side ? g.ny - 1 : 0;
  } else if ((t -= 2 * fy) < 2 * fx) {
    const int64_t side = t / fx, r = t % fx;
    j = r % g.ny; k = r / g.ny; i = side ? g.nx - 1 : 0;
  } else {
    return;
  }
  const double X = i * g.h - 0.5 * g.h * (g.nx - 1);
  const double Y = j * g.h - 0.5 * g.h * (g.ny - 1);
  const double Z = k * g.h - 0.5 * g.h * (g.nz - 1);
  const double r2 = X * X + Y * Y + Z * Z;
  const double r = sqrt(r2);
  const double q = mom[0];
  const double pr = mom[1] * X + mom[2] * Y + mom[3] * Z;
  const double QRR = mom[4] * X * X + mom[5] * Y * Y + mom[6] * Z * Z +
                     2.0 * (mom[7] * X * Y + mom[8] * X * Z + mom[9] * Y * Z);
  const double phi = q / r + pr / (r2 * r) + 0.5 * QRR / (r2 * r2 * r);
  v[i + (int64_t)g.nx * (j + (int64_t)g.ny * k)] = (T)(-phi / (4.0 * M_PI));
}

dim3 sweep_grid(const qm_geom* g) {
  return dim3((g->nx + kBX - 1) / kBX, (g->ny + kBY - 1) / kBY, g->nz);
}

template <typename T>
int sweep(qm_ctx* ctx, hipStream_t s, const qm_geom* g, const T* rho, const T* a, T* b) {
  ProfScope ps(ctx, s, QM_K_JACOBI);
  hipLaunchKernelGGL(k_sweep<T>, sweep_grid(g), dim3(kBX, kBY), 0, s, rho, a, b, g->nx, g->ny, g->nz,
                     (T)g->step);
  QM_CHECK_LAUNCH(ctx);
  return QM_OK;
}

int host_delta(qm_ctx* ctx, hipStream_t s, int dtype, const void* a, const void* b, int64_t n, double* out) {
  double* d = nullptr;
  int rc = qm_ws(ctx, WS_SCALAR, 256, (void**)&d);
  if (rc) return rc;
  rc = qm_reduce_dev(ctx, s, QM_RED_ABSDIFF, dtype, a, b, n, d + 1);
  if (rc) return rc;
  QM_CHECK_HIP(ctx, hipMemcpyAsync(ctx->host_scalars + 1, d + 1, sizeof(double), hipMemcpyDeviceToHost, s));
  QM_CHECK_HIP(ctx, hipStreamSynchronize(s));
  *out = ctx->host_scalars[1];
  return QM_OK;
}

int jacobi_fp64(qm_ctx* ctx, hipStream_t s, const qm_geom* g, const double* rho, double* v, double* tmp,
                int64_t max_sweeps, double tol, int check_every, int64_t* done_out, double* delta_out) {
  const int64_t n = npoints(g);
  int64_t done = 0;
  double delta = -1.0;
  bool converged = false;
  int rc;
  while (done + 2 <= max_sweeps) {
    if ((rc = sweep<double>(ctx, s, g, rho, v, tmp))) return rc;
    if ((rc = sweep<double>(ctx, s, g, rho, tmp, v))) return rc;
    done += 2;
    if (tol >= 0 && check_every > 0 && done % check_every == 0) {
      if ((rc = host_delta(ctx, s, QM_F64, v, tmp, n, &delta))) return rc;
      if (delta <= tol) {
        converged = true;
        break;
      }
    }
  }
  if (!converged && done < max_sweeps) {   // odd count: one more sweep, result copied back
    if ((rc = sweep<double>(ctx, s, g, rho, v, tmp))) return rc;
    QM_CHECK_HIP(ctx, hipMemcpyAsync(v, tmp, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    done += 1;
  }
  if (done_out) *done_out = done;
  if (delta_out) *delta_out = delta;
  if (tol >= 0 && !converged) return qm_fail(ctx, QM_ENOCONV, "Jacobi: %lld sweeps, delta %.3e > %.3e",
                                             (long long)done, delta, tol);
  return QM_OK;
}

int jacobi_fp32(qm_ctx* ctx, hipStream_t s, const qm_geom* g, const double* rho, double* v, int64_t max_sweeps,
                double tol, int block, int64_t* done_out, double* delta_out) {
  const int64_t n = npoints(g);
  float *r = nullptr, *e = nullptr, *t = nullptr;
  int rc;
  if ((rc = qm_ws(ctx, WS_JAC_A, sizeof(float) * n, (void**)&r))) return rc;
  if ((rc = qm_ws(ctx, WS_JAC_B, sizeof(float) * n, (void**)&e))) return rc;
  if ((rc = qm_ws(ctx, WS_JAC_C, sizeof(float) * n, (void**)&t))) return rc;
  const unsigned upd_blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
  int64_t done = 0;
  double delta = -1.0;
  bool converged = false;
  while (done < max_sweeps) {
    const int64_t m = std::min<int64_t>(block, max_sweeps - done);
    hipLaunchKernelGGL(k_residual, sweep_grid(g), dim3(kBX, kBY), 0, s, rho, (const double*)v, r, g->nx, g->ny,
                       g->nz, g->step);
    QM_CHECK_LAUNCH(ctx);
    QM_CHECK_HIP(ctx, hipMemsetAsync(e, 0, sizeof(float) * n, s));
    QM_CHECK_HIP(ctx, hipMemsetAsync(t, 0, sizeof(float) * n, s));
    float* a = e;
    float* b = t;
    for (int64_t j = 0; j < m; ++j) {
      if ((rc = sweep<float>(ctx, s, g, r, a, b))) return rc;
      std::swap(a, b);
    }
    // a = e after m sweeps, b = after m-1 sweeps
    if (tol >= 0) {
      if ((rc = host_delta(ctx, s, QM_F32, a, b, n, &delta))) return rc;
    }
    hipLaunchKernelGGL(k_add_correction, dim3(upd_blocks), dim3(256), 0, s, v, (const float*)a, n);
    QM_CHECK_LAUNCH(ctx);
    done += m;
    if (tol >= 0 && delta <= tol) {
      converged = true;
      break;
    }
  }
  if (done_out) *done_out = done;
  if (delta_out) *delta_out = delta;
  if (tol >= 0 && !converged) return qm_fail(ctx, QM_ENOCONV, "Jacobi (FP32): %lld sweeps, delta %.3e > %.3e",
                                             (long long)done, delta, tol);
  return QM_OK;
}

}  // namespace

extern "C" {

int qm_multipole_boundary(qm_ctx* ctx, void* stream, const qm_geom* g, int dtype, const void* rho, void* v,
                          double moments_out[10]) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  if ((rc = qm_check_geom(ctx, g))) return rc;
  QM_REQUIRE(ctx, rho && v, QM_EINVAL, "multipole: null argument");
  QM_REQUIRE(ctx, dtype == QM_F64 || dtype == QM_F32, QM_EINVAL, "multipole: bad dtype");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = npoints(g);
  int64_t nb = std::min<int64_t>(kMomBlocks, std::max<int64_t>(1, (n + 4095) / 4096));
  const int64_t seg = (n + nb - 1) / nb;
  double* part = nullptr;
  double* mom = nullptr;
  if ((rc = qm_ws(ctx, WS_RED, sizeof(double) * 10 * kMomBlocks, (void**)&part))) return rc;
  if ((rc = qm_ws(ctx, WS_SCALAR, 256, (void**)&mom))) return rc;
  mom += 8;
  const GeomView gv = geom_view(g);
  if (dtype == QM_F64)
    hipLaunchKernelGGL(k_moments_pass1<double>, dim3((unsigned)nb), dim3(kMomThreads), 0, s, (const double*)rho,
                       gv, seg, part);
  else
    hipLaunchKernelGGL(k_moments_pass1<float>, dim3((unsigned)nb), dim3(kMomThreads), 0, s, (const float*)rho,
                       gv, seg, part);
  QM_CHECK_LAUNCH(ctx);
  hipLaunchKernelGGL(k_moments_pass2, dim3(1), dim3(64), 0, s, (const double*)part, (int)nb, mom);
  QM_CHECK_LAUNCH(ctx);
  const int64_t nface = 2 * ((int64_t)g->nx * g->ny + (int64_t)g->nx * g->nz + (int64_t)g->ny * g->nz);
  const unsigned fb = (unsigned)((nface + 255) / 256);
  if (dtype == QM_F64)
    hipLaunchKernelGGL(k_boundary<double>, dim3(fb), dim3(256), 0, s, (const double*)mom, gv, (double*)v);
  else
    hipLaunchKernelGGL(k_boundary<float>, dim3(fb), dim3(256), 0, s, (const double*)mom, gv, (float*)v);
  QM_CHECK_LAUNCH(ctx);
  if (moments_out) {
    QM_CHECK_HIP(ctx, hipMemcpyAsync(ctx->host_scalars + 8, mom, 10 * sizeof(double), hipMemcpyDeviceToHost, s));
    QM_CHECK_HIP(ctx, hipStreamSynchronize(s));
    for (int c = 0; c < 10; ++c) moments_out[c] = ctx->host_scalars[8 + c];
  }
  return QM_OK;
}

int qm_jacobi(qm_ctx* ctx, void* stream, const qm_geom* g, int precision, const void* rho, void* v, void* tmp,
              int64_t max_sweeps, double tolerance, int check_every, int64_t* sweeps_done, double* last_delta) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  if ((rc = qm_check_geom(ctx, g))) return rc;
  QM_REQUIRE(ctx, rho && v, QM_EINVAL, "jacobi: null argument");
  QM_REQUIRE(ctx, max_sweeps >= 0, QM_EINVAL, "jacobi: negative sweep count");
  hipStream_t s = (hipStream_t)stream;
  if (precision == QM_PREC_FP64) {
    QM_REQUIRE(ctx, tmp != nullptr, QM_EINVAL, "jacobi: FP64 path needs a second buffer");
    QM_REQUIRE(ctx, check_every >= 0 && check_every % 2 == 0, QM_EINVAL, "jacobi: check_every must be even");
    return jacobi_fp64(ctx, s, g, (const double*)rho, (double*)v, (double*)tmp, max_sweeps, tolerance,
                       check_every, sweeps_done, last_delta);
  }
  if (precision == QM_PREC_FP32) {
    QM_REQUIRE(ctx, check_every > 0, QM_EINVAL, "jacobi: FP32 path needs a residual interval");
    return jacobi_fp32(ctx, s, g, (const double*)rho, (double*)v, max_sweeps, tolerance, check_every,
                       sweeps_done, last_delta);
  }
  return qm_fail(ctx, QM_EINVAL, "jacobi: bad precision %d", precision);
}

}  // extern "C"

// ===========================================================================
// Batched Jacobi: many grids per launch, fixed sweep counts, hipGraph replay.
//
// Each item is one grid (its own size and step).  A sweep over the whole batch
// is one launch: workgroups find their item by binary search on the item
// table's first-block numbers.  A block of sweeps (FP64: 2 K sweeps; FP32:
// residual, clear, K sweeps, correction) is captured once into a hipGraph on a
// context-owned stream and replayed; the caller's stream is ordered before and
// after with events.  Buffers are cleared with a fill kernel, not a memset
// node.
namespace {

struct JItem {
  int nx, ny, nz;
  double h;
  const void* rho;    // FP64 rho (FP64 path) / FP32 residual (FP32 path)
  void* a;            // V (FP64) / e (FP32)
  void* b;            // tmp (FP64) / t (FP32)
  const double* rho64;
  double* v64;
  float* r32;
  int blk0, bx, by;
};

__device__ __forceinline__ int find_item(const JItem* __restrict__ it, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (it[mid].blk0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ bool item_point(const JItem& q, int local, int& i, int& j, int& k) {
  const int bxi = local % q.bx;
  const int rest = local / q.bx;
  const int byi = rest % q.by;
  k = rest / q.by;
  i = bxi * kBX + threadIdx.x;
  j = byi * kBY + threadIdx.y;
  return i < q.nx && j < q.ny;
}

template <typename T>
__global__ __launch_bounds__(kBX* kBY) void kb_sweep(const JItem* __restrict__ items, int n, int dir) {
  const int it = find_item(items, n, blockIdx.x);
  const JItem& q = items[it];
  int i, j, k;
  if (!item_point(q, blockIdx.x - q.blk0, i, j, k)) return;
  const T* rho = (const T*)q.rho;
  const T* v = (const T*)(dir ? q.b : q.a);
  T* vout = (T*)(dir ? q.a : q.b);
  const int64_t sy = q.nx, sz = (int64_t)q.nx * q.ny;
  const int64_t idx = i + sy * j + sz * k;
  if (i == 0 || j == 0 || k == 0 || i == q.nx - 1 || j == q.ny - 1 || k == q.nz - 1) {
    vout[idx] = v[idx];
    return;
  }
  T s = v[idx - 1] + v[idx + 1];
  s += v[idx + sy];
  s += v[idx - sy];
  s += v[idx + sz];
  s += v[idx - sz];
  s -= rho[idx] / (T)q.h;
  vout[idx] = s / (T)6;
}

__global__ __launch_bounds__(kBX* kBY) void kb_residual(const JItem* __restrict__ items, int n) {
  const int it = find_item(items, n, blockIdx.x);
  const JItem& q = items[it];
  int i, j, k;
  if (!item_point(q, blockIdx.x - q.blk0, i, j, k)) return;
  const int64_t sy = q.nx, sz = (int64_t)q.nx * q.ny;
  const int64_t idx = i + sy * j + sz * k;
  float* e = (float*)q.a;
  float* t = (float*)q.b;
  e[idx] = 0.f;
  t[idx] = 0.f;
  if (i == 0 || j == 0 || k == 0 || i == q.nx - 1 || j == q.ny - 1 || k == q.nz - 1) {
    q.r32[idx] = 0.f;
    return;
  }
  const double* v = q.v64;
  const double c = v[idx];
  double s = v[idx - 1] + v[idx + 1];
  s += v[idx + sy];
  s += v[idx - sy];
  s += v[idx + sz];
  s += v[idx - sz];
  q.r32[idx] = (float)(q.rho64[idx] - q.h * (s - 6.0 * c));
}

__global__ __launch_bounds__(kBX* kBY) void kb_add(const JItem* __restrict__ items, int n, int from_b) {
  const int it = find_item(items, n, blockIdx.x);
  const JItem& q = items[it];
  int i, j, k;
  if (!item_point(q, blockIdx.x - q.blk0, i, j, k)) return;
  const int64_t idx = i + (int64_t)q.nx * j + (int64_t)q.nx * q.ny * k;
  q.v64[idx] += (double)((const float*)(from_b ? q.b : q.a))[idx];
}

__global__ __launch_bounds__(kBX* kBY) void kb_copy_back(const JItem* __restrict__ items, int n) {
  const int it = find_item(items, n, blockIdx.x);
  const JItem& q = items[it];
  int i, j, k;
  if (!item_point(q, blockIdx.x - q.blk0, i, j, k)) return;
  const int64_t idx = i + (int64_t)q.nx * j + (int64_t)q.nx * q.ny * k;
  ((double*)q.a)[idx] = ((const double*)q.b)[idx];
}

}  // namespace

// destroy the graphs whose last replay has finished (all of them if wait)
void qm_reclaim_graphs(qm_ctx* ctx, bool wait) {
  auto& v = ctx->graphs;
  size_t k = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    if (wait) (void)hipEventSynchronize(v[i].done);
    if (hipEventQuery(v[i].done) == hipSuccess) {
      (void)hipGraphExecDestroy(v[i].x);
      (void)hipGraphDestroy(v[i].g);
      (void)hipEventDestroy(v[i].done);
    } else {
      v[k++] = v[i];
    }
  }
  v.resize(k);
}

extern "C" int qm_jacobi_batch(qm_ctx* ctx, void* stream, int n, const qm_geom* geoms, const void* const* rho,
                               void* const* v, void* const* tmp, int precision, int64_t nsweeps, int block,
                               int use_graph) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  QM_REQUIRE(ctx, n >= 0 && (n == 0 || (geoms && rho && v)), QM_EINVAL, "jacobi_batch: bad arguments");
  QM_REQUIRE(ctx, nsweeps >= 0 && block >= 2 && block % 2 == 0, QM_EINVAL,
             "jacobi_batch: sweeps >= 0 and an even block >= 2 required");
  QM_REQUIRE(ctx, precision == QM_PREC_FP64 || precision == QM_PREC_FP32, QM_EINVAL, "jacobi_batch: bad precision");
  if (n == 0 || nsweeps == 0) return QM_OK;
  qm_reclaim_graphs(ctx, false);
  const bool f64 = precision == QM_PREC_FP64;
  std::vector<JItem> items(n);
  int blocks = 0;
  size_t f32_elems = 0;
  for (int i = 0; i < n; ++i) {
    if ((rc = qm_check_geom(ctx, &geoms[i]))) return rc;
    QM_REQUIRE(ctx, rho[i] && v[i] && (!f64 || (tmp && tmp[i])), QM_EINVAL, "jacobi_batch: item %d incomplete", i);
    JItem& q = items[i];
    q.nx = geoms[i].nx;
    q.ny = geoms[i].ny;
    q.nz = geoms[i].nz;
    q.h = geoms[i].step;
    q.bx = (q.nx + kBX - 1) / kBX;
    q.by = (q.ny + kBY - 1) / kBY;
    q.blk0 = blocks;
    const int64_t nb = (int64_t)q.bx * q.by * q.nz;
    QM_REQUIRE(ctx, blocks + nb < 0x7fffffff, QM_ELIMIT, "jacobi_batch: batch too large");
    blocks += (int)nb;
    if (f64) {
      q.rho = rho[i];
      q.a = v[i];
      q.b = tmp[i];
    } else {
      q.rho64 = (const double*)rho[i];
      q.v64 = (double*)v[i];
      f32_elems += (size_t)npoints(&geoms[i]);
    }
  }
  hipStream_t s = (hipStream_t)stream;
  if (!f64) {
    float* w = nullptr;
    if ((rc = qm_ws(ctx, WS_JAC_A, sizeof(float) * 3 * f32_elems, (void**)&w))) return rc;
    size_t off = 0;
    for (int i = 0; i < n; ++i) {
      const size_t N = (size_t)npoints(&geoms[i]);
      JItem& q = items[i];
      q.r32 = w + off;
      q.a = w + off + N;
      q.b = w + off + 2 * N;
      q.rho = q.r32;
      off += 3 * N;
    }
  }
  void* dtab = nullptr;
  if ((rc = qm_ws(ctx, WS_BATCH, sizeof(JItem) * items.size(), &dtab))) return rc;
  QM_CHECK_HIP(ctx, hipMemcpyAsync(dtab, items.data(), sizeof(JItem) * items.size(), hipMemcpyHostToDevice, s));
  QM_CHECK_HIP(ctx, hipStreamSynchronize(s));   // pageable source
  const JItem* dt = (const JItem*)dtab;
  const dim3 grid(blocks), thr(kBX, kBY);
  // one block of work on stream t; kf = sweeps in the block
  auto enqueue_block = [&](hipStream_t t, int64_t kf) -> int {
    if (f64) {
      for (int64_t q = 0; q < kf; ++q) {
        hipLaunchKernelGGL(kb_sweep<double>, grid, thr, 0, t, dt, n, (int)(q & 1));
      }
      if (kf & 1) hipLaunchKernelGGL(kb_copy_back, grid, thr, 0, t, dt, n);
    } else {
      hipLaunchKernelGGL(kb_residual, grid, thr, 0, t, dt, n);
      for (int64_t q = 0; q < kf; ++q) hipLaunchKernelGGL(kb_sweep<float>, grid, thr, 0, t, dt, n, (int)(q & 1));
      hipLaunchKernelGGL(kb_add, grid, thr, 0, t, dt, n, (int)(kf & 1));
    }
    return hipGetLastError() == hipSuccess ? QM_OK : QM_EHIP;
  };
  const int64_t full = nsweeps / block, rest = nsweeps % block;
  ProfScope ps(ctx, s, QM_K_JACOBI);
  if (use_graph && full > 0) {
    if (!ctx->graph_stream) QM_CHECK_HIP(ctx, hipStreamCreateWithFlags(&ctx->graph_stream, hipStreamNonBlocking));
    hipStream_t gs = ctx->graph_stream;
    hipEvent_t ev;
    QM_CHECK_HIP(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    QM_CHECK_HIP(ctx, hipEventRecord(ev, s));
    QM_CHECK_HIP(ctx, hipStreamWaitEvent(gs, ev, 0));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    QM_CHECK_HIP(ctx, hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal));
    rc = enqueue_block(gs, block);
    hipError_t ce = hipStreamEndCapture(gs, &graph);
    if (rc || ce != hipSuccess) {
      (void)hipEventDestroy(ev);
      if (graph) (void)hipGraphDestroy(graph);
      return qm_fail(ctx, QM_EHIP, "jacobi_batch: graph capture failed: %s", hipGetErrorString(ce));
    }
    hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      (void)hipGraphDestroy(graph);
      (void)hipEventDestroy(ev);
      return qm_fail(ctx, QM_EHIP, "jacobi_batch: graph instantiation failed: %s", hipGetErrorString(ie));
    }
    for (int64_t r = 0; r < full; ++r) QM_CHECK_HIP(ctx, hipGraphLaunch(exec, gs));
    if (rest) {
      if ((rc = enqueue_block(gs, rest))) return qm_fail(ctx, rc, "jacobi_batch: launch failed");
    }
    QM_CHECK_HIP(ctx, hipEventRecord(ev, gs));
    QM_CHECK_HIP(ctx, hipStreamWaitEvent(s, ev, 0));
    ctx->graphs.push_back({graph, exec, ev});   // destroyed once ev has fired
  } else {
    for (int64_t r = 0; r < full; ++r)
      if ((rc = enqueue_block(s, block))) return qm_fail(ctx, rc, "jacobi_batch: launch failed");
    if (rest)
      if ((rc = enqueue_block(s, rest))) return qm_fail(ctx, rc, "jacobi_batch: launch failed");
  }
  return QM_OK;
}
