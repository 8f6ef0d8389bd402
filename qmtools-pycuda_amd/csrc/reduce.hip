// Grid reductions, rescale and the nuclear seed.
//
// Reductions are two-pass and FP64: pass 1 gives each block a fixed
// contiguous segment, pass 2 sums the block partials in a fixed tree order,
// so results are bitwise reproducible (the reference uses FP32 atomicAdd
// across blocks, kernel_utils.cu:97, whose order is not fixed).
#include <mutex>

#include "qm_internal.h"

namespace {

constexpr int kRedThreads = 256;
constexpr int kRedMaxBlocks = 2048;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// block-wide sum; result valid in thread 0
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double part[kRedThreads / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) part[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kRedThreads / 64; ++w) s += part[w];
  }
  return s;
}

template <typename T, int OP>
__device__ __forceinline__ double red_term(const T* a, const T* b, int64_t i) {
  const double x = (double)a[i];
  if (OP == QM_RED_SUM) return x;
  const double y = (double)b[i];
  if (OP == QM_RED_ABSDIFF) return fabs(x - y);
  return fabs(x - y) / y;
}

template <typename T, int OP>
__global__ __launch_bounds__(kRedThreads) void k_reduce_pass1(const T* __restrict__ a,
                                                              const T* __restrict__ b, int64_t n,
                                                              int64_t seg, double* __restrict__ part) {
  const int64_t lo = (int64_t)blockIdx.x * seg;
  const int64_t hi = min(n, lo + seg);
  double acc = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kRedThreads) acc += red_term<T, OP>(a, b, i);
  acc = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kRedThreads) void k_reduce_pass2(const double* __restrict__ part, int nb,
                                                              double* __restrict__ out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kRedThreads) acc += part[i];
  acc = block_sum(acc);
  if (threadIdx.x == 0) out[0] = acc;
}

template <typename T>
__global__ void k_scale(T* __restrict__ a, int64_t n, const double* __restrict__ num, double num_const,
                        const double* __restrict__ den) {
  const double f = (num ? num[0] : num_const) / (den ? den[0] : 1.0);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    a[i] = (T)((double)a[i] * f);
}

// One thread per (atom, corner of the enclosing node cell); reference kernel_utils.cu:5-50.
// The reference stores (not adds) and lets atoms that share a node race; here
// the atom with the highest index wins, deterministically (a thread skips its
// node when a later atom writes the same node), which is the result of writing
// the atoms in order.
__device__ __forceinline__ void seed_node(const double4& at, const GeomView& g, int c, int64_t& ix, int64_t& iy,
                                          int64_t& iz) {
  ix = (int64_t)floor((at.x - g.ox) / g.h) + (c & 1);
  iy = (int64_t)floor((at.y - g.oy) / g.h) + ((c >> 1) & 1);
  iz = (int64_t)floor((at.z - g.oz) / g.h) + ((c >> 2) & 1);
}

template <typename T>
__global__ void k_seed(const double4* __restrict__ atoms, int natoms, GeomView g, T* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= natoms * 8) return;
  const int a = t >> 3, c = t & 7;
  const double4 at = atoms[a];
  int64_t ix, iy, iz;
  seed_node(at, g, c, ix, iy, iz);
  if (ix < 0 || iy < 0 || iz < 0 || ix >= g.nx || iy >= g.ny || iz >= g.nz) return;
  for (int b = a + 1; b < natoms; ++b) {
    const double4 bt = atoms[b];
    for (int d = 0; d < 8; ++d) {
      int64_t jx, jy, jz;
      seed_node(bt, g, d, jx, jy, jz);
      if (jx == ix && jy == iy && jz == iz) return;
    }
  }
  const double ax = at.x - g.ox, ay = at.y - g.oy, az = at.z - g.oz;
  double w = at.w;
  w *= fabs(ax - ix * g.h) / g.h;
  w *= fabs(ay - iy * g.h) / g.h;
  w *= fabs(az - iz * g.h) / g.h;
  out[ix + (int64_t)g.nx * (iy + (int64_t)g.ny * iz)] = (T)w;
}

template <typename T, int OP>
void launch_pass1(hipStream_t s, int nb, const void* a, const void* b, int64_t n, int64_t seg, double* part) {
  hipLaunchKernelGGL((k_reduce_pass1<T, OP>), dim3(nb), dim3(kRedThreads), 0, s, (const T*)a, (const T*)b, n,
                     seg, part);
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of the allocation-free reduction (qm_reduce_noalloc): the block partials
// (kRedMaxBlocks doubles) and the device result, 16.6 KB.  Every GPU the library runs on
// holds its own copy until the process ends.  In host code `scratch` is only a handle
// for hipGetSymbolAddress; its device address, one per GPU, is looked up once when a
// context is created on that GPU (qm_noalloc_setup) and used everywhere else.
constexpr int kNoallocDoubles = kRedMaxBlocks + 32;
static __device__ __align__(256) double scratch[kNoallocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};   // device address of `scratch` on each GPU

double* scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch[device];
}
}  // namespace

int qm_noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return QM_ELIMIT;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return QM_OK;
  void* p = nullptr;   // the caller has made `device` current
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) {
    (void)hipGetLastError();
    return QM_EHIP;
  }
  g_scratch[device] = (double*)p;
  return QM_OK;
}

// Both passes with the block partials in `p` (kRedMaxBlocks doubles) and the result in dev_result.
static int reduce_into(qm_ctx* ctx, hipStream_t s, int op, int dtype, const void* a, const void* b, int64_t n,
                       double* p, double* dev_result) {
  QM_REQUIRE(ctx, a != nullptr && n >= 0, QM_EINVAL, "reduce: bad input");
  QM_REQUIRE(ctx, op == QM_RED_SUM || b != nullptr, QM_EINVAL, "reduce: second operand missing");
  QM_REQUIRE(ctx, dtype == QM_F64 || dtype == QM_F32, QM_EINVAL, "reduce: bad dtype");
  QM_REQUIRE(ctx, op >= QM_RED_SUM && op <= QM_RED_RELDIFF, QM_EINVAL, "reduce: bad op");
  int64_t nb = (n + (int64_t)kRedThreads * 16 - 1) / ((int64_t)kRedThreads * 16);
  if (nb < 1) nb = 1;
  if (nb > kRedMaxBlocks) nb = kRedMaxBlocks;
  int64_t seg = (n + nb - 1) / nb;
  if (dtype == QM_F64) {
    if (op == QM_RED_SUM) launch_pass1<double, QM_RED_SUM>(s, (int)nb, a, b, n, seg, p);
    else if (op == QM_RED_ABSDIFF) launch_pass1<double, QM_RED_ABSDIFF>(s, (int)nb, a, b, n, seg, p);
    else launch_pass1<double, QM_RED_RELDIFF>(s, (int)nb, a, b, n, seg, p);
  } else {
    if (op == QM_RED_SUM) launch_pass1<float, QM_RED_SUM>(s, (int)nb, a, b, n, seg, p);
    else if (op == QM_RED_ABSDIFF) launch_pass1<float, QM_RED_ABSDIFF>(s, (int)nb, a, b, n, seg, p);
    else launch_pass1<float, QM_RED_RELDIFF>(s, (int)nb, a, b, n, seg, p);
  }
  QM_CHECK_LAUNCH(ctx);
  hipLaunchKernelGGL(k_reduce_pass2, dim3(1), dim3(kRedThreads), 0, s, (const double*)p, (int)nb, dev_result);
  QM_CHECK_LAUNCH(ctx);
  return QM_OK;
}

int qm_reduce_dev(qm_ctx* ctx, hipStream_t s, int op, int dtype, const void* a, const void* b, int64_t n,
                  double* dev_result) {
  void* part = nullptr;
  int rc = qm_ws(ctx, s, WS_RED, sizeof(double) * kRedMaxBlocks, &part);
  if (rc) return rc;
  return reduce_into(ctx, s, op, dtype, a, b, n, (double*)part, dev_result);
}

int qm_scale_dev(qm_ctx* ctx, hipStream_t s, int dtype, void* a, int64_t n, const double* dev_num,
                 double num_const, const double* dev_den) {
  QM_REQUIRE(ctx, a != nullptr && n >= 0, QM_EINVAL, "scale: bad input");
  if (n == 0) return QM_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (dtype == QM_F64)
    hipLaunchKernelGGL(k_scale<double>, dim3((unsigned)blocks), dim3(256), 0, s, (double*)a, n, dev_num,
                       num_const, dev_den);
  else if (dtype == QM_F32)
    hipLaunchKernelGGL(k_scale<float>, dim3((unsigned)blocks), dim3(256), 0, s, (float*)a, n, dev_num,
                       num_const, dev_den);
  else
    return qm_fail(ctx, QM_EINVAL, "scale: bad dtype");
  QM_CHECK_LAUNCH(ctx);
  return QM_OK;
}

extern "C" {

int qm_reduce(qm_ctx* ctx, void* stream, int op, int dtype, const void* a, const void* b, int64_t n,
              double* result) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  QM_REQUIRE(ctx, result != nullptr, QM_EINVAL, "reduce: null result");
  hipStream_t s = (hipStream_t)stream;
  void* dres = nullptr;
  rc = qm_ws(ctx, s, WS_SCALAR, 256, &dres);
  if (rc) return rc;
  rc = qm_reduce_dev(ctx, s, op, dtype, a, b, n, (double*)dres);
  if (rc) return rc;
  QM_CHECK_HIP(ctx, hipMemcpyAsync(ctx->host_scalars, dres, sizeof(double), hipMemcpyDeviceToHost, s));
  QM_CHECK_HIP(ctx, hipStreamSynchronize(s));
  *result = ctx->host_scalars[0];
  return QM_OK;
}

int qm_reduce_noalloc(qm_ctx* ctx, void* stream, int op, int dtype, const void* a, const void* b, int64_t n,
                      double* result) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  QM_REQUIRE(ctx, result != nullptr, QM_EINVAL, "reduce_noalloc: null result");
  QM_REQUIRE(ctx, n < ((int64_t)1 << 34), QM_ELIMIT,
             "reduce_noalloc: %lld elements is beyond the 2^34 limit of the library's grids", (long long)n);
  double* sc = scratch_of(ctx->device);
  QM_REQUIRE(ctx, sc != nullptr, QM_EHIP, "reduce_noalloc: the scratch address of device %d was not found",
             ctx->device);
  hipStream_t s = (hipStream_t)stream;
  // kernels and copies take the device address of the scratch, never its host-side symbol
  rc = reduce_into(ctx, s, op, dtype, a, b, n, sc, sc + kRedMaxBlocks);
  if (rc) return rc;
  QM_CHECK_HIP(ctx, hipMemcpyAsync(ctx->host_scalars, sc + kRedMaxBlocks, sizeof(double), hipMemcpyDeviceToHost, s));
  QM_CHECK_HIP(ctx, hipStreamSynchronize(s));
  *result = ctx->host_scalars[0];
  return QM_OK;
}

int qm_scale(qm_ctx* ctx, void* stream, int dtype, void* a, int64_t n, double factor) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  return qm_scale_dev(ctx, (hipStream_t)stream, dtype, a, n, nullptr, factor, nullptr);
}

int qm_seed(qm_ctx* ctx, qm_mol* mol, void* stream, const qm_geom* g, int dtype, void* out) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  rc = qm_check_geom(ctx, g);
  if (rc) return rc;
  QM_REQUIRE(ctx, mol && out, QM_EINVAL, "seed: null argument");
  const int nt = mol->natoms * 8;
  const GeomView gv = geom_view(g);
  hipStream_t s = (hipStream_t)stream;
  if (dtype == QM_F64)
    hipLaunchKernelGGL(k_seed<double>, dim3((nt + 255) / 256), dim3(256), 0, s, (const double4*)mol->atoms,
                       mol->natoms, gv, (double*)out);
  else if (dtype == QM_F32)
    hipLaunchKernelGGL(k_seed<float>, dim3((nt + 255) / 256), dim3(256), 0, s, (const double4*)mol->atoms,
                       mol->natoms, gv, (float*)out);
  else
    return qm_fail(ctx, QM_EINVAL, "seed: bad dtype");
  QM_CHECK_LAUNCH(ctx);
  return QM_OK;
}

}  // extern "C"
