// Classical nuclear potential averaged over voxel sub-points (FP64).
//
// Reference kernel_vne_classic.cu:13-84:
//   V(v) = [sum_{s in ndiv^3} sum_A Z_A / |r_s - R_A|] / (cnt h^3)
//   r_s = origin + i h + (s + 1/2) h / ndiv,  cnt = #(s, A) with r_s != R_A.
// One thread per voxel; nuclei are wave-uniform reads.
#include <algorithm>
#include <cmath>

#include "qm_internal.h"

namespace {

template <typename T>
__global__ __launch_bounds__(256) void k_vne(const double4* __restrict__ atoms, int natoms, GeomView g, int ndiv,
                                             T* __restrict__ out) {
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int64_t i = v % g.nx, j = (v / g.nx) % g.ny, k = v / ((int64_t)g.nx * g.ny);
  const double dx = g.h / ndiv;
  const double x0 = g.ox + i * g.h, y0 = g.oy + j * g.h, z0 = g.oz + k * g.h;
  double acc = 0.0;
  int64_t cnt = 0;
  for (int a = 0; a < ndiv; ++a) {
    const double x = x0 + (a + 0.5) * dx;
    for (int b = 0; b < ndiv; ++b) {
      const double y = y0 + (b + 0.5) * dx;
      for (int c = 0; c < ndiv; ++c) {
        const double z = z0 + (c + 0.5) * dx;
        for (int at = 0; at < natoms; ++at) {
          const double4 A = atoms[at];
          const double ex = x - A.x, ey = y - A.y, ez = z - A.z;
          const double d2 = ex * ex + ey * ey + ez * ez;
          if (d2 != 0.0) {
            acc += A.w / sqrt(d2);
            ++cnt;
          }
        }
      }
    }
  }
  out[v] = (T)(acc / (double)cnt / (g.h * g.h * g.h));
}

}  // namespace

extern "C" int qm_vne_classic(qm_ctx* ctx, qm_mol* mol, void* stream, const qm_geom* g, int ndiv, int dtype,
                              void* out) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  if ((rc = qm_check_geom(ctx, g))) return rc;
  QM_REQUIRE(ctx, mol && out, QM_EINVAL, "vne: null argument");
  QM_REQUIRE(ctx, ndiv >= 1 && ndiv <= 64, QM_EINVAL, "vne: divisions must be 1..64");
  const int64_t n = npoints(g);
  const unsigned blocks = (unsigned)((n + 255) / 256);
  const GeomView gv = geom_view(g);
  hipStream_t s = (hipStream_t)stream;
  if (dtype == QM_F64)
    hipLaunchKernelGGL(k_vne<double>, dim3(blocks), dim3(256), 0, s, (const double4*)mol->atoms, mol->natoms, gv,
                       ndiv, (double*)out);
  else if (dtype == QM_F32)
    hipLaunchKernelGGL(k_vne<float>, dim3(blocks), dim3(256), 0, s, (const double4*)mol->atoms, mol->natoms, gv,
                       ndiv, (float*)out);
  else
    return qm_fail(ctx, QM_EINVAL, "vne: bad dtype");
  QM_CHECK_LAUNCH(ctx);
  return QM_OK;
}

// ===========================================================================
// Diffusion / adsorption pseudo-nuclear field (reference kernel_vne.cu:13-128,
// loop qmtools.py:2677-2745).  One sweep, for voxel r with value c:
//   out = sum_{26 neighbours, loop x, y, z = -1..1} DIF V(r+n) / sqrt(|nx|+|ny|+|nz|)
//       + (1 - DIF * 16.794682450997072) c - c ADS
//       + sum_A [d_A < sqrt(3) h] Z_A exp(-d_A),  d_A from the voxel centre (r + 1/2) h
//   out = 0 on the outer voxel layer.
// The reference wraps neighbour indices periodically; only outer-layer voxels
// reach a wrapped neighbour and their value is replaced by 0, so reading
// clamped indices gives the same field.
//
// Kernel: a workgroup of 64 x 4 threads owns a 64 x 8 xy tile (2 rows per
// thread) and marches through zs planes.  Planes k-1, k, k+1 of the tile with
// its one-voxel halo sit in three of four rotating LDS slots; the loads of
// plane k+2 are issued before plane k is computed and stored into the fourth
// slot afterwards, so each plane costs one barrier.  Nuclei within reach of
// the workgroup's box are listed once per workgroup (ballot compaction keeps
// the atom order).  The second sweep of a pair also writes per-workgroup
// partial sums of |c - out|; a one-block kernel adds them in a fixed order
// and applies the loop's stop rule on the device, so later sweeps that were
// already queued return at once and the host reads the state only every few
// dozen pairs.
namespace {

constexpr double kVneSelf = 16.794682450997072;   // kernel_vne.cu:84
constexpr int kVX = 64, kVY = 8, kVTY = 4;
constexpr int kVW = kVX + 2, kVH = kVY + 2, kVPlane = kVW * kVH;
constexpr int kVLoads = (kVPlane + 255) / 256;
constexpr int kVMaxNear = 64;

struct VneState {
  int stop;            // 1: no more sweeps
  int not_converged;   // stopped by the pair cap
  long long pairs;     // pairs of sweeps done
  double delta;        // last sum |res - out| / npts
};

__device__ __forceinline__ double vne_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <typename T>
__global__ __launch_bounds__(256) void k_vne_sweep(const T* __restrict__ v, T* __restrict__ out,
                                                   const double4* __restrict__ atoms, int natoms, GeomView g, int zs,
                                                   double ads, double dif, double* __restrict__ part,
                                                   const VneState* __restrict__ st) {
  if (st->stop) return;   // same value for every thread
  __shared__ T pl[4][kVPlane];
  __shared__ int near[kVMaxNear];
  __shared__ int nnear;
  __shared__ double red[kVTY];
  const int tx = threadIdx.x, ty = threadIdx.y, t = tx + kVX * ty;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int x0 = blockIdx.x * kVX, y0 = blockIdx.y * kVY;
  const int z0 = blockIdx.z * zs, z1 = min(nz, z0 + zs);
  const double h = g.h;
  const double rcut = sqrt(3.0) * h;
  // nuclei that can reach a voxel centre of this workgroup (conservative box test)
  if (ty == 0) {
    const double lx = (x0 + 0.5) * h, hx = (min(x0 + kVX, nx) - 0.5) * h;
    const double ly = (y0 + 0.5) * h, hy = (min(y0 + kVY, ny) - 0.5) * h;
    const double lz = (z0 + 0.5) * h, hz = (z1 - 0.5) * h;
    const double lim = rcut * (1.0 + 1e-9) + 1e-12;
    int cnt = 0;
    for (int a0 = 0; a0 < natoms; a0 += 64) {
      const int a = a0 + tx;
      bool hit = false;
      if (a < natoms) {
        const double4 A = atoms[a];
        const double ax = A.x - g.ox, ay = A.y - g.oy, az = A.z - g.oz;
        const double dx = fmax(0.0, fmax(lx - ax, ax - hx));
        const double dy = fmax(0.0, fmax(ly - ay, ay - hy));
        const double dz = fmax(0.0, fmax(lz - az, az - hz));
        hit = dx * dx + dy * dy + dz * dz <= lim * lim;
      }
      const unsigned long long m = __ballot(hit);
      if (hit) {
        const int pos = cnt + __popcll(m & ((1ull << tx) - 1ull));
        if (pos < kVMaxNear) near[pos] = a;
      }
      cnt += __popcll(m);
    }
    if (tx == 0) nnear = cnt;
  }
  // per-thread plane offsets of its share of the 66 x 10 halo tile (clamped into the grid)
  int64_t off[kVLoads];
  int slot_e[kVLoads];
#pragma unroll
  for (int u = 0; u < kVLoads; ++u) {
    const int e = t + 256 * u;
    slot_e[u] = e;
    const int r = min(e, kVPlane - 1) / kVW, c = min(e, kVPlane - 1) % kVW;
    const int gx = min(max(x0 - 1 + c, 0), nx - 1), gy = min(max(y0 - 1 + r, 0), ny - 1);
    off[u] = gx + (int64_t)nx * gy;
  }
  const int64_t sz = (int64_t)nx * ny;
  auto plane_at = [&](int k) -> int64_t { return sz * min(max(k, 0), nz - 1); };
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const int64_t pk = plane_at(z0 - 1 + p);
#pragma unroll
    for (int u = 0; u < kVLoads; ++u)
      if (slot_e[u] < kVPlane) pl[(z0 - 1 + p) & 3][slot_e[u]] = v[pk + off[u]];
  }
  __syncthreads();
  const int n_near = nnear;
  const bool all_atoms = n_near > kVMaxNear;
  const int nlist = all_atoms ? natoms : n_near;
  const double w1 = dif, w2 = dif / sqrt(2.0), w3 = dif / sqrt(3.0);
  const double self = 1.0 - dif * kVneSelf;
  const int gx = x0 + tx;
  double acc = 0.0;
  // the thread's two points (tile rows 2 ty + 1, 2 ty + 2, column tx + 1) need rows 2 ty .. 2 ty + 3
  // and columns tx .. tx + 2 of planes k-1, k, k+1: kept in registers and shifted along z, so each
  // plane's 12 values are read from LDS once instead of 27 reads per point
  T blk[3][4][3];
  auto read_block = [&](const T* P, T (&b)[4][3]) __attribute__((always_inline)) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) b[r][c] = P[(2 * ty + r) * kVW + tx + c];
  };
  read_block(pl[(z0 - 1) & 3], blk[0]);
  read_block(pl[z0 & 3], blk[1]);
  // Plane loads run two planes ahead into ping-pong registers: iteration k (parity b) requests
  // plane k+3 into nxt[b] and stores plane k+2, requested one iteration earlier, into its LDS
  // slot, so no thread waits for loads it has just issued.  The loads are unconditional (threads
  // past the tile's last slot read a clamped address and store nothing): a load under a branch
  // makes the compiler wait for every outstanding load before the LDS store.
  T nxt[2][kVLoads];
  {
    const int64_t pk2 = plane_at(z0 + 2);
#pragma unroll
    for (int u = 0; u < kVLoads; ++u) nxt[1][u] = v[pk2 + off[u]];
  }
  auto step = [&](int k, auto buf) __attribute__((always_inline)) {
    constexpr int b = decltype(buf)::value;
    const int64_t pk3 = plane_at(k + 3);
#pragma unroll
    for (int u = 0; u < kVLoads; ++u) nxt[b][u] = v[pk3 + off[u]];   // off[u] is clamped: always in the grid
    read_block(pl[(k + 1) & 3], blk[2]);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int gy = y0 + 2 * ty + q;
      const double c = (double)blk[1][q + 1][1];
      double sum = 0.0;
#pragma unroll
      for (int x = -1; x <= 1; ++x)
#pragma unroll
        for (int y = -1; y <= 1; ++y)
#pragma unroll
          for (int z = -1; z <= 1; ++z) {
            const int d = (x != 0) + (y != 0) + (z != 0);
            if (d == 0) continue;
            const double w = d == 1 ? w1 : (d == 2 ? w2 : w3);
            sum = fma(w, (double)blk[z + 1][q + 1 + y][1 + x], sum);
          }
      // multiply-adds written as FMAs: the unrolled copies of this body round alike
      sum = fma(self, c, sum);
      sum = fma(-c, ads, sum);
      const double px = (gx + 0.5) * h, py = (gy + 0.5) * h, pz = (k + 0.5) * h;
      for (int idx = 0; idx < nlist; ++idx) {
        const double4 A = atoms[all_atoms ? idx : near[idx]];
        const double ex = px - (A.x - g.ox), ey = py - (A.y - g.oy), ez = pz - (A.z - g.oz);
        const double d = sqrt(fma(ex, ex, fma(ey, ey, ez * ez)));
        if (d < rcut) sum = fma(A.w, exp(-d), sum);
      }
      if (gx == 0 || gx == nx - 1 || gy == 0 || gy == ny - 1 || k == 0 || k == nz - 1) sum = 0.0;
      if (gx < nx && gy < ny) {
        const T o = (T)sum;
        out[sz * k + (int64_t)nx * gy + gx] = o;
        acc += fabs(c - (double)o);
      }
    }
    // plane k+2 goes to the slot of plane k-2, last read before the previous barrier
#pragma unroll
    for (int u = 0; u < kVLoads; ++u)
      if (slot_e[u] < kVPlane) pl[(k + 2) & 3][slot_e[u]] = nxt[1 - b][u];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        blk[0][r][c] = blk[1][r][c];
        blk[1][r][c] = blk[2][r][c];
      }
    __syncthreads();
  };
  int k = z0;
  for (; k + 1 < z1; k += 2) {
    step(k, std::integral_constant<int, 0>());
    step(k + 1, std::integral_constant<int, 1>());
  }
  if (k < z1) step(k, std::integral_constant<int, 0>());
  if (part) {
    acc = vne_wave_sum(acc);
    if (tx == 0) red[ty] = acc;
    __syncthreads();
    if (t == 0) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < kVTY; ++w) s += red[w];
      part[blockIdx.x + gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z)] = s;
    }
  }
}

__global__ void k_vne_init(VneState* st, int stop) {
  st->stop = stop;
  st->not_converged = 0;
  st->pairs = 0;
  st->delta = 1.0;
}

// sums the partials of a pair's second sweep and applies the stop rule of qmtools.py:2715-2741
__global__ __launch_bounds__(256) void k_vne_finish(const double* __restrict__ part, int nb, double npts, double tol,
                                                    double maxiters, VneState* st) {
  if (st->stop) return;
  __shared__ double red[4];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += part[b];
  s = vne_wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double delta = (red[0] + red[1] + red[2] + red[3]) / npts;
    st->delta = delta;
    st->pairs += 1;
    if ((double)st->pairs > maxiters) {
      st->stop = 1;
      st->not_converged = 1;
    } else if (!(delta > tol)) {
      st->stop = 1;
    }
  }
}

template <typename T>
int vne_diffusion(qm_ctx* ctx, qm_mol* mol, hipStream_t s, const qm_geom* g, double ads, double dif, double tol,
                  double maxiters, T* v, T* tmp, int64_t* pairs_out, double* delta_out, int* conv_out) {
  const int64_t n = npoints(g);
  const unsigned tbx = (unsigned)((g->nx + kVX - 1) / kVX), tby = (unsigned)((g->ny + kVY - 1) / kVY);
  const int64_t tiles = (int64_t)tbx * tby;
  const int64_t tz = std::max<int64_t>(1, (2048 + tiles - 1) / tiles);
  const int zs = (int)std::max<int64_t>(4, (g->nz + tz - 1) / tz);
  const unsigned tbz = (unsigned)((g->nz + zs - 1) / zs);
  const int64_t nb = tiles * tbz;
  QM_REQUIRE(ctx, nb < (1 << 30), QM_ELIMIT, "vne: grid too large");
  void* w = nullptr;
  int rc;
  if ((rc = qm_ws(ctx, s, WS_VNE, sizeof(double) * (size_t)(nb + 8), &w))) return rc;
  VneState* st = (VneState*)w;
  double* part = (double*)w + 8;
  QM_CHECK_HIP(ctx, hipMemsetAsync(v, 0, sizeof(T) * n, s));
  QM_CHECK_HIP(ctx, hipMemsetAsync(tmp, 0, sizeof(T) * n, s));
  hipLaunchKernelGGL(k_vne_init, dim3(1), dim3(1), 0, s, st, (1.0 > tol) ? 0 : 1);
  QM_CHECK_LAUNCH(ctx);
  const GeomView gv = geom_view(g);
  const dim3 grid(tbx, tby, tbz), block(kVX, kVTY);
  const double4* atoms = (const double4*)mol->atoms;
  // the loop runs at least one pair and stops once the count exceeds maxiters:
  // at most max(1, floor(maxiters) + 1) pairs, and nothing is queued beyond that
  const int64_t cap = std::max<int64_t>(1, (int64_t)std::floor(std::max(maxiters, -1.0)) + 1);
  int64_t queued = 0;
  int64_t chunk = 8;
  VneState hs{};
  for (;;) {
    const int64_t m = std::min(chunk, cap - queued);
    if (m <= 0) break;
    for (int64_t p = 0; p < m; ++p) {
      ProfScope ps(ctx, s, QM_K_VNE);
      hipLaunchKernelGGL(k_vne_sweep<T>, grid, block, 0, s, (const T*)v, tmp, atoms, mol->natoms, gv, zs, ads, dif,
                         (double*)nullptr, (const VneState*)st);
      hipLaunchKernelGGL(k_vne_sweep<T>, grid, block, 0, s, (const T*)tmp, v, atoms, mol->natoms, gv, zs, ads, dif,
                         part, (const VneState*)st);
      hipLaunchKernelGGL(k_vne_finish, dim3(1), dim3(256), 0, s, (const double*)part, (int)nb, (double)n, tol,
                         maxiters, st);
    }
    QM_CHECK_LAUNCH(ctx);
    queued += m;
    QM_CHECK_HIP(ctx, hipMemcpyAsync(&hs, st, sizeof(VneState), hipMemcpyDeviceToHost, s));
    QM_CHECK_HIP(ctx, hipStreamSynchronize(s));
    if (hs.stop) break;
    chunk = std::min<int64_t>(chunk * 2, 64);
  }

  if (pairs_out) *pairs_out = hs.pairs;
  if (delta_out) *delta_out = hs.delta;
  if (conv_out) *conv_out = hs.not_converged ? 0 : 1;
  return QM_OK;
}

}  // namespace

extern "C" int qm_vne_diffusion(qm_ctx* ctx, qm_mol* mol, void* stream, const qm_geom* g, int dtype, double adsorb,
                                double diff, double tolerance, double maxiters, void* v, void* tmp,
                                int64_t* pairs, double* delta, int* converged) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  if ((rc = qm_check_geom(ctx, g))) return rc;
  QM_REQUIRE(ctx, mol && v && tmp && v != tmp, QM_EINVAL, "vne_diffusion: null or aliased buffers");
  QM_REQUIRE(ctx, std::isfinite(adsorb) && std::isfinite(diff) && std::isfinite(maxiters) && maxiters < 1e12,
             QM_EINVAL, "vne_diffusion: adsorb, diff and a finite maxiters required");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == QM_F64)
    return vne_diffusion<double>(ctx, mol, s, g, adsorb, diff, tolerance, maxiters, (double*)v, (double*)tmp, pairs,
                                 delta, converged);
  if (dtype == QM_F32)
    return vne_diffusion<float>(ctx, mol, s, g, adsorb, diff, tolerance, maxiters, (float*)v, (float*)tmp, pairs,
                                delta, converged);
  return qm_fail(ctx, QM_EINVAL, "vne_diffusion: bad dtype");
}
