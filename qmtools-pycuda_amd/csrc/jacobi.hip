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

#include <cstring>

#include "qm_internal.h"

namespace {

constexpr int kBX = 64, kBY = 4;

// Sweep kernel: each wave owns a strip of 64 columns x kSR rows and marches
// through zs planes in z.  A lane holds its column's rows j0-1 .. j0+kSR of
// the planes k-1, k, k+1 in registers; x neighbours come from the adjacent
// lanes (lanes 0 and 63 load the one column beyond the strip), y neighbours
// from the same lane's registers.  Waves never wait for each other (no LDS,
// no barrier), and the loads of plane k+2, with the halo columns and rho of
// plane k+1, are issued before plane k is computed, so one plane of memory
// latency is hidden behind the arithmetic.  Each V value is read from memory
// (kSR+2)/kSR times per sweep, rho once, V' written once.
// Summation order: (x-1 + x+1) + y+1 + y-1 + z+1 + z-1 - rho/h, then / 6.
constexpr int kSW = 64;                    // columns per strip (one lane each)
constexpr int kSR = 8;                     // rows per strip
constexpr int kSWaves = 4;                 // strips per workgroup, stacked in y
constexpr int kTileY = kSR * kSWaves;      // rows per workgroup tile

template <typename T>
__device__ __forceinline__ void strip_body(int bx, int by, int bz, int zs, const T* __restrict__ rho,
                                           const T* __restrict__ v, T* __restrict__ vout, int nx, int ny, int nz,
                                           T h) {
  const int lane = threadIdx.x;
  const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);   // one strip per wave
  const int i = bx * kSW + lane;
  const int j0 = by * kTileY + wy * kSR;
  if (j0 >= ny) return;                    // whole wave idle; the kernel has no barrier
  const int z0 = bz * zs, z1 = min(nz, z0 + zs);
  const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
  // Indices are clamped into the grid: a value read at a clamped index only
  // ever feeds an outer-layer point (copied, not relaxed) or a point outside
  // the grid (not stored), so every load is unconditional and in bounds.
  const unsigned ic = (unsigned)min(i, nx - 1);
  const unsigned ihc = (unsigned)min(max(lane == 0 ? i - 1 : i + 1, 0), nx - 1);
  auto row = [&](int j) -> size_t { return sy * (size_t)min(max(j, 0), ny - 1); };
  auto pln = [&](int k) -> size_t { return sz * (size_t)min(max(k, 0), nz - 1); };
  const bool edge = lane == 0 || lane == kSW - 1;
  // multiplications by rounded reciprocals instead of IEEE divisions (about 10 instructions each)
  const T hinv = (T)1 / h;
  T pm[kSR + 2], pc[kSR + 2], pp[kSR + 2], hc[kSR], rr[kSR];
#pragma unroll
  for (int q = 0; q < kSR + 2; ++q) {
    pm[q] = v[pln(z0 - 1) + row(j0 - 1 + q) + ic];
    pc[q] = v[pln(z0) + row(j0 - 1 + q) + ic];
    pp[q] = v[pln(z0 + 1) + row(j0 - 1 + q) + ic];
  }
#pragma unroll
  for (int r = 0; r < kSR; ++r) {
    rr[r] = rho[pln(z0) + row(j0 + r) + ic];
    hc[r] = (T)0;
  }
  if (edge) {
#pragma unroll
    for (int r = 0; r < kSR; ++r) hc[r] = v[pln(z0) + row(j0 + r) + ihc];
  }
  for (int k = z0; k < z1; ++k) {
    T pn[kSR + 2], hn[kSR], rn[kSR];
    const size_t pk2 = pln(k + 2), pk1 = pln(k + 1), pk = sz * (size_t)k;
#pragma unroll
    for (int q = 0; q < kSR + 2; ++q) pn[q] = v[pk2 + row(j0 - 1 + q) + ic];
#pragma unroll
    for (int r = 0; r < kSR; ++r) {
      rn[r] = rho[pk1 + row(j0 + r) + ic];
      hn[r] = (T)0;
    }
    if (edge) {
#pragma unroll
      for (int r = 0; r < kSR; ++r) hn[r] = v[pk1 + row(j0 + r) + ihc];
    }
    const bool zb = k == 0 || k == nz - 1;
#pragma unroll
    for (int r = 0; r < kSR; ++r) {
      const int j = j0 + r;
      const T c = pc[r + 1];
      T xm = __shfl_up(c, 1, kSW);
      T xp = __shfl_down(c, 1, kSW);
      if (lane == 0) xm = hc[r];
      if (lane == kSW - 1) xp = hc[r];
      T sum = xm + xp;
      sum += pc[r + 2];
      sum += pc[r];
      sum += pp[r + 1];
      sum += pm[r + 1];
      sum = fma(-rr[r], hinv, sum);   // explicit: every copy of this code rounds alike
      const bool outer = zb || i == 0 || i == nx - 1 || j == 0 || j == ny - 1;
      const T val = outer ? c : sum * (T)(1.0 / 6.0);
      if (i < nx && j < ny) vout[pk + sy * (size_t)j + (unsigned)i] = val;
    }
#pragma unroll
    for (int q = 0; q < kSR + 2; ++q) {
      pm[q] = pc[q];
      pc[q] = pp[q];
      pp[q] = pn[q];
    }
#pragma unroll
    for (int r = 0; r < kSR; ++r) {
      hc[r] = hn[r];
      rr[r] = rn[r];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kSW* kSWaves) void k_sweep_strip(const T* __restrict__ rho, const T* __restrict__ v,
                                                              T* __restrict__ vout, int nx, int ny, int nz, int zs,
                                                              T h) {
  strip_body<T>(blockIdx.x, blockIdx.y, blockIdx.z, zs, rho, v, vout, nx, ny, nz, h);
}

// Two sweeps per pass (V -> V''): the same strips with a second level.  Lane L
// holds column bx W - 1 + L (W = 62 output columns per strip) and rows
// j0-2 .. j0+R+1 of V; the first sweep is formed on rows j0-1 .. j0+R at all
// 64 lanes (lanes 0 and 63 take the column beyond the strip from an extra
// load), the second on rows j0 .. j0+R-1 at lanes 1 .. 62 from the first-sweep
// values of its neighbours.  In z the second sweep runs one plane behind the
// first.  Each point's arithmetic, summation order included, is that of two
// single sweeps, so the result is bitwise the same.  Per output point and pass
// V is read (R+4)/R x 64/62 times, rho (R+2)/R x 64/62 times, V'' written once.
template <int R>
struct Strip2 {
  static constexpr int W = kSW - 2;                 // output columns per strip
  static constexpr int TY = R * kSWaves;            // output rows per workgroup
};

template <typename T, int R>
__device__ __forceinline__ void strip2_body(int bx, int by, int bz, int zs, const T* __restrict__ rho,
                                            const T* __restrict__ v, T* __restrict__ vout, int nx, int ny, int nz,
                                            T h) {
  constexpr int R0 = R + 4, R1 = R + 2;
  const int lane = threadIdx.x;
  const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int i = bx * Strip2<R>::W - 1 + lane;
  const int j0 = by * Strip2<R>::TY + wy * R;
  if (j0 >= ny) return;                    // whole wave idle; the kernel has no barrier
  const int z0 = bz * zs, z1 = min(nz, z0 + zs);
  const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
  // clamped indices: values read at them only feed outer-layer points (copied) or points outside the grid
  const unsigned ic = (unsigned)min(max(i, 0), nx - 1);
  const unsigned ihc = (unsigned)min(max(lane == 0 ? i - 1 : i + 1, 0), nx - 1);
  auto row = [&](int j) -> size_t { return sy * (size_t)min(max(j, 0), ny - 1); };
  auto pln = [&](int k) -> size_t { return sz * (size_t)min(max(k, 0), nz - 1); };
  const bool edge = lane == 0 || lane == kSW - 1;
  // multiplications by rounded reciprocals instead of IEEE divisions (about 10 instructions each)
  const T hinv = (T)1 / h;
  const bool xb = i <= 0 || i >= nx - 1;            // outer layer (or outside) in x
  T am[R0], ac[R0], ap[R0];         // V on planes m-1, m, m+1, rows j0-2+q
  T bmm[R1], bm[R1];                // first sweep on planes m-2, m-1, rows j0-1+p
  T hx[R1], rc[R1], rm[R1];         // halo column of plane m; rho of planes m and m-1 (rows j0-1+p)
  // Loads run two planes ahead, into ping-pong buffers: iteration m (parity b) requests V
  // plane m+3, rho and halo of plane m+2 into buffer b and moves the data requested one
  // iteration earlier (buffer 1-b) into place, so no wave waits for loads it has just issued.
  T an[2][R0], rn[2][R1], hn[2][R1];
  const int m0 = z0 - 1;
  auto request = [&](int m, auto buf) __attribute__((always_inline)) {
    constexpr int b = decltype(buf)::value;
    const size_t pv = pln(m + 3), pr = pln(m + 2);
#pragma unroll
    for (int q = 0; q < R0; ++q) an[b][q] = v[pv + row(j0 - 2 + q) + ic];
#pragma unroll
    for (int p = 0; p < R1; ++p) {
      rn[b][p] = rho[pr + row(j0 - 1 + p) + ic];
      hn[b][p] = (T)0;
    }
    if (edge) {
#pragma unroll
      for (int p = 0; p < R1; ++p) hn[b][p] = v[pr + row(j0 - 1 + p) + ihc];
    }
  };
#pragma unroll
  for (int q = 0; q < R0; ++q) {
    am[q] = v[pln(m0 - 1) + row(j0 - 2 + q) + ic];
    ac[q] = v[pln(m0) + row(j0 - 2 + q) + ic];
    ap[q] = v[pln(m0 + 1) + row(j0 - 2 + q) + ic];
    an[1][q] = v[pln(m0 + 2) + row(j0 - 2 + q) + ic];
  }
#pragma unroll
  for (int p = 0; p < R1; ++p) {
    rc[p] = rho[pln(m0) + row(j0 - 1 + p) + ic];
    rn[1][p] = rho[pln(m0 + 1) + row(j0 - 1 + p) + ic];
    rm[p] = (T)0;
    hx[p] = (T)0;
    hn[1][p] = (T)0;
    bmm[p] = (T)0;
    bm[p] = (T)0;
  }
  if (edge) {
#pragma unroll
    for (int p = 0; p < R1; ++p) {
      hx[p] = v[pln(m0) + row(j0 - 1 + p) + ihc];
      hn[1][p] = v[pln(m0 + 1) + row(j0 - 1 + p) + ihc];
    }
  }
  auto step = [&](int m, auto buf) __attribute__((always_inline)) {
    constexpr int b = decltype(buf)::value;
    request(m, buf);
    // first sweep on plane m
    const bool zb1 = m <= 0 || m >= nz - 1;
    T bc[R1];
#pragma unroll
    for (int p = 0; p < R1; ++p) {
      const int j = j0 - 1 + p;
      const T c = ac[p + 1];
      T xm = __shfl_up(c, 1, kSW);
      T xp = __shfl_down(c, 1, kSW);
      if (lane == 0) xm = hx[p];
      if (lane == kSW - 1) xp = hx[p];
      T sum = xm + xp;
      sum += ac[p + 2];
      sum += ac[p];
      sum += ap[p + 1];
      sum += am[p + 1];
      sum = fma(-rc[p], hinv, sum);   // explicit: every copy of this code rounds alike
      const bool outer = zb1 || xb || j <= 0 || j >= ny - 1;
      bc[p] = outer ? c : sum * (T)(1.0 / 6.0);
    }
    // second sweep on plane m-1
    const int k = m - 1;
    if (k >= z0) {
      const bool zb2 = k == 0 || k == nz - 1;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        const T c = bm[r + 1];
        const T xm = __shfl_up(c, 1, kSW);
        const T xp = __shfl_down(c, 1, kSW);
        T sum = xm + xp;
        sum += bm[r + 2];
        sum += bm[r];
        sum += bc[r + 1];
        sum += bmm[r + 1];
        sum = fma(-rm[r + 1], hinv, sum);
        const bool outer = zb2 || xb || j == 0 || j == ny - 1;
        const T val = outer ? c : sum * (T)(1.0 / 6.0);
        if (lane >= 1 && lane <= kSW - 2 && i < nx && j < ny) vout[sz * (size_t)k + sy * (size_t)j + (unsigned)i] = val;
      }
    }
#pragma unroll
    for (int q = 0; q < R0; ++q) {
      am[q] = ac[q];
      ac[q] = ap[q];
      ap[q] = an[1 - b][q];
    }
#pragma unroll
    for (int p = 0; p < R1; ++p) {
      bmm[p] = bm[p];
      bm[p] = bc[p];
      rm[p] = rc[p];
      rc[p] = rn[1 - b][p];
      hx[p] = hn[1 - b][p];
    }
  };
  int m = m0;
  for (; m + 1 <= z1; m += 2) {
    step(m, std::integral_constant<int, 0>());
    step(m + 1, std::integral_constant<int, 1>());
  }
  if (m <= z1) step(m, std::integral_constant<int, 0>());
}

constexpr int kR2D = 4, kR2F = 8;   // rows per strip of the two-sweep kernel: FP64, FP32
template <typename T>
constexpr int r2() { return sizeof(T) == 8 ? kR2D : kR2F; }

template <typename T>
__global__ __launch_bounds__(kSW* kSWaves) void k_sweep_strip2(const T* __restrict__ rho, const T* __restrict__ v,
                                                               T* __restrict__ vout, int nx, int ny, int nz, int zs,
                                                               T h) {
  strip2_body<T, r2<T>()>(blockIdx.x, blockIdx.y, blockIdx.z, zs, rho, v, vout, nx, ny, nz, h);
}

// planes per workgroup of the two-sweep kernel: at most 1024 workgroups (two rounds of the
// 512 that fit the chip at two waves per SIMD), at least 4 planes each; a positive `fixed`
// (the tuning knob) overrides
template <typename T>
inline int pick_zs2(int nx, int ny, int nz, int fixed) {
  if (fixed > 0) return std::min(fixed, nz);
  const int tiles = ((nx + Strip2<r2<T>()>::W - 1) / Strip2<r2<T>()>::W) *
                    ((ny + Strip2<r2<T>()>::TY - 1) / Strip2<r2<T>()>::TY);
  const int tz = std::max(1, 1024 / tiles);
  return std::max(4, (nz + tz - 1) / tz);
}

// FP32 two-sweep pass with two adjacent columns per lane (8-byte loads and stores, half the
// lane shuffles per point).  A strip spans 128 columns, lane L holding columns i and i + 1 with
// i = 124 bx - 2 + 2L; the first sweep is formed on all 128 columns and the second on columns
// 2 ... 125, the strip's 124 outputs.  The first sweep at the two outer columns sees a clamped
// or the lane's own value as its missing neighbour; it feeds only the second sweep at columns 1
// and 126, which is not stored, so no halo column is loaded.  nx must be even: pairs start at
// even columns and lie wholly inside or outside the grid.  Each point's arithmetic and summation
// order are those of the one-column kernels, so the results are bitwise the same.
constexpr int kWP = 2 * kSW - 4;   // output columns per strip
constexpr int kR2P = 4;            // rows per strip
template <int R>
struct Strip2P {
  static constexpr int W = kWP;
  static constexpr int TY = R * kSWaves;
};

template <int R>
__device__ __forceinline__ void strip2p_body(int bx, int by, int bz, int zs, const float* __restrict__ rho,
                                             const float* __restrict__ v, float* __restrict__ vout, int nx, int ny,
                                             int nz, float h) {
  constexpr int R0 = R + 4, R1 = R + 2;
  const int lane = threadIdx.x;
  const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int i = bx * kWP - 2 + 2 * lane;   // first column of the lane's pair (even)
  const int j0 = by * Strip2P<R>::TY + wy * R;
  if (j0 >= ny) return;                    // whole wave idle; the kernel has no barrier
  const int z0 = bz * zs, z1 = min(nz, z0 + zs);
  const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
  // clamped pair: values read at it only feed outer-layer points (copied) or points outside the grid
  const unsigned ic = (unsigned)min(max(i, 0), nx - 2);
  auto row = [&](int j) -> size_t { return sy * (size_t)min(max(j, 0), ny - 1); };
  auto pln = [&](int k) -> size_t { return sz * (size_t)min(max(k, 0), nz - 1); };
  auto ld = [](const float* __restrict__ p, size_t off) -> float2 { return *(const float2*)(p + off); };
  const float hinv = 1.0f / h;
  const bool xb0 = i <= 0 || i >= nx - 1, xb1 = i + 1 <= 0 || i + 1 >= nx - 1;
  const bool store = lane >= 1 && lane <= kSW - 2 && i < nx;
  float2 am[R0], ac[R0], ap[R0];   // V on planes m-1, m, m+1, rows j0-2+q
  float2 bmm[R1], bm[R1];          // first sweep on planes m-2, m-1, rows j0-1+p
  float2 rc[R1], rm[R1];           // rho of planes m and m-1 (rows j0-1+p)
  // loads two planes ahead into ping-pong buffers, as in the one-column kernel
  float2 an[2][R0], rn[2][R1];
  const int m0 = z0 - 1;
  auto request = [&](int m, auto buf) __attribute__((always_inline)) {
    constexpr int b = decltype(buf)::value;
    const size_t pv = pln(m + 3), pr = pln(m + 2);
#pragma unroll
    for (int q = 0; q < R0; ++q) an[b][q] = ld(v, pv + row(j0 - 2 + q) + ic);
#pragma unroll
    for (int p = 0; p < R1; ++p) rn[b][p] = ld(rho, pr + row(j0 - 1 + p) + ic);
  };
#pragma unroll
  for (int q = 0; q < R0; ++q) {
    am[q] = ld(v, pln(m0 - 1) + row(j0 - 2 + q) + ic);
    ac[q] = ld(v, pln(m0) + row(j0 - 2 + q) + ic);
    ap[q] = ld(v, pln(m0 + 1) + row(j0 - 2 + q) + ic);
    an[1][q] = ld(v, pln(m0 + 2) + row(j0 - 2 + q) + ic);
  }
#pragma unroll
  for (int p = 0; p < R1; ++p) {
    rc[p] = ld(rho, pln(m0) + row(j0 - 1 + p) + ic);
    rn[1][p] = ld(rho, pln(m0 + 1) + row(j0 - 1 + p) + ic);
    rm[p] = make_float2(0.f, 0.f);
    bmm[p] = make_float2(0.f, 0.f);
    bm[p] = make_float2(0.f, 0.f);
  }
  // V' at the pair from its x neighbours (l, c.x) and (c.y, r) and the four others
  auto relax = [&](float2 c, float l, float r, float2 yp, float2 ym, float2 zp, float2 zm, float2 q, bool o0,
                   bool o1) __attribute__((always_inline)) -> float2 {
    float s0 = l + c.y;
    s0 += yp.x;
    s0 += ym.x;
    s0 += zp.x;
    s0 += zm.x;
    s0 = fmaf(-q.x, hinv, s0);
    float s1 = c.x + r;
    s1 += yp.y;
    s1 += ym.y;
    s1 += zp.y;
    s1 += zm.y;
    s1 = fmaf(-q.y, hinv, s1);
    return make_float2(o0 ? c.x : s0 * (float)(1.0 / 6.0), o1 ? c.y : s1 * (float)(1.0 / 6.0));
  };
  auto step = [&](int m, auto buf) __attribute__((always_inline)) {
    constexpr int b = decltype(buf)::value;
    request(m, buf);
    // first sweep on plane m
    const bool zb1 = m <= 0 || m >= nz - 1;
    float2 bc[R1];
#pragma unroll
    for (int p = 0; p < R1; ++p) {
      const int j = j0 - 1 + p;
      const float2 c = ac[p + 1];
      const float l = __shfl_up(c.y, 1, kSW);
      const float r = __shfl_down(c.x, 1, kSW);
      const bool o = zb1 || j <= 0 || j >= ny - 1;
      bc[p] = relax(c, l, r, ac[p + 2], ac[p], ap[p + 1], am[p + 1], rc[p], o || xb0, o || xb1);
    }
    // second sweep on plane m-1
    const int k = m - 1;
    if (k >= z0) {
      const bool zb2 = k == 0 || k == nz - 1;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        const float2 c = bm[r + 1];
        const float l = __shfl_up(c.y, 1, kSW);
        const float rr = __shfl_down(c.x, 1, kSW);
        const bool o = zb2 || j == 0 || j == ny - 1;
        const float2 val = relax(c, l, rr, bm[r + 2], bm[r], bc[r + 1], bmm[r + 1], rm[r + 1], o || xb0, o || xb1);
        if (store && j < ny) *(float2*)(vout + sz * (size_t)k + sy * (size_t)j + (unsigned)i) = val;
      }
    }
#pragma unroll
    for (int q = 0; q < R0; ++q) {
      am[q] = ac[q];
      ac[q] = ap[q];
      ap[q] = an[1 - b][q];
    }
#pragma unroll
    for (int p = 0; p < R1; ++p) {
      bmm[p] = bm[p];
      bm[p] = bc[p];
      rm[p] = rc[p];
      rc[p] = rn[1 - b][p];
    }
  };
  int m = m0;
  for (; m + 1 <= z1; m += 2) {
    step(m, std::integral_constant<int, 0>());
    step(m + 1, std::integral_constant<int, 1>());
  }
  if (m <= z1) step(m, std::integral_constant<int, 0>());
}

template <int R>
__global__ __launch_bounds__(kSW* kSWaves) void k_sweep_strip2p(const float* __restrict__ rho,
                                                                const float* __restrict__ v, float* __restrict__ vout,
                                                                int nx, int ny, int nz, int zs, float h) {
  strip2p_body<R>(blockIdx.x, blockIdx.y, blockIdx.z, zs, rho, v, vout, nx, ny, nz, h);
}

// planes per workgroup of the paired two-sweep kernel (as pick_zs2)
// the paired kernel's grids: even nx
inline bool pairs_fit(int nx, int ny) { return nx % 2 == 0 && ny > 0; }

inline int pick_zs2p(int nx, int ny, int nz, int fixed) {
  if (fixed > 0) return std::min(fixed, nz);
  const int tiles = ((nx + kWP - 1) / kWP) * ((ny + Strip2P<kR2P>::TY - 1) / Strip2P<kR2P>::TY);
  const int tz = std::max(1, 1024 / tiles);
  return std::max(4, (nz + tz - 1) / tz);
}

// planes per workgroup: about 1024 workgroups (4096 waves) for the chip, at least 4 planes each;
// a positive `fixed` (the tuning knob) overrides
inline int pick_zs(int nx, int ny, int nz, int fixed) {
  if (fixed > 0) return std::min(fixed, nz);
  const int tiles = ((nx + kSW - 1) / kSW) * ((ny + kTileY - 1) / kTileY);
  const int tz = std::max(1, (1024 + tiles - 1) / tiles);
  return std::max(4, (nz + tz - 1) / tz);
}

// r = rho - h (sum_6 V - 6 V) on interior points (FP64), 0 on the outer layer, stored FP32
// The first FP32 sweep of the correction from e = 0: every neighbour sum is +0, so the
// sweep's value is fma(-r, 1/h, +0) * (1/6) on interior points and 0 on the outer layer,
// bitwise what the sweep kernels compute (hinv is the same rounded 1/h).
__device__ __forceinline__ float first_correction(float r, float hinv) { return fmaf(-r, hinv, 0.f) * (float)(1.0 / 6.0); }

// e1 (optional): also store the first sweep of the correction
__global__ __launch_bounds__(kBX* kBY) void k_residual(const double* __restrict__ rho,
                                                       const double* __restrict__ v, float* __restrict__ r,
                                                       float* __restrict__ e1, int nx, int ny, int nz, double h) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= nx || j >= ny) return;
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  const int64_t idx = i + sy * j + sz * k;
  if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) {
    r[idx] = 0.f;
    if (e1) e1[idx] = 0.f;
    return;
  }
  const double c = v[idx];
  double s = v[idx - 1] + v[idx + 1];
  s += v[idx + sy];
  s += v[idx - sy];
  s += v[idx + sz];
  s += v[idx - sz];
  const float rr = (float)(rho[idx] - h * (s - 6.0 * c));
  r[idx] = rr;
  if (e1) e1[idx] = first_correction(rr, 1.0f / (float)h);   // 1/h rounded as in the sweep kernels
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
int sweep(qm_ctx* ctx, hipStream_t s, const qm_geom* g, const T* rho, const T* a, T* b, bool prof = true) {
  ProfScope ps(ctx, s, QM_K_JACOBI, 1, prof);
  const int zs = pick_zs(g->nx, g->ny, g->nz, ctx->jacobi_zs);
  const dim3 grid((g->nx + kSW - 1) / kSW, (g->ny + kTileY - 1) / kTileY, (g->nz + zs - 1) / zs);
  hipLaunchKernelGGL(k_sweep_strip<T>, grid, dim3(kSW, kSWaves), 0, s, rho, a, b, g->nx, g->ny, g->nz, zs,
                     (T)g->step);
  QM_CHECK_LAUNCH(ctx);
  return QM_OK;
}

// two sweeps a -> b in one pass (bitwise the same as sweep(a -> t), sweep(t -> b))
template <typename T>
int sweep2(qm_ctx* ctx, hipStream_t s, const qm_geom* g, const T* rho, const T* a, T* b, bool prof = true) {
  ProfScope ps(ctx, s, QM_K_JACOBI, 2, prof);
  if (sizeof(T) == 4 && ctx->jacobi_pairs && pairs_fit(g->nx, g->ny)) {
    const int zs = pick_zs2p(g->nx, g->ny, g->nz, ctx->jacobi_zs);
    const dim3 grid((g->nx + kWP - 1) / kWP, (g->ny + Strip2P<kR2P>::TY - 1) / Strip2P<kR2P>::TY,
                    (g->nz + zs - 1) / zs);
    hipLaunchKernelGGL(k_sweep_strip2p<kR2P>, grid, dim3(kSW, kSWaves), 0, s, (const float*)rho, (const float*)a,
                       (float*)b, g->nx, g->ny, g->nz, zs, (float)g->step);
    QM_CHECK_LAUNCH(ctx);
    return QM_OK;
  }
  constexpr int R = r2<T>();
  const int zs = pick_zs2<T>(g->nx, g->ny, g->nz, ctx->jacobi_zs);
  const dim3 grid((g->nx + Strip2<R>::W - 1) / Strip2<R>::W, (g->ny + Strip2<R>::TY - 1) / Strip2<R>::TY,
                  (g->nz + zs - 1) / zs);
  hipLaunchKernelGGL(k_sweep_strip2<T>, grid, dim3(kSW, kSWaves), 0, s, rho, a, b, g->nx, g->ny, g->nz, zs,
                     (T)g->step);
  QM_CHECK_LAUNCH(ctx);
  return QM_OK;
}

// A block of n sweeps from buffer a that ends in a, with its last two sweeps
// single (so a and the other buffer hold consecutive iterates): P two-sweep
// passes (P even) and S = n - 2P >= 2 single sweeps.
inline void split_block(int64_t n, int64_t& P, int64_t& S) {
  P = n >= 2 ? 2 * ((n - 2) / 4) : 0;
  S = n - 2 * P;
}

int host_delta(qm_ctx* ctx, hipStream_t s, int dtype, const void* a, const void* b, int64_t n, double* out) {
  double* d = nullptr;
  int rc = qm_ws(ctx, s, WS_SCALAR, 256, (void**)&d);
  if (rc) return rc;
  rc = qm_reduce_dev(ctx, s, QM_RED_ABSDIFF, dtype, a, b, n, d + 1);
  if (rc) return rc;
  QM_CHECK_HIP(ctx, hipMemcpyAsync(ctx->host_scalars + 1, d + 1, sizeof(double), hipMemcpyDeviceToHost, s));
  QM_CHECK_HIP(ctx, hipStreamSynchronize(s));
  *out = ctx->host_scalars[1];
  return QM_OK;
}

// A block of work captured once into a hipGraph on the context's graph stream
// and replayed there; the caller's stream is ordered before and after the
// replays with events.  Captured work records no profiling events and no
// memset nodes.  The graph is kept in the context until its last replay has
// finished.
struct GraphBlock {
  qm_ctx* ctx;
  hipStream_t s, gs = nullptr;
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  hipEvent_t ev = nullptr;
  GraphBlock(qm_ctx* c, hipStream_t st) : ctx(c), s(st) {}
  template <typename F>
  int capture(F enqueue) {
    if (!ctx->graph_stream) QM_CHECK_HIP(ctx, hipStreamCreateWithFlags(&ctx->graph_stream, hipStreamNonBlocking));
    gs = ctx->graph_stream;
    QM_CHECK_HIP(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    QM_CHECK_HIP(ctx, hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(gs);
    hipError_t ce = hipStreamEndCapture(gs, &g);
    if (rc || ce != hipSuccess) return qm_fail(ctx, QM_EHIP, "graph capture failed: %s", hipGetErrorString(ce));
    QM_CHECK_HIP(ctx, hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
    return QM_OK;
  }
  // replay `times` times on the graph stream, after everything already queued on s
  int replay(int64_t times, int prof_weight) {
    QM_CHECK_HIP(ctx, hipEventRecord(ev, s));
    QM_CHECK_HIP(ctx, hipStreamWaitEvent(gs, ev, 0));
    {
      ProfScope ps(ctx, gs, QM_K_JACOBI, (int)(prof_weight * times));
      for (int64_t r = 0; r < times; ++r) QM_CHECK_HIP(ctx, hipGraphLaunch(x, gs));
    }
    QM_CHECK_HIP(ctx, hipEventRecord(ev, gs));
    QM_CHECK_HIP(ctx, hipStreamWaitEvent(s, ev, 0));
    return QM_OK;
  }
  ~GraphBlock() {
    if (x && ev) {
      ctx->graphs.push_back({g, x, ev});
    } else {
      if (x) (void)hipGraphExecDestroy(x);
      if (g) (void)hipGraphDestroy(g);
      if (ev) (void)hipEventDestroy(ev);
    }
  }
};

int jacobi_fp64(qm_ctx* ctx, hipStream_t s, const qm_geom* g, const double* rho, double* v, double* tmp,
                int64_t max_sweeps, double tol, int check_every, int64_t* done_out, double* delta_out) {
  const int64_t n = npoints(g);
  const bool checks = tol >= 0 && check_every > 0;
  // sweeps per graph replay: the check interval, or a fixed block without checks
  const int64_t block = checks ? check_every : 200;
  int64_t done = 0;
  double delta = -1.0;
  bool converged = false;
  int rc;
  if (max_sweeps >= block) {
    GraphBlock gb(ctx, s);
    int64_t P, S;
    split_block(block, P, S);
    if ((rc = gb.capture([&](hipStream_t t) -> int {
           int r1;
           for (int64_t q = 0; q < P; q += 2) {
             if ((r1 = sweep2<double>(ctx, t, g, rho, v, tmp, false))) return r1;
             if ((r1 = sweep2<double>(ctx, t, g, rho, tmp, v, false))) return r1;
           }
           for (int64_t q = 0; q < S; q += 2) {
             if ((r1 = sweep<double>(ctx, t, g, rho, v, tmp, false))) return r1;
             if ((r1 = sweep<double>(ctx, t, g, rho, tmp, v, false))) return r1;
           }
           return QM_OK;
         })))
      return rc;
    while (done + block <= max_sweeps) {
      // without checks replay all blocks at once
      const int64_t times = checks ? 1 : (max_sweeps - done) / block;
      if ((rc = gb.replay(times, (int)block))) return rc;
      done += times * block;
      if (checks) {
        if ((rc = host_delta(ctx, s, QM_F64, v, tmp, n, &delta))) return rc;
        if (delta <= tol) {
          converged = true;
          break;
        }
      }
    }
  }
  while (!converged && done + 2 <= max_sweeps) {
    if ((rc = sweep<double>(ctx, s, g, rho, v, tmp))) return rc;
    if ((rc = sweep<double>(ctx, s, g, rho, tmp, v))) return rc;
    done += 2;
    if (checks && done % check_every == 0) {
      if ((rc = host_delta(ctx, s, QM_F64, v, tmp, n, &delta))) return rc;
      if (delta <= tol) converged = true;
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

__global__ void k_fill0(float* __restrict__ a, float* __restrict__ b, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    a[i] = 0.f;
    b[i] = 0.f;
  }
}

int jacobi_fp32(qm_ctx* ctx, hipStream_t s, const qm_geom* g, const double* rho, double* v, int64_t max_sweeps,
                double tol, int block, int64_t* done_out, double* delta_out) {
  const int64_t n = npoints(g);
  float *r = nullptr, *e = nullptr, *t = nullptr;
  int rc;
  if ((rc = qm_ws(ctx, s, WS_JAC_A, sizeof(float) * n, (void**)&r))) return rc;
  if ((rc = qm_ws(ctx, s, WS_JAC_B, sizeof(float) * n, (void**)&e))) return rc;
  if ((rc = qm_ws(ctx, s, WS_JAC_C, sizeof(float) * n, (void**)&t))) return rc;
  const unsigned upd_blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
  const bool checks = tol >= 0;
  // one residual block of m sweeps: residual, m FP32 sweeps, correction (kernels only).
  // With convergence checks the correction starts from cleared buffers and the block ends in
  // two single sweeps (e last, t the one before); without, the residual kernel also forms the
  // first sweep (into t) and the other m - 1 run as two-sweep passes (a single one if odd).
  auto enqueue = [&](hipStream_t q, int64_t m, bool prof) -> int {
    float* a = e;
    float* b = t;
    int64_t P, S;
    if (checks) {
      hipLaunchKernelGGL(k_residual, sweep_grid(g), dim3(kBX, kBY), 0, q, rho, (const double*)v, r, nullptr, g->nx,
                         g->ny, g->nz, g->step);
      hipLaunchKernelGGL(k_fill0, dim3(upd_blocks), dim3(256), 0, q, e, t, n);
      split_block(m, P, S);
    } else {
      hipLaunchKernelGGL(k_residual, sweep_grid(g), dim3(kBX, kBY), 0, q, rho, (const double*)v, r, t, g->nx,
                         g->ny, g->nz, g->step);
      std::swap(a, b);
      P = (m - 1) / 2;
      S = (m - 1) % 2;
    }
    for (int64_t j = 0; j < P; ++j) {
      int r1 = sweep2<float>(ctx, q, g, r, a, b, prof);
      if (r1) return r1;
      std::swap(a, b);
    }
    for (int64_t j = 0; j < S; ++j) {
      int r1 = sweep<float>(ctx, q, g, r, a, b, prof);
      if (r1) return r1;
      std::swap(a, b);
    }
    hipLaunchKernelGGL(k_add_correction, dim3(upd_blocks), dim3(256), 0, q, v, (const float*)a, n);
    return hipGetLastError() == hipSuccess ? QM_OK : QM_EHIP;
  };
  int64_t done = 0;
  double delta = -1.0;
  bool converged = false;
  if (max_sweeps >= block) {
    GraphBlock gb(ctx, s);
    if ((rc = gb.capture([&](hipStream_t q) { return enqueue(q, block, false); }))) return rc;
    while (done + block <= max_sweeps) {
      const int64_t times = checks ? 1 : (max_sweeps - done) / block;
      if ((rc = gb.replay(times, block))) return rc;
      done += times * block;
      if (checks) {
        // after a replay, e holds the last sweep and t the one before (block is even)
        if ((rc = host_delta(ctx, s, QM_F32, e, t, n, &delta))) return rc;
        if (delta <= tol) {
          converged = true;
          break;
        }
      }
    }
  }
  if (!converged && done < max_sweeps) {
    const int64_t m = max_sweeps - done;
    if ((rc = enqueue(s, m, true))) return qm_fail(ctx, rc, "Jacobi (FP32): launch failed");
    if (checks) {
      float* last = (m & 1) ? t : e;
      float* prev = (m & 1) ? e : t;
      if ((rc = host_delta(ctx, s, QM_F32, last, prev, n, &delta))) return rc;
      converged = delta <= tol;
    }
    done = max_sweeps;
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
  if ((rc = qm_ws(ctx, s, WS_RED, sizeof(double) * 10 * kMomBlocks, (void**)&part))) return rc;
  if ((rc = qm_ws(ctx, s, WS_SCALAR, 256, (void**)&mom))) return rc;
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
  qm_reclaim_graphs(ctx, false);
  hipStream_t s = (hipStream_t)stream;
  if (precision == QM_PREC_FP64) {
    QM_REQUIRE(ctx, tmp != nullptr, QM_EINVAL, "jacobi: FP64 path needs a second buffer");
    QM_REQUIRE(ctx, check_every >= 0 && check_every % 2 == 0, QM_EINVAL, "jacobi: check_every must be even");
    return jacobi_fp64(ctx, s, g, (const double*)rho, (double*)v, (double*)tmp, max_sweeps, tolerance,
                       check_every, sweeps_done, last_delta);
  }
  if (precision == QM_PREC_FP32) {
    QM_REQUIRE(ctx, check_every > 0 && check_every % 2 == 0, QM_EINVAL,
               "jacobi: FP32 path needs an even residual interval");
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
  int blk0, bx, by;          // pointwise kernels: 64 x 4 tiles, one z plane per block
  int blk0s, tbx, tby, zs;   // sweep: 64 x 32 tiles (four 64 x 8 strips), zs planes per block
  int blk0t, tbx2, tby2, zs2;  // two-sweep pass: 62-column strips (124 with paired columns)
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

__device__ __forceinline__ int find_item_s(const JItem* __restrict__ it, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (it[mid].blk0s <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int find_item_t(const JItem* __restrict__ it, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (it[mid].blk0t <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T>
__global__ __launch_bounds__(kSW* kSWaves) void kb_sweep2(const JItem* __restrict__ items, int n, int dir) {
  const JItem& q = items[find_item_t(items, n, blockIdx.x)];
  const int local = blockIdx.x - q.blk0t;
  const int bx = local % q.tbx2;
  const int rest = local / q.tbx2;
  const int by = rest % q.tby2;
  const int bz = rest / q.tby2;
  strip2_body<T, r2<T>()>(bx, by, bz, q.zs2, (const T*)q.rho, (const T*)(dir ? q.b : q.a), (T*)(dir ? q.a : q.b),
                          q.nx, q.ny, q.nz, (T)q.h);
}

template <int R>
__global__ __launch_bounds__(kSW* kSWaves) void kb_sweep2p(const JItem* __restrict__ items, int n, int dir) {
  const JItem& q = items[find_item_t(items, n, blockIdx.x)];
  const int local = blockIdx.x - q.blk0t;
  const int bx = local % q.tbx2;
  const int rest = local / q.tbx2;
  const int by = rest % q.tby2;
  const int bz = rest / q.tby2;
  strip2p_body<R>(bx, by, bz, q.zs2, (const float*)q.rho, (const float*)(dir ? q.b : q.a), (float*)(dir ? q.a : q.b),
                  q.nx, q.ny, q.nz, (float)q.h);
}

template <typename T>
__global__ __launch_bounds__(kSW* kSWaves) void kb_sweep(const JItem* __restrict__ items, int n, int dir) {
  const JItem& q = items[find_item_s(items, n, blockIdx.x)];
  const int local = blockIdx.x - q.blk0s;
  const int bx = local % q.tbx;
  const int rest = local / q.tbx;
  const int by = rest % q.tby;
  const int bz = rest / q.tby;
  strip_body<T>(bx, by, bz, q.zs, (const T*)q.rho, (const T*)(dir ? q.b : q.a), (T*)(dir ? q.a : q.b), q.nx, q.ny,
                  q.nz, (T)q.h);
}

__global__ __launch_bounds__(kBX* kBY) void kb_residual(const JItem* __restrict__ items, int n) {
  const int it = find_item(items, n, blockIdx.x);
  const JItem& q = items[it];
  int i, j, k;
  if (!item_point(q, blockIdx.x - q.blk0, i, j, k)) return;
  const int64_t sy = q.nx, sz = (int64_t)q.nx * q.ny;
  const int64_t idx = i + sy * j + sz * k;
  // the first sweep of the correction goes to b (every point, so a needs no clearing)
  float* t = (float*)q.b;
  if (i == 0 || j == 0 || k == 0 || i == q.nx - 1 || j == q.ny - 1 || k == q.nz - 1) {
    q.r32[idx] = 0.f;
    t[idx] = 0.f;
    return;
  }
  const double* v = q.v64;
  const double c = v[idx];
  double s = v[idx - 1] + v[idx + 1];
  s += v[idx + sy];
  s += v[idx - sy];
  s += v[idx + sz];
  s += v[idx - sz];
  const float rr = (float)(q.rho64[idx] - q.h * (s - 6.0 * c));
  q.r32[idx] = rr;
  t[idx] = first_correction(rr, 1.0f / (float)q.h);   // 1/h rounded as in the sweep kernels
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

// destroy the graphs (and events) whose last replay has finished (all of them if wait); an entry
// with no graph only carries the event of a replay of the cached batch graph
void qm_reclaim_graphs(qm_ctx* ctx, bool wait) {
  auto& v = ctx->graphs;
  size_t k = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    if (wait) (void)hipEventSynchronize(v[i].done);
    if (hipEventQuery(v[i].done) == hipSuccess) {
      if (v[i].x) (void)hipGraphExecDestroy(v[i].x);
      if (v[i].g) (void)hipGraphDestroy(v[i].g);
      (void)hipEventDestroy(v[i].done);
    } else {
      v[k++] = v[i];
    }
  }
  v.resize(k);
}

// Copies a host item table to dtab on stream s through the context's pinned staging buffer: the
// copy is stream-ordered and the host only waits (if at all) for the previous upload to have left
// the staging buffer.
static int upload_table(qm_ctx* ctx, hipStream_t s, void* dtab, const void* table, size_t bytes) {
  if (!ctx->table_uploaded) QM_CHECK_HIP(ctx, hipEventCreateWithFlags(&ctx->table_uploaded, hipEventDisableTiming));
  else QM_CHECK_HIP(ctx, hipEventSynchronize(ctx->table_uploaded));
  if (ctx->table_pinned_bytes < bytes) {
    if (ctx->table_pinned) QM_CHECK_HIP(ctx, hipHostFree(ctx->table_pinned));
    ctx->table_pinned = nullptr;
    ctx->table_pinned_bytes = 0;
    const size_t want = std::max<size_t>(bytes, 64 << 10);
    QM_CHECK_HIP(ctx, hipHostMalloc(&ctx->table_pinned, want, hipHostMallocDefault));
    ctx->table_pinned_bytes = want;
  }
  std::memcpy(ctx->table_pinned, table, bytes);
  QM_CHECK_HIP(ctx, hipMemcpyAsync(dtab, ctx->table_pinned, bytes, hipMemcpyHostToDevice, s));
  QM_CHECK_HIP(ctx, hipEventRecord(ctx->table_uploaded, s));
  return QM_OK;
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
  int blocks = 0, sblocks = 0, tblocks = 0;
  size_t f32_elems = 0;
  constexpr int RD = kR2D, RF = kR2F;
  // FP32 two-sweep passes take the paired-column kernel when every grid has an even nx
  bool pairs = !f64 && ctx->jacobi_pairs;
  for (int i = 0; i < n && pairs; ++i) pairs = pairs_fit(geoms[i].nx, geoms[i].ny);
  const int W = pairs ? kWP : Strip2<RD>::W;
  const int TY = f64 ? Strip2<RD>::TY : pairs ? Strip2P<kR2P>::TY : Strip2<RF>::TY;
  // two-sweep passes: one slab length for the whole batch, for about 1024 workgroups in all
  // (per-grid slabs of small grids would be a few planes, each re-reading 4 halo planes)
  int64_t tile_planes = 0;
  for (int i = 0; i < n; ++i) {
    if ((rc = qm_check_geom(ctx, &geoms[i]))) return rc;
    tile_planes += (int64_t)((geoms[i].nx + W - 1) / W) * ((geoms[i].ny + TY - 1) / TY) * geoms[i].nz;
  }
  const int zs2_batch = ctx->jacobi_zs > 0 ? ctx->jacobi_zs : (int)std::max<int64_t>(4, (tile_planes + 1023) / 1024);
  for (int i = 0; i < n; ++i) {
    QM_REQUIRE(ctx, rho[i] && v[i] && (!f64 || (tmp && tmp[i])), QM_EINVAL, "jacobi_batch: item %d incomplete", i);
    JItem& q = items[i];
    q.nx = geoms[i].nx;
    q.ny = geoms[i].ny;
    q.nz = geoms[i].nz;
    q.h = geoms[i].step;
    q.bx = (q.nx + kBX - 1) / kBX;
    q.by = (q.ny + kBY - 1) / kBY;
    q.blk0 = blocks;
    q.zs = pick_zs(q.nx, q.ny, q.nz, ctx->jacobi_zs);
    q.tbx = (q.nx + kSW - 1) / kSW;
    q.tby = (q.ny + kTileY - 1) / kTileY;
    q.blk0s = sblocks;
    sblocks += q.tbx * q.tby * ((q.nz + q.zs - 1) / q.zs);
    {
      q.zs2 = std::min(zs2_batch, q.nz);
      q.tbx2 = (q.nx + W - 1) / W;
      q.tby2 = (q.ny + TY - 1) / TY;
      q.blk0t = tblocks;
      tblocks += q.tbx2 * q.tby2 * ((q.nz + q.zs2 - 1) / q.zs2);
    }
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
    if ((rc = qm_ws(ctx, s, WS_JAC_A, sizeof(float) * 3 * f32_elems, (void**)&w))) return rc;
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
  const size_t tbytes = sizeof(JItem) * items.size();
  if ((rc = qm_ws(ctx, s, WS_BATCH, tbytes, &dtab))) return rc;
  if ((rc = upload_table(ctx, s, dtab, items.data(), tbytes))) return rc;
  const JItem* dt = (const JItem*)dtab;
  const dim3 grid(blocks), sgrid(sblocks), tgrid(tblocks), thr(kBX, kBY), sthr(kSW, kSWaves);
  // one block of work on stream t; kf = sweeps in the block
  auto enqueue_block = [&](hipStream_t t, int64_t kf) -> int {
    // kf / 2 two-sweep passes and kf % 2 single sweeps; dir = 1 while the result is in b
    int dir = 0;
    if (f64) {
      for (int64_t q = 0; q < kf / 2; ++q, dir ^= 1)
        hipLaunchKernelGGL(kb_sweep2<double>, tgrid, sthr, 0, t, dt, n, dir);
      if (kf & 1) {
        hipLaunchKernelGGL(kb_sweep<double>, sgrid, sthr, 0, t, dt, n, dir);
        dir ^= 1;
      }
      if (dir) hipLaunchKernelGGL(kb_copy_back, grid, thr, 0, t, dt, n);
    } else {
      // the residual kernel forms the first sweep of the correction (into b); kf - 1 follow
      hipLaunchKernelGGL(kb_residual, grid, thr, 0, t, dt, n);
      dir = 1;
      for (int64_t q = 0; q < (kf - 1) / 2; ++q, dir ^= 1) {
        if (pairs) hipLaunchKernelGGL(kb_sweep2p<kR2P>, tgrid, sthr, 0, t, dt, n, dir);
        else hipLaunchKernelGGL(kb_sweep2<float>, tgrid, sthr, 0, t, dt, n, dir);
      }
      if ((kf - 1) & 1) {
        hipLaunchKernelGGL(kb_sweep<float>, sgrid, sthr, 0, t, dt, n, dir);
        dir ^= 1;
      }
      hipLaunchKernelGGL(kb_add, grid, thr, 0, t, dt, n, dir);
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
    // every kernel argument and grid of the block follows from the item table, its address, the
    // block length and the precision
    std::vector<char> key(sizeof(int) * 4 + sizeof(void*) + tbytes);
    {
      const int hdr[4] = {n, block, f64 ? 1 : 0, pairs ? 1 : 0};
      std::memcpy(key.data(), hdr, sizeof(hdr));
      std::memcpy(key.data() + sizeof(hdr), &dtab, sizeof(void*));
      std::memcpy(key.data() + sizeof(hdr) + sizeof(void*), items.data(), tbytes);
    }
    if (!ctx->jgraph_exec || key != ctx->jgraph_key) {
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
      if (ctx->jgraph_exec) {   // the replaced graph goes once its replays so far have run
        hipEvent_t old;
        QM_CHECK_HIP(ctx, hipEventCreateWithFlags(&old, hipEventDisableTiming));
        QM_CHECK_HIP(ctx, hipEventRecord(old, gs));
        ctx->graphs.push_back({ctx->jgraph, ctx->jgraph_exec, old});
      }
      ctx->jgraph = graph;
      ctx->jgraph_exec = exec;
      ctx->jgraph_key.swap(key);
    }
    for (int64_t r = 0; r < full; ++r) QM_CHECK_HIP(ctx, hipGraphLaunch(ctx->jgraph_exec, gs));
    if (rest) {
      if ((rc = enqueue_block(gs, rest))) return qm_fail(ctx, rc, "jacobi_batch: launch failed");
    }
    QM_CHECK_HIP(ctx, hipEventRecord(ev, gs));
    QM_CHECK_HIP(ctx, hipStreamWaitEvent(s, ev, 0));
    ctx->graphs.push_back({nullptr, nullptr, ev});   // the event is destroyed once it has fired
  } else {
    for (int64_t r = 0; r < full; ++r)
      if ((rc = enqueue_block(s, block))) return qm_fail(ctx, rc, "jacobi_batch: launch failed");
    if (rest)
      if ((rc = enqueue_block(s, rest))) return qm_fail(ctx, rc, "jacobi_batch: launch failed");
  }
  return QM_OK;
}
