// FFT Poisson solve of the total charge: hand-written mixed-radix Stockham FFT.
//
// Reference kernel_fft.cu:23-106 + cuFFT (qmtools.py:2578-2639):
//   rhs = -( -Q/h^3 + sum_A Z_A (2 pi s^2)^-3/2 exp(-r^2/2s^2) [r^2 <= (cutoff s)^2] ) / eps0
//   (Angstrom / volt units; voxel centres (i+1/2) h + origin; nuclei at coords * 0.529177211;
//   Q read x-fastest, rhs written C order [i][j][k])
//   periodic:  V^ = rhs^ * (-1/(4 pi^2 |k|^2)), V^(0) = 0, k in cycles/Angstrom
//   free space (no reference counterpart): V = rhs (*) G on the 2x zero-padded grid,
//     G(d) = -h^3/(4 pi |d|), G(0) = -h^3 C0/(4 pi h), C0 = 2 (3 asinh(1/sqrt2) - pi/4)
//     (the average of 1/|u| over the voxel), G^ cached per geometry.
//
// Layout: complex buffer in C order [n0][n1][n2] (axis 2 contiguous).  One
// kernel per axis; a workgroup loads B lines into LDS (coalesced: for axis 2
// the lines are contiguous, otherwise the B lines are adjacent along axis 2),
// runs the Stockham stages out of place between two LDS planes, and writes
// back.  Twiddles exp(-2 pi i m/n) come from a per-length table in LDS; a
// radix-R stage uses the same table for its R-point DFT.  Zero padding is
// exploited by transforming only the lines that can be non-zero.
#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <vector>

#include "qm_internal.h"

namespace {

constexpr double kRBohr = 0.529177211;       // kernel_fft.cu:15
constexpr double kEps0 = 0.0055263494;       // kernel_fft.cu:16, e/(V Angstrom)
constexpr double kAng2Bor = 1.8897259886;
constexpr int kThreads = 256;
constexpr int kFftThreads = 512;   // threads of an axis-pass workgroup
// points per workgroup (B lines of n): 8 per thread for complex128, 16 for complex64, so that a
// workgroup's LDS is about the same and a strided axis reads 128 contiguous bytes (B adjacent lines)
// (power-of-two lengths only: the mixed-radix kernel's register staging would spill)
template <typename R, bool POW2>
struct FftPts {
  static constexpr int kMax = sizeof(R) == 4 && POW2 ? 8192 : 4096;
};
constexpr int kMaxStages = 24;

template <typename R>
struct Cx {
  R x, y;
};

template <typename R>
__device__ __forceinline__ Cx<R> cmul(Cx<R> a, Cx<R> b) {
  return Cx<R>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// LDS index skew: one empty slot after every W complex values, W = 256 bytes / value size
// (16 for complex128, 32 for complex64), so that a W-value run spans the 64 banks once and
// the strided accesses of the Stockham stages (stride RR in the first stage's writes) and of
// the twiddle table land on distinct banks.  Lines are LS slots apart, LS = W/8 mod W, so the
// 8 lines that one LDS cycle serves in the load and store phases of a strided axis (8 lines
// x W/8 consecutive points) do not share banks either.
template <typename R>
struct Skew {
  static constexpr int S = sizeof(R) == 8 ? 4 : 5;   // log2 W
};
template <typename R>
__host__ __device__ __forceinline__ int sk(int i) { return i + (i >> Skew<R>::S); }
template <typename R>
__host__ __device__ __forceinline__ int line_slots(int n) {
  constexpr int W = 1 << Skew<R>::S;
  return (n + (n >> Skew<R>::S) + W - 1) / W * W + W / 8;
}
template <typename R>
__host__ __device__ __forceinline__ int table_slots(int n) { return n + (n >> Skew<R>::S) + 1; }

// Division by a run-time divisor 1 <= d < 2^16 of dividends x < 2^16: q = umulhi(x, ceil(2^32 / d))
// is exact (the product overshoots x / d by less than x / 2^32 < 2^-16 <= 1 / d, which cannot reach
// the next integer since frac(x / d) <= 1 - 1 / d); d = 1 is stored as 0.  The mixed-radix passes
// divide every element and butterfly index by the length, the lines per workgroup, the butterfly
// count and the span; these take one multiply instead of a division sequence.
struct FastDiv {
  uint32_t mul;
  __host__ static FastDiv of(int d) {
    FastDiv f;
    f.mul = d <= 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + (uint64_t)d - 1) / (uint64_t)d);
    return f;
  }
};
__device__ __forceinline__ int fdiv(int x, FastDiv f) { return f.mul ? (int)__umulhi((uint32_t)x, f.mul) : x; }

struct AxisPlan {
  int n;              // transform length
  int nstage;
  int radix[kMaxStages];
  int span[kMaxStages];
  FastDiv ndiv;                 // by n
  FastDiv mdiv[kMaxStages];     // by n / radix (butterflies per line)
  FastDiv pdiv[kMaxStages];     // by the span
};

struct AxisLines {
  int64_t stride;     // distance between consecutive elements of a line
  int nl0, nl1;       // lines: l0 in [0, nl0), l1 in [0, nl1)
  int64_t s0, s1;     // line base = l0 * s0 + l1 * s1
  int B;              // lines per workgroup (consecutive l1)
  FastDiv bdiv;       // by B
  int nin;            // positions >= nin of a line are zero on input (not loaded)
  int nout;           // positions >= nout are not needed after the pass (not stored)
};

// Transform of the free-space Green's function for one padded shape, computed at unit grid step:
// the transform is real and even in every axis, so one octant is kept, o1 = m1 / 2 + 1 and
// o2 = m2 / 2 + 1 wide; the value at grid step h is g[min(i, m0 - i)][min(j, m1 - j)][min(k, m2 - k)]
// times scale = h^2 (Angstrom), since g(d) = 1 / |d| scales as 1 / h and the voxel volume as h^3.
template <typename R>
struct GreenView {
  const R* g;
  int m0, m1, m2;
  int o1, o2;
  R scale;
};

// R-point DFT of x (forward sign through the table; the inverse uses a conjugated table).
// w1 = table index step of exp(-2 pi i / R).
template <typename R, int RR>
__device__ __forceinline__ void dft(Cx<R>* x, const Cx<R>* __restrict__ tw, int w1) {
  if (RR == 2) {
    const Cx<R> a = x[0], b = x[1];
    x[0] = Cx<R>{a.x + b.x, a.y + b.y};
    x[1] = Cx<R>{a.x - b.x, a.y - b.y};
  } else if (RR == 4) {
    const Cx<R> a0{x[0].x + x[2].x, x[0].y + x[2].y}, a1{x[0].x - x[2].x, x[0].y - x[2].y};
    const Cx<R> b0{x[1].x + x[3].x, x[1].y + x[3].y}, b1{x[1].x - x[3].x, x[1].y - x[3].y};
    const Cx<R> b1w = cmul(b1, tw[sk<R>(w1)]);   // exp(-+ i pi/2)
    x[0] = Cx<R>{a0.x + b0.x, a0.y + b0.y};
    x[1] = Cx<R>{a1.x + b1w.x, a1.y + b1w.y};
    x[2] = Cx<R>{a0.x - b0.x, a0.y - b0.y};
    x[3] = Cx<R>{a1.x - b1w.x, a1.y - b1w.y};
  } else if (RR == 8) {
    // X_2r = DFT4(x_m + x_m+4)_r, X_2r+1 = DFT4((x_m - x_m+4) W8^m)_r; W8^m = tw[m w1], W4 = W8^2
    Cx<R> a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = Cx<R>{x[q].x + x[q + 4].x, x[q].y + x[q + 4].y};
      b[q] = Cx<R>{x[q].x - x[q + 4].x, x[q].y - x[q + 4].y};
    }
    b[1] = cmul(b[1], tw[sk<R>(w1)]);
    b[2] = cmul(b[2], tw[sk<R>(2 * w1)]);
    b[3] = cmul(b[3], tw[sk<R>(3 * w1)]);
    const Cx<R> w4 = tw[sk<R>(2 * w1)];
    auto dft4 = [&](const Cx<R>* c, Cx<R>& o0, Cx<R>& o1, Cx<R>& o2, Cx<R>& o3) {
      const Cx<R> d0{c[0].x + c[2].x, c[0].y + c[2].y}, d1{c[0].x - c[2].x, c[0].y - c[2].y};
      const Cx<R> d2{c[1].x + c[3].x, c[1].y + c[3].y};
      const Cx<R> d3 = cmul(Cx<R>{c[1].x - c[3].x, c[1].y - c[3].y}, w4);
      o0 = Cx<R>{d0.x + d2.x, d0.y + d2.y};
      o1 = Cx<R>{d1.x + d3.x, d1.y + d3.y};
      o2 = Cx<R>{d0.x - d2.x, d0.y - d2.y};
      o3 = Cx<R>{d1.x - d3.x, d1.y - d3.y};
    };
    dft4(a, x[0], x[2], x[4], x[6]);
    dft4(b, x[1], x[3], x[5], x[7]);
  } else {
    Cx<R> y[RR];
#pragma unroll
    for (int j = 0; j < RR; ++j) {
      Cx<R> acc = x[0];
#pragma unroll
      for (int q = 1; q < RR; ++q) {
        const Cx<R> t = cmul(x[q], tw[sk<R>(((j * q) % RR) * w1)]);
        acc.x += t.x;
        acc.y += t.y;
      }
      y[j] = acc;
    }
#pragma unroll
    for (int j = 0; j < RR; ++j) x[j] = y[j];
  }
}

// One radix-RR Stockham stage in place: all inputs of the thread's butterflies
// are read into registers, then (after a barrier) the outputs are written.
// POW2: n, p and RR are powers of two, so the line / butterfly / twiddle indices come from
// shifts and masks instead of integer divisions by run-time values.
template <typename R, int RR, bool POW2 = false>
__device__ __forceinline__ void stage_inplace(Cx<R>* buf, const Cx<R>* __restrict__ tw, int n, int B, int p,
                                              FastDiv mdiv = FastDiv{0}, FastDiv pdiv = FastDiv{0}) {
  const int LS = line_slots<R>(n);
  constexpr int MAXB = (FftPts<R, POW2>::kMax / kFftThreads + RR - 1) / RR;   // butterflies per thread (B n <= kMax)
  const int m = n / RR;
  const int tstep = n / (p * RR);
  const int total = B * m;
  const int lm = POW2 ? __builtin_ctz(m) : 0;
  auto split = [&](int bf, int& l, int& i, int& k) __attribute__((always_inline)) {
    if (POW2) {
      l = bf >> lm;
      i = bf & (m - 1);
      k = i & (p - 1);
    } else {   // bf < B n <= 4096
      l = fdiv(bf, mdiv);
      i = bf - l * m;
      k = i - fdiv(i, pdiv) * p;
    }
  };
  Cx<R> x[MAXB][RR];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int bf = threadIdx.x + b * kFftThreads;
    if (bf < total) {
      int l, i, k;
      split(bf, l, i, k);
      const Cx<R>* src = buf + l * LS;
#pragma unroll
      for (int j = 0; j < RR; ++j) {
        Cx<R> v = src[sk<R>(i + j * m)];
        // j k tstep <= (RR - 1)(p - 1) n / (p RR) < n: the twiddle index needs no reduction
        if (k) v = cmul(v, tw[sk<R>(j * k * tstep)]);
        x[b][j] = v;
      }
      dft<R, RR>(x[b], tw, m);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int bf = threadIdx.x + b * kFftThreads;
    if (bf < total) {
      int l, i, k;
      split(bf, l, i, k);
      Cx<R>* dst = buf + l * LS;
      const int o = (i - k) * RR + k;
#pragma unroll
      for (int j = 0; j < RR; ++j) dst[sk<R>(o + j * p)] = x[b][j];
    }
  }
  __syncthreads();
}

// Any radix, out of place between two LDS planes (used when a length has a prime factor > 7).
// Indices as in stage_inplace (quotients by multiply-high, twiddle index below n); the DFT table
// index (j q mod RR) m is carried from one q to the next.
template <typename R>
__device__ __forceinline__ void stage_generic(const Cx<R>* in, Cx<R>* out, const Cx<R>* __restrict__ tw, int n,
                                              int B, int RR, int p, FastDiv mdiv, FastDiv pdiv) {
  const int m = n / RR;
  const int tstep = n / (p * RR);
  const int LS = line_slots<R>(n);
  for (int bf = threadIdx.x; bf < B * m; bf += kFftThreads) {
    const int l = fdiv(bf, mdiv), i = bf - l * m, k = i - fdiv(i, pdiv) * p;
    const Cx<R>* src = in + l * LS;
    Cx<R>* dst = out + l * LS;
    const int o = (i - k) * RR + k;
    for (int j = 0; j < RR; ++j) {
      Cx<R> acc{0, 0};
      int jq = 0;   // j q mod RR
      for (int q = 0; q < RR; ++q) {
        Cx<R> v = src[sk<R>(i + q * m)];
        if (k) v = cmul(v, tw[sk<R>(q * k * tstep)]);
        const Cx<R> t = cmul(v, tw[sk<R>(jq * m)]);
        acc.x += t.x;
        acc.y += t.y;
        jq += j;
        if (jq >= RR) jq -= RR;
      }
      dst[sk<R>(o + j * p)] = acc;
    }
  }
  __syncthreads();
}

// The stages of one transform of a workgroup's B lines, from buf0 (buf1: the second plane of the
// generic stages); returns the plane that holds the result.
template <typename R, bool POW2>
__device__ __forceinline__ Cx<R>* run_stages(const AxisPlan& plan, Cx<R>* buf0, Cx<R>* buf1, const Cx<R>* tw, int n,
                                             int B, int inplace) {
  if (POW2) {
    for (int s = 0; s < plan.nstage; ++s) {
      const int p = plan.span[s];
      switch (plan.radix[s]) {
        case 2: stage_inplace<R, 2, true>(buf0, tw, n, B, p); break;
        case 4: stage_inplace<R, 4, true>(buf0, tw, n, B, p); break;
        default: stage_inplace<R, 8, true>(buf0, tw, n, B, p); break;
      }
    }
    return buf0;
  }
  if (inplace) {
    for (int s = 0; s < plan.nstage; ++s) {
      const int p = plan.span[s];
      const FastDiv md = plan.mdiv[s], pd = plan.pdiv[s];
      switch (plan.radix[s]) {
        case 2: stage_inplace<R, 2>(buf0, tw, n, B, p, md, pd); break;
        case 3: stage_inplace<R, 3>(buf0, tw, n, B, p, md, pd); break;
        case 4: stage_inplace<R, 4>(buf0, tw, n, B, p, md, pd); break;
        case 5: stage_inplace<R, 5>(buf0, tw, n, B, p, md, pd); break;
        case 7: stage_inplace<R, 7>(buf0, tw, n, B, p, md, pd); break;
        default: stage_inplace<R, 8>(buf0, tw, n, B, p, md, pd); break;
      }
    }
    return buf0;
  }
  Cx<R>* in = buf0;
  Cx<R>* out = buf1;
  for (int s = 0; s < plan.nstage; ++s) {
    stage_generic<R>(in, out, tw, n, B, plan.radix[s], plan.span[s], plan.mdiv[s], plan.pdiv[s]);
    Cx<R>* t = in;
    in = out;
    out = t;
  }
  return in;
}

// One FFT pass along one axis.  INV: inverse (conjugated) transform, unnormalised.
// inplace: all radices <= 8 (LDS = table + B lines); otherwise two planes of B lines.
// POW2: the length is a power of two (radices 8, 4, 2 only), which keeps the
// register footprint of the butterflies, and so the kernel's, small.
// MUL (axis 0 of the free-space solve only): forward transform, product with the Green's
// function transform, inverse transform, all on the lines held in LDS.
template <typename R, bool INV, bool POW2, bool MUL = false>
__global__ __launch_bounds__(kFftThreads) void k_fft_axis(Cx<R>* __restrict__ data, const Cx<R>* __restrict__ twg,
                                                       AxisPlan plan, AxisLines ln, int inplace, GreenView<R> gv) {
  extern __shared__ unsigned char smem[];
  const int n = plan.n;
  const int B = ln.B;
  const int LS = line_slots<R>(n);
  Cx<R>* tw = (Cx<R>*)smem;
  Cx<R>* buf0 = tw + table_slots<R>(n);
  Cx<R>* buf1 = buf0 + (int64_t)B * LS;
  const int lines_per_row = (ln.nl1 + B - 1) / B;
  const int64_t wg = blockIdx.x;
  const int64_t l0 = wg / lines_per_row;
  const int l1b = (int)(wg % lines_per_row) * B;
  for (int q = threadIdx.x; q < n; q += kFftThreads) {
    Cx<R> w = twg[q];
    if (INV) w.y = -w.y;
    tw[sk<R>(q)] = w;
  }
  const int64_t base0 = l0 * ln.s0 + (int64_t)l1b * ln.s1;
  const int nb = min(B, ln.nl1 - l1b);
  // element e of the workgroup's B lines: line l, position a (line-major on the contiguous axis,
  // line-minor otherwise, so a load instruction reads B adjacent lines); POW2 plans have B a
  // power of two as well, and the split is shifts and masks
  const int lgn = POW2 ? __builtin_ctz(n) : 0, lgb = POW2 ? __builtin_ctz(B) : 0;
  auto split = [&](int e) __attribute__((always_inline)) {   // (line, position); e < B n <= 8192
    if (ln.stride == 1) {
      const int l = POW2 ? e >> lgn : fdiv(e, plan.ndiv);
      return make_int2(l, e - l * n);
    }
    const int a = POW2 ? e >> lgb : fdiv(e, ln.bdiv);
    return make_int2(POW2 ? e & (B - 1) : e - a * B, a);
  };
  // all of a thread's loads are issued before any is used (B n <= kMax: 8 or 16 per thread)
  constexpr int kPer = FftPts<R, POW2>::kMax / kFftThreads;
  Cx<R> ld[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int e = threadIdx.x + q * kFftThreads;
    ld[q] = Cx<R>{0, 0};
    if (e < B * n) {
      const int2 la = split(e);
      if (la.x < nb && la.y < ln.nin) ld[q] = data[base0 + (int64_t)la.x * ln.s1 + (int64_t)la.y * ln.stride];
    }
  }
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int e = threadIdx.x + q * kFftThreads;
    if (e < B * n) {
      const int2 la = split(e);
      buf0[la.x * LS + sk<R>(la.y)] = ld[q];
    }
  }
  __syncthreads();
  Cx<R>* res = run_stages<R, POW2>(plan, buf0, buf1, tw, n, B, inplace);
  if (MUL) {
    // line (j, k) = (l0, l1b + l), position i; the octant index folds each of them
    const int j = (int)l0, jj = min(j, gv.m1 - j);
    for (int e = threadIdx.x; e < B * n; e += kFftThreads) {
      const int2 la = split(e);
      if (la.x < nb) {
        const int k = l1b + la.x, i = la.y;
        const R f = gv.g[((int64_t)min(i, gv.m0 - i) * gv.o1 + jj) * gv.o2 + min(k, gv.m2 - k)] * gv.scale;
        Cx<R>& v = res[la.x * LS + sk<R>(la.y)];
        v.x *= f;
        v.y *= f;
      }
    }
    for (int q = threadIdx.x; q < n; q += kFftThreads) tw[sk<R>(q)].y = -tw[sk<R>(q)].y;
    __syncthreads();
    res = run_stages<R, POW2>(plan, res, res == buf0 ? buf1 : buf0, tw, n, B, inplace);
  }
  for (int e = threadIdx.x; e < B * n; e += kFftThreads) {
    const int2 la = split(e);
    if (la.x < nb && la.y < ln.nout)
      data[base0 + (int64_t)la.x * ln.s1 + (int64_t)la.y * ln.stride] = res[la.x * LS + sk<R>(la.y)];
  }
}

// rhs, C order, into a [m0][m1][m2] complex buffer (m >= n: zero padding already in place).
// Tile: 32 i x 32 k at fixed j; Q is read along i (x-fastest) and written along k.
template <typename R, typename TI>
__global__ __launch_bounds__(kThreads) void k_rhs(const TI* __restrict__ q, const double4* __restrict__ atoms,
                                                  int natoms, GeomView g, double sigma, double cut2, int m1, int m2,
                                                  Cx<R>* __restrict__ out) {
  __shared__ double tile[32][33];
  const int i0 = blockIdx.x * 32, k0 = blockIdx.z * 32, j = blockIdx.y;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, k = k0 + r;
    double v = 0.0;
    if (i < g.nx && k < g.nz) v = (double)q[i + (int64_t)g.nx * (j + (int64_t)g.ny * k)];
    tile[r][tx] = v;
  }
  __syncthreads();
  const double step = g.h / kAng2Bor;
  const double norm = 1.0 / (sigma * sigma * sigma * 2.0 * M_PI * sqrt(2.0 * M_PI));
  const double inv2s2 = 1.0 / (2.0 * sigma * sigma);
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + r, k = k0 + tx;
    if (i >= g.nx || k >= g.nz) continue;
    double rho = tile[tx][r] / (-(step * step * step));
    const double x = g.ox / kAng2Bor + (i + 0.5) * step;
    const double y = g.oy / kAng2Bor + (j + 0.5) * step;
    const double z = g.oz / kAng2Bor + (k + 0.5) * step;
    for (int a = 0; a < natoms; ++a) {
      const double4 A = atoms[a];
      const double dx = x - A.x * kRBohr, dy = y - A.y * kRBohr, dz = z - A.z * kRBohr;
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 <= cut2) rho += A.w * norm * exp(-r2 * inv2s2);
    }
    out[((int64_t)i * m1 + j) * m2 + k] = Cx<R>{(R)(-rho / kEps0), (R)0};
  }
}

// periodic Poisson in k space, normalisation of the inverse transform folded in (kernel_fft.cu:82-106)
template <typename R>
__global__ void k_poisson_periodic(Cx<R>* __restrict__ a, int n0, int n1, int n2, double step, double scale) {
  const int64_t N = (int64_t)n0 * n1 * n2;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  const int k = (int)(idx % n2), j = (int)((idx / n2) % n1), i = (int)(idx / ((int64_t)n1 * n2));
  if (idx == 0) {
    a[0] = Cx<R>{0, 0};
    return;
  }
  const double kx = (i > n0 / 2 - 1 ? (double)(i - n0) : (double)i) / (n0 * step);
  const double ky = (j > n1 / 2 - 1 ? (double)(j - n1) : (double)j) / (n1 * step);
  const double kz = (k > n2 / 2 - 1 ? (double)(k - n2) : (double)k) / (n2 * step);
  const double f = -1.0 / (4.0 * M_PI * M_PI * (kx * kx + ky * ky + kz * kz)) * scale;
  Cx<R> v = a[idx];
  v.x = (R)(v.x * f);
  v.y = (R)(v.y * f);
  a[idx] = v;
}

// g(d) at unit grid step on the m0 x m1 x m2 cyclic grid (m >= 2n), C order: |d_i| = min(i, m - i),
// g(0) = C0.  Distances between two voxels of the grid are below n, so only |d| < n is ever used.
template <typename R>
__global__ void k_green(Cx<R>* __restrict__ a, int m0, int m1, int m2) {
  const int64_t N = (int64_t)m0 * m1 * m2;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  const int k = (int)(idx % m2), j = (int)((idx / m2) % m1), i = (int)(idx / ((int64_t)m1 * m2));
  const double x = min(i, m0 - i);
  const double y = min(j, m1 - j);
  const double z = min(k, m2 - k);
  const double r = sqrt(x * x + y * y + z * z);
  const double c0 = 2.0 * (3.0 * asinh(1.0 / sqrt(2.0)) - M_PI / 4.0);
  const double g = (idx == 0) ? c0 : 1.0 / r;
  a[idx] = Cx<R>{(R)g, (R)0};
}

// the octant i <= m0 / 2, j <= m1 / 2, k <= m2 / 2 of the transform's real part, times scale
template <typename R>
__global__ void k_green_store(const Cx<R>* __restrict__ a, R* __restrict__ gh, int m1, int m2, int o0, int o1,
                              int o2, double scale) {
  const int64_t N = (int64_t)o0 * o1 * o2;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  const int k = (int)(idx % o2), j = (int)((idx / o2) % o1), i = (int)(idx / ((int64_t)o1 * o2));
  gh[idx] = (R)((double)a[((int64_t)i * m1 + j) * m2 + k].x * scale);
}

template <typename R, typename TO>
__global__ void k_extract(const Cx<R>* __restrict__ a, int m1, int m2, int n0, int n1, int n2, TO* __restrict__ out) {
  const int64_t N = (int64_t)n0 * n1 * n2;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  const int k = (int)(idx % n2), j = (int)((idx / n2) % n1), i = (int)(idx / ((int64_t)n1 * n2));
  out[idx] = (TO)a[((int64_t)i * m1 + j) * m2 + k].x;
}

// ------------------------------------------------------------------ host: plans
std::vector<int> factorize(int n) {
  std::vector<int> f;
  for (int r : {8, 4, 2, 3, 5, 7}) {
    while (n % r == 0) {
      f.push_back(r);
      n /= r;
    }
  }
  for (int p = 11; n > 1; p += 2)
    while (n % p == 0) {
      f.push_back(p);
      n /= p;
    }
  return f;
}

}  // namespace

// Transforms of the free-space Green's function at unit step, one octant per (padded shape,
// precision), are kept up to kGreenCacheBytes per context; beyond that the least recently used
// are released (stream-ordered, from the stream-ordered allocator).
constexpr size_t kGreenCacheBytes = (size_t)4 << 30;

struct FftPlanCache {
  std::map<std::pair<int, int>, void*> twiddles;            // (n, bytes per real) -> device table
  size_t twiddle_bytes = 0;
  struct Green {
    int m0, m1, m2;   // padded shape
    int rbytes;
    void* gh;
    size_t bytes;
    uint64_t used;   // last use (cache clock)
  };
  std::vector<Green> greens;
  size_t green_bytes = 0;
  uint64_t clock = 0;
};

size_t qm_fft_cache_bytes(const FftPlanCache* c) { return c ? c->twiddle_bytes + c->green_bytes : 0; }

void qm_fft_cache_free(FftPlanCache* c) {
  if (!c) return;
  for (auto& kv : c->twiddles) (void)hipFree(kv.second);
  for (auto& g : c->greens) (void)hipFreeAsync(g.gh, nullptr);
  (void)hipStreamSynchronize(nullptr);
  delete c;
}

namespace {

template <typename R>
int twiddle_table(qm_ctx* ctx, int n, const Cx<R>** out) {
  if (!ctx->fft) ctx->fft = new FftPlanCache();
  auto key = std::make_pair(n, (int)sizeof(R));
  auto it = ctx->fft->twiddles.find(key);
  if (it != ctx->fft->twiddles.end()) {
    *out = (const Cx<R>*)it->second;
    return QM_OK;
  }
  std::vector<Cx<R>> h(n);
  for (int m = 0; m < n; ++m) {
    // exp(-2 pi i m / n) with the angle reduced exactly in integers
    const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
    h[m] = Cx<R>{(R)std::cos(ang), (R)std::sin(ang)};
  }
  void* d = nullptr;
  QM_CHECK_HIP(ctx, hipMalloc(&d, sizeof(Cx<R>) * n));
  QM_CHECK_HIP(ctx, hipMemcpy(d, h.data(), sizeof(Cx<R>) * n, hipMemcpyHostToDevice));
  ctx->fft->twiddles[key] = d;
  ctx->fft->twiddle_bytes += sizeof(Cx<R>) * n;
  *out = (const Cx<R>*)d;
  return QM_OK;
}

// Free-space padding: the smallest length >= 2n with no prime factor above 7.  A cyclic length of
// at least 2n - 1 holds the linear convolution of two n-point signals exactly, and such a length
// keeps every pass on the in-place radices (2n = 304 = 16 x 19 would take the generic stage; 315 = 3^2 5 7).
int freespace_len(int n) {
  for (int v = 2 * n;; ++v) {
    int x = v;
    for (int r : {2, 3, 5, 7})
      while (x % r == 0) x /= r;
    if (x == 1) return v;
  }
}

// one pass along `axis` of a C-order [m0][m1][m2] buffer, on the lines whose
// other two indices are below (lim_a, lim_b) (in increasing axis order); positions >= nin
// are read as zero and positions >= nout are not written.  gv: forward transform, product
// with the Green's function transform and inverse transform in one pass (axis 0 only).
template <typename R, bool INV>
int fft_pass(qm_ctx* ctx, hipStream_t s, Cx<R>* data, const int m[3], int axis, int lim_a, int lim_b, int nin,
             int nout, const GreenView<R>* gv = nullptr) {
  const int n = m[axis];
  const Cx<R>* tw = nullptr;
  int rc = twiddle_table<R>(ctx, n, &tw);
  if (rc) return rc;
  AxisPlan plan{};
  plan.n = n;
  std::vector<int> f = factorize(n);
  QM_REQUIRE(ctx, (int)f.size() <= kMaxStages, QM_ELIMIT, "fft: too many stages for length %d", n);
  int span = 1;
  plan.nstage = (int)f.size();
  plan.ndiv = FastDiv::of(n);
  for (size_t q = 0; q < f.size(); ++q) {
    plan.radix[q] = f[q];
    plan.span[q] = span;
    plan.mdiv[q] = FastDiv::of(n / f[q]);
    plan.pdiv[q] = FastDiv::of(span);
    span *= f[q];
  }
  const int64_t st[3] = {(int64_t)m[1] * m[2], (int64_t)m[2], 1};
  AxisLines ln{};
  ln.stride = st[axis];
  int oa = axis == 0 ? 1 : 0;            // other axes, oa < ob
  int ob = axis == 2 ? 1 : 2;
  ln.nl0 = lim_a;
  ln.nl1 = lim_b;
  ln.s0 = st[oa];
  ln.s1 = st[ob];
  ln.nin = nin;
  ln.nout = nout;
  QM_REQUIRE(ctx, nin >= 1 && nin <= n && nout >= 1 && nout <= n && (!gv || axis == 0), QM_EINVAL,
             "fft: bad pass");
  bool inplace = true;
  for (int r : f) inplace = inplace && r <= 8;
  const int planes = inplace ? 1 : 2;
  // B lines of n points per workgroup: B * n <= kMax (the register staging bound of stage_inplace)
  const bool pow2 = (n & (n - 1)) == 0;
  const int kMaxPts = pow2 ? FftPts<R, true>::kMax : FftPts<R, false>::kMax;
  QM_REQUIRE(ctx, n <= 4096, QM_ELIMIT, "fft: length %d exceeds 4096", n);
  int B = std::max(1, std::min(16, kMaxPts / n));
  if (!inplace) B = std::max(1, std::min(B, 2048 / n));
  B = std::max(1, std::min(B, ln.nl1));
  if (pow2)
    while (B & (B - 1)) B &= B - 1;   // the power-of-two kernel splits indices by shifts
  ln.B = B;
  ln.bdiv = FastDiv::of(B);
  const int lines_per_row = (ln.nl1 + B - 1) / B;
  const int64_t nwg = (int64_t)ln.nl0 * lines_per_row;
  const size_t shmem = sizeof(Cx<R>) * ((size_t)table_slots<R>(n) + (size_t)planes * B * line_slots<R>(n));
  QM_REQUIRE(ctx, shmem <= (size_t)128 * 1024, QM_ELIMIT, "fft: length %d needs %zu bytes of LDS per workgroup", n,
             shmem);
  QM_REQUIRE(ctx, nwg <= 0x7fffffff, QM_ELIMIT, "fft: too many lines");
  if (nwg == 0) return QM_OK;
  static bool attr_set = false;
  if (!attr_set) {
    const void* fns[4] = {(const void*)k_fft_axis<R, INV, false>, (const void*)k_fft_axis<R, INV, true>,
                          (const void*)k_fft_axis<R, false, false, true>, (const void*)k_fft_axis<R, false, true, true>};
    for (const void* f : fns)
      QM_CHECK_HIP(ctx, hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024));
    attr_set = true;
  }
  ProfScope ps(ctx, s, QM_K_FFT_AXIS);
  const GreenView<R> g0 = gv ? *gv : GreenView<R>{};
  const dim3 grid((unsigned)nwg), thr(kFftThreads);
  const int ip = inplace ? 1 : 0;
  if (gv) {
    if (pow2)
      hipLaunchKernelGGL((k_fft_axis<R, false, true, true>), grid, thr, shmem, s, data, tw, plan, ln, 1, g0);
    else
      hipLaunchKernelGGL((k_fft_axis<R, false, false, true>), grid, thr, shmem, s, data, tw, plan, ln, ip, g0);
  } else if (pow2) {
    hipLaunchKernelGGL((k_fft_axis<R, INV, true>), grid, thr, shmem, s, data, tw, plan, ln, 1, g0);
  } else {
    hipLaunchKernelGGL((k_fft_axis<R, INV, false>), grid, thr, shmem, s, data, tw, plan, ln, ip, g0);
  }
  QM_CHECK_LAUNCH(ctx);
  return QM_OK;
}

template <typename R>
int green_hat(qm_ctx* ctx, hipStream_t s, const int m[3], Cx<R>* work, const R** out) {
  if (!ctx->fft) ctx->fft = new FftPlanCache();
  FftPlanCache& fc = *ctx->fft;
  for (auto& gr : fc.greens)
    if (gr.m0 == m[0] && gr.m1 == m[1] && gr.m2 == m[2] && gr.rbytes == (int)sizeof(R)) {
      gr.used = ++fc.clock;
      *out = (const R*)gr.gh;
      return QM_OK;
    }
  const int64_t M = (int64_t)m[0] * m[1] * m[2];
  hipLaunchKernelGGL(k_green<R>, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, work, m[0], m[1], m[2]);
  QM_CHECK_LAUNCH(ctx);
  int rc;
  if ((rc = fft_pass<R, false>(ctx, s, work, m, 2, m[0], m[1], m[2], m[2]))) return rc;
  if ((rc = fft_pass<R, false>(ctx, s, work, m, 1, m[0], m[2], m[1], m[1]))) return rc;
  if ((rc = fft_pass<R, false>(ctx, s, work, m, 0, m[1], m[2], m[0], m[0]))) return rc;
  const int o[3] = {m[0] / 2 + 1, m[1] / 2 + 1, m[2] / 2 + 1};
  const int64_t O = (int64_t)o[0] * o[1] * o[2];
  // make room: release the least recently used transforms (work queued earlier on s still reads them)
  const size_t bytes = sizeof(R) * (size_t)O;
  while (!fc.greens.empty() && fc.green_bytes + bytes > kGreenCacheBytes) {
    size_t lru = 0;
    for (size_t i = 1; i < fc.greens.size(); ++i)
      if (fc.greens[i].used < fc.greens[lru].used) lru = i;
    QM_CHECK_HIP(ctx, hipFreeAsync(fc.greens[lru].gh, s));
    fc.green_bytes -= fc.greens[lru].bytes;
    fc.greens.erase(fc.greens.begin() + lru);
  }
  void* gh = nullptr;
  QM_CHECK_HIP(ctx, hipMallocAsync(&gh, bytes, s));
  // V = -(h^3 / 4 pi) IFFT(rhs^ g^) / M   (rhs = -rho/eps0, g = 1/r); the step enters as h^2 at use
  const double scale = -1.0 / (4.0 * M_PI) / (double)M;
  hipLaunchKernelGGL(k_green_store<R>, dim3((unsigned)((O + 255) / 256)), dim3(256), 0, s, (const Cx<R>*)work,
                     (R*)gh, m[1], m[2], o[0], o[1], o[2], scale);
  QM_CHECK_LAUNCH(ctx);
  fc.greens.push_back({m[0], m[1], m[2], (int)sizeof(R), gh, bytes, ++fc.clock});
  fc.green_bytes += bytes;
  *out = (const R*)gh;
  return QM_OK;
}

template <typename R, typename TIO>
int hartree_fft(qm_ctx* ctx, qm_mol* mol, hipStream_t s, const qm_geom* g, const TIO* q, double sigma,
                double cutoff, int mode, TIO* out) {
  const bool fs = mode == QM_FFT_FREESPACE;
  const int n[3] = {g->nx, g->ny, g->nz};
  const int m[3] = {fs ? freespace_len(n[0]) : n[0], fs ? freespace_len(n[1]) : n[1],
                    fs ? freespace_len(n[2]) : n[2]};
  const int64_t M = (int64_t)m[0] * m[1] * m[2];
  Cx<R>* work = nullptr;
  int rc = qm_ws(ctx, s, WS_FFT_A, sizeof(Cx<R>) * (size_t)M, (void**)&work);
  if (rc) return rc;
  GreenView<R> gv{};
  if (fs) {
    if ((rc = green_hat<R>(ctx, s, m, work, &gv.g))) return rc;
    const double h_ang = g->step / kAng2Bor;
    gv.m0 = m[0];
    gv.m1 = m[1];
    gv.m2 = m[2];
    gv.o1 = m[1] / 2 + 1;
    gv.o2 = m[2] / 2 + 1;
    gv.scale = (R)(h_ang * h_ang);
  }
  const GeomView gview = geom_view(g);
  const double cut2 = (cutoff * sigma) * (cutoff * sigma);
  // rhs on [i < n0][j < n1][k < n2]; every pass reads only positions already written (the
  // padding is never cleared: it is read as zero) and writes only positions read later
  hipLaunchKernelGGL((k_rhs<R, TIO>), dim3((n[0] + 31) / 32, n[1], (n[2] + 31) / 32), dim3(kThreads), 0, s, q,
                     (const double4*)mol->atoms, mol->natoms, gview, sigma, cut2, m[1], m[2], work);
  QM_CHECK_LAUNCH(ctx);
  // forward: z on lines (i < n0, j < n1), y on (i < n0, all k), x on all (j, k); inverse in reverse
  if ((rc = fft_pass<R, false>(ctx, s, work, m, 2, n[0], n[1], n[2], m[2]))) return rc;
  if ((rc = fft_pass<R, false>(ctx, s, work, m, 1, n[0], m[2], n[1], m[1]))) return rc;
  if (fs) {
    // x: forward, times G^, inverse, keeping i < n0
    if ((rc = fft_pass<R, false>(ctx, s, work, m, 0, m[1], m[2], n[0], n[0], &gv))) return rc;
  } else {
    if ((rc = fft_pass<R, false>(ctx, s, work, m, 0, m[1], m[2], m[0], m[0]))) return rc;
    const unsigned blocks = (unsigned)((M + 255) / 256);
    const double step_ang = g->step / kAng2Bor;
    hipLaunchKernelGGL(k_poisson_periodic<R>, dim3(blocks), dim3(256), 0, s, work, m[0], m[1], m[2], step_ang,
                       1.0 / (double)M);
    QM_CHECK_LAUNCH(ctx);
    if ((rc = fft_pass<R, true>(ctx, s, work, m, 0, m[1], m[2], m[0], m[0]))) return rc;
  }
  if ((rc = fft_pass<R, true>(ctx, s, work, m, 1, n[0], m[2], m[1], n[1]))) return rc;
  if ((rc = fft_pass<R, true>(ctx, s, work, m, 2, n[0], n[1], m[2], n[2]))) return rc;
  const int64_t N = (int64_t)n[0] * n[1] * n[2];
  hipLaunchKernelGGL((k_extract<R, TIO>), dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, (const Cx<R>*)work,
                     m[1], m[2], n[0], n[1], n[2], out);
  QM_CHECK_LAUNCH(ctx);
  return QM_OK;
}

}  // namespace

extern "C" int qm_hartree_fft(qm_ctx* ctx, qm_mol* mol, void* stream, const qm_geom* g, int precision, int dtype,
                              const void* q, double sigma, double cutoff, int mode, void* out) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  if ((rc = qm_check_geom(ctx, g))) return rc;
  QM_REQUIRE(ctx, mol && q && out, QM_EINVAL, "fft: null argument");
  QM_REQUIRE(ctx, std::isfinite(sigma) && sigma > 0 && std::isfinite(cutoff) && cutoff > 0, QM_EINVAL,
             "fft: sigma and cutoff must be positive");
  QM_REQUIRE(ctx, mode == QM_FFT_PERIODIC || mode == QM_FFT_FREESPACE, QM_EINVAL, "fft: bad mode");
  QM_REQUIRE(ctx, g->nx >= 2 && g->ny >= 2 && g->nz >= 2, QM_EINVAL, "fft: grid too small");
  hipStream_t s = (hipStream_t)stream;
  if (precision == QM_PREC_FP64) {
    if (dtype == QM_F64)
      return hartree_fft<double, double>(ctx, mol, s, g, (const double*)q, sigma, cutoff, mode, (double*)out);
    if (dtype == QM_F32)
      return hartree_fft<double, float>(ctx, mol, s, g, (const float*)q, sigma, cutoff, mode, (float*)out);
  } else if (precision == QM_PREC_FP32) {
    if (dtype == QM_F64)
      return hartree_fft<float, double>(ctx, mol, s, g, (const double*)q, sigma, cutoff, mode, (double*)out);
    if (dtype == QM_F32)
      return hartree_fft<float, float>(ctx, mol, s, g, (const float*)q, sigma, cutoff, mode, (float*)out);
  } else {
    return qm_fail(ctx, QM_EINVAL, "fft: bad precision %d", precision);
  }
  return qm_fail(ctx, QM_EINVAL, "fft: bad dtype");
}
