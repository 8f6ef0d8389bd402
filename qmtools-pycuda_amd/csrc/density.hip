// Electron density on the grid: basis-on-grid + MFMA contraction.
//
// The reference (kernel_density.cu:34-155) recomputes phi_p and phi_q for
// every (p, q) pair at every voxel.  Here the work is split in two kernels
// per chunk of points:
//
//   1. k_phi_*:      Phi[t][f] for all points t of the chunk and basis
//                    functions f (one thread per point, shells are wave
//                    uniform, writes coalesced along t);
//   2. k_contract_*: rho_t = sum_n Phi[t][n] (Phi P)[t][n], the diagonal of
//                    Phi P Phi^T, as an MFMA GEMM with the row-dot fused
//                    into the epilogue (no X = Phi P round trip to HBM).
//
// P is the total density matrix 2*(tril(dm) + tril(dm,-1)^T): the reference
// reads only the lower triangle of dm and doubles off-diagonal terms
// (kernel_density.cu:74-80,138) and multiplies by 2 for the closed shell (:147).
//
// Layouts:
//   FP64:   Phi^T[f * ld + t]            (f64 MFMA 16x16x4: A lane l = (t: l%16, k: l/16);
//                                         D lane l, register r = (t: 4r + l/16, n: l%16))
//   F16/BF16: K8-interleaved [(f/8) * ld + t] * 8 + f%8, hi and lo planes
//           (MFMA 16x16x32: A lane l = (t: l%16, k: 8*(l/16)+0..7), one 16-byte load)
//   P uses the same layouts with nbf_pad in place of ld (P is symmetric).
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "qm_internal.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));

constexpr int kPhiThreads = 256;
constexpr int kTileM = 128;      // points per contraction workgroup (4 waves x 32 rows)
constexpr int kThreads = 256;

// Real solid harmonic prefactors of the reference (kernel_density.cu:17-28).
constexpr double C00 = 0.28209479177387814;
constexpr double C1 = 0.4886025119029199;
constexpr double C2A = 1.0925484305920792;
constexpr double C20 = 0.31539156525252005;
constexpr double C22 = 0.5462742152960396;

// Roundings fixed by explicit FMAs (no choice left to contraction), so that every kernel that
// evaluates FP32 basis values (written to memory or evaluated inside the contraction) gets the
// same bits: squared distance in FP64, 2z^2 - x^2 - y^2 as (z-x)(z+x) + (z-y)(z+y), radial sums.
__device__ __forceinline__ double dist2(double ex, double ey, double ez) {
  return __builtin_fma(ez, ez, __builtin_fma(ey, ey, ex * ex));
}
__device__ __forceinline__ float d0_factor(float x, float y, float z) {
  return __builtin_fmaf(z - x, z + x, (z - y) * (z + y));
}
__device__ __forceinline__ double d0_factor(double x, double y, double z) {
  return __builtin_fma(z - x, z + x, (z - y) * (z + y));
}

// Values of the 2L+1 functions of a shell, m order (0, +1, -1, +2, -2).
template <typename R>
__device__ __forceinline__ int shell_values(int L, R x, R y, R z, R radial, R* v) {
  if (L == 0) {
    v[0] = (R)C00 * radial;
    return 1;
  }
  if (L == 1) {
    v[0] = (R)C1 * z * radial;
    v[1] = (R)C1 * x * radial;
    v[2] = (R)C1 * y * radial;
    return 3;
  }
  // 2z^2 - x^2 - y^2 and x^2 - y^2 as products of sums and differences: no cancellation
  // where they pass through zero (in FP32 the plain forms lose up to ulp(r^2) there)
  v[0] = (R)C20 * d0_factor(x, y, z) * radial;
  v[1] = (R)C2A * x * z * radial;
  v[2] = (R)C2A * z * y * radial;
  v[3] = (R)C22 * ((x - y) * (x + y)) * radial;
  v[4] = (R)C2A * x * y * radial;
  return 5;
}

// Calls emit(c, value) for the 2L+1 values of a shell with constant indices (L is
// uniform over the wave), so no value is picked out of a register array at run time.
template <typename R, typename Emit>
__device__ __forceinline__ void shell_emit(int L, R x, R y, R z, R radial, Emit&& emit) {
  R v[5];
  if (L == 0) {
    shell_values<R>(0, x, y, z, radial, v);
    emit(0, v[0]);
  } else if (L == 1) {
    shell_values<R>(1, x, y, z, radial, v);
    emit(0, v[0]);
    emit(1, v[1]);
    emit(2, v[2]);
  } else {
    shell_values<R>(2, x, y, z, radial, v);
    emit(0, v[0]);
    emit(1, v[1]);
    emit(2, v[2]);
    emit(3, v[3]);
    emit(4, v[4]);
  }
}

// Position of chunk point t: sub-point s = t / nvox of voxel v0 + t % nvox (cell centred).
__device__ __forceinline__ void chunk_point(const GeomView& g, int64_t v0, int nvox, int nsub, int t, double& x,
                                            double& y, double& z) {
  const int s = t / nvox;
  const int64_t v = v0 + (t - (int64_t)s * nvox);
  const int64_t i = v % g.nx;
  const int64_t j = (v / g.nx) % g.ny;
  const int64_t k = v / ((int64_t)g.nx * g.ny);
  const int sx = s % nsub, sy = (s / nsub) % nsub, sz = s / (nsub * nsub);
  const double inv = 1.0 / nsub;
  x = g.ox + ((double)i + (sx + 0.5) * inv) * g.h;
  y = g.oy + ((double)j + (sy + 0.5) * inv) * g.h;
  z = g.oz + ((double)k + (sz + 0.5) * inv) * g.h;
}

// exp(-a r2) == 0 in double for a r2 > 745.2: a shell whose smallest exponent
// passes that bound contributes exact zeros, so skipping it changes nothing.
// exponent beyond which a primitive's value (e^-50 ~ 2e-22 of its coefficient) is not evaluated
constexpr double kExpCut = kPhiExpCut;

__device__ __forceinline__ void phi_f64_body(int blk, const MolView& m, const GeomView& g, int64_t v0, int nvox,
                                             int nsub, int npts, const double* __restrict__ pts, int ld,
                                             int nbf_pad, double* __restrict__ phi) {
  const int t = blk * kPhiThreads + threadIdx.x;
  if (t >= ld) return;
  if (t >= npts) {
    for (int f = 0; f < nbf_pad; ++f) phi[(int64_t)f * ld + t] = 0.0;
    return;
  }
  double x, y, z;
  if (pts) {
    x = pts[3 * (int64_t)t];
    y = pts[3 * (int64_t)t + 1];
    z = pts[3 * (int64_t)t + 2];
  } else {
    chunk_point(g, v0, nvox, nsub, t, x, y, z);
  }
  // shells come grouped by atom: the offset to the nucleus is formed once per atom (uniform test)
  int atom = -1;
  double dx = 0, dy = 0, dz = 0, r2 = 0;
  for (int s = 0; s < m.nshell; ++s) {
    const DevShell sh = m.shells[s];
    if (sh.atom != atom) {
      atom = sh.atom;
      const double4 a = m.atoms[atom];
      dx = x - a.x;
      dy = y - a.y;
      dz = z - a.z;
      r2 = dx * dx + dy * dy + dz * dz;
    }
    double radial = 0.0;
    if (r2 <= (double)sh.rcut2) {
      for (int p = 0; p < sh.nprim; ++p) {
        const double2 pc = m.prims[sh.poff + p];
        const double e = pc.x * r2;
        if (e < kExpCut) radial += pc.y * exp(-e);   // beyond: below 2e-22 of the coefficient
      }
    }
    double* __restrict__ col = phi + (int64_t)sh.bf * ld + t;
    shell_emit<double>(sh.L, dx, dy, dz, radial, [&](int c, double val) { col[(int64_t)c * ld] = val; });
  }
  for (int f = m.nbf; f < nbf_pad; ++f) phi[(int64_t)f * ld + t] = 0.0;
}

__global__ __launch_bounds__(kPhiThreads) void k_phi_f64(MolView m, GeomView g, int64_t v0, int nvox, int nsub,
                                                         int npts, const double* __restrict__ pts, int ld,
                                                         int nbf_pad, double* __restrict__ phi) {
  phi_f64_body(blockIdx.x, m, g, v0, nvox, nsub, npts, pts, ld, nbf_pad, phi);
}

template <typename T16>
__device__ __forceinline__ T16 to16(float x);
template <>
__device__ __forceinline__ _Float16 to16<_Float16>(float x) { return (_Float16)x; }
template <>
__device__ __forceinline__ __bf16 to16<__bf16>(float x) { return (__bf16)x; }

// FP32 basis values Phi^T[f * ld + t] (same layout as FP64), single pass, and
// the per-point power of two pscale[t] that brings the point's largest |phi|
// to [1, 2): the split contraction multiplies by it before rounding to 16 bits,
// which keeps points far from every nucleus out of the FP16 subnormal range.
// exp(-a r^2) is evaluated as exp2(-(a log2 e) r^2) with a log2 e rounded once
// on the host (prim_l2e), so the argument carries a single FP32 rounding.
__device__ __forceinline__ void phi_f32_body(int blk, const MolView& m, const float2* __restrict__ prim_l2e,
                                             const GeomView& g, int64_t v0, int nvox, int nsub, int npts, int ld,
                                             int nbf_pad, float* __restrict__ phi, float* __restrict__ pscale) {
  const int t = blk * kPhiThreads + threadIdx.x;
  if (t >= ld) return;
  if (t >= npts) {
    for (int f = 0; f < nbf_pad; ++f) phi[(int64_t)f * ld + t] = 0.f;
    pscale[t] = 1.f;
    return;
  }
  double x, y, z;
  chunk_point(g, v0, nvox, nsub, t, x, y, z);
  // offsets to the nucleus are formed in FP64 and rounded once: rounding x and R
  // separately to FP32 costs ~1e-6 bohr, i.e. ~1e-4 relative error in tight primitives
  float vmax = 0.f;
  int atom = -1;   // shells come grouped by atom: offsets once per atom
  float dx = 0.f, dy = 0.f, dz = 0.f, r2 = 0.f;
  for (int s = 0; s < m.nshell; ++s) {
    const DevShell sh = m.shells[s];
    if (sh.atom != atom) {   // r^2 in FP64, rounded once: exp(-a r^2) amplifies its error by a r^2
      atom = sh.atom;
      const double4 a = m.atoms[atom];
      const double ex = x - a.x, ey = y - a.y, ez = z - a.z;
      dx = (float)ex;
      dy = (float)ey;
      dz = (float)ez;
      r2 = (float)dist2(ex, ey, ez);
    }
    float radial = 0.f;
    if (r2 <= sh.rcut2) {
      for (int p = 0; p < sh.nprim; ++p) {
        const float2 pc = prim_l2e[sh.poff + p];
        radial = __builtin_fmaf(pc.y, __builtin_amdgcn_exp2f(-pc.x * r2), radial);
      }
    }
    float* __restrict__ col = phi + (int64_t)sh.bf * ld + t;
    shell_emit<float>(sh.L, dx, dy, dz, radial, [&](int c, float val) {
      col[(int64_t)c * ld] = val;
      vmax = fmaxf(vmax, fabsf(val));
    });
  }
  for (int f = m.nbf; f < nbf_pad; ++f) phi[(int64_t)f * ld + t] = 0.f;
  int e = 0;
  if (vmax > 0.f) frexpf(vmax, &e);   // vmax = f 2^e, f in [0.5, 1)
  pscale[t] = ldexpf(1.0f, 1 - e);
}

__global__ __launch_bounds__(kPhiThreads) void k_phi_f32(MolView m, const float2* __restrict__ prim_l2e,
                                                         GeomView g, int64_t v0, int nvox, int nsub, int npts,
                                                         int ld, int nbf_pad, float* __restrict__ phi,
                                                         float* __restrict__ pscale) {
  phi_f32_body(blockIdx.x, m, prim_l2e, g, v0, nvox, nsub, npts, ld, nbf_pad, phi, pscale);
}

// Basis values already scaled and split for the 16-bit contraction: the hi and
// lo planes of (phi S) in the K8-interleaved operand layout, element (f, t) at
// ((f / 8) ld + t) 8 + f % 8, in one pass.  S is one power of two per molecule
// (split_phi_scale) that keeps every |phi S| <= 2^14, so hi never overflows FP16
// and values down to 2^-14 / S keep the full 22 bits of hi + lo; smaller ones
// (points where every basis function is that small) lose low bits to FP16
// subnormals, far below the density's relative tolerance.  The (hi, lo) pair of
// function f goes to LDS word [f % 8][thread] (consecutive lanes, consecutive
// words: no bank conflicts, no branches on f); after every 8th function the lane
// reads its 8 words back and stores 16 bytes per plane.  Each lane only reads
// what it wrote, so no barrier is needed.  Same arithmetic for the values as
// phi_f32_body.
template <typename T16>
__device__ __forceinline__ uint32_t bits16(T16 v) { return (uint32_t)__builtin_bit_cast(uint16_t, v); }

template <typename T16>
__device__ __forceinline__ void phi_split_body(int blk, const MolView& m, const float2* __restrict__ prim_l2e,
                                               const GeomView& g, int64_t v0, int nvox, int nsub, int npts, int ld,
                                               int nbf_pad, float S, T16* __restrict__ hi, T16* __restrict__ lo) {
  // no contraction: hi is the 16-bit rounding of the FP32 value (a product folded into the
  // conversion would round once from the exact product), identical in the fused kernel
#pragma clang fp contract(off)
  __shared__ uint32_t stage[8][kPhiThreads];
  const int t = blk * kPhiThreads + threadIdx.x;
  if (t >= ld) return;
  double x = 0, y = 0, z = 0;
  if (t < npts) chunk_point(g, v0, nvox, nsub, t, x, y, z);
  const float sc = t < npts ? S : 0.f;   // padding points store zeros through the same (uniform) loop
  uint32_t* __restrict__ mine = &stage[0][threadIdx.x];
  int f = 0;   // same sequence in every lane
  auto push = [&](float xs) __attribute__((always_inline)) {   // xs: the value already times sc
    const T16 hq = (T16)xs;
    const T16 lq = (T16)(xs - (float)hq);
    mine[(f & 7) * kPhiThreads] = bits16(hq) | (bits16(lq) << 16);
    if ((f & 7) == 7) {
      uint32_t w[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) w[q] = mine[q * kPhiThreads];
      // hi halves of words (2j, 2j+1) -> H[j], lo halves -> L[j]
      uint4 H, L;
      H.x = __builtin_amdgcn_perm(w[1], w[0], 0x05040100u);
      H.y = __builtin_amdgcn_perm(w[3], w[2], 0x05040100u);
      H.z = __builtin_amdgcn_perm(w[5], w[4], 0x05040100u);
      H.w = __builtin_amdgcn_perm(w[7], w[6], 0x05040100u);
      L.x = __builtin_amdgcn_perm(w[1], w[0], 0x07060302u);
      L.y = __builtin_amdgcn_perm(w[3], w[2], 0x07060302u);
      L.z = __builtin_amdgcn_perm(w[5], w[4], 0x07060302u);
      L.w = __builtin_amdgcn_perm(w[7], w[6], 0x07060302u);
      const int64_t o = ((int64_t)(f >> 3) * ld + t) * 8;
      *(uint4*)(hi + o) = H;
      *(uint4*)(lo + o) = L;
    }
    ++f;
  };
  int atom = -1;   // shells come grouped by atom: offsets once per atom
  float dx = 0.f, dy = 0.f, dz = 0.f, r2 = 0.f;
  for (int s = 0; s < m.nshell; ++s) {
    const DevShell sh = m.shells[s];
    if (sh.atom != atom) {   // r^2 in FP64, rounded once: exp(-a r^2) amplifies its error by a r^2
      atom = sh.atom;
      const double4 a = m.atoms[atom];
      const double ex = x - a.x, ey = y - a.y, ez = z - a.z;
      dx = (float)ex;
      dy = (float)ey;
      dz = (float)ez;
      r2 = (float)dist2(ex, ey, ez);
    }
    float radial = 0.f;
    if (r2 <= sh.rcut2) {
      for (int p = 0; p < sh.nprim; ++p) {
        const float2 pc = prim_l2e[sh.poff + p];
        radial = __builtin_fmaf(pc.y, __builtin_amdgcn_exp2f(-pc.x * r2), radial);
      }
    }
    // S is a power of two: scaling the radial part scales every value exactly
    shell_emit<float>(sh.L, dx, dy, dz, radial * sc, [&](int, float val) { push(val); });
  }
  while (f < nbf_pad) push(0.f);
}

template <typename T16>
__global__ __launch_bounds__(kPhiThreads) void k_phi_split(MolView m, const float2* __restrict__ prim_l2e, GeomView g,
                                                           int64_t v0, int nvox, int nsub, int npts, int ld,
                                                           int nbf_pad, float S, T16* __restrict__ hi,
                                                           T16* __restrict__ lo) {
  phi_split_body<T16>(blockIdx.x, m, prim_l2e, g, v0, nvox, nsub, npts, ld, nbf_pad, S, hi, lo);
}

// sum over the 16 lanes that share lane/16 (the n index of the MFMA D layout)
__device__ __forceinline__ double sum16(double v) {
  v += __shfl_xor(v, 8, 64);
  v += __shfl_xor(v, 4, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 1, 64);
  return v;
}
__device__ __forceinline__ float sum16f(float v) {
  v += __shfl_xor(v, 8, 64);
  v += __shfl_xor(v, 4, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 1, 64);
  return v;
}

template <typename TO>
__device__ __forceinline__ void store_rho(TO* out, int64_t idx, double v) { out[idx] = (TO)v; }

// ---------------------------------------------------------------------------
// Contraction kernels.  A workgroup (8 waves) owns TM points and every column
// of one column pass; the accumulators of X = Phi P for all its columns live
// in registers, so the Phi tile is read from HBM once for the GEMM (staged in
// double-buffered LDS, shared by the 8 waves) and once more for the row-dot
// epilogue.  P (L2-resident) is streamed straight into MFMA operands; wave w
// owns columns [w NT 16, (w+1) NT 16) of the pass.  Per-row partial sums of
// each wave go to LDS and are added in a fixed order (deterministic).
constexpr int kWaves = 8;
constexpr int kWG = kWaves * 64;

// FP64: MFMA f64 16x16x4.  A lane l = (t: l%16, k: l/16), B lane l = (k: l/16, n: l%16),
// D lane l, register r = (t: 4r + l/16, n: l%16).
template <int NT, typename TO>
__device__ __forceinline__ void contract_f64_body(int blk, const double* __restrict__ phi, int ld,
                                                  const double* __restrict__ M, int nbf_pad, int npts,
                                                  double scale, TO* __restrict__ out, int64_t v0) {
  // M is upper triangular (M = D + 2U of the symmetric P), so column tile c
  // (columns 16c .. 16c+15) only needs the K stages s <= c (16 k values each).
  // Wave w owns the column tiles ct0 + w + 8j of a pass: cyclic, so every
  // wave keeps about the same number of live tiles while the triangle narrows.
  constexpr int TM = 64, KS = 16, PAD = 16;
  __shared__ double As[2][KS][TM + PAD];
  __shared__ double red[kWaves][TM];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = lane & 15, lk = lane >> 4;
  const int row0 = blk * TM;
  for (int e = threadIdx.x; e < kWaves * TM; e += kWG) red[e / TM][e % TM] = 0.0;
  const int nks = nbf_pad / KS;
  for (int c0 = 0; c0 < nbf_pad; c0 += kWaves * NT * 16) {
    const int ct0 = c0 / 16;
    const int nks_pass = min(nks, ct0 + kWaves * NT);
    int tile[NT];
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) tile[ct] = ct0 + wave + kWaves * ct;
    d4 acc[4][NT];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) acc[rt][ct] = d4{0, 0, 0, 0};
    // stage 0
    double pre[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = threadIdx.x + q * kWG, kk = e / TM, t = e % TM;
      pre[q] = phi[(int64_t)kk * ld + row0 + t];
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = threadIdx.x + q * kWG;
      As[0][e / TM][e % TM] = pre[q];
    }
    double bq[NT];
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
      const int n = tile[ct] * 16 + li;
      bq[ct] = M[n < nbf_pad ? (int64_t)lk * nbf_pad + n : 0];
    }
    __syncthreads();
    for (int s = 0; s < nks_pass; ++s) {
      const int k0 = s * KS;
      const bool more = s + 1 < nks_pass;
      if (more) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int e = threadIdx.x + q * kWG, kk = e / TM, t = e % TM;
          pre[q] = phi[(int64_t)(k0 + KS + kk) * ld + row0 + t];
        }
      }
      const double(*A)[TM + PAD] = As[s & 1];
#pragma unroll
      for (int ks = 0; ks < KS; ks += 4) {
        // prefetch the M operands of the next k-substep (possibly in the next stage)
        const int kn = k0 + ks + 4;
        // Unconditional loads (a branch around a load would leave the wait counts
        // unknown after it): a tile past the matrix or past its last stage reads
        // element 0 instead, and that value is never used.
        double bn[NT];
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
          const int n = tile[ct] * 16 + li;
          const bool live = n < nbf_pad && kn < 16 * (tile[ct] + 1) && kn < nbf_pad;
          bn[ct] = M[live ? (int64_t)(kn + lk) * nbf_pad + n : 0];
        }
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
          const double a = A[ks + lk][rt * 16 + li];
#pragma unroll
          for (int ct = 0; ct < NT; ++ct)
            if (tile[ct] >= s && tile[ct] * 16 < nbf_pad)
              acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bq[ct], acc[rt][ct], 0, 0, 0);
        }
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) bq[ct] = bn[ct];
      }
      if (more) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int e = threadIdx.x + q * kWG;
          As[(s + 1) & 1][e / TM][e % TM] = pre[q];
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      double part[4] = {0, 0, 0, 0};
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) {
        const int n = tile[ct] * 16 + li;
        if (tile[ct] * 16 < nbf_pad) {
          const double* ph = phi + (int64_t)n * ld + row0 + rt * 16 + lk;
#pragma unroll
          for (int r = 0; r < 4; ++r) part[r] += acc[rt][ct][r] * ph[4 * r];
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double v = sum16(part[r]);
        if (li == 0) red[wave][rt * 16 + 4 * r + lk] += v;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < TM) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) v += red[w][threadIdx.x];
    const int t = row0 + threadIdx.x;
    if (t < npts) store_rho<TO>(out, v0 + t, v * scale);
  }
}

// Barrier-free variant: every wave loads its own A fragments (the 64-row basis
// tile, shared by the 8 waves through L1/L2) and P fragments one k-step ahead;
// waves never wait for each other until the final row reduction.
template <int NT, typename TO>
__device__ __forceinline__ void contract_f64_direct_body(int blk, const double* __restrict__ phi, int ld,
                                                         const double* __restrict__ P, int nbf_pad, int npts,
                                                         double scale, TO* __restrict__ out, int64_t v0) {
  constexpr int TM = 64;
  __shared__ double red[kWaves][TM];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int row0 = blk * TM;
  for (int e = threadIdx.x; e < kWaves * TM; e += kWG) red[e / TM][e % TM] = 0.0;
  __syncthreads();
  const double* arow = phi + (int64_t)lk * ld + row0 + li;
  for (int c0 = 0; c0 < nbf_pad; c0 += kWaves * NT * 16) {
    const int wc0 = c0 + wave * NT * 16;
    if (wc0 >= nbf_pad) continue;
    d4 acc[4][NT];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) acc[rt][ct] = d4{0, 0, 0, 0};
    double aq[4], bq[NT];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) aq[rt] = arow[rt * 16];
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
      const int n = wc0 + ct * 16 + li;
      bq[ct] = n < nbf_pad ? P[(int64_t)lk * nbf_pad + n] : 0.0;
    }
    for (int k0 = 0; k0 < nbf_pad; k0 += 4) {
      const int kn = k0 + 4;
      double an[4], bn[NT];
      if (kn < nbf_pad) {
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) an[rt] = arow[(int64_t)kn * ld + rt * 16];
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
          const int n = wc0 + ct * 16 + li;
          bn[ct] = n < nbf_pad ? P[(int64_t)(kn + lk) * nbf_pad + n] : 0.0;
        }
      } else {
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) an[rt] = 0.0;
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) bn[ct] = 0.0;
      }
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct)
          acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(aq[rt], bq[ct], acc[rt][ct], 0, 0, 0);
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) aq[rt] = an[rt];
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) bq[ct] = bn[ct];
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      double part[4] = {0, 0, 0, 0};
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) {
        const int n = wc0 + ct * 16 + li;
        if (n < nbf_pad) {
          const double* ph = phi + (int64_t)n * ld + row0 + rt * 16 + lk;
#pragma unroll
          for (int r = 0; r < 4; ++r) part[r] += acc[rt][ct][r] * ph[4 * r];
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double v = sum16(part[r]);
        if (li == 0) red[wave][rt * 16 + 4 * r + lk] += v;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < TM) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) v += red[w][threadIdx.x];
    const int t = row0 + threadIdx.x;
    if (t < npts) store_rho<TO>(out, v0 + t, v * scale);
  }
}

template <int NT, typename TO>
__global__ __launch_bounds__(kWG) void k_contract_f64_direct(const double* __restrict__ phi, int ld,
                                                             const double* __restrict__ P, int nbf_pad, int npts,
                                                             double scale, TO* __restrict__ out, int64_t v0) {
  contract_f64_direct_body<NT, TO>(blockIdx.x, phi, ld, P, nbf_pad, npts, scale, out, v0);
}

template <int NT, typename TO>
__global__ __launch_bounds__(kWG) void k_contract_f64(const double* __restrict__ phi, int ld,
                                                      const double* __restrict__ P, int nbf_pad, int npts,
                                                      double scale, TO* __restrict__ out, int64_t v0) {
  contract_f64_body<NT, TO>(blockIdx.x, phi, ld, P, nbf_pad, npts, scale, out, v0);
}

template <typename T16>
struct Vec8;
template <>
struct Vec8<_Float16> {
  typedef h8 type;
  static __device__ __forceinline__ f4 mma(const h8& a, const h8& b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct Vec8<__bf16> {
  typedef b8 type;
  static __device__ __forceinline__ f4 mma(const b8& a, const b8& b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

// Split precision: A = pscale_t Phi[t][k] is rounded to 16-bit hi + lo while it is
// staged into LDS (once per workgroup, shared by the 8 waves), then
// X = Ahi Bhi + Ahi Blo + Alo Bhi (FP32 accumulate), and
// rho_t = scale / pscale_t * sum_n Phi[t][n] X[t][n] with the FP32 Phi (row-dot in FP64).
// MFMA 16x16x32: A lane l = (t: l%16, k: 8(l/16)+0..7), B likewise, D lane l, reg r = (t: 4(l/16)+r, n: l%16).
// Staging: thread (row, k-group) loads 8 values of one row (each load instruction covers 64
// consecutive rows, 256 B) and writes one 16-byte hi and lo word to LDS.
// ---------------------------------------------------------------------------
// Tiled FP64 contraction (variant 2).  A workgroup of 8 waves owns a tile of
// 128 points x 128 columns (column block cb) and forms X = Phi M over
// k < min(nbf_pad, 128 (cb + 1)) only (M = D + 2U is upper triangular); waves
// are arranged 4 x 2 and own 32 x 64 sub-tiles (8 f64 16x16x4 MFMA tiles).
// Both operands go through double-buffered LDS in stages of 16 k; the global
// loads of stage s+1 are issued before stage s is computed.  The workgroup
// writes the partial row-dot sum_{n in block} Phi_tn X_tn of its 128 points to
// part[cb][t]; k_tiles_finish adds the blocks in a fixed order.  Small
// accumulators and 72 KB of LDS let two workgroups share a CU.
// Workgroups are numbered so that the column blocks of one row tile are
// consecutive on the same XCD (workgroup b runs on XCD b mod 8) and share its
// L2 copy of the row tile's basis values; the heaviest block goes first.
constexpr int kTT = 128, kTKS = 16, kTPad = 8, kTWG = 512;

__device__ __forceinline__ bool tile_coords(int b, int nrt, int ncb, int& rt, int& cb) {
  const int xcd = b & 7, slot = b >> 3;
  const int rl = slot / ncb;
  cb = ncb - 1 - (slot - rl * ncb);
  rt = rl * 8 + xcd;
  return rt < nrt;
}

inline int tile_grid(int nrt, int ncb) { return (nrt + 7) / 8 * 8 * ncb; }

// LDS DMA of 16 bytes per lane: global_load_lds_dwordx4 writes lane i's 16 bytes
// at M0 + 16 i, so a wave fills 1 KB of LDS starting at lds_wave_base (uniform)
// without passing through registers.  Issued from inline assembly so that the
// compiler does not add a wait for every outstanding load before each LDS read;
// callers wait with an explicit s_waitcnt vmcnt before the barrier that
// publishes the data.  No other code in these kernels uses M0.
__device__ __forceinline__ void lds_dma16(const void* gsrc, const void* lds_wave_base) {
  const unsigned l = __builtin_amdgcn_readfirstlane(
      (unsigned)(uintptr_t)(__attribute__((address_space(3))) const void*)lds_wave_base);
  asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(l), "v"(gsrc) : "memory");
}

// saddr form of lds_dma16: wave-uniform 64-bit base (SGPRs) plus a 32-bit byte offset per lane;
// lds_addr is the wave's LDS byte address (M0).
__device__ __forceinline__ void lds_dma16_s(const void* sbase, unsigned voff, unsigned lds_addr) {
  asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, %2" ::"s"(lds_addr), "v"(voff), "s"(sbase) : "memory");
}

// k per stage of the DMA-fed FP64 tiles (ring of 32 / kF64Ks stages).  Stages of 8 k in a ring
// of 4 (one more stage of lookahead, twice the barriers) measured 4% slower than 16 k in a ring of 2.
constexpr int kF64Ks = 16;

// Row-dot of the transposed FP64 tiles (M was the first MFMA operand): lane
// (li, lk) register q of tile (r, c) is X[t = li][n = lk + 4 q]; 16 lanes read 16
// consecutive points of one function (128 contiguous bytes); the 4 lane groups
// that share a point are added by two butterflies.
__device__ __forceinline__ void f64_rowdot(const d4 (&acc)[2][4], const double* __restrict__ phi, int ld, int np,
                                           int row0, int c0, int wr, int wc, int li, int lk, double (*red)[kTT]) {
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int tl = wr * 32 + r * 16 + li;
    double p = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int n = c0 + wc * 64 + c * 16 + lk + 4 * q;   // c0 < 0 only for right-aligned blocks
        if (n >= 0 && n < np) p += acc[r][c][q] * phi[(int64_t)n * ld + row0 + tl];
      }
    }
    p += __shfl_xor(p, 16);
    p += __shfl_xor(p, 32);
    if (lk == 0) red[wc][tl] = p;
  }
}

__device__ __forceinline__ void tiled_f64_body(int rt, int cb, const double* __restrict__ phi, int ld,
                                               const double* __restrict__ M, int np, double* __restrict__ part,
                                               int ldp) {
  __shared__ double As[2][kTKS][kTT + kTPad];
  __shared__ double Bs[2][kTKS][kTT + kTPad];
  __shared__ double red[2][kTT];
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  const int row0 = rt * kTT, c0 = cb * kTT;
  const int nst = min(np, c0 + kTT) / kTKS;
  constexpr int Q = kTKS * kTT / kTWG;          // 4 values of each operand per thread and stage
  constexpr int KQ = kTWG / kTT;                // k rows covered by one q step
  const int kk0 = t / kTT, ii = t % kTT;
  // per-thread offsets; the per-stage and per-q parts of the addresses are uniform
  const unsigned aoff = (unsigned)(kk0 * ld + ii);
  const unsigned boff = (unsigned)(kk0 * np + min(c0 + ii, np - 1));   // columns past the matrix are never used
  const double* abase = phi + row0;
  double pa[Q], pb[Q];
  auto load = [&](int st) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int k = st * kTKS + KQ * q;
      pa[q] = (abase + (int64_t)k * ld)[aoff];
      pb[q] = (M + (int64_t)k * np)[boff];
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      As[buf][kk0 + KQ * q][ii] = pa[q];
      Bs[buf][kk0 + KQ * q][ii] = pb[q];
    }
  };
  d4 acc[2][4];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = d4{0, 0, 0, 0};
  load(0);
  store(0);
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const bool more = st + 1 < nst;
    if (more) load(st + 1);
    const int buf = st & 1;
#pragma unroll
    for (int ks = 0; ks < kTKS; ks += 4) {
      double a[2], b[4];
#pragma unroll
      for (int r = 0; r < 2; ++r) a[r] = As[buf][ks + lk][wr * 32 + r * 16 + li];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = Bs[buf][ks + lk][wc * 64 + c * 16 + li];
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c)   // M first: tile (r, c) holds X^T, lane (li, lk) register q = X[t = li][n = lk + 4 q]
          acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[c], a[r], acc[r][c], 0, 0, 0);
    }
    if (more) store((st + 1) & 1);
    __syncthreads();
  }
  f64_rowdot(acc, phi, ld, np, row0, c0, wr, wc, li, lk, red);
  __syncthreads();
  if (t < kTT) part[(int64_t)cb * ldp + row0 + t] = red[0][t] + red[1][t];
}

// Tiled FP64 contraction fed by LDS DMA (variant 4): the 128 x 128 workgroup
// tiles and per-tile MFMA order of tiled_f64_body, with waves of 16 rows x all
// 128 columns (so that every wave skips the same tiles of the triangle and the
// diagonal stages have compile-time tiles, with no runtime tile selection and no
// cross-wave row-dot reduction); both operands go global -> LDS by
// global_load_lds_dwordx4 (lane i of a wave brings points / columns 2i, 2i+1 of
// one k row, so a wave fills one padded LDS row) into a two-stage ring, issued
// right after the barrier that opens the previous stage.  The loads come from
// inline assembly (lds_dma16), the wait before each barrier is explicit.
// KS k per stage in a ring of NB stages (NB KS = 32: the same 72 KB of LDS); stage st + NB - 1 is
// requested right after the barrier that opens stage st.  The row-dot is summed in the K loop from
// the LDS copy of the basis values of each diagonal stage (below), not re-read from global memory.
template <int KS, int NB>
__device__ __forceinline__ void tiled_f64_dma_body(int rt, int cb, const double* __restrict__ phi, int ld,
                                                   const double* __restrict__ M, int np, int nbf, int ncb,
                                                   double* __restrict__ part, int ldp,
                                                   const uint32_t* __restrict__ amask = nullptr, int nw = 0) {
  static_assert(KS % 4 == 0 && (2 * KS) % 8 == 0, "stages of whole MFMA k-steps, rows split over 8 waves");
  static_assert(KS == 16, "the diagonal stages below are the block's 8 tiles of 16 columns");
  constexpr int RPW = 2 * KS / 8;   // rows of one operand fetched per wave and stage
  __shared__ __align__(16) double ring[NB][2][KS][kTT + kTPad];
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int li = lane & 15, lk = lane >> 4;
  // Column blocks are right-aligned on ne = nbf rounded up to 16: block cb covers columns
  // c0 .. c0 + 127 with c0 = 128 cb - shift, so only block 0 (the cheapest, k < c0 + 128) holds
  // padding, as columns below 0; rows and columns of M from nbf on are zero and are not visited.
  const int ne = (nbf + 15) / 16 * 16;
  const int row0 = rt * kTT, c0 = cb * kTT - (ncb * kTT - ne);
  const int jd = c0 / KS;   // c0 is a multiple of 16: stage jd + j is the block's diagonal stage j
  // wave w brings k rows RPW (w % 4) .. RPW (w % 4) + RPW - 1 of operand w / 4 (0: basis values, 1: M)
  const int op = wave >> 2, k0 = RPW * (wave & 3);
  const double* src = op == 0 ? phi + row0 + 2 * lane : M + min(max(c0 + 2 * lane, 0), np - 2);   // clamped: unused
  const int64_t lds_ld = op == 0 ? ld : np;

  // Wave w owns rows 16 w .. 16 w + 15 and all 8 tiles of 16 columns of the block, so every wave
  // drops the same tiles out of the triangle and completes the same tile in each diagonal stage.
  // Groups of 16 functions whose basis values are all zero on this row tile (activity bit clear,
  // k_tile_mask) add only zero products.  A block with no active group among its own columns writes
  // zeros before its first load; otherwise the stages below the block's diagonal run only for the
  // active groups (their list is compacted into LDS), loads included, and the eight diagonal stages
  // always run.  Every decision is uniform per workgroup.
  const uint32_t* __restrict__ mk = amask ? amask + (int64_t)rt * nw : nullptr;
  auto act = [&](int st) __attribute__((always_inline)) -> bool {
    return !mk || ((mk[st >> 5] >> (st & 31)) & 1u);
  };
  if (mk) {
    bool any = false;
    for (int j = 0; j < 8; ++j) any = any || (jd + j >= 0 && act(jd + j));
    if (!any) {
      if (lk == 0) part[(int64_t)cb * ldp + row0 + 16 * wave + li] = 0.0;
      return;
    }
  }
  constexpr int kSeqMax = 256;   // below-diagonal stages listed in LDS (16 kSeqMax functions)
  __shared__ int seq[kSeqMax + 1];
  const bool compact = mk && jd <= kSeqMax;
  int nloop = jd > 0 ? jd : 0;
  if (compact) {
    if (wave == 0) {   // wave 0 lists the active stages below the diagonal in order
      int cnt = 0;
      for (int b = 0; b < jd; b += 64) {
        const int st = b + lane;
        const bool a = st < jd && act(st);
        const uint64_t bal = __ballot(a);
        if (a) seq[cnt + __popcll(bal & ((1ull << lane) - 1ull))] = st;
        cnt += __popcll(bal);
      }
      if (lane == 0) seq[kSeqMax] = cnt;
    }
    __syncthreads();
    nloop = __builtin_amdgcn_readfirstlane(seq[kSeqMax]);
  }
  // the ring's stage sequence: position p < nloop is a listed stage below the diagonal, then the
  // diagonal stages jd + j0 .. jd + 7 (j0 > 0 only for block 0, whose first columns are padding)
  const int j0 = jd < 0 ? -jd : 0;
  const int npos = nloop + 8 - j0;
  auto stage_of = [&](int q) __attribute__((always_inline)) -> int {
    return q < nloop ? (compact ? __builtin_amdgcn_readfirstlane(seq[q]) : q) : jd + j0 + (q - nloop);
  };
  auto issue = [&](int q) __attribute__((always_inline)) {
    const int st = stage_of(q);
#pragma unroll
    for (int j = 0; j < RPW; ++j) {
      const int k = k0 + j;
      lds_dma16(src + (int64_t)(st * KS + k) * lds_ld, &ring[q % NB][op][k][0]);
    }
  };
  d4 acc[8];
  double pr = 0.0;   // this lane's row-dot partial sum (row 16 w + li)
#pragma unroll
  for (int c = 0; c < 8; ++c) acc[c] = d4{0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < NB - 1; ++q)
    if (q < npos) issue(q);
  // opens position q: this thread's loads of q are done once at most RPW (NB - 2) newer ones are
  // pending, everyone's after the barrier, and everyone is done with q - 1, whose slot is refilled
  auto open = [&](int q) __attribute__((always_inline)) {
    const int newer = min(NB - 2, npos - 1 - q);
    if (RPW == 2 && newer >= 2) asm volatile("s_waitcnt vmcnt(4)" ::: "memory");
    else if (RPW == 2 && newer == 1) asm volatile("s_waitcnt vmcnt(2)" ::: "memory");
    else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (q + NB - 1 < npos) issue(q + NB - 1);
  };
  // MFMAs of one stage for the tiles first .. 7 (transposed tiles: M first)
  auto body = [&](int buf, auto first) __attribute__((always_inline)) {
    constexpr int cs = decltype(first)::value;
#pragma unroll
    for (int ks = 0; ks < KS; ks += 4) {
      double b[8];
      const double a = ring[buf][0][ks + lk][16 * wave + li];
#pragma unroll
      for (int c = cs; c < 8; ++c) b[c] = ring[buf][1][ks + lk][c * 16 + li];
#pragma unroll
      for (int c = cs; c < 8; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[c], a, acc[c], 0, 0, 0);
    }
  };
  // Row-dot of the tile a diagonal stage completes: tile c (columns n0 = c0 + 16 c ..) gets its last
  // products in the stage with KS st = n0, whose A operand in the ring slot holds the basis values of
  // exactly its columns (k = n - n0 = lk + 4 q for lane (li, lk), register q).  Tiles are summed in
  // column order.
  auto rowdot = [&](int buf, auto tile) __attribute__((always_inline)) {
    constexpr int c = decltype(tile)::value;
#pragma unroll
    for (int q = 0; q < 4; ++q) pr = fma(acc[c][q], ring[buf][0][lk + 4 * q][16 * wave + li], pr);
  };
  // stages below the block's diagonal (none in block 0): every tile, no row-dot
  for (int q = 0; q < nloop; ++q) {
    open(q);
    if (compact || act(q)) body(q % NB, std::integral_constant<int, 0>());
  }
  // The eight diagonal stages.  M is zero below its diagonal: in diagonal stage j tile c still needs
  // products while KS j <= 16 c + 15, i.e. c >= j, and tile j gets its last ones.  Block 0 starts at
  // stage 0 = diagonal stage -jd; its tiles below column 0 are those below -jd, skipped too.
  auto diag = [&](auto jc) __attribute__((always_inline)) {
    constexpr int J = decltype(jc)::value;
    if (J < j0) return;
    const int q = nloop + J - j0;
    open(q);
    body(q % NB, jc);
    rowdot(q % NB, jc);
  };
  diag(std::integral_constant<int, 0>());
  diag(std::integral_constant<int, 1>());
  diag(std::integral_constant<int, 2>());
  diag(std::integral_constant<int, 3>());
  diag(std::integral_constant<int, 4>());
  diag(std::integral_constant<int, 5>());
  diag(std::integral_constant<int, 6>());
  diag(std::integral_constant<int, 7>());
  pr += __shfl_xor(pr, 16);
  pr += __shfl_xor(pr, 32);
  if (lk == 0) part[(int64_t)cb * ldp + row0 + 16 * wave + li] = pr;
}

// Tiled 16-bit split contraction: the tiling, triangle, workgroup order and
// partial sums of the FP64 tiled kernel, with stages of 32 k.  The M planes are
// copied as stored (K8-interleaved); X = Ah Bh + Ah Bl + Al Bh in FP32.
// PRE false: FP32 basis values, scaled per point and rounded to hi + lo while
//   they are staged into LDS; row-dot with the FP32 values in FP64.
// PRE true: values already scaled by the molecule's S and split (k_phi_split:
//   hi plane, then lo plane, np ld values each, K8-interleaved), staged as plain
//   16-byte copies; the MFMAs take M as their first operand, so each 16 x 16
//   tile of X comes out transposed (lane (li, lk), register q = X[t = li][n =
//   4 lk + q]) and the row-dot reads its 4 consecutive functions of one point
//   as a single 8-byte word per plane (presplit_rowdot).  The finishing pass
//   divides by S^2 (folded into its scale) instead of the per-point pscale.
template <typename T16, int RT, int CT, int TRW, int TCW, int WC>
__device__ __forceinline__ void presplit_rowdot(const f4 (&acc)[RT][CT], const T16* __restrict__ a_hi,
                                                const T16* __restrict__ a_lo, int ld, int np, int row0, int c0,
                                                int wr, int wc, int li, int lk, double (*red)[kTT]) {
  typedef _Float16 h4 __attribute__((ext_vector_type(4)));
  typedef __bf16 b4 __attribute__((ext_vector_type(4)));
  typedef typename std::conditional<std::is_same<T16, _Float16>::value, h4, b4>::type V4;
#pragma unroll
  for (int r = 0; r < RT; ++r) {
    const int tl = wr * TRW + r * 16 + li;
    double p = 0.0;
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const int n0 = c0 + wc * TCW + c * 16 + 4 * lk;   // a multiple of 4: n0 .. n0+3 share one 8-group
      if (n0 >= 0 && n0 < np) {   // c0 < 0 only for right-aligned blocks
        // element (n, t) of a plane sits at ((n / 8) ld + t) 8 + n % 8
        const int64_t o = ((int64_t)(n0 >> 3) * ld + row0 + tl) * 8 + (n0 & 7);
        const V4 h = __builtin_bit_cast(V4, *(const uint2*)(a_hi + o));
        const V4 l = __builtin_bit_cast(V4, *(const uint2*)(a_lo + o));
#pragma unroll
        for (int q = 0; q < 4; ++q) p += (double)(acc[r][c][q] * ((float)h[q] + (float)l[q]));
      }
    }
    p += __shfl_xor(p, 16);
    p += __shfl_xor(p, 32);
    if (lk == 0) red[wc][tl] = p;
  }
}

template <typename T16, int WR, int WC, bool PRE = false>
__device__ __forceinline__ void tiled_split_body(int rt, int cb, const void* __restrict__ phiv, int ld,
                                                 const float* __restrict__ pscale, const T16* __restrict__ p_hi,
                                                 const T16* __restrict__ p_lo, int np, double* __restrict__ part,
                                                 int ldp) {
  const float* __restrict__ phi = (const float*)phiv;
  const T16* __restrict__ a_hi = (const T16*)phiv;
  const T16* __restrict__ a_lo = a_hi + (int64_t)np * ld;
  // WR x WC waves; wave (wr, wc) owns (128/WR) x (128/WC) outputs
  typedef typename Vec8<T16>::type V8;
  constexpr int NTH = 64 * WR * WC;
  constexpr int TRW = kTT / WR, TCW = kTT / WC, RT = TRW / 16, CT = TCW / 16;
  constexpr int NA = 4 * kTT / NTH;        // (row, k group) pairs staged per thread and stage
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  __shared__ u32x4 Ah[2][4][kTT], Al[2][4][kTT], Bh[2][4][kTT], Bl[2][4][kTT];
  __shared__ double red[WC][kTT];
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wr = wave / WC, wc = wave % WC;
  const int li = lane & 15, lk = lane >> 4;
  const int row0 = rt * kTT, c0 = cb * kTT;
  const int nst = min(np, c0 + kTT) / 32;
  // staging: pair e = t + NTH j -> row / column ii = e % 128 of the tile and k group g = e / 128
  float sc[NA];
  const float* asrc[NA];
  int64_t bcol[NA];
#pragma unroll
  for (int j = 0; j < NA; ++j) {
    const int e = t + NTH * j, ii = e % kTT, g = e / kTT;
    if (!PRE) sc[j] = pscale[row0 + ii];
    asrc[j] = phi + (int64_t)(8 * g) * ld + row0 + ii;
    bcol[j] = min(c0 + ii, np - 1);   // columns past the matrix are never used
  }
  float pa[PRE ? 1 : NA * 8];
  u32x4 pah[NA], pal[NA], pbh[NA], pbl[NA];
  auto load = [&](int st) __attribute__((always_inline)) {
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      const int e = t + NTH * j, ii = e % kTT, g = e / kTT;
      if (PRE) {
        const int64_t oa = ((int64_t)(st * 4 + g) * ld + row0 + ii) * 8;
        pah[j] = *(const u32x4*)(a_hi + oa);
        pal[j] = *(const u32x4*)(a_lo + oa);
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) pa[8 * j + q] = asrc[j][(int64_t)(st * 32 + q) * ld];
      }
      const int64_t o = ((int64_t)(st * 4 + g) * np + bcol[j]) * 8;
      pbh[j] = *(const u32x4*)(p_hi + o);
      pbl[j] = *(const u32x4*)(p_lo + o);
    }
  };
  auto store = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      const int e = t + NTH * j, ii = e % kTT, g = e / kTT;
      if (PRE) {
        Ah[buf][g][ii] = pah[j];
        Al[buf][g][ii] = pal[j];
      } else {
        V8 h, l;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const float x = pa[8 * j + q] * sc[j];
          const T16 hq = (T16)x;
          h[q] = hq;
          l[q] = (T16)(x - (float)hq);
        }
        Ah[buf][g][ii] = __builtin_bit_cast(u32x4, h);
        Al[buf][g][ii] = __builtin_bit_cast(u32x4, l);
      }
      Bh[buf][g][ii] = pbh[j];
      Bl[buf][g][ii] = pbl[j];
    }
  };
  f4 acc[RT][CT];
#pragma unroll
  for (int r = 0; r < RT; ++r)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[r][c] = f4{0, 0, 0, 0};
  load(0);
  store(0);
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const bool more = st + 1 < nst;
    if (more) load(st + 1);
    const int buf = st & 1;
    V8 ah[RT], al[RT], bh[CT], bl[CT];
#pragma unroll
    for (int r = 0; r < RT; ++r) {
      ah[r] = __builtin_bit_cast(V8, Ah[buf][lk][wr * TRW + r * 16 + li]);
      al[r] = __builtin_bit_cast(V8, Al[buf][lk][wr * TRW + r * 16 + li]);
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      bh[c] = __builtin_bit_cast(V8, Bh[buf][lk][wc * TCW + c * 16 + li]);
      bl[c] = __builtin_bit_cast(V8, Bl[buf][lk][wc * TCW + c * 16 + li]);
    }
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        if (PRE) {   // transposed tiles: M first
          acc[r][c] = Vec8<T16>::mma(bh[c], al[r], acc[r][c]);
          acc[r][c] = Vec8<T16>::mma(bl[c], ah[r], acc[r][c]);
          acc[r][c] = Vec8<T16>::mma(bh[c], ah[r], acc[r][c]);
        } else {
          acc[r][c] = Vec8<T16>::mma(al[r], bh[c], acc[r][c]);
          acc[r][c] = Vec8<T16>::mma(ah[r], bl[c], acc[r][c]);
          acc[r][c] = Vec8<T16>::mma(ah[r], bh[c], acc[r][c]);
        }
      }
    if (more) store((st + 1) & 1);
    __syncthreads();
  }
  if (PRE) {
    presplit_rowdot<T16, RT, CT, TRW, TCW, WC>(acc, a_hi, a_lo, ld, np, row0, c0, wr, wc, li, lk, red);
  } else {
#pragma unroll
    for (int r = 0; r < RT; ++r) {
      double p4[4] = {0, 0, 0, 0};
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        const int n = c0 + wc * TCW + c * 16 + li;
        if (n < np) {
          const float4 v = *(const float4*)(phi + (int64_t)n * ld + row0 + wr * TRW + r * 16 + 4 * lk);
          p4[0] += (double)(acc[r][c][0] * v.x);
          p4[1] += (double)(acc[r][c][1] * v.y);
          p4[2] += (double)(acc[r][c][2] * v.z);
          p4[3] += (double)(acc[r][c][3] * v.w);
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double v = sum16(p4[q]);
        if (li == 0) red[wc][wr * TRW + r * 16 + 4 * lk + q] = v;
      }
    }
  }
  __syncthreads();
  for (int e = t; e < kTT; e += NTH) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < WC; ++w) v += red[w][e];
    part[(int64_t)cb * ldp + row0 + e] = v;
  }
}

template <typename T16, int WR, int WC, bool PRE = false>
__global__ __launch_bounds__(64 * WR * WC) void k_contract_split_tiles(const void* __restrict__ phi, int ld,
                                                                      const float* __restrict__ pscale,
                                                                      const T16* __restrict__ p_hi,
                                                                      const T16* __restrict__ p_lo, int np, int nrt,
                                                                      int ncb, double* __restrict__ part) {
  int rt, cb;
  if (!tile_coords(blockIdx.x, nrt, ncb, rt, cb)) return;
  tiled_split_body<T16, WR, WC, PRE>(rt, cb, phi, ld, pscale, p_hi, p_lo, np, part, ld);
}

// Tiled split contraction on pre-split basis values with both operands
// fetched by LDS DMA (global_load_lds_dwordx4: each lane's 16 bytes land at
// M0 + 16 lane, so a wave fills 1 KB of LDS without touching registers) into a
// two-stage ring of 32 k: stage st + 1 is requested right after the barrier
// that opens stage st, so its loads run behind stage st's MFMAs while the
// workgroup's registers stay free (two workgroups share a CU).  The ring slot
// written at stage st was last read at stage st - 1, which every wave has
// finished once it passes stage st's barrier.  The loads are issued from inline
// assembly (M0 holds the LDS address) so that the compiler does not wait for
// them before each LDS read; the explicit s_waitcnt vmcnt before the barrier
// waits for them instead.  No other code in this kernel uses M0.  The transposed
// MFMA order per tile is that of tiled_split_body<PRE>; the row-dot sums are
// formed here in the K loop, from the LDS copy of the basis values (below).
//
// The stage loop is split so that its code has no runtime tile selection: the
// stages below the block's diagonal run every tile with no row-dot, and the four
// diagonal stages are unrolled with compile-time first tile and row-dot tile.
// The two waves of a SIMD take interleaved 16-column tiles, so both drop out of
// the triangle alike and no diagonal stage branches on the wave.  (A version that
// selected tiles at run time spent 43 VALU instructions per 24 MFMAs in a full
// stage, 16 of them register moves; this one spends 4, with 80 VGPRs.)  Measured
// and dropped: persistent workgroups that prefetch the next tile's first stage
// (slower: the row tiles' column blocks drift apart); a three- or four-slot ring
// at one workgroup per CU (two waves per SIMD: 35% slower); warming L2 with a
// dummy DMA of stage st + 2 (10% slower).

template <typename T16>
__device__ __forceinline__ void tiled_dma_body(int rt, int cb, const T16* __restrict__ a_hi, int ld,
                                               const T16* __restrict__ p_hi, const T16* __restrict__ p_lo, int np,
                                               int nbf, int ncb, double* __restrict__ part, int ldp) {
  typedef typename Vec8<T16>::type V8;
  constexpr int WR = 4, WC = 2, TRW = kTT / WR, TCW = kTT / WC, RT = TRW / 16, CT = TCW / 16;
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  // ring slot: 4 operand planes (A hi, A lo, B hi, B lo) of [4 k groups][128] 16-byte words
  __shared__ u32x4 ring[2][4][4][kTT];
  __shared__ double red[WC][kTT];
  const char* __restrict__ a_hi_b = (const char*)a_hi;
  const char* __restrict__ a_lo_b = (const char*)(a_hi + (int64_t)np * ld);
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wr = wave & 3, wc = wave >> 2;   // waves w, w + 4 share a SIMD
  const int li = lane & 15, lk = lane >> 4;
  // column blocks right-aligned on ne = nbf rounded up to 32 (as in tiled_f64_dma_body): only
  // block 0 holds padding, as columns below 0, skipped like tiles below the diagonal
  const int ne = (nbf + 31) / 32 * 32;
  const int row0 = rt * kTT, c0 = cb * kTT - (ncb * kTT - ne);
  const int nst = (c0 + kTT) / 32;
  const int jd = c0 / 32;   // c0 is a multiple of 32: stage jd + j is the block's diagonal stage j
  // thread t fetches word t of each plane: k group g = t / 128, row / column ii = t % 128.  The
  // stage's base address is wave-uniform (SGPRs, saddr form); the lane's byte offset within the
  // stage is at most 64 ld < 2^32 (a plane of basis values is at most 2 GiB).
  const int g = t / kTT, ii = t % kTT;
  const int bcol = min(max(c0 + ii, 0), np - 1);   // clamped columns are never used
  const unsigned oa = (unsigned)(g * ld + ii) * 16u;
  const unsigned ob = (unsigned)(g * np + bcol) * 16u;
  const unsigned lring = (unsigned)(uintptr_t)(__attribute__((address_space(3))) void*)&ring[0][0][0][0] + 1024u * wave;
  auto issue = [&](int st) __attribute__((always_inline)) {
    const unsigned l = lring + (unsigned)(st & 1) * (unsigned)sizeof(ring[0]);
    // 4 k groups of 16 bytes per stage; the tile's first row goes into the uniform base
    const int64_t sa = (int64_t)st * 64 * ld + (int64_t)row0 * 16, sb = (int64_t)st * 64 * np;
    lds_dma16_s(a_hi_b + sa, oa, l);
    lds_dma16_s(a_lo_b + sa, oa, l + (unsigned)sizeof(ring[0][0]));
    lds_dma16_s((const char*)p_hi + sb, ob, l + 2u * (unsigned)sizeof(ring[0][0]));
    lds_dma16_s((const char*)p_lo + sb, ob, l + 3u * (unsigned)sizeof(ring[0][0]));
  };
  f4 acc[RT][CT];
  double pr[RT];   // this lane's row-dot partial sums (rows wr TRW + 16 r + li)
  // Wave (wr, wc) owns rows wr TRW .. + TRW - 1 and the interleaved 16-column tiles 2 c + wc, c =
  // 0 .. CT - 1, of the block (columns c0 + 32 c + 16 wc ..), so that both waves of a SIMD drop
  // tiles out of the triangle alike.  MFMAs of one stage for the tiles first .. CT - 1:
  auto body = [&](int slot, auto first) __attribute__((always_inline)) {
    constexpr int cs = decltype(first)::value;
    V8 ah[RT], al[RT], bh[CT], bl[CT];
#pragma unroll
    for (int r = 0; r < RT; ++r) {
      ah[r] = __builtin_bit_cast(V8, ring[slot][0][lk][wr * TRW + r * 16 + li]);
      al[r] = __builtin_bit_cast(V8, ring[slot][1][lk][wr * TRW + r * 16 + li]);
    }
#pragma unroll
    for (int c = cs; c < CT; ++c) {
      bh[c] = __builtin_bit_cast(V8, ring[slot][2][lk][32 * c + 16 * wc + li]);
      bl[c] = __builtin_bit_cast(V8, ring[slot][3][lk][32 * c + 16 * wc + li]);
    }
    // transposed tiles (M first); each product runs over all tiles before the next, so consecutive
    // MFMAs are independent (a tile's three products keep their order)
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
      for (int c = cs; c < CT; ++c) acc[r][c] = Vec8<T16>::mma(bh[c], al[r], acc[r][c]);
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
      for (int c = cs; c < CT; ++c) acc[r][c] = Vec8<T16>::mma(bl[c], ah[r], acc[r][c]);
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
      for (int c = cs; c < CT; ++c) acc[r][c] = Vec8<T16>::mma(bh[c], ah[r], acc[r][c]);
  };
  // Row-dot of the tile a stage completes: the wave's tile c (columns n0 = c0 + 32 c + 16 wc ..)
  // gets its last products in stage (n0 + 15) / 32 = diagonal stage c, whose A operand in the ring
  // slot holds the basis values of exactly its columns (k = n).  Lane (li, lk), register q of a
  // tile is X[t = li][n = 4 lk + q]: the 4 values of one point sit in one 8-byte word of each
  // plane.  Each wave sums its tiles in column order.
  auto rowdot = [&](int slot, auto tile) __attribute__((always_inline)) {
    constexpr int c = decltype(tile)::value;
    typedef _Float16 h4 __attribute__((ext_vector_type(4)));
    typedef __bf16 b4 __attribute__((ext_vector_type(4)));
    typedef typename std::conditional<std::is_same<T16, _Float16>::value, h4, b4>::type V4;
    const int kg = 2 * wc + (lk >> 1);   // k group of n - 32 st = 16 wc + 4 lk + q
#pragma unroll
    for (int r = 0; r < RT; ++r) {
      const int row = wr * TRW + r * 16 + li;
      const V4 h = __builtin_bit_cast(V4, ((const uint2*)&ring[slot][0][kg][row])[lk & 1]);
      const V4 l = __builtin_bit_cast(V4, ((const uint2*)&ring[slot][1][kg][row])[lk & 1]);
#pragma unroll
      for (int q = 0; q < 4; ++q) pr[r] += (double)(acc[r][c][q] * ((float)h[q] + (float)l[q]));
    }
  };
#pragma unroll
  for (int r = 0; r < RT; ++r) {
    pr[r] = 0.0;
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[r][c] = f4{0, 0, 0, 0};
  }
  // opens stage st: this thread's loads of st have landed, every thread's have after the barrier,
  // and every wave is done with st - 1, whose slot receives st + 1
  auto open = [&](int st) __attribute__((always_inline)) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (st + 1 < nst) issue(st + 1);
  };
  issue(0);
  // stages below the block's diagonal (none in block 0): every tile, no row-dot
  for (int st = 0; st < jd; ++st) {
    open(st);
    body(st & 1, std::integral_constant<int, 0>());
  }
  // The four diagonal stages, with compile-time tiles.  M is zero below its diagonal: in diagonal
  // stage j, the wave's tile c (columns c0 + 32 c + 16 wc ..) still needs products while 32 j <
  // 32 c + 16 wc + 16, i.e. c >= j, and tile j gets its last ones.  Block 0 starts at stage 0 =
  // diagonal stage -jd; its tiles below column 0 are those below -jd, skipped too.
  auto diag = [&](auto jc) __attribute__((always_inline)) {
    constexpr int J = decltype(jc)::value;
    const int st = jd + J;
    if (st < 0) return;
    open(st);
    body(st & 1, jc);
    __builtin_amdgcn_sched_barrier(0);
    rowdot(st & 1, jc);
    __builtin_amdgcn_sched_barrier(0);
  };
  diag(std::integral_constant<int, 0>());
  diag(std::integral_constant<int, 1>());
  diag(std::integral_constant<int, 2>());
  diag(std::integral_constant<int, 3>());
#pragma unroll
  for (int r = 0; r < RT; ++r) {
    double p = pr[r];
    p += __shfl_xor(p, 16);
    p += __shfl_xor(p, 32);
    if (lk == 0) red[wc][wr * TRW + r * 16 + li] = p;
  }
  __syncthreads();
  if (t < kTT) part[(int64_t)cb * ldp + row0 + t] = red[0][t] + red[1][t];
}

template <typename T16>
__global__ __launch_bounds__(512, 4) void k_contract_split_dma(const void* __restrict__ phi, int ld,
                                                            const T16* __restrict__ p_hi, const T16* __restrict__ p_lo,
                                                            int np, int nbf, int nrt, int ncb,
                                                            double* __restrict__ part) {
  int rt, cb;
  if (!tile_coords(blockIdx.x, nrt, ncb, rt, cb)) return;
  tiled_dma_body<T16>(rt, cb, (const T16*)phi, ld, p_hi, p_lo, np, nbf, ncb, part, ld);
}

// ---------------------------------------------------------------------------
// Fused split density (split variant 4): the basis values are
// evaluated inside the contraction kernel, stage by stage, straight into LDS,
// so they never travel through HBM (the two-kernel path writes and re-reads
// 4 bytes per function and point).
//
// A workgroup of 8 waves owns 128 points and one column group of 320 columns
// (20 tiles of 16), right-aligned on nbf rounded up to 32 like the DMA tiles
// (only group 0 holds padding, as columns below 0).  K runs in stages of 32
// over k < c0 + 320 (M = D + 2U is upper triangular).  The waves split by role:
//
//   * waves 4..7 (producers, one per SIMD) evaluate the basis values of the
//     next stage: wave 4 + g takes the 8 functions of k group g for all 128
//     points (two per lane), so the shells it visits are wave-uniform; a
//     per-molecule table gives the shell and m of the group's first function.
//     Values are scaled by S, split into 16-bit hi + lo with the roundings of
//     phi_split_body, and stored as one 16-byte word per plane and point into
//     the next slot of the basis-value ring.  The producers also bring the
//     next stage of M (hi and lo, 32 k x 320 columns) into the M ring, loading
//     it into registers before the evaluation and storing it after;
//   * waves 0..3 (consumers, one per SIMD) run the MFMAs of the current
//     stage: wave w owns the 5 tiles w, w + 4, .., w + 16 of the group (the
//     column tiles are dealt round-robin so that the tiles that drop out of
//     the triangle leave every wave evenly) and all 8 row tiles, 160 FP32
//     accumulators.  X = Mh Phl + Ml Phh + Mh Phh per tile (M first: the tile
//     holds X transposed), and the row-dot of a tile is summed in the stage
//     that completes it, from the basis values in the ring (as in
//     tiled_dma_body).
//
// One barrier per stage hands the two ring slots over; a producer and a
// consumer share each SIMD, so the evaluation (VALU, exp2) of stage st + 1
// runs beside the MFMAs of stage st.  With one producer wave per SIMD nothing
// hides its load latency, so the molecule's tables (nuclei, shells, primitives,
// function groups) are copied into LDS when the workgroup starts and every
// table read during the stages is an LDS read.
// LDS: 32 KB basis values + 80 KB M + 4 KB row-dot + up to 40 KB of tables.
//
// Measured on the 60-atom, 256^3 workload: 68 ms per grid against 27 ms for
// basis values written to memory plus the DMA tiles (variant 3).  The single
// producer wave per SIMD is the bound: its evaluation (LDS table reads in a
// dependent chain per shell, FP64 offsets, per-slot selection of the harmonic)
// takes about 3 us per stage, against well under 1 us of MFMAs; with the
// evaluation switched off the kernel takes 24 ms.  Kept as a variant (no
// basis-value buffer at all) and timed by the autotuner.
constexpr int kFP = 128;          // points per workgroup
constexpr int kFCT = 20;          // 16-column tiles per column group
constexpr int kFC = 16 * kFCT;    // columns per column group
constexpr int kFChunk = 1 << 24;  // points per launch (partial sums: 8 bytes x column groups each)
constexpr int kFTab = 40960;      // bytes of LDS for the molecule's tables

inline int fused_groups(int nbf) { return ((nbf + 31) / 32 * 32 + kFC - 1) / kFC; }

// Byte offsets of the molecule tables in LDS: shells (int4 atom, L, nprim, first primitive),
// nuclei (3 doubles), primitives (float2 alpha log2 e, coeff), function groups (int2), cutoffs.
struct FusedTab {
  int at, pr, kg, rc, bytes;
};
__host__ __device__ inline FusedTab fused_tab(int natoms, int nshell, int nprim, int nkg) {
  FusedTab t;
  t.at = 16 * nshell;
  t.pr = t.at + 24 * natoms;
  t.kg = t.pr + 8 * nprim;
  t.rc = t.kg + 8 * nkg;
  t.bytes = (t.rc + 4 * nshell + 15) / 16 * 16;
  return t;
}
inline bool fused_fits(const qm_mol* m) {
  return fused_tab(m->natoms, m->nshell, m->nprim, m->nbf_pad / 8).bytes <= kFTab;
}

// Solid-harmonic factor of function m of an L shell (m order 0, +1, -1, +2, -2): the
// factor that shell_values multiplies by the radial part, with the same operations.
__device__ __forceinline__ float harm_factor(int L, int m, float x, float y, float z) {
  if (L == 0) return (float)C00;
  if (L == 1) return m == 0 ? (float)C1 * z : (m == 1 ? (float)C1 * x : (float)C1 * y);
  switch (m) {
    case 0: return (float)C20 * d0_factor(x, y, z);
    case 1: return (float)C2A * x * z;
    case 2: return (float)C2A * z * y;
    case 3: return (float)C22 * ((x - y) * (x + y));
    default: return (float)C2A * x * y;
  }
}

__device__ __forceinline__ int cdiv_floor0(int a, int b) {   // max(0, ceil(a / b)) for b > 0
  return a <= 0 ? 0 : (a + b - 1) / b;
}

// Copies the molecule's tables into the workgroup's LDS (all 512 threads; caller synchronises).
__device__ __forceinline__ void fused_load_tables(unsigned char* tab, const MolView& m,
                                                  const float2* __restrict__ l2e, const int2* __restrict__ kgt,
                                                  int nkg) {
  const FusedTab o = fused_tab(m.natoms, m.nshell, m.nprim, nkg);
  for (int i = threadIdx.x; i < m.nshell; i += blockDim.x) {
    const DevShell sh = m.shells[i];
    ((int4*)tab)[i] = make_int4(sh.atom, sh.L, sh.nprim, sh.poff);
    ((float*)(tab + o.rc))[i] = sh.rcut2;
  }
  for (int i = threadIdx.x; i < m.natoms; i += blockDim.x) {
    const double4 a = m.atoms[i];
    double* d = (double*)(tab + o.at) + 3 * i;
    d[0] = a.x;
    d[1] = a.y;
    d[2] = a.z;
  }
  for (int i = threadIdx.x; i < m.nprim; i += blockDim.x) ((float2*)(tab + o.pr))[i] = l2e[i];
  for (int i = threadIdx.x; i < nkg; i += blockDim.x) ((int2*)(tab + o.kg))[i] = kgt[i];
}

template <typename T16>
__device__ __forceinline__ void fused_tile(int rt, int cg, const unsigned char* tab, const MolView& m,
                                           const GeomView& g, int64_t v0, int nvox, int nsub, int npts,
                                           const T16* __restrict__ p_hi, const T16* __restrict__ p_lo, int np,
                                           int nbf, int ncg, float S, double* __restrict__ part, int ldp) {
  typedef typename Vec8<T16>::type V8;
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  __shared__ u32x4 phs[2][2][4][kFP];   // basis values: [slot][hi, lo][k group][point]
  __shared__ u32x4 mrs[2][2][4][kFC];   // M: [slot][hi, lo][k group][column]
  __shared__ double red[4][kFP];
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int ne = (nbf + 31) / 32 * 32;
  const int c0 = ne - (ncg - cg) * kFC;   // first column of the group (< 0 only in group 0)
  const int nst = (c0 + kFC) / 32;
  const int row0 = rt * kFP;
  const int li = lane & 15, lk = lane >> 4;

  if (wave >= 4) {
    // ------------------------------------------------------------ producers
    const FusedTab o = fused_tab(m.natoms, m.nshell, m.nprim, np / 8);
    const int4* tsh = (const int4*)tab;
    const double* tat = (const double*)(tab + o.at);
    const float2* tpr = (const float2*)(tab + o.pr);
    const int2* tkg = (const int2*)(tab + o.kg);
    const float* trc = (const float*)(tab + o.rc);
    const int pw = wave - 4;   // k group of every stage handled by this wave
    double px[2], py[2], pz[2];
    float sc[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int tt = row0 + lane + 64 * i;
      px[i] = py[i] = pz[i] = 0.0;
      if (tt < npts) chunk_point(g, v0, nvox, nsub, tt, px[i], py[i], pz[i]);
      sc[i] = tt < npts ? S : 0.f;   // padding points store zeros
    }
    // M of a stage goes through registers, requested two stages ahead (one register set per stage
    // in flight, set st & 1), so its load latency has the evaluation of two stages to complete
    // (the stage loop below is unrolled by two so that the set index is a constant)
    u32x4 mst[2][2][kFC / 64];
    auto load_m = [&](int st, auto set) __attribute__((always_inline)) {
      constexpr int R = decltype(set)::value;
      if (st >= nst) return;
      const int64_t krow = (int64_t)(4 * st + pw) * np;
#pragma unroll
      for (int c = 0; c < kFC / 64; ++c) {
        const int col = max(c0 + 64 * c + lane, 0);   // columns below 0 belong to skipped tiles
        mst[R][0][c] = *(const u32x4*)(p_hi + (krow + col) * 8);
        mst[R][1][c] = *(const u32x4*)(p_lo + (krow + col) * 8);
      }
    };
    auto store_m = [&](int st, auto set) __attribute__((always_inline)) {
      constexpr int R = decltype(set)::value;
      const int slot = st & 1;
#pragma unroll
      for (int c = 0; c < kFC / 64; ++c) {
        mrs[slot][0][pw][64 * c + lane] = mst[R][0][c];
        mrs[slot][1][pw][64 * c + lane] = mst[R][1][c];
      }
    };
    const std::integral_constant<int, 0> set0;
    const std::integral_constant<int, 1> set1;
    auto make_phi = [&](int st) __attribute__((always_inline)) {
#pragma clang fp contract(off)   // the roundings of phi_split_body
      const int slot = st & 1;
      const int2 e = tkg[4 * st + pw];
      int s = __builtin_amdgcn_readfirstlane(e.x), mm = __builtin_amdgcn_readfirstlane(e.y);
      int L = 0, atom = -1;
      float dx[2], dy[2], dz[2], r2[2], rad[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) dx[i] = dy[i] = dz[i] = r2[i] = rad[i] = 0.f;
      auto open_shell = [&]() __attribute__((always_inline)) {
        const int4 sh = tsh[s];
        const int sa = __builtin_amdgcn_readfirstlane(sh.x);
        const int np_ = __builtin_amdgcn_readfirstlane(sh.z), po = __builtin_amdgcn_readfirstlane(sh.w);
        const float rc = trc[s];
        L = __builtin_amdgcn_readfirstlane(sh.y);
        if (sa != atom) {   // r^2 in FP64, rounded once (as phi_split_body)
          atom = sa;
          const double ax = tat[3 * atom], ay = tat[3 * atom + 1], az = tat[3 * atom + 2];
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            const double ex = px[i] - ax, ey = py[i] - ay, ez = pz[i] - az;
            dx[i] = (float)ex;
            dy[i] = (float)ey;
            dz[i] = (float)ez;
            r2[i] = (float)dist2(ex, ey, ez);
          }
        }
        // primitives outer, the two points inner; the next primitive is read while this one is used
        float radial[2] = {0.f, 0.f};
        if (r2[0] <= rc || r2[1] <= rc) {
          float2 pc = tpr[po];
          for (int p = 0; p < np_; ++p) {
            const float2 nx = tpr[po + min(p + 1, np_ - 1)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
              radial[i] = __builtin_fmaf(pc.y, __builtin_amdgcn_exp2f(-pc.x * r2[i]), radial[i]);
            pc = nx;
          }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) rad[i] = (r2[i] <= rc ? radial[i] : 0.f) * sc[i];   // S: a power of two
      };
      if (s < m.nshell) open_shell();
      uint32_t hw[2][4], lw[2][4];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (s < m.nshell && mm == 2 * L + 1) {
          ++s;
          mm = 0;
          if (s < m.nshell) open_shell();
        }
        const bool live = s < m.nshell;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const float xs = live ? harm_factor(L, mm, dx[i], dy[i], dz[i]) * rad[i] : 0.f;
          const T16 hq = (T16)xs;
          const T16 lq = (T16)(xs - (float)hq);
          if (j & 1) {
            hw[i][j >> 1] |= bits16(hq) << 16;
            lw[i][j >> 1] |= bits16(lq) << 16;
          } else {
            hw[i][j >> 1] = bits16(hq);
            lw[i][j >> 1] = bits16(lq);
          }
        }
        ++mm;
      }
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        phs[slot][0][pw][lane + 64 * i] = u32x4{hw[i][0], hw[i][1], hw[i][2], hw[i][3]};
        phs[slot][1][pw][lane + 64 * i] = u32x4{lw[i][0], lw[i][1], lw[i][2], lw[i][3]};
      }
    };
    load_m(0, set0);
    load_m(1, set1);
    make_phi(0);
    store_m(0, set0);
    load_m(2, set0);   // into the register set stage 0 has released
    __syncthreads();
    for (int st = 0; st < nst; st += 2) {   // stage st + 1 uses set 1, stage st + 2 set 0
      if (st + 1 < nst) {
        make_phi(st + 1);
        store_m(st + 1, set1);
        load_m(st + 3, set1);
      }
      __syncthreads();
      if (st + 1 < nst) {
        if (st + 2 < nst) {
          make_phi(st + 2);
          store_m(st + 2, set0);
          load_m(st + 4, set0);
        }
        __syncthreads();
      }
    }
    __syncthreads();   // the consumers' row-dot partials are in red
  } else {
    // ------------------------------------------------------------ consumers
    const int w = wave;
    f4 acc[8][5];
    double pr[8];   // row-dot partial sums of this lane: point 16 r + li, summed over its columns
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      pr[r] = 0.0;
#pragma unroll
      for (int c = 0; c < 5; ++c) acc[r][c] = f4{0, 0, 0, 0};
    }
    // tile c of this wave: columns n0(c) = c0 + 16 (w + 4 c) .. + 15; it takes stage st while
    // 32 st <= n0 + 15, and never if n0 < 0 (padding)
    const int cpad = min(cdiv_floor0(-c0 - 16 * w, 64), 5);
    // first live tile in stage st; non-decreasing in st, so the stages run in 5 phases of fixed
    // first tile (a loop per phase keeps the accumulators in place across stages)
    auto first_live = [&](int st) { return min(max(cdiv_floor0(32 * st - c0 - 15 - 16 * w, 64), cpad), 5); };
    auto body = [&](int st, auto first) __attribute__((always_inline)) {
      constexpr int CS = decltype(first)::value;
      const int slot = st & 1;
      // the first live tile completes in this stage when its columns lie in 32 st .. 32 st + 31
      const int nfirst = c0 + 16 * (w + 4 * CS);
      const bool done = nfirst < 32 * st + 32;
      const int dc = (nfirst - 32 * st) >> 4;   // which half of the stage's k range holds its columns
      V8 mh[5], ml[5];
#pragma unroll
      for (int c = CS; c < 5; ++c) {
        mh[c] = __builtin_bit_cast(V8, mrs[slot][0][lk][16 * (w + 4 * c) + li]);
        ml[c] = __builtin_bit_cast(V8, mrs[slot][1][lk][16 * (w + 4 * c) + li]);
      }
      // basis-value fragments one row tile ahead; the scheduling barrier after each row tile keeps
      // the compiler from hoisting further loads (registers are taken by the 160 accumulators)
      V8 ah = __builtin_bit_cast(V8, phs[slot][0][lk][li]);
      V8 al = __builtin_bit_cast(V8, phs[slot][1][lk][li]);
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        V8 ahn = ah, aln = al;
        if (r + 1 < 8) {
          ahn = __builtin_bit_cast(V8, phs[slot][0][lk][16 * (r + 1) + li]);
          aln = __builtin_bit_cast(V8, phs[slot][1][lk][16 * (r + 1) + li]);
        }
        // transposed tiles (M first); each product runs over the tiles before the next, so
        // consecutive MFMAs are independent (a tile's three products keep their order)
#pragma unroll
        for (int c = CS; c < 5; ++c) acc[r][c] = Vec8<T16>::mma(mh[c], al, acc[r][c]);
#pragma unroll
        for (int c = CS; c < 5; ++c) acc[r][c] = Vec8<T16>::mma(ml[c], ah, acc[r][c]);
#pragma unroll
        for (int c = CS; c < 5; ++c) acc[r][c] = Vec8<T16>::mma(mh[c], ah, acc[r][c]);
        __builtin_amdgcn_sched_barrier(0);
        ah = ahn;
        al = aln;
      }
      if (done) {
        // lane (li, lk), register q of the tile is X[t = 16 r + li][n = nfirst + 4 lk + q]; the 4 basis
        // values of that point sit in one 8-byte word of each plane of k group 2 dc + lk / 2
        typedef _Float16 h4 __attribute__((ext_vector_type(4)));
        typedef __bf16 b4 __attribute__((ext_vector_type(4)));
        typedef typename std::conditional<std::is_same<T16, _Float16>::value, h4, b4>::type V4;
        const int kg = 2 * dc + (lk >> 1);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const V4 h = __builtin_bit_cast(V4, ((const uint2*)&phs[slot][0][kg][16 * r + li])[lk & 1]);
          const V4 l = __builtin_bit_cast(V4, ((const uint2*)&phs[slot][1][kg][16 * r + li])[lk & 1]);
#pragma unroll
          for (int q = 0; q < 4; ++q) pr[r] += (double)(acc[r][CS][q] * ((float)h[q] + (float)l[q]));
        }
      }
    };
    __syncthreads();
    int st = 0;
    auto phase = [&](auto first) __attribute__((always_inline)) {
      for (; st < nst && first_live(st) == decltype(first)::value; ++st) {
        body(st, first);
        __syncthreads();
      }
    };
    phase(std::integral_constant<int, 0>());
    phase(std::integral_constant<int, 1>());
    phase(std::integral_constant<int, 2>());
    phase(std::integral_constant<int, 3>());
    phase(std::integral_constant<int, 4>());
    for (; st < nst; ++st) __syncthreads();   // no live tile left in this wave
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      double p = pr[r];
      p += __shfl_xor(p, 16);
      p += __shfl_xor(p, 32);
      if (lk == 0) red[w][16 * r + li] = p;
    }
    __syncthreads();
  }
  if (t < kFP) part[(int64_t)cg * ldp + row0 + t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
}

template <typename T16>
__global__ __launch_bounds__(512) void k_density_fused(MolView m, const float2* __restrict__ l2e,
                                                       const int2* __restrict__ kgt, GeomView g, int64_t v0, int nvox,
                                                       int nsub, int npts, const T16* __restrict__ p_hi,
                                                       const T16* __restrict__ p_lo, int np, int nbf, int nrt, int ncg,
                                                       float S, double* __restrict__ part, int ldp) {
  __shared__ __attribute__((aligned(16))) unsigned char tab[kFTab];
  int rt, cg;
  if (!tile_coords(blockIdx.x, nrt, ncg, rt, cg)) return;
  fused_load_tables(tab, m, l2e, kgt, np / 8);
  __syncthreads();
  fused_tile<T16>(rt, cg, tab, m, g, v0, nvox, nsub, npts, p_hi, p_lo, np, nbf, ncg, S, part, ldp);
}

// ---------------------------------------------------------------------------
// Wide split contraction (split variant 5): basis values written pre-split by
// k_phi_split, a workgroup of 8 waves owns 256 points and one column group of
// 320 columns (right-aligned like the fused kernel's groups), and every wave
// owns 128 points x 5 tiles of 16 columns: wave 4 h + w takes rows 128 h ..
// 128 h + 127 and the tiles w, w + 4, .., w + 16 of the group (dealt
// round-robin so the triangle leaves the waves evenly), 160 FP32 accumulators
// each, two waves per SIMD.  Per stage of 32 k a wave reads 16 KB of basis-
// value fragments and 10 KB of M fragments from LDS for 120 MFMAs (0.22 KB per
// MFMA, against 0.5 KB for the 32 x 64 wave tiles of variant 3).  Both operands
// come by LDS DMA (global_load_lds_dwordx4) into a two-stage ring: per stage
// 32 KB of basis values (hi, lo) and 40 KB of M; the loads of stage st + 1 are
// issued right after the barrier that opens stage st.  MFMA order, products
// and row-dot are those of the fused kernel's consumers.  LDS: 144 KB + 8 KB.
//
// Measured on the 60-atom, 256^3 workload: the results of variant 3 to 2.5e-16 relative (the
// row-dot terms are grouped differently), but 30.4 ms of
// contraction against 17.7 ms.  Counters (rocprofv3, same MFMA busy cycles in both): it waits
// on LDS instruction issue 4.5x less than variant 3, but with one workgroup (two waves) per
// SIMD pair the barrier and the fragment loads at the start of each stage leave the MFMA pipes
// idle (35% busy against 52% for the four waves per SIMD of variant 3).  Kept as a variant.
constexpr int kWP = 256;   // points per workgroup

inline int wide_groups(int nbf) { return fused_groups(nbf); }

template <typename T16>
__device__ __forceinline__ void wide_tile(int rt, int cg, const T16* __restrict__ a_hi, int ld,
                                          const T16* __restrict__ p_hi, const T16* __restrict__ p_lo, int np,
                                          int nbf, int ncg, double* __restrict__ part, int ldp) {
  typedef typename Vec8<T16>::type V8;
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  __shared__ u32x4 phs[2][2][4][kWP];   // basis values: [slot][hi, lo][k group][point]
  __shared__ u32x4 mrs[2][2][4][kFC];   // M: [slot][hi, lo][k group][column]
  __shared__ double red[2][4][128];
  const T16* __restrict__ a_lo = a_hi + (int64_t)np * ld;
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int hr = wave >> 2, w = wave & 3;   // row half, column slice
  const int ne = (nbf + 31) / 32 * 32;
  const int c0 = ne - (ncg - cg) * kFC;
  const int nst = (c0 + kFC) / 32;
  const int row0 = rt * kWP;
  const int li = lane & 15, lk = lane >> 4;
  // per stage: 32 DMA instructions of basis values (plane, k group, 64-point chunk) and 40 of M
  // (plane, k group, 64-column chunk); wave v issues instructions v, v + 8, .., < 72
  auto issue = [&](int st) __attribute__((always_inline)) {
    const int slot = st & 1;
    for (int i = wave; i < 72; i += 8) {
      if (i < 32) {
        const int pl = i >> 4, kg = (i >> 2) & 3, ch = i & 3;
        const T16* src = pl ? a_lo : a_hi;
        lds_dma16(src + ((int64_t)(4 * st + kg) * ld + row0 + 64 * ch + lane) * 8, &phs[slot][pl][kg][64 * ch]);
      } else {
        const int j = i - 32, pl = j / 20, kg = (j % 20) / 5, ch = j % 5;
        const T16* src = pl ? p_lo : p_hi;
        const int col = max(c0 + 64 * ch + lane, 0);   // columns below 0 belong to skipped tiles
        lds_dma16(src + ((int64_t)(4 * st + kg) * np + col) * 8, &mrs[slot][pl][kg][64 * ch]);
      }
    }
  };
  f4 acc[8][5];
  double pr[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    pr[r] = 0.0;
#pragma unroll
    for (int c = 0; c < 5; ++c) acc[r][c] = f4{0, 0, 0, 0};
  }
  const int cpad = min(cdiv_floor0(-c0 - 16 * w, 64), 5);
  auto first_live = [&](int st) { return min(max(cdiv_floor0(32 * st - c0 - 15 - 16 * w, 64), cpad), 5); };
  const int rbase = 128 * hr;
  auto body = [&](int st, auto first) __attribute__((always_inline)) {
    constexpr int CS = decltype(first)::value;
    const int slot = st & 1;
    const int nfirst = c0 + 16 * (w + 4 * CS);
    const bool done = nfirst < 32 * st + 32;
    const int dc = (nfirst - 32 * st) >> 4;
    V8 mh[5], ml[5];
#pragma unroll
    for (int c = CS; c < 5; ++c) {
      mh[c] = __builtin_bit_cast(V8, mrs[slot][0][lk][16 * (w + 4 * c) + li]);
      ml[c] = __builtin_bit_cast(V8, mrs[slot][1][lk][16 * (w + 4 * c) + li]);
    }
    V8 ah = __builtin_bit_cast(V8, phs[slot][0][lk][rbase + li]);
    V8 al = __builtin_bit_cast(V8, phs[slot][1][lk][rbase + li]);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      V8 ahn = ah, aln = al;
      if (r + 1 < 8) {
        ahn = __builtin_bit_cast(V8, phs[slot][0][lk][rbase + 16 * (r + 1) + li]);
        aln = __builtin_bit_cast(V8, phs[slot][1][lk][rbase + 16 * (r + 1) + li]);
      }
      // transposed tiles (M first); each product runs over the tiles before the next, so consecutive
      // MFMAs are independent (a tile's three products still accumulate in the same order)
#pragma unroll
      for (int c = CS; c < 5; ++c) acc[r][c] = Vec8<T16>::mma(mh[c], al, acc[r][c]);
#pragma unroll
      for (int c = CS; c < 5; ++c) acc[r][c] = Vec8<T16>::mma(ml[c], ah, acc[r][c]);
#pragma unroll
      for (int c = CS; c < 5; ++c) acc[r][c] = Vec8<T16>::mma(mh[c], ah, acc[r][c]);
      __builtin_amdgcn_sched_barrier(0);
      ah = ahn;
      al = aln;
    }
    if (done) {
      typedef _Float16 h4 __attribute__((ext_vector_type(4)));
      typedef __bf16 b4 __attribute__((ext_vector_type(4)));
      typedef typename std::conditional<std::is_same<T16, _Float16>::value, h4, b4>::type V4;
      const int kg = 2 * dc + (lk >> 1);
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const V4 h = __builtin_bit_cast(V4, ((const uint2*)&phs[slot][0][kg][rbase + 16 * r + li])[lk & 1]);
        const V4 l = __builtin_bit_cast(V4, ((const uint2*)&phs[slot][1][kg][rbase + 16 * r + li])[lk & 1]);
#pragma unroll
        for (int q = 0; q < 4; ++q) pr[r] += (double)(acc[r][CS][q] * ((float)h[q] + (float)l[q]));
      }
    }
  };
  issue(0);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  int st = 0;
  auto step = [&](auto first) __attribute__((always_inline)) {
    if (st + 1 < nst) issue(st + 1);   // into the slot every wave finished with at the last barrier
    body(st, first);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // this wave's loads of stage st + 1
    __syncthreads();
  };
  auto phase = [&](auto first) __attribute__((always_inline)) {
    for (; st < nst && first_live(st) == decltype(first)::value; ++st) step(first);
  };
  phase(std::integral_constant<int, 0>());
  phase(std::integral_constant<int, 1>());
  phase(std::integral_constant<int, 2>());
  phase(std::integral_constant<int, 3>());
  phase(std::integral_constant<int, 4>());
  for (; st < nst; ++st) {   // no live tile left in this wave: it still issues its loads
    if (st + 1 < nst) issue(st + 1);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    double p = pr[r];
    p += __shfl_xor(p, 16);
    p += __shfl_xor(p, 32);
    if (lk == 0) red[hr][w][16 * r + li] = p;
  }
  __syncthreads();
  if (t < kWP) {
    const int h = t >> 7, tt = t & 127;
    part[(int64_t)cg * ldp + row0 + t] = ((red[h][0][tt] + red[h][1][tt]) + red[h][2][tt]) + red[h][3][tt];
  }
}

template <typename T16>
__global__ __launch_bounds__(512) void k_contract_split_wide(const void* __restrict__ phi, int ld,
                                                             const T16* __restrict__ p_hi,
                                                             const T16* __restrict__ p_lo, int np, int nbf, int nrt,
                                                             int ncg, double* __restrict__ part) {
  int rt, cg;
  if (!tile_coords(blockIdx.x, nrt, ncg, rt, cg)) return;
  wide_tile<T16>(rt, cg, (const T16*)phi, ld, p_hi, p_lo, np, nbf, ncg, part, ld);
}

// Variant 3: tiles of 128 points x 64 columns with waves stacked 8 x 1 (wave w
// owns rows 16w .. 16w+15 and all 64 columns), so every wave has the same
// columns and, in the stages that cross the block's diagonal, all waves skip
// the same MFMAs: tile column c (columns cs .. cs+15) takes the k-substep
// starting at k only if k <= cs + 15 (M is zero below the diagonal).  The
// narrower blocks also follow the triangle more closely; row-dots complete
// inside a wave.  Column blocks of 64: ncb = nbf_pad / 64.
constexpr int kRN = 64;

__device__ __forceinline__ void tiled_f64_rows_body(int rt, int cb, const double* __restrict__ phi, int ld,
                                                    const double* __restrict__ M, int np, double* __restrict__ part,
                                                    int ldp) {
  constexpr int CT = kRN / 16;
  __shared__ double As[2][kTKS][kTT + kTPad];
  __shared__ double Bs[2][kTKS][kRN + kTPad];
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int li = lane & 15, lk = lane >> 4;
  const int row0 = rt * kTT, c0 = cb * kRN;
  const int nst = min(np, c0 + kRN) / kTKS;
  const int nfull = c0 / kTKS;                   // stages below the block: every tile takes them
  constexpr int QA = kTKS * kTT / kTWG;          // 4 basis values per thread and stage
  constexpr int KA = kTWG / kTT;
  const int kka = t / kTT, ia = t % kTT;
  const int kkb = t / kRN, ib = t % kRN;         // 2 x 16 x 64: threads < 1024 -> one M value each (t < 1024 always)
  const unsigned aoff = (unsigned)(kka * ld + ia);
  const double* abase = phi + row0;
  const int bcol = min(c0 + ib, np - 1);         // columns past the matrix are never used
  double pa[QA], pb[2];
  auto load = [&](int st) {
#pragma unroll
    for (int q = 0; q < QA; ++q) pa[q] = (abase + (int64_t)(st * kTKS + KA * q) * ld)[aoff];
#pragma unroll
    for (int q = 0; q < 2; ++q) pb[q] = M[(int64_t)(st * kTKS + kkb + 8 * q) * np + bcol];
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < QA; ++q) As[buf][kka + KA * q][ia] = pa[q];
#pragma unroll
    for (int q = 0; q < 2; ++q) Bs[buf][kkb + 8 * q][ib] = pb[q];
  };
  d4 acc[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) acc[c] = d4{0, 0, 0, 0};
  // one stage on the tiles c >= C (straight-line code: no branch around an MFMA)
  auto stage = [&](auto cfirst, int buf) {
    constexpr int C = decltype(cfirst)::value;
#pragma unroll
    for (int ks = 0; ks < kTKS; ks += 4) {
      const double a = As[buf][ks + lk][wave * 16 + li];
      double b[CT];
#pragma unroll
      for (int c = C; c < CT; ++c) b[c] = Bs[buf][ks + lk][c * 16 + li];
#pragma unroll
      for (int c = C; c < CT; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[c], acc[c], 0, 0, 0);
    }
  };
  load(0);
  store(0);
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const bool more = st + 1 < nst;
    if (more) load(st + 1);
    const int buf = st & 1;
    // stage j = st - nfull of the diagonal: tiles c < j lie below it for every k of the stage
    switch (st < nfull ? 0 : st - nfull) {
      case 0: stage(std::integral_constant<int, 0>{}, buf); break;
      case 1: stage(std::integral_constant<int, 1>{}, buf); break;
      case 2: stage(std::integral_constant<int, 2>{}, buf); break;
      default: stage(std::integral_constant<int, 3>{}, buf); break;
    }
    if (more) store((st + 1) & 1);
    __syncthreads();
  }
  double p4[4] = {0, 0, 0, 0};
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const int n = c0 + c * 16 + li;
    if (n < np) {
      const double* ph = phi + (int64_t)n * ld + row0 + wave * 16 + lk;
#pragma unroll
      for (int q = 0; q < 4; ++q) p4[q] += acc[c][q] * ph[4 * q];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double v = sum16(p4[q]);
    if (li == 0) part[(int64_t)cb * ldp + row0 + wave * 16 + 4 * q + lk] = v;
  }
}

__global__ __launch_bounds__(kTWG) void k_contract_f64_rows(const double* __restrict__ phi, int ld,
                                                            const double* __restrict__ M, int np, int nrt, int ncb,
                                                            double* __restrict__ part) {
  int rt, cb;
  if (!tile_coords(blockIdx.x, nrt, ncb, rt, cb)) return;
  tiled_f64_rows_body(rt, cb, phi, ld, M, np, part, ld);
}

template <bool DMA>
__global__ __launch_bounds__(kTWG) void k_contract_f64_tiles(const double* __restrict__ phi, int ld,
                                                             const double* __restrict__ M, int np, int nbf, int nrt,
                                                             int ncb, double* __restrict__ part,
                                                             const uint32_t* __restrict__ amask, int nw) {
  int rt, cb;
  if (!tile_coords(blockIdx.x, nrt, ncb, rt, cb)) return;
  if (DMA) tiled_f64_dma_body<kF64Ks, 32 / kF64Ks>(rt, cb, phi, ld, M, np, nbf, ncb, part, ld, amask, nw);
  else tiled_f64_body(rt, cb, phi, ld, M, np, part, ld);
}

// Activity of each row tile's groups of 16 basis functions: bit q of mask[rt nw + q / 32] is set
// when some point of the tile is within the cutoff (r^2 <= rcut2) of a shell with a function in
// group q.  A clear bit means the tile's basis values in the group are all exactly zero (the
// basis-value kernels write 0 beyond a shell's cutoff).  Conservative: the distance is taken to the
// bounding box of the tile's points, against the cutoff with a relative margin of 1e-6.
__device__ __forceinline__ void tile_mask_body(const MolView& m, const GeomView& g, int64_t v0, int nvox, int nsub,
                                               int npts, int rt, int nw, uint32_t* __restrict__ mask) {
  __shared__ double box[kTT / 64][6];
  __shared__ uint32_t words[64];
  const int t = rt * kTT + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};   // lo xyz, hi xyz
  if (t < npts) {
    double x, y, z;
    chunk_point(g, v0, nvox, nsub, t, x, y, z);
    b[0] = b[3] = x;
    b[1] = b[4] = y;
    b[2] = b[5] = z;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      b[d] = fmin(b[d], __shfl_xor(b[d], off, 64));
      b[3 + d] = fmax(b[3 + d], __shfl_xor(b[3 + d], off, 64));
    }
  if (lane == 0)
    for (int d = 0; d < 6; ++d) box[wave][d] = b[d];
  for (int w = threadIdx.x; w < nw; w += kTT) words[w] = 0u;
  __syncthreads();
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = box[0][d];
    hi[d] = box[0][3 + d];
    for (int w = 1; w < kTT / 64; ++w) {
      lo[d] = fmin(lo[d], box[w][d]);
      hi[d] = fmax(hi[d], box[w][3 + d]);
    }
  }
  if (lo[0] <= hi[0]) {   // the tile holds at least one point
    for (int sidx = threadIdx.x; sidx < m.nshell; sidx += kTT) {
      const DevShell sh = m.shells[sidx];
      const double4 a = m.atoms[sh.atom];
      const double c[3] = {a.x, a.y, a.z};
      double d2 = 0.0;
      for (int d = 0; d < 3; ++d) {
        const double e = fmax(0.0, fmax(lo[d] - c[d], c[d] - hi[d]));
        d2 += e * e;
      }
      if (d2 <= (double)sh.rcut2 * (1.0 + 1e-6))
        for (int q = sh.bf >> 4; q <= (sh.bf + 2 * sh.L) >> 4; ++q) atomicOr(&words[q >> 5], 1u << (q & 31));
    }
  }
  __syncthreads();
  for (int w = threadIdx.x; w < nw; w += kTT) mask[(int64_t)rt * nw + w] = words[w];
}

__global__ __launch_bounds__(kTT) void k_tile_mask(MolView m, GeomView g, int64_t v0, int nvox, int nsub, int npts,
                                                   int nw, uint32_t* __restrict__ mask) {
  tile_mask_body(m, g, v0, nvox, nsub, npts, blockIdx.x, nw, mask);
}

// rho_t = scale * sum_cb part[cb][t] (fixed order), optionally / pscale_t
template <typename TO>
__global__ void k_tiles_finish(const double* __restrict__ part, int ldp, int ncb, int npts, double scale,
                               const float* __restrict__ pscale, TO* __restrict__ out, int64_t v0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts) return;
  double v = 0.0;
  for (int cb = 0; cb < ncb; ++cb) v += part[(int64_t)cb * ldp + t];
  if (pscale) v /= (double)pscale[t];
  store_rho<TO>(out, v0 + t, v * scale);
}

template <typename T16, int NT, int TM, typename TO>
__device__ __forceinline__ void contract_split_body(int blk, const float* __restrict__ phi, int ld,
                                                    const float* __restrict__ pscale,
                                                    const T16* __restrict__ p_hi, const T16* __restrict__ p_lo,
                                                    int nbf_pad, int npts, double scale, TO* __restrict__ out,
                                                    int64_t v0) {
  // p_hi / p_lo hold M = D + 2U of the symmetric P (upper triangular), so
  // column tile c only needs the stages s with 32 s <= 16 c + 15, i.e. 2 s <= c.
  // Wave w owns the column tiles ct0 + w + 8j of a pass (cyclic, balanced).
  typedef typename Vec8<T16>::type V8;
  constexpr int KG = 4;             // 4 groups of 8 k = 32 k per stage
  constexpr int RT = TM / 16;       // row tiles
  __shared__ uint4 Ah[2][KG][TM], Al[2][KG][TM];
  __shared__ double red[kWaves][TM];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = lane & 15, lk = lane >> 4;
  const int row0 = blk * TM;
  for (int e = threadIdx.x; e < kWaves * TM; e += kWG) red[e / TM][e % TM] = 0.0;
  const int nks = nbf_pad / 32;
  const bool stager = threadIdx.x < KG * TM;
  const int sr = threadIdx.x % TM, sg = threadIdx.x / TM;     // staged row and k-group
  const float sc = stager ? pscale[row0 + sr] : 1.f;
  const float* src = phi + (int64_t)(sg * 8) * ld + row0 + sr;
  auto stage_store = [&](int buf, const float* v) {
    V8 h, l;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float x = v[q] * sc;
      const T16 hq = (T16)x;
      h[q] = hq;
      l[q] = (T16)(x - (float)hq);
    }
    Ah[buf][sg][sr] = *(const uint4*)&h;
    Al[buf][sg][sr] = *(const uint4*)&l;
  };
  for (int c0 = 0; c0 < nbf_pad; c0 += kWaves * NT * 16) {
    const int ct0 = c0 / 16;
    const int nks_pass = min(nks, (ct0 + kWaves * NT + 1) / 2);
    int tile[NT];
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) tile[ct] = ct0 + wave + kWaves * ct;
    f4 acc[RT][NT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) acc[rt][ct] = f4{0, 0, 0, 0};
    float pre[8];
    if (stager) {
#pragma unroll
      for (int q = 0; q < 8; ++q) pre[q] = src[(int64_t)q * ld];
      stage_store(0, pre);
    }
    __syncthreads();
    // M operands of a stage: unconditional loads (a tile past the matrix or past
    // its last stage reads element 0, unused), one stage ahead of their use
    V8 bh[NT], bl[NT];
    auto load_b = [&](int st, V8* h, V8* l) {
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) {
        const int n = tile[ct] * 16 + li;
        const bool live = n < nbf_pad && 2 * st <= tile[ct] && st < nks_pass;
        const int64_t o = live ? ((int64_t)(st * KG + lk) * nbf_pad + n) * 8 : 0;
        h[ct] = *(const V8*)(p_hi + o);
        l[ct] = *(const V8*)(p_lo + o);
      }
    };
    load_b(0, bh, bl);
    for (int s = 0; s < nks_pass; ++s) {
      const bool more = s + 1 < nks_pass;
      if (more && stager) {
        const float* nsrc = src + (int64_t)((s + 1) * 32) * ld;
#pragma unroll
        for (int q = 0; q < 8; ++q) pre[q] = nsrc[(int64_t)q * ld];
      }
      V8 nh[NT], nl[NT];
      load_b(s + 1, nh, nl);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        const uint4 uh = Ah[s & 1][lk][rt * 16 + li];
        const uint4 ul = Al[s & 1][lk][rt * 16 + li];
        const V8 ah = *(const V8*)&uh;
        const V8 al = *(const V8*)&ul;
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
          if (2 * s <= tile[ct] && tile[ct] * 16 < nbf_pad) {
            acc[rt][ct] = Vec8<T16>::mma(al, bh[ct], acc[rt][ct]);
            acc[rt][ct] = Vec8<T16>::mma(ah, bl[ct], acc[rt][ct]);
            acc[rt][ct] = Vec8<T16>::mma(ah, bh[ct], acc[rt][ct]);
          }
        }
      }
      if (more && stager) stage_store((s + 1) & 1, pre);
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) {
        bh[ct] = nh[ct];
        bl[ct] = nl[ct];
      }
      __syncthreads();
    }
    {
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        double part[4] = {0, 0, 0, 0};
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
          const int n = tile[ct] * 16 + li;
          if (tile[ct] * 16 < nbf_pad) {
            const float4 v = *(const float4*)(phi + (int64_t)n * ld + row0 + rt * 16 + 4 * lk);
            part[0] += (double)(acc[rt][ct][0] * v.x);
            part[1] += (double)(acc[rt][ct][1] * v.y);
            part[2] += (double)(acc[rt][ct][2] * v.z);
            part[3] += (double)(acc[rt][ct][3] * v.w);
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double v = sum16(part[r]);
          if (li == 0) red[wave][rt * 16 + 4 * lk + r] += v;
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < TM) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) v += red[w][threadIdx.x];
    const int t = row0 + threadIdx.x;
    if (t < npts) store_rho<TO>(out, v0 + t, v / (double)pscale[t] * scale);
  }
}

template <typename T16, int NT, int TM, typename TO>
__global__ __launch_bounds__(kWG) void k_contract_split(const float* __restrict__ phi, int ld,
                                                        const float* __restrict__ pscale,
                                                        const T16* __restrict__ p_hi, const T16* __restrict__ p_lo,
                                                        int nbf_pad, int npts, double scale, TO* __restrict__ out,
                                                        int64_t v0) {
  contract_split_body<T16, NT, TM, TO>(blockIdx.x, phi, ld, pscale, p_hi, p_lo, nbf_pad, npts, scale, out, v0);
}

// Sum the nsub^3 sub-point densities of each voxel: out[v0 + v] = sum_s rho[s * nvox + v].
template <typename TO>
__global__ void k_subgrid_sum(const double* __restrict__ rho, int nvox, int nsub3, TO* __restrict__ out,
                              int64_t v0) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  double acc = 0.0;
  for (int s = 0; s < nsub3; ++s) acc += rho[(int64_t)s * nvox + v];
  out[v0 + v] = (TO)acc;
}

}  // namespace

// --------------------------------------------------------------------------- host side
namespace {

// M (upper triangle of P, off-diagonal doubled) -> 16-bit hi/lo planes in the
// K8-interleaved layout, scaled by pfac.
template <typename T16>
__global__ void k_split_P(const double* __restrict__ P64, int npad, float pfac, T16* __restrict__ hi,
                          T16* __restrict__ lo) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)npad * npad) return;
  const int k = (int)(idx / npad), n = (int)(idx % npad);
  const float x = (float)(P64[idx] * (double)pfac);
  const T16 h = to16<T16>(x);
  const T16 l = to16<T16>(x - (float)h);
  const int64_t o = ((int64_t)(k >> 3) * npad + n) * 8 + (k & 7);
  hi[o] = h;
  lo[o] = l;
}

int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

constexpr size_t kPhiBudget = (size_t)4 << 30;   // bytes of basis values per chunk

}  // namespace

int qm_density_upload_P(qm_ctx* ctx, qm_mol* mol, hipStream_t s) {
  const int n = mol->nbf, np = mol->nbf_pad;
  std::vector<double> P((size_t)np * np, 0.0);
  double pmax = 0.0;
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < n; ++q) {
      const double v = mol->h_P[(size_t)p * n + q];
      P[(size_t)p * np + q] = v;
      pmax = std::max(pmax, std::fabs(v));
    }
  int e = 0;
  if (pmax > 0) std::frexp(pmax, &e);
  mol->pfac = std::ldexp(1.0, 1 - e);   // largest |P| scaled into [1, 2)
  const int64_t tot = (int64_t)np * np;
  if (!mol->P64) {
    if (hipMalloc(&mol->P64, sizeof(double) * tot) != hipSuccess || hipMalloc(&mol->M64, sizeof(double) * tot) != hipSuccess ||
        hipMalloc(&mol->P16, sizeof(uint16_t) * 2 * tot) != hipSuccess ||
        hipMalloc(&mol->Pb16, sizeof(uint16_t) * 2 * tot) != hipSuccess) {
      (void)hipGetLastError();
      return qm_fail(ctx, QM_ENOMEM, "density matrix: device allocation failed");
    }
  }
  // ordered on s: kernels queued earlier on s still read the previous matrix
  QM_CHECK_HIP(ctx, hipMemcpyAsync(mol->P64, P.data(), sizeof(double) * tot, hipMemcpyHostToDevice, s));
  // M = D + 2U (upper triangle of P, off-diagonal doubled): phi^T P phi = phi^T M phi
  std::vector<double> Mh((size_t)np * np, 0.0);
  for (int p = 0; p < np; ++p)
    for (int q = p; q < np; ++q) Mh[(size_t)p * np + q] = (p == q ? 1.0 : 2.0) * P[(size_t)p * np + q];
  QM_CHECK_HIP(ctx, hipMemcpyAsync(mol->M64, Mh.data(), sizeof(double) * tot, hipMemcpyHostToDevice, s));
  const unsigned blocks = (unsigned)((tot + 255) / 256);
  hipLaunchKernelGGL(k_split_P<_Float16>, dim3(blocks), dim3(256), 0, s, (const double*)mol->M64, np,
                     (float)mol->pfac, (_Float16*)mol->P16, (_Float16*)mol->P16 + tot);
  QM_CHECK_LAUNCH(ctx);
  hipLaunchKernelGGL(k_split_P<__bf16>, dim3(blocks), dim3(256), 0, s, (const double*)mol->M64, np,
                     (float)mol->pfac, (__bf16*)mol->Pb16, (__bf16*)mol->Pb16 + tot);
  QM_CHECK_LAUNCH(ctx);
  QM_CHECK_HIP(ctx, hipStreamSynchronize(s));   // the host matrices are pageable and go out of scope
  return QM_OK;
}

namespace {

// columns per wave per pass: NT tiles of 16, NT = min(5, ceil(nbf_pad / 128))
inline int pick_nt(int np) { return std::min(5, (np + 127) / 128); }

template <typename TO>
void launch_contract_f64(hipStream_t s, int ld, const double* phi, const double* P, const double* M, int np,
                         int npts, double scale, TO* out, int64_t v0, int direct) {
  const dim3 grid(ld / 64), block(kWG);
  if (direct) {
    switch (pick_nt(np)) {
      case 1: hipLaunchKernelGGL((k_contract_f64_direct<1, TO>), grid, block, 0, s, phi, ld, P, np, npts, scale, out, v0); return;
      case 2: hipLaunchKernelGGL((k_contract_f64_direct<2, TO>), grid, block, 0, s, phi, ld, P, np, npts, scale, out, v0); return;
      case 3: hipLaunchKernelGGL((k_contract_f64_direct<3, TO>), grid, block, 0, s, phi, ld, P, np, npts, scale, out, v0); return;
      case 4: hipLaunchKernelGGL((k_contract_f64_direct<4, TO>), grid, block, 0, s, phi, ld, P, np, npts, scale, out, v0); return;
      default: hipLaunchKernelGGL((k_contract_f64_direct<5, TO>), grid, block, 0, s, phi, ld, P, np, npts, scale, out, v0); return;
    }
  }
  switch (pick_nt(np)) {
    case 1: hipLaunchKernelGGL((k_contract_f64<1, TO>), grid, block, 0, s, phi, ld, M, np, npts, scale, out, v0); break;
    case 2: hipLaunchKernelGGL((k_contract_f64<2, TO>), grid, block, 0, s, phi, ld, M, np, npts, scale, out, v0); break;
    case 3: hipLaunchKernelGGL((k_contract_f64<3, TO>), grid, block, 0, s, phi, ld, M, np, npts, scale, out, v0); break;
    case 4: hipLaunchKernelGGL((k_contract_f64<4, TO>), grid, block, 0, s, phi, ld, M, np, npts, scale, out, v0); break;
    default: hipLaunchKernelGGL((k_contract_f64<5, TO>), grid, block, 0, s, phi, ld, M, np, npts, scale, out, v0); break;
  }
}

template <typename T16, typename TO>
void launch_contract_split(hipStream_t s, int ld, const float* phi, const float* pscale, const T16* Ph,
                           const T16* Pl, int np, int npts, double scale, TO* out, int64_t v0, int tm) {
  const dim3 grid(ld / tm), block(kWG);
#define QM_SPLIT_CASE(NT)                                                                                        \
  if (tm == 64)                                                                                                  \
    hipLaunchKernelGGL((k_contract_split<T16, NT, 64, TO>), grid, block, 0, s, phi, ld, pscale, Ph, Pl, np, npts,  \
                       scale, out, v0);                                                                          \
  else                                                                                                           \
    hipLaunchKernelGGL((k_contract_split<T16, NT, 128, TO>), grid, block, 0, s, phi, ld, pscale, Ph, Pl, np, npts, \
                       scale, out, v0)
  switch (pick_nt(np)) {
    case 1: QM_SPLIT_CASE(1); break;
    case 2: QM_SPLIT_CASE(2); break;
    case 3: QM_SPLIT_CASE(3); break;
    case 4: QM_SPLIT_CASE(4); break;
    default: QM_SPLIT_CASE(5); break;
  }
#undef QM_SPLIT_CASE
}

template <typename TO>
int density_fp64(qm_ctx* ctx, qm_mol* mol, hipStream_t s, const qm_geom* g, int nsub, TO* out) {
  const int np = mol->nbf_pad;
  const int nsub3 = nsub * nsub * nsub;
  const int64_t nvox_tot = npoints(g);
  int64_t maxpts = (int64_t)(kPhiBudget / ((size_t)np * sizeof(double)));
  maxpts = std::max<int64_t>(maxpts / kTileM * kTileM, kTileM);
  int64_t nvox = std::max<int64_t>(maxpts / nsub3, 1);
  nvox = std::min<int64_t>(nvox, nvox_tot);
  nvox = std::min<int64_t>(nvox, ((int64_t)1 << 30) / nsub3);
  const int64_t ldmax = round_up(nvox * nsub3, kTileM);
  void* phi = nullptr;
  int rc = qm_ws(ctx, s, WS_PHI, sizeof(double) * (size_t)np * ldmax, &phi);
  if (rc) return rc;
  double* rho = nullptr;
  if (nsub3 > 1) {
    rc = qm_ws(ctx, s, WS_RHO, sizeof(double) * ldmax, (void**)&rho);
    if (rc) return rc;
  }
  const bool tiled = ctx->f64_variant >= 2;   // 2 tiles, 3 rows, 4 tiles fed by LDS DMA
  const int ncb = ctx->f64_variant == 3 ? np / kRN : (np + kTT - 1) / kTT;
  double* part = nullptr;
  if (tiled && (rc = qm_ws(ctx, s, WS_PART, sizeof(double) * (size_t)ncb * ldmax, (void**)&part))) return rc;
  // activity bits of the row tiles' 16-function groups (the LDS-DMA tile kernel skips zero groups)
  const int nw = (np / 16 + 31) / 32;
  uint32_t* mask = nullptr;
  if (ctx->f64_variant == 4 && ctx->tile_skip && nw <= 64 &&
      (rc = qm_ws(ctx, s, WS_MASK, sizeof(uint32_t) * (size_t)nw * (ldmax / kTT), (void**)&mask)))
    return rc;
  const GeomView gv = geom_view(g);
  const MolView mv = mol->view();
  const double h3 = g->step * g->step * g->step;
  for (int64_t v0 = 0; v0 < nvox_tot; v0 += nvox) {
    const int nv = (int)std::min<int64_t>(nvox, nvox_tot - v0);
    const int npts = nv * nsub3;
    const int ld = (int)round_up(npts, kTileM);
    {
      ProfScope ps(ctx, s, QM_K_PHI);
      hipLaunchKernelGGL(k_phi_f64, dim3((ld + kPhiThreads - 1) / kPhiThreads), dim3(kPhiThreads), 0, s, mv, gv,
                         v0, nv, nsub, npts, (const double*)nullptr, ld, np, (double*)phi);
    }
    QM_CHECK_LAUNCH(ctx);
    ProfScope pc(ctx, s, QM_K_DENSITY_CONTRACT);
    if (tiled) {
      const int nrt = ld / kTT;
      if (ctx->f64_variant == 3)
        hipLaunchKernelGGL(k_contract_f64_rows, dim3(tile_grid(nrt, ncb)), dim3(kTWG), 0, s, (const double*)phi, ld,
                           (const double*)mol->M64, np, nrt, ncb, part);
      else if (ctx->f64_variant == 4) {
        if (mask)
          hipLaunchKernelGGL(k_tile_mask, dim3(nrt), dim3(kTT), 0, s, mv, gv, v0, nv, nsub, npts, nw, mask);
        hipLaunchKernelGGL(k_contract_f64_tiles<true>, dim3(tile_grid(nrt, ncb)), dim3(kTWG), 0, s,
                           (const double*)phi, ld, (const double*)mol->M64, np, mol->nbf, nrt, ncb, part,
                           (const uint32_t*)mask, nw);
      } else {
        hipLaunchKernelGGL(k_contract_f64_tiles<false>, dim3(tile_grid(nrt, ncb)), dim3(kTWG), 0, s,
                           (const double*)phi, ld, (const double*)mol->M64, np, mol->nbf, nrt, ncb, part,
                           (const uint32_t*)nullptr, 0);
      }
      QM_CHECK_LAUNCH(ctx);
      if (nsub3 == 1) {
        hipLaunchKernelGGL(k_tiles_finish<TO>, dim3((npts + 255) / 256), dim3(256), 0, s, (const double*)part, ld, ncb,
                           npts, h3, (const float*)nullptr, out, v0);
      } else {
        hipLaunchKernelGGL(k_tiles_finish<double>, dim3((npts + 255) / 256), dim3(256), 0, s, (const double*)part, ld,
                           ncb, npts, h3 / nsub3, (const float*)nullptr, rho, (int64_t)0);
        QM_CHECK_LAUNCH(ctx);
        hipLaunchKernelGGL(k_subgrid_sum<TO>, dim3((nv + 255) / 256), dim3(256), 0, s, (const double*)rho, nv,
                           nsub3, out, v0);
      }
    } else if (nsub3 == 1) {
      launch_contract_f64<TO>(s, ld, (const double*)phi, (const double*)mol->P64, (const double*)mol->M64, np, npts, h3, out, v0,
                              ctx->f64_variant);
    } else {
      launch_contract_f64<double>(s, ld, (const double*)phi, (const double*)mol->P64, (const double*)mol->M64, np, npts, h3 / nsub3, rho,
                                  (int64_t)0, ctx->f64_variant);
      QM_CHECK_LAUNCH(ctx);
      hipLaunchKernelGGL(k_subgrid_sum<TO>, dim3((nv + 255) / 256), dim3(256), 0, s, (const double*)rho, nv,
                         nsub3, out, v0);
    }
    QM_CHECK_LAUNCH(ctx);
  }
  return QM_OK;
}

// Power of two S with |phi S| <= 2^14 for every basis function at every point:
// a contracted shell is bounded by sum_p |c_p| g_L(alpha_p), with the largest
// angular factor times max_r r^L exp(-a r^2): g_0 = C00, g_1 = C1 / sqrt(2 e a),
// g_2 = max(2 C20, C22, C2A / 2) / (e a).
float split_phi_scale(const qm_mol* m) {
  const double g2 = std::max(std::max(2.0 * C20, C22), 0.5 * C2A);
  double bound = 0.0;
  for (const DevShell& sh : m->h_shells) {
    double b = 0.0;
    for (int p = 0; p < sh.nprim; ++p) {
      const double a = m->h_alpha[sh.poff + p], c = std::fabs(m->h_coeff[sh.poff + p]);
      b += c * (sh.L == 0 ? C00 : sh.L == 1 ? C1 / std::sqrt(2.0 * M_E * a) : g2 / (M_E * a));
    }
    bound = std::max(bound, b);
  }
  if (!(bound > 0.0) || !std::isfinite(bound)) return 1.0f;
  int e = 0;
  std::frexp(bound, &e);   // bound < 2^e
  return std::ldexp(1.0f, std::max(-100, std::min(100, 14 - e)));
}

// Split variants of the contraction: 0 / 2 tiles of 8 (4 x 2) / 4 (2 x 2) waves on FP32
// basis values; 1 tiles of 8 waves on pre-split values staged through registers;
// 3 the same fed by LDS DMA; 64 / 128 all-columns kernels.
inline bool split_tiled(int v) { return v >= 0 && v <= 3; }
inline bool split_pre(int v) { return v == 1 || v == 3; }

template <typename T16>
void launch_split_tiles(hipStream_t s, int v, int blocks, const void* phi, int ld, const float* pscale,
                        const T16* Ph, const T16* Pl, int np, int nbf, int nrt, int ncb, double* part) {
  switch (v) {
    case 1:
      hipLaunchKernelGGL((k_contract_split_tiles<T16, 4, 2, true>), dim3(blocks), dim3(512), 0, s, phi, ld, pscale,
                         Ph, Pl, np, nrt, ncb, part);
      break;
    case 2:
      hipLaunchKernelGGL((k_contract_split_tiles<T16, 2, 2>), dim3(blocks), dim3(256), 0, s, phi, ld, pscale, Ph, Pl,
                         np, nrt, ncb, part);
      break;
    case 3:
      hipLaunchKernelGGL(k_contract_split_dma<T16>, dim3(blocks), dim3(512), 0, s, phi, ld, Ph, Pl, np, nbf, nrt, ncb,
                         part);
      break;
    default:
      hipLaunchKernelGGL((k_contract_split_tiles<T16, 4, 2>), dim3(blocks), dim3(512), 0, s, phi, ld, pscale, Ph, Pl,
                         np, nrt, ncb, part);
      break;
  }
}

// Split variant 5: basis values written pre-split, contracted by k_contract_split_wide (workgroups
// of 256 points x 320 columns); chunks of points padded to a multiple of 256.
template <typename T16, typename TO>
int density_wide(qm_ctx* ctx, qm_mol* mol, hipStream_t s, const qm_geom* g, int nsub, TO* out) {
  const int np = mol->nbf_pad;
  const int nsub3 = nsub * nsub * nsub;
  const int64_t nvox_tot = npoints(g);
  int64_t maxpts = (int64_t)(kPhiBudget / ((size_t)np * sizeof(float)));
  maxpts = std::max<int64_t>(maxpts / kWP * kWP, kWP);
  int64_t nvox = std::max<int64_t>(maxpts / nsub3, 1);
  nvox = std::min<int64_t>(nvox, nvox_tot);
  nvox = std::min<int64_t>(nvox, ((int64_t)1 << 30) / nsub3);
  const int64_t ldmax = round_up(nvox * nsub3, kWP);
  const int ncg = wide_groups(mol->nbf);
  void* phi = nullptr;
  int rc = qm_ws(ctx, s, WS_PHI, sizeof(float) * (size_t)np * ldmax, &phi);
  if (rc) return rc;
  double* rho = nullptr;
  if (nsub3 > 1 && (rc = qm_ws(ctx, s, WS_RHO, sizeof(double) * ldmax, (void**)&rho))) return rc;
  double* part = nullptr;
  if ((rc = qm_ws(ctx, s, WS_PART, sizeof(double) * (size_t)ncg * ldmax, (void**)&part))) return rc;
  const T16* Phi = (const T16*)(std::is_same<T16, _Float16>::value ? mol->P16 : mol->Pb16);
  const T16* Plo = Phi + (int64_t)np * np;
  const float S = split_phi_scale(mol);
  const GeomView gv = geom_view(g);
  const MolView mv = mol->view();
  const double sc = g->step * g->step * g->step / mol->pfac / ((double)S * S);
  for (int64_t v0 = 0; v0 < nvox_tot; v0 += nvox) {
    const int nv = (int)std::min<int64_t>(nvox, nvox_tot - v0);
    const int npts = nv * nsub3;
    const int ld = (int)round_up(npts, kWP);
    {
      ProfScope ps(ctx, s, QM_K_PHI_SPLIT);
      hipLaunchKernelGGL(k_phi_split<T16>, dim3((ld + kPhiThreads - 1) / kPhiThreads), dim3(kPhiThreads), 0, s, mv,
                         (const float2*)mol->prims_l2e, gv, v0, nv, nsub, npts, ld, np, S, (T16*)phi,
                         (T16*)phi + (int64_t)np * ld);
    }
    QM_CHECK_LAUNCH(ctx);
    const int nrt = ld / kWP;
    {
      ProfScope pc(ctx, s, QM_K_DENSITY_SPLIT);
      hipLaunchKernelGGL(k_contract_split_wide<T16>, dim3(tile_grid(nrt, ncg)), dim3(512), 0, s, phi, ld, Phi, Plo,
                         np, mol->nbf, nrt, ncg, part);
    }
    QM_CHECK_LAUNCH(ctx);
    if (nsub3 == 1) {
      hipLaunchKernelGGL(k_tiles_finish<TO>, dim3((npts + 255) / 256), dim3(256), 0, s, (const double*)part, ld, ncg,
                         npts, sc, (const float*)nullptr, out, v0);
    } else {
      hipLaunchKernelGGL(k_tiles_finish<double>, dim3((npts + 255) / 256), dim3(256), 0, s, (const double*)part, ld,
                         ncg, npts, sc / nsub3, (const float*)nullptr, rho, (int64_t)0);
      QM_CHECK_LAUNCH(ctx);
      hipLaunchKernelGGL(k_subgrid_sum<TO>, dim3((nv + 255) / 256), dim3(256), 0, s, (const double*)rho, nv, nsub3,
                         out, v0);
    }
    QM_CHECK_LAUNCH(ctx);
  }
  return QM_OK;
}

// Split variant 4: basis values evaluated inside the contraction (k_density_fused), no basis-value
// workspace; launches of up to kFChunk points, partial sums per column group.
template <typename T16, typename TO>
int density_fused(qm_ctx* ctx, qm_mol* mol, hipStream_t s, const qm_geom* g, int nsub, TO* out) {
  const int np = mol->nbf_pad;
  const int nsub3 = nsub * nsub * nsub;
  const int64_t nvox_tot = npoints(g);
  int64_t nvox = std::max<int64_t>(kFChunk / nsub3, 1);
  nvox = std::min<int64_t>(nvox, nvox_tot);
  const int64_t ldmax = round_up(nvox * nsub3, kFP);
  const int ncg = fused_groups(mol->nbf);
  int rc = QM_OK;
  double* rho = nullptr;
  if (nsub3 > 1 && (rc = qm_ws(ctx, s, WS_RHO, sizeof(double) * ldmax, (void**)&rho))) return rc;
  double* part = nullptr;
  if ((rc = qm_ws(ctx, s, WS_PART, sizeof(double) * (size_t)ncg * ldmax, (void**)&part))) return rc;
  const T16* Phi = (const T16*)(std::is_same<T16, _Float16>::value ? mol->P16 : mol->Pb16);
  const T16* Plo = Phi + (int64_t)np * np;
  const float S = split_phi_scale(mol);
  const GeomView gv = geom_view(g);
  const MolView mv = mol->view();
  // both factors of the row-dot carry S; divided out with the volume element
  const double sc = g->step * g->step * g->step / mol->pfac / ((double)S * S);
  for (int64_t v0 = 0; v0 < nvox_tot; v0 += nvox) {
    const int nv = (int)std::min<int64_t>(nvox, nvox_tot - v0);
    const int npts = nv * nsub3;
    const int ld = (int)round_up(npts, kFP);
    const int nrt = ld / kFP;
    {
      ProfScope pc(ctx, s, QM_K_DENSITY_SPLIT);
      hipLaunchKernelGGL(k_density_fused<T16>, dim3(tile_grid(nrt, ncg)), dim3(512), 0, s, mv,
                         (const float2*)mol->prims_l2e, (const int2*)mol->kgroups, gv, v0, nv, nsub, npts, Phi, Plo,
                         np, mol->nbf, nrt, ncg, S, part, ld);
    }
    QM_CHECK_LAUNCH(ctx);
    if (nsub3 == 1) {
      hipLaunchKernelGGL(k_tiles_finish<TO>, dim3((npts + 255) / 256), dim3(256), 0, s, (const double*)part, ld, ncg,
                         npts, sc, (const float*)nullptr, out, v0);
    } else {
      hipLaunchKernelGGL(k_tiles_finish<double>, dim3((npts + 255) / 256), dim3(256), 0, s, (const double*)part, ld,
                         ncg, npts, sc / nsub3, (const float*)nullptr, rho, (int64_t)0);
      QM_CHECK_LAUNCH(ctx);
      hipLaunchKernelGGL(k_subgrid_sum<TO>, dim3((nv + 255) / 256), dim3(256), 0, s, (const double*)rho, nv, nsub3,
                         out, v0);
    }
    QM_CHECK_LAUNCH(ctx);
  }
  return QM_OK;
}

template <typename T16, typename TO>
int density_split(qm_ctx* ctx, qm_mol* mol, hipStream_t s, const qm_geom* g, int nsub, TO* out) {
  if (ctx->split_tm == 4 && fused_fits(mol)) return density_fused<T16, TO>(ctx, mol, s, g, nsub, out);
  // tables too large for the fused kernel's LDS: basis values written pre-split, LDS DMA tiles
  const int variant = ctx->split_tm == 4 ? 3 : ctx->split_tm;
  if (variant == 5) return density_wide<T16, TO>(ctx, mol, s, g, nsub, out);
  const int np = mol->nbf_pad;
  const int nsub3 = nsub * nsub * nsub;
  const int64_t nvox_tot = npoints(g);
  int64_t maxpts = (int64_t)(kPhiBudget / ((size_t)np * sizeof(float)));
  maxpts = std::max<int64_t>(maxpts / kTileM * kTileM, kTileM);
  int64_t nvox = std::max<int64_t>(maxpts / nsub3, 1);
  nvox = std::min<int64_t>(nvox, nvox_tot);
  nvox = std::min<int64_t>(nvox, ((int64_t)1 << 30) / nsub3);
  const int64_t ldmax = round_up(nvox * nsub3, kTileM);
  void* phi = nullptr;
  int rc = qm_ws(ctx, s, WS_PHI, sizeof(float) * (size_t)np * ldmax, &phi);
  if (rc) return rc;
  float* pscale = nullptr;
  rc = qm_ws(ctx, s, WS_PHI2, sizeof(float) * ldmax, (void**)&pscale);
  if (rc) return rc;
  double* rho = nullptr;
  if (nsub3 > 1) {
    rc = qm_ws(ctx, s, WS_RHO, sizeof(double) * ldmax, (void**)&rho);
    if (rc) return rc;
  }
  const T16* Phi = (const T16*)(std::is_same<T16, _Float16>::value ? mol->P16 : mol->Pb16);
  const T16* Plo = Phi + (int64_t)np * np;
  const bool pre = split_pre(variant);
  const bool tiled = split_tiled(variant);
  const float S = split_phi_scale(mol);
  const int ncb = (np + kTT - 1) / kTT;
  double* part = nullptr;
  if (tiled && (rc = qm_ws(ctx, s, WS_PART, sizeof(double) * (size_t)ncb * ldmax, (void**)&part))) return rc;
  const GeomView gv = geom_view(g);
  const MolView mv = mol->view();
  const double h3 = g->step * g->step * g->step / mol->pfac;
  for (int64_t v0 = 0; v0 < nvox_tot; v0 += nvox) {
    const int nv = (int)std::min<int64_t>(nvox, nvox_tot - v0);
    const int npts = nv * nsub3;
    const int ld = (int)round_up(npts, kTileM);
    {
      ProfScope ps(ctx, s, QM_K_PHI_SPLIT);
      // both layouts take 4 bytes per (function, point): FP32, or 16-bit hi + 16-bit lo
      if (pre)
        hipLaunchKernelGGL(k_phi_split<T16>, dim3((ld + kPhiThreads - 1) / kPhiThreads), dim3(kPhiThreads), 0, s, mv,
                           (const float2*)mol->prims_l2e, gv, v0, nv, nsub, npts, ld, np, S, (T16*)phi,
                           (T16*)phi + (int64_t)np * ld);
      else
        hipLaunchKernelGGL(k_phi_f32, dim3((ld + kPhiThreads - 1) / kPhiThreads), dim3(kPhiThreads), 0, s, mv,
                           (const float2*)mol->prims_l2e, gv, v0, nv, nsub, npts, ld, np, (float*)phi, pscale);
    }
    QM_CHECK_LAUNCH(ctx);
    ProfScope pc(ctx, s, QM_K_DENSITY_SPLIT);
    if (tiled) {
      const int nrt = ld / kTT;
      launch_split_tiles<T16>(s, variant, tile_grid(nrt, ncb), phi, ld, pscale, Phi, Plo, np, mol->nbf, nrt, ncb,
                              part);
      QM_CHECK_LAUNCH(ctx);
      // pre-split: both factors of the row-dot carry S, divided out here with the volume element
      const float* ps = pre ? nullptr : (const float*)pscale;
      const double sc = pre ? h3 / ((double)S * S) : h3;
      if (nsub3 == 1) {
        hipLaunchKernelGGL(k_tiles_finish<TO>, dim3((npts + 255) / 256), dim3(256), 0, s, (const double*)part, ld, ncb,
                           npts, sc, ps, out, v0);
      } else {
        hipLaunchKernelGGL(k_tiles_finish<double>, dim3((npts + 255) / 256), dim3(256), 0, s, (const double*)part, ld,
                           ncb, npts, sc / nsub3, ps, rho, (int64_t)0);
        QM_CHECK_LAUNCH(ctx);
        hipLaunchKernelGGL(k_subgrid_sum<TO>, dim3((nv + 255) / 256), dim3(256), 0, s, (const double*)rho, nv,
                           nsub3, out, v0);
      }
    } else if (nsub3 == 1) {
      launch_contract_split<T16, TO>(s, ld, (const float*)phi, pscale, Phi, Plo, np, npts, h3, out, v0,
                                     variant);
    } else {
      launch_contract_split<T16, double>(s, ld, (const float*)phi, pscale, Phi, Plo, np, npts, h3 / nsub3, rho,
                                         (int64_t)0, variant);
      QM_CHECK_LAUNCH(ctx);
      hipLaunchKernelGGL(k_subgrid_sum<TO>, dim3((nv + 255) / 256), dim3(256), 0, s, (const double*)rho, nv,
                         nsub3, out, v0);
    }
    QM_CHECK_LAUNCH(ctx);
  }
  return QM_OK;
}

}  // namespace

extern "C" {

int qm_density(qm_ctx* ctx, qm_mol* mol, void* stream, const qm_geom* g, int subgrid, int precision,
               int out_dtype, void* out, int normalize, double* raw_total) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  rc = qm_check_geom(ctx, g);
  if (rc) return rc;
  QM_REQUIRE(ctx, mol && out, QM_EINVAL, "density: null argument");
  QM_REQUIRE(ctx, mol->has_dm, QM_EINVAL, "density: molecule has no density matrix");
  QM_REQUIRE(ctx, subgrid >= 1 && subgrid <= 16, QM_EINVAL, "density: subgrid must be 1..16");
  QM_REQUIRE(ctx, out_dtype == QM_F64 || out_dtype == QM_F32, QM_EINVAL, "density: bad output dtype");
  hipStream_t s = (hipStream_t)stream;
  const bool f64 = out_dtype == QM_F64;
  switch (precision) {
    case QM_PREC_FP64:
      rc = f64 ? density_fp64<double>(ctx, mol, s, g, subgrid, (double*)out)
               : density_fp64<float>(ctx, mol, s, g, subgrid, (float*)out);
      break;
    case QM_PREC_F16X3:
      rc = f64 ? density_split<_Float16, double>(ctx, mol, s, g, subgrid, (double*)out)
               : density_split<_Float16, float>(ctx, mol, s, g, subgrid, (float*)out);
      break;
    case QM_PREC_BF16X3:
      rc = f64 ? density_split<__bf16, double>(ctx, mol, s, g, subgrid, (double*)out)
               : density_split<__bf16, float>(ctx, mol, s, g, subgrid, (float*)out);
      break;
    default:
      return qm_fail(ctx, QM_EINVAL, "density: bad precision %d", precision);
  }
  if (rc) return rc;
  if (normalize || raw_total) {
    double* dtot = nullptr;
    rc = qm_ws(ctx, s, WS_SCALAR, 256, (void**)&dtot);
    if (rc) return rc;
    rc = qm_reduce_dev(ctx, s, QM_RED_SUM, out_dtype, out, nullptr, npoints(g), dtot);
    if (rc) return rc;
    if (raw_total) {
      QM_CHECK_HIP(ctx, hipMemcpyAsync(ctx->host_scalars, dtot, sizeof(double), hipMemcpyDeviceToHost, s));
    }
    if (normalize) {
      rc = qm_scale_dev(ctx, s, out_dtype, out, npoints(g), nullptr, mol->qtot, dtot);
      if (rc) return rc;
    }
    if (raw_total) {
      QM_CHECK_HIP(ctx, hipStreamSynchronize(s));
      *raw_total = ctx->host_scalars[0];
    }
  }
  return QM_OK;
}

int qm_basis_points(qm_ctx* ctx, qm_mol* mol, void* stream, const double* pts, int64_t npts, double* phi) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  QM_REQUIRE(ctx, mol && pts && phi, QM_EINVAL, "basis_points: null argument");
  QM_REQUIRE(ctx, npts > 0 && npts < ((int64_t)1 << 30), QM_ELIMIT, "basis_points: bad point count");
  // writes phi[f * npts + t] for f < nbf (ld = npts); rows beyond nbf are not stored
  const int n = (int)npts;
  const MolView mv = mol->view();
  GeomView gv{};
  gv.nx = 1; gv.ny = 1; gv.nz = 1; gv.h = 1.0;
  hipLaunchKernelGGL(k_phi_f64, dim3((n + kPhiThreads - 1) / kPhiThreads), dim3(kPhiThreads), 0,
                     (hipStream_t)stream, mv, gv, (int64_t)0, n, 1, n, pts, n, mol->nbf, phi);
  QM_CHECK_LAUNCH(ctx);
  return QM_OK;
}

int qm_tune_set(qm_ctx* ctx, int knob, int value) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  if (knob == QM_TUNE_SPLIT_VARIANT) {
    QM_REQUIRE(ctx, split_tiled(value) || value == 4 || value == 5 || value == 64 || value == 128, QM_EINVAL,
               "tune: split variant must be 0..3 (tiles), 4 (fused), 5 (wide tiles) or 64 / 128 (rows per "
               "workgroup)");
    ctx->split_tm = value;
    return QM_OK;
  }
  if (knob == QM_TUNE_F64_VARIANT) {
    QM_REQUIRE(ctx, value >= 0 && value <= 4, QM_EINVAL, "tune: f64 variant must be 0..4");
    ctx->f64_variant = value;
    return QM_OK;
  }
  if (knob == QM_TUNE_JACOBI_ZS) {
    QM_REQUIRE(ctx, value >= 0 && value <= 4096, QM_EINVAL, "tune: Jacobi planes must be 0..4096");
    ctx->jacobi_zs = value;
    return QM_OK;
  }
  if (knob == QM_TUNE_JACOBI_PAIRS) {
    QM_REQUIRE(ctx, value == 0 || value == 1, QM_EINVAL, "tune: Jacobi pairs must be 0 or 1");
    ctx->jacobi_pairs = value;
    return QM_OK;
  }
  if (knob == QM_TUNE_TILE_SKIP) {
    QM_REQUIRE(ctx, value == 0 || value == 1, QM_EINVAL, "tune: tile skip must be 0 or 1");
    ctx->tile_skip = value;
    return QM_OK;
  }
  return qm_fail(ctx, QM_EINVAL, "tune: unknown knob %d", knob);
}

int qm_tune_get(qm_ctx* ctx, int knob, int* value) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  QM_REQUIRE(ctx, value != nullptr, QM_EINVAL, "tune: null output");
  if (knob == QM_TUNE_SPLIT_VARIANT) {
    *value = ctx->split_tm;
    return QM_OK;
  }
  if (knob == QM_TUNE_F64_VARIANT) {
    *value = ctx->f64_variant;
    return QM_OK;
  }
  if (knob == QM_TUNE_JACOBI_ZS) {
    *value = ctx->jacobi_zs;
    return QM_OK;
  }
  if (knob == QM_TUNE_JACOBI_PAIRS) {
    *value = ctx->jacobi_pairs;
    return QM_OK;
  }
  if (knob == QM_TUNE_TILE_SKIP) {
    *value = ctx->tile_skip;
    return QM_OK;
  }
  return qm_fail(ctx, QM_EINVAL, "tune: unknown knob %d", knob);
}

int qm_autotune_density(qm_ctx* ctx, qm_mol* mol, void* stream, const qm_geom* g, int precision, double* best_ms) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  if ((rc = qm_check_geom(ctx, g))) return rc;
  QM_REQUIRE(ctx, mol && mol->has_dm, QM_EINVAL, "autotune: molecule without density matrix");
  hipStream_t s = (hipStream_t)stream;
  // a slab of whole xy planes, about 1.6M points
  qm_geom slab = *g;
  const int64_t plane = (int64_t)g->nx * g->ny;
  slab.nz = (int)std::max<int64_t>(1, std::min<int64_t>(g->nz, ((int64_t)1 << 21) / std::max<int64_t>(plane, 1)));
  void* out = nullptr;
  if ((rc = qm_ws(ctx, s, WS_MISC, sizeof(double) * (size_t)npoints(&slab), &out))) return rc;
  const bool split = precision == QM_PREC_F16X3 || precision == QM_PREC_BF16X3;
  std::vector<int> cands = split ? std::vector<int>{0, 1, 2, 3, 4, 5, 64, 128} : std::vector<int>{0, 1, 2, 3, 4};
  int& knob = split ? ctx->split_tm : ctx->f64_variant;
  hipEvent_t a, b;
  QM_CHECK_HIP(ctx, hipEventCreate(&a));
  QM_CHECK_HIP(ctx, hipEventCreate(&b));
  double best = 1e300;
  int best_v = knob;
  for (int v : cands) {
    knob = v;
    double t = 1e300;
    for (int rep = 0; rep < 3; ++rep) {   // first repetition warms caches and workspaces
      QM_CHECK_HIP(ctx, hipEventRecord(a, s));
      rc = qm_density(ctx, mol, stream, &slab, 1, precision, split ? QM_F32 : QM_F64, out, 0, nullptr);
      if (rc) break;
      QM_CHECK_HIP(ctx, hipEventRecord(b, s));
      QM_CHECK_HIP(ctx, hipEventSynchronize(b));
      float ms = 0.f;
      QM_CHECK_HIP(ctx, hipEventElapsedTime(&ms, a, b));
      if (rep > 0) t = std::min(t, (double)ms);
    }
    if (rc) break;
    if (t < best) {
      best = t;
      best_v = v;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  knob = best_v;
  if (rc) return rc;
  if (best_ms) *best_ms = best;
  return QM_OK;
}

}  // extern "C"

// ===========================================================================
// Batched density: molecules of uneven size in fused launches.
//
// Every molecule's grid is cut into chunks of at most kBatchChunk points; the
// chunks are packed into rounds whose basis values fit one Phi workspace.  A
// round is one basis-value launch over all its chunks plus one contraction
// launch per column-tile class NT (chunks sorted by NT, so waves are not left
// idle by small molecules).  A workgroup finds its chunk by binary search on
// the chunk table's first-block numbers and runs the same per-tile code as the
// single-molecule kernels, so a molecule's density is identical alone and in
// a batch.
namespace {

constexpr int kBatchChunk = 1 << 20;

struct BChunk {
  MolView mol;
  const float2* l2e;
  GeomView g;
  int64_t v0;
  int nvox;
  int ld;
  int nbf_pad;
  int nt;
  void* phi;
  float* pscale;
  const void* ph;       // P64 / M64 (FP64) or 16-bit hi plane
  const void* pl;       // 16-bit lo plane
  void* out;
  double scale;
  double* part;         // tiled kernels: per-column-block partial sums [ncb][ld]
  float phis;           // pre-split basis values: the molecule's scale S
  int ncb;
  int blk0_phi;         // first block in the round's basis-value (and tiled finish) launch
  int blk0;             // first block in its contraction launch
  const int2* kg;       // fused kernel: the molecule's function-group table
  uint32_t* mask;       // FP64 LDS-DMA tiles: activity bits of the row tiles' 16-function groups (or null)
  int nw;               // mask words per row tile
};

__device__ __forceinline__ int find_chunk(const BChunk* __restrict__ c, int n, int b, bool phi_launch) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int first = phi_launch ? c[mid].blk0_phi : c[mid].blk0;
    if (first <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kPhiThreads) void kb_phi_f64(const BChunk* __restrict__ c, int n) {
  const int i = find_chunk(c, n, blockIdx.x, true);
  const BChunk& q = c[i];
  phi_f64_body(blockIdx.x - q.blk0_phi, q.mol, q.g, q.v0, q.nvox, 1, q.nvox, nullptr, q.ld, q.nbf_pad,
               (double*)q.phi);
}

__global__ __launch_bounds__(kPhiThreads) void kb_phi_f32(const BChunk* __restrict__ c, int n) {
  const int i = find_chunk(c, n, blockIdx.x, true);
  const BChunk& q = c[i];
  phi_f32_body(blockIdx.x - q.blk0_phi, q.mol, q.l2e, q.g, q.v0, q.nvox, 1, q.nvox, q.ld, q.nbf_pad, (float*)q.phi,
               q.pscale);
}

template <typename T16>
__global__ __launch_bounds__(kPhiThreads) void kb_phi_split(const BChunk* __restrict__ c, int n) {
  const int i = find_chunk(c, n, blockIdx.x, true);
  const BChunk& q = c[i];
  T16* hi = (T16*)q.phi;
  phi_split_body<T16>(blockIdx.x - q.blk0_phi, q.mol, q.l2e, q.g, q.v0, q.nvox, 1, q.nvox, q.ld, q.nbf_pad, q.phis,
                      hi, hi + (int64_t)q.nbf_pad * q.ld);
}

template <int NT, typename TO>
__global__ __launch_bounds__(kWG) void kb_contract_f64(const BChunk* __restrict__ c, int n) {
  const int i = find_chunk(c, n, blockIdx.x, false);
  const BChunk& q = c[i];
  contract_f64_body<NT, TO>(blockIdx.x - q.blk0, (const double*)q.phi, q.ld, (const double*)q.ph, q.nbf_pad,
                            q.nvox, q.scale, (TO*)q.out, q.v0);
}

template <typename T16, int NT, int TM, typename TO>
__global__ __launch_bounds__(kWG) void kb_contract_split(const BChunk* __restrict__ c, int n) {
  const int i = find_chunk(c, n, blockIdx.x, false);
  const BChunk& q = c[i];
  contract_split_body<T16, NT, TM, TO>(blockIdx.x - q.blk0, (const float*)q.phi, q.ld, q.pscale,
                                       (const T16*)q.ph, (const T16*)q.pl, q.nbf_pad, q.nvox, q.scale,
                                       (TO*)q.out, q.v0);
}

template <int NT, typename TO>
__global__ __launch_bounds__(kWG) void kb_contract_f64_direct(const BChunk* __restrict__ c, int n) {
  const int i = find_chunk(c, n, blockIdx.x, false);
  const BChunk& q = c[i];
  contract_f64_direct_body<NT, TO>(blockIdx.x - q.blk0, (const double*)q.phi, q.ld, (const double*)q.ph, q.nbf_pad,
                                   q.nvox, q.scale, (TO*)q.out, q.v0);
}

template <typename TO>
void launch_batch_f64(hipStream_t s, int nt, const BChunk* c, int n, int blocks, int direct) {
  if (direct) {
    switch (nt) {
      case 1: hipLaunchKernelGGL((kb_contract_f64_direct<1, TO>), dim3(blocks), dim3(kWG), 0, s, c, n); return;
      case 2: hipLaunchKernelGGL((kb_contract_f64_direct<2, TO>), dim3(blocks), dim3(kWG), 0, s, c, n); return;
      case 3: hipLaunchKernelGGL((kb_contract_f64_direct<3, TO>), dim3(blocks), dim3(kWG), 0, s, c, n); return;
      case 4: hipLaunchKernelGGL((kb_contract_f64_direct<4, TO>), dim3(blocks), dim3(kWG), 0, s, c, n); return;
      default: hipLaunchKernelGGL((kb_contract_f64_direct<5, TO>), dim3(blocks), dim3(kWG), 0, s, c, n); return;
    }
  }
  switch (nt) {
    case 1: hipLaunchKernelGGL((kb_contract_f64<1, TO>), dim3(blocks), dim3(kWG), 0, s, c, n); break;
    case 2: hipLaunchKernelGGL((kb_contract_f64<2, TO>), dim3(blocks), dim3(kWG), 0, s, c, n); break;
    case 3: hipLaunchKernelGGL((kb_contract_f64<3, TO>), dim3(blocks), dim3(kWG), 0, s, c, n); break;
    case 4: hipLaunchKernelGGL((kb_contract_f64<4, TO>), dim3(blocks), dim3(kWG), 0, s, c, n); break;
    default: hipLaunchKernelGGL((kb_contract_f64<5, TO>), dim3(blocks), dim3(kWG), 0, s, c, n); break;
  }
}

template <typename T16, typename TO>
void launch_batch_split(hipStream_t s, int nt, int tm, const BChunk* c, int n, int blocks) {
#define QM_BSPLIT(NT)                                                                                    \
  if (tm == 64)                                                                                          \
    hipLaunchKernelGGL((kb_contract_split<T16, NT, 64, TO>), dim3(blocks), dim3(kWG), 0, s, c, n);      \
  else                                                                                                   \
    hipLaunchKernelGGL((kb_contract_split<T16, NT, 128, TO>), dim3(blocks), dim3(kWG), 0, s, c, n)
  switch (nt) {
    case 1: QM_BSPLIT(1); break;
    case 2: QM_BSPLIT(2); break;
    case 3: QM_BSPLIT(3); break;
    case 4: QM_BSPLIT(4); break;
    default: QM_BSPLIT(5); break;
  }
#undef QM_BSPLIT
}

__global__ __launch_bounds__(kTWG) void kb_contract_f64_rows(const BChunk* __restrict__ c, int n) {
  const BChunk& q = c[find_chunk(c, n, blockIdx.x, false)];
  int rt, cb;
  if (!tile_coords(blockIdx.x - q.blk0, q.ld / kTT, q.ncb, rt, cb)) return;
  tiled_f64_rows_body(rt, cb, (const double*)q.phi, q.ld, (const double*)q.ph, q.nbf_pad, q.part, q.ld);
}

// activity bits of every chunk of a round, numbered as the basis-value launch (two 128-point tiles
// per block of that launch); chunks without a mask are left alone
__global__ __launch_bounds__(kTT) void kb_tile_mask(const BChunk* __restrict__ c, int n) {
  const int pb = blockIdx.x >> 1;
  const BChunk& q = c[find_chunk(c, n, pb, true)];
  const int rt = (pb - q.blk0_phi) * 2 + (blockIdx.x & 1);
  if (!q.mask || rt >= q.ld / kTT) return;   // uniform per workgroup
  tile_mask_body(q.mol, q.g, q.v0, q.nvox, 1, q.nvox, rt, q.nw, q.mask);
}

template <bool DMA>
__global__ __launch_bounds__(kTWG) void kb_contract_f64_tiles(const BChunk* __restrict__ c, int n) {
  const BChunk& q = c[find_chunk(c, n, blockIdx.x, false)];
  int rt, cb;
  if (!tile_coords(blockIdx.x - q.blk0, q.ld / kTT, q.ncb, rt, cb)) return;
  if (DMA)
    tiled_f64_dma_body<kF64Ks, 32 / kF64Ks>(rt, cb, (const double*)q.phi, q.ld, (const double*)q.ph, q.nbf_pad,
                                            q.mol.nbf, q.ncb, q.part, q.ld, q.mask, q.nw);
  else tiled_f64_body(rt, cb, (const double*)q.phi, q.ld, (const double*)q.ph, q.nbf_pad, q.part, q.ld);
}

template <typename T16, int WR, int WC, bool PRE = false>
__global__ __launch_bounds__(64 * WR * WC) void kb_contract_split_tiles(const BChunk* __restrict__ c, int n) {
  const BChunk& q = c[find_chunk(c, n, blockIdx.x, false)];
  int rt, cb;
  if (!tile_coords(blockIdx.x - q.blk0, q.ld / kTT, q.ncb, rt, cb)) return;
  tiled_split_body<T16, WR, WC, PRE>(rt, cb, q.phi, q.ld, q.pscale, (const T16*)q.ph, (const T16*)q.pl, q.nbf_pad,
                                     q.part, q.ld);
}

template <typename T16>
__global__ __launch_bounds__(512, 4) void kb_contract_split_dma(const BChunk* __restrict__ c, int n) {
  const BChunk& q = c[find_chunk(c, n, blockIdx.x, false)];
  int rt, cb;
  if (!tile_coords(blockIdx.x - q.blk0, q.ld / kTT, q.ncb, rt, cb)) return;
  tiled_dma_body<T16>(rt, cb, (const T16*)q.phi, q.ld, (const T16*)q.ph, (const T16*)q.pl, q.nbf_pad, q.mol.nbf, q.ncb,
                      q.part, q.ld);
}

template <typename T16>
void launch_batch_split_tiles(hipStream_t s, int v, int blocks, const BChunk* c, int n) {
  switch (v) {
    case 1: hipLaunchKernelGGL((kb_contract_split_tiles<T16, 4, 2, true>), dim3(blocks), dim3(512), 0, s, c, n); break;
    case 2: hipLaunchKernelGGL((kb_contract_split_tiles<T16, 2, 2>), dim3(blocks), dim3(256), 0, s, c, n); break;
    case 3: hipLaunchKernelGGL(kb_contract_split_dma<T16>, dim3(blocks), dim3(512), 0, s, c, n); break;
    default: hipLaunchKernelGGL((kb_contract_split_tiles<T16, 4, 2>), dim3(blocks), dim3(512), 0, s, c, n); break;
  }
}

template <typename T16>
__global__ __launch_bounds__(512) void kb_density_fused(const BChunk* __restrict__ c, int n) {
  __shared__ __attribute__((aligned(16))) unsigned char tab[kFTab];
  const BChunk& q = c[find_chunk(c, n, blockIdx.x, false)];
  int rt, cg;
  if (!tile_coords(blockIdx.x - q.blk0, q.ld / kFP, q.ncb, rt, cg)) return;
  fused_load_tables(tab, q.mol, q.l2e, q.kg, q.nbf_pad / 8);
  __syncthreads();
  fused_tile<T16>(rt, cg, tab, q.mol, q.g, q.v0, q.nvox, 1, q.nvox, (const T16*)q.ph, (const T16*)q.pl, q.nbf_pad,
                  q.mol.nbf, q.ncb, q.phis, q.part, q.ld);
}

// same sums as k_tiles_finish; one thread per point, blocks numbered like the basis-value launch
template <typename TO>
__global__ __launch_bounds__(kPhiThreads) void kb_tiles_finish(const BChunk* __restrict__ c, int n, int split) {
  const BChunk& q = c[find_chunk(c, n, blockIdx.x, true)];
  const int t = (blockIdx.x - q.blk0_phi) * kPhiThreads + threadIdx.x;
  if (t >= q.nvox) return;
  double v = 0.0;
  for (int cb = 0; cb < q.ncb; ++cb) v += q.part[(int64_t)cb * q.ld + t];
  if (split) v /= (double)q.pscale[t];
  store_rho<TO>((TO*)q.out, q.v0 + t, v * q.scale);
}

struct Round {
  size_t first, count;                      // chunk range in the table (sorted by NT within the round)
  std::vector<std::pair<int, std::pair<size_t, size_t>>> groups;   // (nt, [first, count))
  std::vector<int> group_blocks;
  int phi_blocks;
};

// Charge normalisation of every molecule of a batch: Q <- Q sum Z / sum Q (device-side factors).
int batch_normalize(qm_ctx* ctx, hipStream_t s, int nmol, qm_mol* const* mols, const qm_geom* geoms,
                    void* const* outs, int out_dtype, int normalize) {
  if (!normalize) return QM_OK;
  double* tot = nullptr;
  int rc = QM_OK;
  if ((rc = qm_ws(ctx, s, WS_BATCH2, sizeof(double) * (size_t)nmol, (void**)&tot))) return rc;
  for (int i = 0; i < nmol; ++i) {
    if ((rc = qm_reduce_dev(ctx, s, QM_RED_SUM, out_dtype, outs[i], nullptr, npoints(&geoms[i]), tot + i))) return rc;
    if ((rc = qm_scale_dev(ctx, s, out_dtype, outs[i], npoints(&geoms[i]), nullptr, mols[i]->qtot, tot + i)))
      return rc;
  }
  return QM_OK;
}

// Batched split variant 4: chunks of every molecule in rounds whose partial sums fit kFPartBudget;
// one fused launch and one finish launch per round.
constexpr size_t kFPartBudget = (size_t)1 << 30;

template <typename T16>
int density_batch_fused(qm_ctx* ctx, hipStream_t s, int nmol, qm_mol* const* mols, const qm_geom* geoms,
                        void* const* outs, int out_dtype) {
  const bool half = std::is_same<T16, _Float16>::value;
  std::vector<BChunk> table;
  std::vector<std::pair<size_t, size_t>> rounds;   // [first, count) of each round
  std::vector<std::pair<int, int>> round_blocks;   // (fused blocks, finish blocks)
  size_t part = 0, first = 0, max_part = 0;
  int blocks = 0, fblocks = 0;
  auto flush = [&]() {
    if (table.size() == first) return;
    rounds.push_back({first, table.size() - first});
    round_blocks.push_back({blocks, fblocks});
    max_part = std::max(max_part, part);
    first = table.size();
    part = 0;
    blocks = 0;
    fblocks = 0;
  };
  for (int i = 0; i < nmol; ++i) {
    qm_mol* m = mols[i];
    const int np = m->nbf_pad;
    const int64_t N = npoints(&geoms[i]);
    const float S = split_phi_scale(m);
    const int ncg = fused_groups(m->nbf);
    const uint16_t* base = (const uint16_t*)(half ? m->P16 : m->Pb16);
    const double h3 = geoms[i].step * geoms[i].step * geoms[i].step;
    for (int64_t v0 = 0; v0 < N; v0 += kBatchChunk) {
      BChunk c{};
      c.mol = m->view();
      c.l2e = m->prims_l2e;
      c.kg = m->kgroups;
      c.g = geom_view(&geoms[i]);
      c.v0 = v0;
      c.nvox = (int)std::min<int64_t>(kBatchChunk, N - v0);
      c.ld = (int)round_up(c.nvox, kFP);
      c.nbf_pad = np;
      c.ncb = ncg;
      c.ph = base;
      c.pl = base + (size_t)np * np;
      c.phis = S;
      c.scale = h3 / m->pfac / ((double)S * S);
      c.out = outs[i];
      const size_t bytes = sizeof(double) * (size_t)ncg * c.ld;
      if (part + bytes > kFPartBudget) flush();
      c.part = (double*)(uintptr_t)part;   // offset for now, rebased below
      part += bytes;
      c.blk0 = blocks;
      blocks += tile_grid(c.ld / kFP, ncg);
      c.blk0_phi = fblocks;
      fblocks += (c.ld + kPhiThreads - 1) / kPhiThreads;
      table.push_back(c);
    }
  }
  flush();
  void *pt = nullptr, *dtab = nullptr;
  int rc = QM_OK;
  if ((rc = qm_ws(ctx, s, WS_PART, std::max<size_t>(max_part, 16), &pt))) return rc;
  if ((rc = qm_ws(ctx, s, WS_BATCH, sizeof(BChunk) * table.size(), &dtab))) return rc;
  for (BChunk& c : table) c.part = (double*)((char*)pt + (size_t)(uintptr_t)c.part);
  QM_CHECK_HIP(ctx, hipMemcpyAsync(dtab, table.data(), sizeof(BChunk) * table.size(), hipMemcpyHostToDevice, s));
  QM_CHECK_HIP(ctx, hipStreamSynchronize(s));    // the host table is pageable and goes out of scope
  const BChunk* dt = (const BChunk*)dtab;
  for (size_t r = 0; r < rounds.size(); ++r) {
    const BChunk* rc0 = dt + rounds[r].first;
    const int rn = (int)rounds[r].second;
    {
      ProfScope pc(ctx, s, QM_K_DENSITY_SPLIT);
      hipLaunchKernelGGL(kb_density_fused<T16>, dim3(round_blocks[r].first), dim3(512), 0, s, rc0, rn);
    }
    QM_CHECK_LAUNCH(ctx);
    if (out_dtype == QM_F64)
      hipLaunchKernelGGL(kb_tiles_finish<double>, dim3(round_blocks[r].second), dim3(kPhiThreads), 0, s, rc0, rn, 0);
    else
      hipLaunchKernelGGL(kb_tiles_finish<float>, dim3(round_blocks[r].second), dim3(kPhiThreads), 0, s, rc0, rn, 0);
    QM_CHECK_LAUNCH(ctx);
  }
  return QM_OK;
}

}  // namespace

extern "C" int qm_density_batch(qm_ctx* ctx, void* stream, int nmol, qm_mol* const* mols, const qm_geom* geoms,
                                void* const* outs, int precision, int out_dtype, int normalize) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  QM_REQUIRE(ctx, nmol >= 0 && (nmol == 0 || (mols && geoms && outs)), QM_EINVAL, "density_batch: bad arguments");
  QM_REQUIRE(ctx, precision == QM_PREC_FP64 || precision == QM_PREC_F16X3 || precision == QM_PREC_BF16X3,
             QM_EINVAL, "density_batch: bad precision");
  QM_REQUIRE(ctx, out_dtype == QM_F64 || out_dtype == QM_F32, QM_EINVAL, "density_batch: bad output dtype");
  if (nmol == 0) return QM_OK;
  for (int i = 0; i < nmol; ++i) {
    QM_REQUIRE(ctx, mols[i] && mols[i]->has_dm && outs[i], QM_EINVAL, "density_batch: item %d incomplete", i);
    if ((rc = qm_check_geom(ctx, &geoms[i]))) return rc;
  }
  hipStream_t s = (hipStream_t)stream;
  const bool f64 = precision == QM_PREC_FP64;
  bool all_fit = true;
  for (int i = 0; i < nmol; ++i) all_fit = all_fit && fused_fits(mols[i]);
  if (!f64 && ctx->split_tm == 4 && all_fit) {
    rc = precision == QM_PREC_F16X3 ? density_batch_fused<_Float16>(ctx, s, nmol, mols, geoms, outs, out_dtype)
                                    : density_batch_fused<__bf16>(ctx, s, nmol, mols, geoms, outs, out_dtype);
    if (rc) return rc;
    return batch_normalize(ctx, s, nmol, mols, geoms, outs, out_dtype, normalize);
  }
  const size_t elt = f64 ? sizeof(double) : sizeof(float);
  const int sv = ctx->split_tm == 4 ? 3 : ctx->split_tm;   // fused variant without room for the tables
  const int tm = f64 ? 64 : sv;
  const bool tiled = f64 ? ctx->f64_variant >= 2 : split_tiled(sv);
  // basis values written scaled and split (same bytes as FP32)
  const bool pre = !f64 && split_pre(sv);
  // chunks per molecule, packed into rounds of at most kPhiBudget bytes of basis values
  std::vector<BChunk> table;
  std::vector<Round> rounds;
  std::vector<BChunk> cur;
  size_t cur_bytes = 0, cur_ps = 0, cur_part = 0;
  size_t mask_bytes = 0;   // activity bits of every chunk (FP64 LDS-DMA tiles), kept for the whole call
  auto flush = [&]() {
    if (cur.empty()) return;
    Round r;
    r.first = table.size();
    r.count = cur.size();
    std::stable_sort(cur.begin(), cur.end(), [](const BChunk& a, const BChunk& b) { return a.nt < b.nt; });
    int pb = 0;
    for (auto& c : cur) {
      c.blk0_phi = pb;
      pb += (c.ld + kPhiThreads - 1) / kPhiThreads;
    }
    r.phi_blocks = pb;
    size_t k = 0;
    while (k < cur.size()) {
      size_t e = k;
      int blocks = 0;
      while (e < cur.size() && cur[e].nt == cur[k].nt) {
        cur[e].blk0 = blocks;
        blocks += tiled ? tile_grid(cur[e].ld / kTT, cur[e].ncb) : cur[e].ld / tm;   // tiled: multiples of 8
        ++e;
      }
      r.groups.push_back({cur[k].nt, {r.first + k, e - k}});
      r.group_blocks.push_back(blocks);
      k = e;
    }
    table.insert(table.end(), cur.begin(), cur.end());
    rounds.push_back(r);
    cur.clear();
    cur_bytes = 0;
    cur_ps = 0;
    cur_part = 0;
  };
  for (int i = 0; i < nmol; ++i) {
    qm_mol* m = mols[i];
    const int np = m->nbf_pad;
    const int64_t N = npoints(&geoms[i]);
    const int64_t cap = std::max<int64_t>(kTileM, std::min<int64_t>(kBatchChunk,
                                          (int64_t)(kPhiBudget / ((size_t)np * elt)) / kTileM * kTileM));
    for (int64_t v0 = 0; v0 < N; v0 += cap) {
      BChunk c{};
      c.mol = m->view();
      c.l2e = m->prims_l2e;
      c.g = geom_view(&geoms[i]);
      c.v0 = v0;
      c.nvox = (int)std::min<int64_t>(cap, N - v0);
      c.ld = (int)round_up(c.nvox, kTileM);
      c.nbf_pad = np;
      c.nt = tiled ? 0 : pick_nt(np);     // tiled chunks form one launch per round
      c.ncb = (f64 && ctx->f64_variant == 3) ? np / kRN : (np + kTT - 1) / kTT;
      if (f64 && ctx->f64_variant == 4 && ctx->tile_skip && (np / 16 + 31) / 32 <= 64) {
        c.nw = (np / 16 + 31) / 32;
        c.mask = (uint32_t*)(uintptr_t)mask_bytes;   // offset for now, rebased below
        mask_bytes += sizeof(uint32_t) * (size_t)c.nw * (c.ld / kTT);
      }
      const size_t bytes = (size_t)np * c.ld * elt;
      if (cur_bytes + bytes > kPhiBudget) flush();
      c.phi = (void*)(uintptr_t)cur_bytes;            // offsets for now, rebased below
      c.pscale = (float*)(uintptr_t)cur_ps;
      c.part = (double*)(uintptr_t)cur_part;
      cur_bytes += bytes;
      cur_ps += (size_t)c.ld * sizeof(float);
      if (tiled) cur_part += sizeof(double) * (size_t)c.ncb * c.ld;
      const double h3 = geoms[i].step * geoms[i].step * geoms[i].step;
      if (f64) {
        c.ph = ctx->f64_variant == 1 ? m->P64 : m->M64;   // the LDS-staged kernel takes the triangular form
        c.pl = nullptr;
        c.scale = h3;
      } else {
        const bool half = precision == QM_PREC_F16X3;
        const uint16_t* base = (const uint16_t*)(half ? m->P16 : m->Pb16);
        c.ph = base;
        c.pl = base + (size_t)np * np;
        c.phis = split_phi_scale(m);
        c.scale = pre ? h3 / m->pfac / ((double)c.phis * c.phis) : h3 / m->pfac;
      }
      c.out = outs[i];
      cur.push_back(c);
    }
  }
  flush();
  // workspaces: basis values of one round, point scales, and the chunk table
  size_t max_bytes = 0, max_ps = 0, max_part = 0;
  for (const Round& r : rounds) {
    size_t b = 0, p = 0, q = 0;
    for (size_t k = 0; k < r.count; ++k) {
      const BChunk& c = table[r.first + k];
      b = std::max(b, (size_t)(uintptr_t)c.phi + (size_t)c.nbf_pad * c.ld * elt);
      p = std::max(p, (size_t)(uintptr_t)c.pscale + (size_t)c.ld * sizeof(float));
      if (tiled) q = std::max(q, (size_t)(uintptr_t)c.part + sizeof(double) * (size_t)c.ncb * c.ld);
    }
    max_bytes = std::max(max_bytes, b);
    max_ps = std::max(max_ps, p);
    max_part = std::max(max_part, q);
  }
  void *phi = nullptr, *ps = nullptr, *dtab = nullptr, *pt = nullptr;
  if ((rc = qm_ws(ctx, s, WS_PHI, max_bytes, &phi))) return rc;
  if ((rc = qm_ws(ctx, s, WS_PHI2, std::max<size_t>(max_ps, 16), &ps))) return rc;
  if (tiled && (rc = qm_ws(ctx, s, WS_PART, max_part, &pt))) return rc;
  if ((rc = qm_ws(ctx, s, WS_BATCH, sizeof(BChunk) * table.size(), &dtab))) return rc;
  void* mk = nullptr;
  if (mask_bytes && (rc = qm_ws(ctx, s, WS_MASK, mask_bytes, &mk))) return rc;
  for (BChunk& c : table) {
    c.mask = c.nw ? (uint32_t*)((char*)mk + (size_t)(uintptr_t)c.mask) : nullptr;
    c.phi = (char*)phi + (size_t)(uintptr_t)c.phi;
    c.pscale = (float*)((char*)ps + (size_t)(uintptr_t)c.pscale);
    c.part = tiled ? (double*)((char*)pt + (size_t)(uintptr_t)c.part) : nullptr;
  }
  QM_CHECK_HIP(ctx, hipMemcpyAsync(dtab, table.data(), sizeof(BChunk) * table.size(), hipMemcpyHostToDevice, s));
  QM_CHECK_HIP(ctx, hipStreamSynchronize(s));    // the host table is pageable and goes out of scope
  const BChunk* dt = (const BChunk*)dtab;
  for (const Round& r : rounds) {
    {
      ProfScope pp(ctx, s, f64 ? QM_K_PHI : QM_K_PHI_SPLIT);
      if (f64) {
        hipLaunchKernelGGL(kb_phi_f64, dim3(r.phi_blocks), dim3(kPhiThreads), 0, s, dt + r.first, (int)r.count);
        if (mask_bytes)
          hipLaunchKernelGGL(kb_tile_mask, dim3(2 * r.phi_blocks), dim3(kTT), 0, s, dt + r.first, (int)r.count);
      } else if (pre && precision == QM_PREC_F16X3)
        hipLaunchKernelGGL(kb_phi_split<_Float16>, dim3(r.phi_blocks), dim3(kPhiThreads), 0, s, dt + r.first,
                           (int)r.count);
      else if (pre)
        hipLaunchKernelGGL(kb_phi_split<__bf16>, dim3(r.phi_blocks), dim3(kPhiThreads), 0, s, dt + r.first,
                           (int)r.count);
      else
        hipLaunchKernelGGL(kb_phi_f32, dim3(r.phi_blocks), dim3(kPhiThreads), 0, s, dt + r.first, (int)r.count);
    }
    QM_CHECK_LAUNCH(ctx);
    for (size_t gi = 0; gi < r.groups.size(); ++gi) {
      const int nt = r.groups[gi].first;
      const BChunk* gc = dt + r.groups[gi].second.first;
      const int gn = (int)r.groups[gi].second.second;
      const int blocks = r.group_blocks[gi];
      ProfScope pc(ctx, s, f64 ? QM_K_DENSITY_CONTRACT : QM_K_DENSITY_SPLIT);
      if (tiled) {
        if (f64 && ctx->f64_variant == 3)
          hipLaunchKernelGGL(kb_contract_f64_rows, dim3(blocks), dim3(kTWG), 0, s, gc, gn);
        else if (f64 && ctx->f64_variant == 4)
          hipLaunchKernelGGL(kb_contract_f64_tiles<true>, dim3(blocks), dim3(kTWG), 0, s, gc, gn);
        else if (f64)
          hipLaunchKernelGGL(kb_contract_f64_tiles<false>, dim3(blocks), dim3(kTWG), 0, s, gc, gn);
        else if (precision == QM_PREC_F16X3)
          launch_batch_split_tiles<_Float16>(s, sv, blocks, gc, gn);
        else
          launch_batch_split_tiles<__bf16>(s, sv, blocks, gc, gn);
        QM_CHECK_LAUNCH(ctx);
        // finish over the whole round: chunk blocks numbered as in the basis-value launch
        const int per_point = !f64 && !pre;   // per-point scales to divide out
        if (out_dtype == QM_F64)
          hipLaunchKernelGGL(kb_tiles_finish<double>, dim3(r.phi_blocks), dim3(kPhiThreads), 0, s, dt + r.first,
                             (int)r.count, per_point);
        else
          hipLaunchKernelGGL(kb_tiles_finish<float>, dim3(r.phi_blocks), dim3(kPhiThreads), 0, s, dt + r.first,
                             (int)r.count, per_point);
      } else if (f64) {
        if (out_dtype == QM_F64) launch_batch_f64<double>(s, nt, gc, gn, blocks, ctx->f64_variant == 1);
        else launch_batch_f64<float>(s, nt, gc, gn, blocks, ctx->f64_variant == 1);
      } else if (precision == QM_PREC_F16X3) {
        if (out_dtype == QM_F64) launch_batch_split<_Float16, double>(s, nt, tm, gc, gn, blocks);
        else launch_batch_split<_Float16, float>(s, nt, tm, gc, gn, blocks);
      } else {
        if (out_dtype == QM_F64) launch_batch_split<__bf16, double>(s, nt, tm, gc, gn, blocks);
        else launch_batch_split<__bf16, float>(s, nt, tm, gc, gn, blocks);
      }
      QM_CHECK_LAUNCH(ctx);
    }
  }
  return batch_normalize(ctx, s, nmol, mols, geoms, outs, out_dtype, normalize);
}
