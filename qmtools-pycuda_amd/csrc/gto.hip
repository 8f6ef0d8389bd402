// Analytic potential of the GTO density plus nuclei (FP64).
//
// Reference kernel_hartree.cu:288-869 evaluates, per grid node C and per
// basis-function pair (p >= q), hand-derived closed forms of
//   int phi_p(r) phi_q(r) / |r - C| d^3r
// for each (La, Lb) class up to D-D.  Here the same integrals come from
// McMurchie-Davidson Hermite expansions (Helgaker, Jorgensen, Olsen,
// "Molecular Electronic-Structure Theory", sec. 9.5, 9.9):
//
//   phi_a phi_b = sum_tuv E^ab_t E^ab_u E^ab_v Lambda_tuv(p, P),
//   int Lambda_tuv / r_C = (2 pi / p) R_tuv(p, P - C).
//
// The density-matrix contraction is done once per molecule on the host: every
// primitive pair becomes a Hermite charge distribution (p, P, D_tuv) with
//   D_tuv = -f (2 pi/p) K c_a c_b sum_{ma,mb} P_(a ma),(b mb) sum_cart E_t E_u E_v
// (f = 2 for distinct shells), so that V_el(C) = sum_dist sum_tuv D_tuv R_tuv.
// The device kernel then only evaluates Boys functions and R_tuv per point.
// This has no singularity at T = 0 (the reference's closed forms divide by
// powers of T = gamma |PC|^2 and lose precision there).
//
// Boys function F_n(T): for T < 40, Taylor expansion (8 terms) around the
// nearest point of a 0.1-spaced table of F_0..F_11 and e^-T (computed on the host
// from the all-positive series e^-T sum (2T)^k / ((2n+1)...(2n+2k+1))), then
// downward recursion F_{n-1} = (2T F_n + e^-T)/(2n-1); for T >= 40,
// F_0 = sqrt(pi/T)/2 and upward recursion (e^-40 is below FP64 resolution of F).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <system_error>
#include <thread>
#include <vector>

#include "qm_internal.h"

namespace {

constexpr int kLmax = 4;                 // La + Lb <= 4
constexpr int kNTUV = 35;                // (t,u,v) with t+u+v <= 4
constexpr int kBoysN = 13;               // F_0 .. F_11 (up to L + 7 for L <= 4), then e^-T
constexpr int kBoysPts = 401;            // T = 0, 0.1, ..., 40
constexpr double kBoysStep = 0.1;
constexpr double kBoysTmax = 40.0;
constexpr int kThreads = 256;
constexpr double kGtoScreen = 1e-17;      // largest dropped contribution of one distribution (hartree)

__host__ __device__ constexpr int ntuv(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
__host__ __device__ constexpr int tuv_index(int t, int u, int v) {
  return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + (u + v) * (u + v + 1) / 2 + v;
}

struct HDist {
  double p, Px, Py, Pz;
  double D[kNTUV];
};

// ------------------------------------------------------------------ device
// 1 / k for the Taylor factors and 1 / (2n - 1) for the recursion: multiplications by
// rounded reciprocals instead of FP64 divisions (each a ~10-instruction sequence)
__device__ constexpr double kInvInt[10] = {0.0, 1.0, 1.0 / 2, 1.0 / 3, 1.0 / 4, 1.0 / 5, 1.0 / 6, 1.0 / 7, 1.0 / 8, 1.0 / 9};

template <int L>
__device__ __forceinline__ void boys(double T, const double* __restrict__ tab, double* F) {
  if (T < kBoysTmax) {
    const int i = (int)(T * (1.0 / kBoysStep) + 0.5);
    const double d = i * kBoysStep - T;       // T = Ti - d, |d| <= 0.05
    const double* row = tab + i * kBoysN;
    // F_L(T) = sum_k F_{L+k}(Ti) d^k / k!
    double f = row[L + 7];
#pragma unroll
    for (int k = 6; k >= 0; --k) f = row[L + k] + f * d * kInvInt[k + 1];
    F[L] = f;
    // e^-T = e^-Ti e^d: the table's last column holds e^-Ti, e^d is its Taylor series to d^9
    // (remainder below 3e-20)
    double ed = 1.0;
#pragma unroll
    for (int k = 9; k >= 1; --k) ed = 1.0 + d * ed * kInvInt[k];
    const double ex = row[kBoysN - 1] * ed;
#pragma unroll
    for (int n = L; n >= 1; --n) F[n - 1] = (2.0 * T * F[n] + ex) * kInvInt[2 * n - 1];
  } else {
    // F_0 = sqrt(pi / T) / 2 from one reciprocal square root (no FP64 division or sqrt), 1/T = r^2
    const double r = rsqrt(T);
    const double inv = r * r;
    F[0] = 0.88622692545275801365 * r;   // sqrt(pi) / 2
#pragma unroll
    for (int n = 1; n <= L; ++n) F[n] = F[n - 1] * ((double)(2 * n - 1) * 0.5 * inv);
  }
}

// sum_tuv D_tuv R^0_tuv for t+u+v <= L, R from the Hermite recursion (Helgaker eq. 9.9.18-20)
template <int L>
__device__ __forceinline__ double hermite_contract(double X, double Y, double Z, const double* rn,
                                                   const double* __restrict__ D) {
  constexpr int N = ntuv(L);
  double cur[N], nxt[N];
  cur[0] = rn[L];
#pragma unroll
  for (int n = L - 1; n >= 0; --n) {
    const int dmax = L - n;
    nxt[0] = rn[n];
#pragma unroll
    for (int d = 1; d <= L; ++d) {
      if (d > dmax) break;
#pragma unroll
      for (int t = d; t >= 0; --t) {
#pragma unroll
        for (int u = d - t; u >= 0; --u) {
          const int v = d - t - u;
          double val;
          if (t > 0) {
            val = X * cur[tuv_index(t - 1, u, v)];
            if (t > 1) val += (double)(t - 1) * cur[tuv_index(t - 2, u, v)];
          } else if (u > 0) {
            val = Y * cur[tuv_index(t, u - 1, v)];
            if (u > 1) val += (double)(u - 1) * cur[tuv_index(t, u - 2, v)];
          } else {
            val = Z * cur[tuv_index(t, u, v - 1)];
            if (v > 1) val += (double)(v - 1) * cur[tuv_index(t, u, v - 2)];
          }
          nxt[tuv_index(t, u, v)] = val;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < N; ++q) cur[q] = nxt[q];
  }
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < N; ++q) acc += D[q] * cur[q];
  return acc;
}

template <int L>
__device__ __forceinline__ double dist_sum(const HDist* __restrict__ ds, int lo, int hi, double cx, double cy,
                                           double cz, const double* __restrict__ tab) {
  double acc = 0.0;
  for (int d = lo; d < hi; ++d) {
    const HDist& h = ds[d];
    const double p = h.p;
    const double X = h.Px - cx, Y = h.Py - cy, Z = h.Pz - cz;
    const double T = p * (X * X + Y * Y + Z * Z);
    double F[L + 1];
    boys<L>(T, tab, F);
    double rn[L + 1];
    double m = 1.0;
#pragma unroll
    for (int n = 0; n <= L; ++n) {
      rn[n] = m * F[n];
      m *= -2.0 * p;
    }
    acc += hermite_contract<L>(X, Y, Z, rn, h.D);
  }
  return acc;
}

struct DistLists {
  const HDist* d[kLmax + 1];
  int n[kLmax + 1];
};

// partial[chunk][pt] = sum over this chunk's slice of every L list
// 512 threads share one 42 KB LDS copy of the table (two workgroups per CU; 4 waves per SIMD)
constexpr int kGtoThreads = 512;
__global__ __launch_bounds__(kGtoThreads) void k_gto(DistLists lists, const double* __restrict__ boys_tab,
                                                     const double* __restrict__ pts, GeomView g, int64_t npts,
                                                     int nchunk, double* __restrict__ partial) {
  __shared__ double tab[kBoysPts * kBoysN];
  for (int i = threadIdx.x; i < kBoysPts * kBoysN; i += kGtoThreads) tab[i] = boys_tab[i];
  __syncthreads();
  const int64_t t = (int64_t)blockIdx.x * kGtoThreads + threadIdx.x;
  const int chunk = blockIdx.y;
  double cx, cy, cz;
  const int64_t tt = t < npts ? t : npts - 1;
  if (pts) {
    cx = pts[3 * tt];
    cy = pts[3 * tt + 1];
    cz = pts[3 * tt + 2];
  } else {
    const int64_t i = tt % g.nx, j = (tt / g.nx) % g.ny, k = tt / ((int64_t)g.nx * g.ny);
    cx = g.ox + i * g.h;
    cy = g.oy + j * g.h;
    cz = g.oz + k * g.h;
  }
  double acc = 0.0;
#define QM_GTO_L(LL)                                                             \
  {                                                                              \
    const int n = lists.n[LL];                                                   \
    const int per = (n + nchunk - 1) / nchunk;                                   \
    const int lo = min(n, chunk * per), hi = min(n, lo + per);                   \
    if (hi > lo) acc += dist_sum<LL>(lists.d[LL], lo, hi, cx, cy, cz, tab);     \
  }
  QM_GTO_L(0)
  QM_GTO_L(1)
  QM_GTO_L(2)
  QM_GTO_L(3)
  QM_GTO_L(4)
#undef QM_GTO_L
  if (t < npts) partial[(int64_t)chunk * npts + t] = acc;
}

// out[t] = sum_chunks partial + sum_A Z_A / |C - R_A| (|C - R_A| > 1e-8, kernel_hartree.cu:851-865)
template <typename TO>
__global__ void k_gto_finish(const double* __restrict__ partial, int nchunk, const double4* __restrict__ atoms,
                             int natoms, const double* __restrict__ pts, GeomView g, int64_t npts,
                             TO* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts) return;
  double cx, cy, cz;
  if (pts) {
    cx = pts[3 * t];
    cy = pts[3 * t + 1];
    cz = pts[3 * t + 2];
  } else {
    const int64_t i = t % g.nx, j = (t / g.nx) % g.ny, k = t / ((int64_t)g.nx * g.ny);
    cx = g.ox + i * g.h;
    cy = g.oy + j * g.h;
    cz = g.oz + k * g.h;
  }
  double v = 0.0;
  for (int c = 0; c < nchunk; ++c) v += partial[(int64_t)c * npts + t];
  for (int a = 0; a < natoms; ++a) {
    const double4 A = atoms[a];
    const double dx = cx - A.x, dy = cy - A.y, dz = cz - A.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r > 1e-8) v += A.w / r;
  }
  out[t] = (TO)v;
}

// ------------------------------------------------------------------ host: tables
std::vector<double> make_boys_table() {
  std::vector<double> tab(kBoysPts * kBoysN);
  for (int i = 0; i < kBoysPts; ++i) {
    const double T = i * kBoysStep;
    for (int n = 0; n < kBoysN - 1; ++n) {
      // e^-T sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1)): positive terms, no cancellation
      long double term = 1.0L / (2 * n + 1), s = term;
      for (int k = 1; k < 2000; ++k) {
        term *= 2.0L * T / (2 * n + 2 * k + 1);
        s += term;
        if (term < s * 1e-21L) break;
      }
      tab[i * kBoysN + n] = (double)(std::exp(-(long double)T) * s);
    }
    tab[i * kBoysN + kBoysN - 1] = (double)std::exp(-(long double)T);
  }
  return tab;
}

// Cartesian expansion of the reference's solid harmonics (kernel_density.cu:14-31), m order (0,+1,-1,+2,-2)
struct CartTerm {
  double c;
  int e[3];
};
constexpr double C00 = 0.28209479177387814, C1 = 0.4886025119029199, C2A = 1.0925484305920792,
                 C20 = 0.31539156525252005, C22 = 0.5462742152960396;

std::vector<CartTerm> cart_terms(int L, int mi) {
  if (L == 0) return {{C00, {0, 0, 0}}};
  if (L == 1) {
    if (mi == 0) return {{C1, {0, 0, 1}}};
    if (mi == 1) return {{C1, {1, 0, 0}}};
    return {{C1, {0, 1, 0}}};
  }
  switch (mi) {
    case 0: return {{2 * C20, {0, 0, 2}}, {-C20, {2, 0, 0}}, {-C20, {0, 2, 0}}};
    case 1: return {{C2A, {1, 0, 1}}};
    case 2: return {{C2A, {0, 1, 1}}};
    case 3: return {{C22, {2, 0, 0}}, {-C22, {0, 2, 0}}};
    default: return {{C2A, {1, 1, 0}}};
  }
}

// E[i][j][t] for one axis, i,j <= 2, t <= 4 (Helgaker eq. 9.5.6-9.5.7), without K
void hermite_E(double a, double b, double XAB, double E[3][3][6]) {
  const double p = a + b;
  const double XPA = -b / p * XAB, XPB = a / p * XAB, i2p = 0.5 / p;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int t = 0; t < 6; ++t) E[i][j][t] = 0.0;
  E[0][0][0] = 1.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        double v;
        if (i > 0) {
          v = XPA * E[i - 1][j][t] + (t + 1) * E[i - 1][j][t + 1];
          if (t > 0) v += i2p * E[i - 1][j][t - 1];
        } else {
          v = XPB * E[i][j - 1][t] + (t + 1) * E[i][j - 1][t + 1];
          if (t > 0) v += i2p * E[i][j - 1][t - 1];
        }
        E[i][j][t] = v;
      }
    }
}

// f(i) for every i in [0, n), taken in increasing i by up to 16 host threads (the calling thread
// included; with fewer threads available it takes what the others leave)
template <typename F>
void host_parallel(int n, F&& f) {
  std::atomic<int> next{0};
  auto runner = [&]() {
    for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) f(i);
  };
  const int nthr = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  try {
    for (int t = 1; t < nthr && t < n; ++t) pool.emplace_back(runner);
  } catch (const std::system_error&) {
  }
  runner();
  for (auto& t : pool) t.join();
}

struct GtoCache {
  HDist* dev = nullptr;
  int off[kLmax + 1] = {0, 0, 0, 0, 0};
  int n[kLmax + 1] = {0, 0, 0, 0, 0};
  double* boys = nullptr;
};

int build_distributions(qm_ctx* ctx, qm_mol* mol, hipStream_t s, GtoCache& c) {
  const int nsh = mol->nshell, nbf = mol->nbf;
  // Cartesian term lists per (L, m index)
  std::vector<CartTerm> terms[3][5];
  for (int L = 0; L <= 2; ++L)
    for (int mi = 0; mi < 2 * L + 1; ++mi) terms[L][mi] = cart_terms(L, mi);
  // distributions of shell A with the shells B <= A, per L, in the order (B, primitive of A,
  // primitive of B); the table is the concatenation over A in increasing order, the same whether
  // the shells are built by one thread or several
  struct ShellOut {
    std::vector<HDist> l[kLmax + 1];
  };
  std::vector<ShellOut> per(nsh);
  auto build_shell = [&](int sa) {
    const DevShell& A = mol->h_shells[sa];
    const double* RA = &mol->xyz[3 * A.atom];
    for (int sb = 0; sb <= sa; ++sb) {
      const DevShell& B = mol->h_shells[sb];
      const double* RB = &mol->xyz[3 * B.atom];
      const double f = (sa == sb) ? 1.0 : 2.0;
      // Cartesian pair weights W[(ex_a),(ex_b)] = sum_{ma,mb} P_ab fa fb over matching exponents
      struct CW {
        int ea[3], eb[3];
        double w;
      };
      std::vector<CW> cw;
      double wmax = 0.0;
      for (int ma = 0; ma < 2 * A.L + 1; ++ma)
        for (int mb = 0; mb < 2 * B.L + 1; ++mb) {
          const double P = mol->h_P[(size_t)(A.bf + ma) * nbf + (B.bf + mb)];
          if (P == 0.0) continue;
          for (const CartTerm& ta : terms[A.L][ma])
            for (const CartTerm& tb : terms[B.L][mb]) {
              const double w = P * ta.c * tb.c;
              bool merged = false;
              for (CW& x : cw)
                if (x.ea[0] == ta.e[0] && x.ea[1] == ta.e[1] && x.ea[2] == ta.e[2] && x.eb[0] == tb.e[0] &&
                    x.eb[1] == tb.e[1] && x.eb[2] == tb.e[2]) {
                  x.w += w;
                  merged = true;
                  break;
                }
              if (!merged) cw.push_back({{ta.e[0], ta.e[1], ta.e[2]}, {tb.e[0], tb.e[1], tb.e[2]}, w});
            }
        }
      for (const CW& x : cw) wmax = std::max(wmax, std::fabs(x.w));
      if (cw.empty() || wmax == 0.0) continue;
      const int L = A.L + B.L;
      const double AB[3] = {RA[0] - RB[0], RA[1] - RB[1], RA[2] - RB[2]};
      const double AB2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
      for (int ia = 0; ia < A.nprim; ++ia) {
        const double a = mol->h_alpha[A.poff + ia], ca = mol->h_coeff[A.poff + ia];
        for (int ib = 0; ib < B.nprim; ++ib) {
          const double b = mol->h_alpha[B.poff + ib], cb = mol->h_coeff[B.poff + ib];
          const double p = a + b;
          const double K = std::exp(-a * b / p * AB2);
          const double pref = -f * 2.0 * M_PI / p * K * ca * cb;
          // exact zero contributions only (exp underflow or zero coefficient)
          if (pref == 0.0) continue;
          double E[3][3][3][6];
          for (int d = 0; d < 3; ++d) hermite_E(a, b, AB[d], E[d]);
          HDist h{};
          h.p = p;
          h.Px = (a * RA[0] + b * RB[0]) / p;
          h.Py = (a * RA[1] + b * RB[1]) / p;
          h.Pz = (a * RA[2] + b * RB[2]) / p;
          for (const CW& x : cw) {
            const double w = pref * x.w;
            for (int t = 0; t <= x.ea[0] + x.eb[0]; ++t) {
              const double et = E[0][x.ea[0]][x.eb[0]][t];
              if (et == 0.0) continue;
              for (int u = 0; u <= x.ea[1] + x.eb[1]; ++u) {
                const double eu = E[1][x.ea[1]][x.eb[1]][u];
                if (eu == 0.0) continue;
                for (int v = 0; v <= x.ea[2] + x.eb[2]; ++v) {
                  const double ev = E[2][x.ea[2]][x.eb[2]][v];
                  h.D[tuv_index(t, u, v)] += w * et * eu * ev;
                }
              }
            }
          }
          // Screening: |sum_tuv D_tuv R_tuv(C)| <= sum_tuv |D_tuv| (2p)^{n/2} 2^n n! for every point C
          // (n = t+u+v; Hermite-function bound on the derivatives of F_0), so a distribution whose
          // bound is below kGtoScreen changes no potential by more than that.
          double bound = 0.0;
          for (int t = 0; t <= L; ++t)
            for (int u = 0; t + u <= L; ++u)
              for (int v = 0; t + u + v <= L; ++v) {
                const int n = t + u + v;
                double g = std::pow(2.0 * p, 0.5 * n);
                for (int k = 1; k <= n; ++k) g *= 2.0 * k;
                bound += std::fabs(h.D[tuv_index(t, u, v)]) * g;
              }
          if (bound < kGtoScreen) continue;
          per[sa].l[L].push_back(h);
        }
      }
    }
  };
  // shells taken largest first (shell A pairs with A + 1 shells)
  host_parallel(nsh, [&](int k) { build_shell(nsh - 1 - k); });
  // table offsets: L-major, shells in increasing order within each L
  std::vector<size_t> at((size_t)(kLmax + 1) * nsh);
  size_t tot = 0;
  for (int L = 0; L <= kLmax; ++L) {
    c.off[L] = (int)tot;
    for (int sa = 0; sa < nsh; ++sa) {
      at[(size_t)L * nsh + sa] = tot;
      tot += per[sa].l[L].size();
    }
    c.n[L] = (int)(tot - c.off[L]);
  }
  // stream-ordered: kernels queued earlier on s still read the previous table
  if (c.dev) {
    QM_CHECK_HIP(ctx, hipFreeAsync(c.dev, s));
    c.dev = nullptr;
  }
  if (tot) {
    // the context's pinned staging buffer (the last upload from it finished: synchronised below)
    const size_t bytes = sizeof(HDist) * tot;
    if (bytes > ctx->gto_staged_bytes) {
      if (ctx->gto_staged) QM_CHECK_HIP(ctx, hipHostFree(ctx->gto_staged));
      ctx->gto_staged = nullptr;
      ctx->gto_staged_bytes = 0;
      QM_CHECK_HIP(ctx, hipHostMalloc(&ctx->gto_staged, bytes, hipHostMallocDefault));
      ctx->gto_staged_bytes = bytes;
    }
    HDist* staged = (HDist*)ctx->gto_staged;
    // the shells' lists go to their places in the staging buffer
    host_parallel(nsh, [&](int sa) {
      for (int L = 0; L <= kLmax; ++L)
        if (!per[sa].l[L].empty())
          std::copy(per[sa].l[L].begin(), per[sa].l[L].end(), staged + at[(size_t)L * nsh + sa]);
    });
    QM_CHECK_HIP(ctx, hipMallocAsync((void**)&c.dev, sizeof(HDist) * tot, s));
    QM_CHECK_HIP(ctx, hipMemcpyAsync(c.dev, staged, bytes, hipMemcpyHostToDevice, s));
    QM_CHECK_HIP(ctx, hipStreamSynchronize(s));   // the staging buffer is rewritten by the next build
  }
  mol->n_gto_builds++;
  if (!c.boys) {
    const std::vector<double> tab = make_boys_table();
    QM_CHECK_HIP(ctx, hipMalloc(&c.boys, sizeof(double) * tab.size()));
    QM_CHECK_HIP(ctx, hipMemcpy(c.boys, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
  }
  return QM_OK;
}

int get_cache(qm_ctx* ctx, qm_mol* mol, hipStream_t s, GtoCache** out) {
  if (!mol->gto) mol->gto = new GtoCache();
  GtoCache* c = (GtoCache*)mol->gto;
  if (mol->gto_dirty) {
    int rc = build_distributions(ctx, mol, s, *c);
    if (rc) return rc;
    mol->gto_dirty = false;
  }
  *out = c;
  return QM_OK;
}

template <typename TO>
int gto_eval(qm_ctx* ctx, qm_mol* mol, hipStream_t s, const double* pts, const GeomView& gv, int64_t npts,
             TO* out) {
  GtoCache* c = nullptr;
  int rc = get_cache(ctx, mol, s, &c);
  if (rc) return rc;
  DistLists dl;
  int total = 0;
  for (int L = 0; L <= kLmax; ++L) {
    dl.d[L] = c->dev ? c->dev + c->off[L] : nullptr;
    dl.n[L] = c->n[L];
    total += c->n[L];
  }
  const int64_t pblocks = (npts + kGtoThreads - 1) / kGtoThreads;
  // enough workgroups for 256 CUs: split the distributions over chunks when points are few
  int nchunk = (int)std::max<int64_t>(1, std::min<int64_t>(256, 2048 / std::max<int64_t>(pblocks, 1)));
  nchunk = std::max(1, std::min(nchunk, std::max(1, total / 64)));
  double* partial = nullptr;
  if ((rc = qm_ws(ctx, s, WS_GTO, sizeof(double) * (size_t)nchunk * npts, (void**)&partial))) return rc;
  QM_REQUIRE(ctx, pblocks <= 0x7fffffff, QM_ELIMIT, "gto: too many points");
  {
    ProfScope ps(ctx, s, QM_K_GTO);
    hipLaunchKernelGGL(k_gto, dim3((unsigned)pblocks, nchunk), dim3(kGtoThreads), 0, s, dl, (const double*)c->boys,
                     pts, gv, npts, nchunk, partial);
  }
  QM_CHECK_LAUNCH(ctx);
  hipLaunchKernelGGL(k_gto_finish<TO>, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s,
                     (const double*)partial, nchunk, (const double4*)mol->atoms, mol->natoms, pts, gv, npts, out);
  QM_CHECK_LAUNCH(ctx);
  return QM_OK;
}

}  // namespace

void qm_gto_cache_free(void* p) {
  GtoCache* c = (GtoCache*)p;
  if (!c) return;
  if (c->dev) {
    (void)hipFreeAsync(c->dev, nullptr);
    (void)hipStreamSynchronize(nullptr);
  }
  if (c->boys) (void)hipFree(c->boys);
  delete c;
}

extern "C" {

int qm_hartree_gto_points(qm_ctx* ctx, qm_mol* mol, void* stream, const double* pts, int64_t npts, double* out) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  QM_REQUIRE(ctx, mol && pts && out, QM_EINVAL, "gto: null argument");
  QM_REQUIRE(ctx, mol->has_dm, QM_EINVAL, "gto: molecule has no density matrix");
  QM_REQUIRE(ctx, npts > 0, QM_EINVAL, "gto: no points");
  GeomView gv{};
  return gto_eval<double>(ctx, mol, (hipStream_t)stream, pts, gv, npts, out);
}

int qm_hartree_gto_grid(qm_ctx* ctx, qm_mol* mol, void* stream, const qm_geom* g, int dtype, void* out) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  if ((rc = qm_check_geom(ctx, g))) return rc;
  QM_REQUIRE(ctx, mol && out, QM_EINVAL, "gto: null argument");
  QM_REQUIRE(ctx, mol->has_dm, QM_EINVAL, "gto: molecule has no density matrix");
  const GeomView gv = geom_view(g);
  if (dtype == QM_F64) return gto_eval<double>(ctx, mol, (hipStream_t)stream, nullptr, gv, npoints(g), (double*)out);
  if (dtype == QM_F32) return gto_eval<float>(ctx, mol, (hipStream_t)stream, nullptr, gv, npoints(g), (float*)out);
  return qm_fail(ctx, QM_EINVAL, "gto: bad dtype");
}

}  // extern "C"
