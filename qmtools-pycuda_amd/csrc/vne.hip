// Classical nuclear potential averaged over voxel sub-points (FP64).
//
// Reference kernel_vne_classic.cu:13-84:
//   V(v) = [sum_{s in ndiv^3} sum_A Z_A / |r_s - R_A|] / (cnt h^3)
//   r_s = origin + i h + (s + 1/2) h / ndiv,  cnt = #(s, A) with r_s != R_A.
// One thread per voxel; nuclei are wave-uniform reads.
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
