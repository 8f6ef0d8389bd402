// Context, workspaces and molecule tables of the qmgrid C-ABI.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "qm_internal.h"

#define QM_VERSION_MAJOR 1
#define QM_VERSION_MINOR 0
#define QM_VERSION_PATCH 0

int qm_fail(qm_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

int qm_enter(qm_ctx* ctx) {
  if (!ctx) return QM_EINVAL;
  ctx->err.clear();
  QM_CHECK_HIP(ctx, hipSetDevice(ctx->device));
  return QM_OK;
}

int qm_ws(qm_ctx* ctx, hipStream_t s, int slot, size_t bytes, void** out) {
  Workspace& w = ctx->ws[slot];
  if (w.bytes < bytes) {
    if (w.ptr) {
      // stream-ordered: work queued before this point on s still reads the old buffer
      QM_CHECK_HIP(ctx, hipFreeAsync(w.ptr, s));
      w.ptr = nullptr;
      w.bytes = 0;
    }
    size_t want = std::max(bytes, (size_t)4096);
    want = (want + 4095) & ~(size_t)4095;
    if (hipMallocAsync(&w.ptr, want, s) != hipSuccess) {
      (void)hipGetLastError();
      w.ptr = nullptr;
      return qm_fail(ctx, QM_ENOMEM, "workspace %d: cannot allocate %zu bytes", slot, want);
    }
    w.bytes = want;
  }
  *out = w.ptr;
  return QM_OK;
}

int qm_check_geom(qm_ctx* ctx, const qm_geom* g) {
  QM_REQUIRE(ctx, g != nullptr, QM_EINVAL, "null geometry");
  QM_REQUIRE(ctx, g->nx > 0 && g->ny > 0 && g->nz > 0, QM_EINVAL, "grid sizes must be positive");
  QM_REQUIRE(ctx, (int64_t)g->nx * g->ny * g->nz < ((int64_t)1 << 34), QM_ELIMIT, "grid too large");
  QM_REQUIRE(ctx, std::isfinite(g->step) && g->step > 0, QM_EINVAL, "grid step must be positive");
  return QM_OK;
}

static void clear_records(qm_ctx* ctx);

extern "C" {

int qm_version(void) { return QM_VERSION_MAJOR * 10000 + QM_VERSION_MINOR * 100 + QM_VERSION_PATCH; }

int qm_ctx_create(int device, qm_ctx** out) {
  if (!out) return QM_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return QM_EHIP;
  }
  if (device < 0 || device >= n) return QM_EINVAL;
  qm_ctx* c = new (std::nothrow) qm_ctx();
  if (!c) return QM_ENOMEM;
  c->device = device;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0) c->ncu = ncu;
  (void)hipGetLastError();
  if (hipSetDevice(device) != hipSuccess ||
      hipHostMalloc((void**)&c->host_scalars, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return QM_EHIP;
  }
  // with the device current: where this GPU holds the allocation-free reduction's scratch
  if (qm_noalloc_setup(device) != QM_OK) {
    (void)hipHostFree(c->host_scalars);
    delete c;
    return QM_EHIP;
  }
  *out = c;
  return QM_OK;
}

int qm_ctx_destroy(qm_ctx* ctx) {
  if (!ctx) return QM_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  for (auto& w : ctx->ws)
    if (w.ptr) (void)hipFreeAsync(w.ptr, nullptr);
  (void)hipStreamSynchronize(nullptr);
  if (ctx->host_scalars) (void)hipHostFree(ctx->host_scalars);
  qm_reclaim_graphs(ctx, true);
  if (ctx->jgraph_exec) (void)hipGraphExecDestroy(ctx->jgraph_exec);
  if (ctx->jgraph) (void)hipGraphDestroy(ctx->jgraph);
  if (ctx->table_uploaded) (void)hipEventDestroy(ctx->table_uploaded);
  if (ctx->table_pinned) (void)hipHostFree(ctx->table_pinned);
  if (ctx->gto_staged) (void)hipHostFree(ctx->gto_staged);
  qm_fft_cache_free(ctx->fft);
  clear_records(ctx);
  if (ctx->graph_stream) (void)hipStreamDestroy(ctx->graph_stream);
  delete ctx;
  return QM_OK;
}

const char* qm_last_error(const qm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int qm_ctx_memory(const qm_ctx* ctx, int64_t* bytes) {
  if (!ctx || !bytes) return QM_EINVAL;
  size_t b = qm_fft_cache_bytes(ctx->fft);
  for (const auto& w : ctx->ws) b += w.bytes;
  *bytes = (int64_t)b;
  return QM_OK;
}

static void clear_records(qm_ctx* ctx) {
  for (auto& r : ctx->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  ctx->recs.clear();
}

int qm_profile(qm_ctx* ctx, int enable) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  if (enable) {
    QM_CHECK_HIP(ctx, hipDeviceSynchronize());
    clear_records(ctx);
  }
  ctx->prof = enable != 0;
  return QM_OK;
}

int qm_profile_read(qm_ctx* ctx, int kernel, double* total_ms, int64_t* launches) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  QM_REQUIRE(ctx, kernel >= 0 && kernel < QM_K_COUNT, QM_EINVAL, "profile: bad kernel id");
  double tot = 0.0;
  int64_t n = 0;
  for (auto& r : ctx->recs) {
    if (r.id != kernel) continue;
    QM_CHECK_HIP(ctx, hipEventSynchronize(r.b));
    float ms = 0.f;
    QM_CHECK_HIP(ctx, hipEventElapsedTime(&ms, r.a, r.b));
    tot += ms;
    n += r.weight;
  }
  if (total_ms) *total_ms = tot;
  if (launches) *launches = n;
  return QM_OK;
}

static void free_mol_device(qm_mol* m) {
  if (m->atoms) (void)hipFree(m->atoms);
  if (m->shells) (void)hipFree(m->shells);
  if (m->prims) (void)hipFree(m->prims);
  if (m->prims_l2e) (void)hipFree(m->prims_l2e);
  m->prims_l2e = nullptr;
  if (m->kgroups) (void)hipFree(m->kgroups);
  m->kgroups = nullptr;
  if (m->P64) (void)hipFree(m->P64);
  if (m->M64) (void)hipFree(m->M64);
  if (m->P16) (void)hipFree(m->P16);
  if (m->Pb16) (void)hipFree(m->Pb16);
  qm_gto_cache_free(m->gto);
  m->gto = nullptr;
  m->atoms = nullptr; m->shells = nullptr; m->prims = nullptr;
  m->P64 = nullptr; m->M64 = nullptr; m->P16 = nullptr; m->Pb16 = nullptr;
}

int qm_mol_create(qm_ctx* ctx, int natoms, const int* Z, const double* xyz_bohr, int nshell,
                  const int* shell_atom, const int* shell_L, const int* shell_nprim,
                  const int* shell_poff, int nprim, const double* alpha, const double* coeff,
                  qm_mol** out) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  QM_REQUIRE(ctx, out != nullptr, QM_EINVAL, "null output handle");
  *out = nullptr;
  QM_REQUIRE(ctx, natoms > 0 && nshell > 0 && nprim > 0, QM_EINVAL, "empty molecule");
  QM_REQUIRE(ctx, Z && xyz_bohr && shell_atom && shell_L && shell_nprim && shell_poff && alpha && coeff,
             QM_EINVAL, "null table pointer");
  qm_mol* m = new (std::nothrow) qm_mol();
  QM_REQUIRE(ctx, m != nullptr, QM_ENOMEM, "host allocation failed");
  m->natoms = natoms;
  m->nshell = nshell;
  m->nprim = nprim;
  m->xyz.assign(xyz_bohr, xyz_bohr + 3 * natoms);
  m->Z.resize(natoms);
  double qtot = 0;
  for (int a = 0; a < natoms; ++a) {
    if (Z[a] <= 0 || !std::isfinite(xyz_bohr[3 * a]) || !std::isfinite(xyz_bohr[3 * a + 1]) ||
        !std::isfinite(xyz_bohr[3 * a + 2])) {
      delete m;
      return qm_fail(ctx, QM_EINVAL, "atom %d: bad charge or coordinates", a);
    }
    m->Z[a] = Z[a];
    qtot += Z[a];
  }
  m->qtot = qtot;
  int bf = 0;
  m->h_shells.resize(nshell);
  for (int s = 0; s < nshell; ++s) {
    DevShell d{};
    d.atom = shell_atom[s];
    d.L = shell_L[s];
    d.nprim = shell_nprim[s];
    d.poff = shell_poff[s];
    d.bf = bf;
    if (d.atom < 0 || d.atom >= natoms || d.L < 0 || d.L > 2 || d.nprim <= 0 || d.poff < 0 ||
        d.poff + d.nprim > nprim) {
      delete m;
      return qm_fail(ctx, QM_EINVAL, "shell %d: bad table entry", s);
    }
    double amin = alpha[d.poff];
    for (int p = 1; p < d.nprim; ++p) amin = std::min(amin, alpha[d.poff + p]);
    // rounded up, so a skipped shell has alpha r^2 > kPhiExpCut for every primitive
    d.rcut2 = amin > 0.0 ? std::nextafter((float)(kPhiExpCut / amin), INFINITY) : INFINITY;
    bf += 2 * d.L + 1;
    m->h_shells[s] = d;
  }
  m->nbf = bf;
  m->nbf_pad = (bf + 63) / 64 * 64;
  m->h_alpha.assign(alpha, alpha + nprim);
  m->h_coeff.assign(coeff, coeff + nprim);

  std::vector<double4> at(natoms);
  for (int a = 0; a < natoms; ++a)
    at[a] = make_double4(xyz_bohr[3 * a], xyz_bohr[3 * a + 1], xyz_bohr[3 * a + 2], (double)Z[a]);
  // where each group of 8 basis functions starts: (shell, m); groups past nbf hold (nshell, 0)
  std::vector<int2> kg(m->nbf_pad / 8, make_int2(nshell, 0));
  for (int s = 0; s < nshell; ++s)
    for (int c = 0; c < 2 * m->h_shells[s].L + 1; ++c) {
      const int f = m->h_shells[s].bf + c;
      if (f % 8 == 0) kg[f / 8] = make_int2(s, c);
    }
  std::vector<double2> pr(nprim);
  std::vector<float2> pr32(nprim);
  for (int p = 0; p < nprim; ++p) {
    pr[p] = make_double2(alpha[p], coeff[p]);
    pr32[p] = make_float2((float)(alpha[p] * 1.4426950408889634074), (float)coeff[p]);
  }

  if (hipMalloc(&m->atoms, sizeof(double4) * natoms) != hipSuccess ||
      hipMalloc(&m->shells, sizeof(DevShell) * nshell) != hipSuccess ||
      hipMalloc(&m->prims, sizeof(double2) * nprim) != hipSuccess ||
      hipMalloc(&m->prims_l2e, sizeof(float2) * nprim) != hipSuccess ||
      hipMalloc(&m->kgroups, sizeof(int2) * kg.size()) != hipSuccess) {
    (void)hipGetLastError();
    free_mol_device(m);
    delete m;
    return qm_fail(ctx, QM_ENOMEM, "molecule tables: device allocation failed");
  }
  if (hipMemcpy(m->atoms, at.data(), sizeof(double4) * natoms, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->shells, m->h_shells.data(), sizeof(DevShell) * nshell, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->prims, pr.data(), sizeof(double2) * nprim, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->prims_l2e, pr32.data(), sizeof(float2) * nprim, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->kgroups, kg.data(), sizeof(int2) * kg.size(), hipMemcpyHostToDevice) != hipSuccess) {
    hipError_t e = hipGetLastError();
    free_mol_device(m);
    delete m;
    return qm_fail(ctx, QM_EHIP, "molecule tables: upload failed: %s", hipGetErrorString(e));
  }
  *out = m;
  return QM_OK;
}

int qm_mol_set_dm(qm_ctx* ctx, qm_mol* mol, void* stream, const double* dm) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  QM_REQUIRE(ctx, mol && dm, QM_EINVAL, "null argument");
  const int n = mol->nbf;
  mol->h_P.assign((size_t)n * n, 0.0);
  for (int p = 0; p < n; ++p) {
    for (int q = 0; q <= p; ++q) {
      double v = dm[(size_t)p * n + q];
      QM_REQUIRE(ctx, std::isfinite(v), QM_EINVAL, "density matrix element (%d,%d) is not finite", p, q);
      mol->h_P[(size_t)p * n + q] = 2.0 * v;
      mol->h_P[(size_t)q * n + p] = 2.0 * v;
    }
  }
  rc = qm_density_upload_P(ctx, mol, (hipStream_t)stream);
  if (rc) return rc;
  mol->has_dm = true;
  mol->n_p_uploads++;
  mol->gto_dirty = true;
  return QM_OK;
}

int qm_mol_set_coords(qm_ctx* ctx, qm_mol* mol, void* stream, const double* xyz_bohr) {
  int rc = qm_enter(ctx);
  if (rc) return rc;
  QM_REQUIRE(ctx, mol && xyz_bohr, QM_EINVAL, "null argument");
  std::vector<double4> at(mol->natoms);
  for (int a = 0; a < mol->natoms; ++a) {
    QM_REQUIRE(ctx, std::isfinite(xyz_bohr[3 * a]) && std::isfinite(xyz_bohr[3 * a + 1]) &&
                        std::isfinite(xyz_bohr[3 * a + 2]),
               QM_EINVAL, "atom %d: coordinates are not finite", a);
    at[a] = make_double4(xyz_bohr[3 * a], xyz_bohr[3 * a + 1], xyz_bohr[3 * a + 2], mol->Z[a]);
  }
  mol->xyz.assign(xyz_bohr, xyz_bohr + 3 * mol->natoms);
  // ordered on the stream: kernels queued earlier on it read the old table, later ones the new
  hipStream_t s = (hipStream_t)stream;
  QM_CHECK_HIP(ctx, hipMemcpyAsync(mol->atoms, at.data(), sizeof(double4) * mol->natoms, hipMemcpyHostToDevice, s));
  QM_CHECK_HIP(ctx, hipStreamSynchronize(s));   // the host table is pageable and goes out of scope
  mol->gto_dirty = true;
  mol->n_coord_updates++;
  return QM_OK;
}

int qm_mol_stats(const qm_mol* mol, int64_t* counts) {
  if (!mol || !counts) return QM_EINVAL;
  counts[0] = mol->n_p_uploads;
  counts[1] = mol->n_coord_updates;
  counts[2] = mol->n_gto_builds;
  return QM_OK;
}

int qm_mol_nbf(const qm_mol* mol) { return mol ? mol->nbf : -1; }

int qm_mol_destroy(qm_ctx* ctx, qm_mol* mol) {
  if (!mol) return QM_OK;
  if (ctx) (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  free_mol_device(mol);
  delete mol;
  return QM_OK;
}

}  // extern "C"
