// capi.hip -- the extern "C" entry points of libporturb (declared in include/porturb.h).
#include "pu_common.h"
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <algorithm>

namespace pu {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
int fail(int code, const std::string &msg) { g_last_error = msg; return code; }

void collect_timers(pu_ctx *ctx) {
  for (auto &kv : ctx->timers) {
    for (auto &ev : kv.second.pending) {
      hipEventSynchronize(ev.second);
      float ms = 0;
      hipEventElapsedTime(&ms, ev.first, ev.second);
      kv.second.total_ms += ms;
      kv.second.count += 1;
      hipEventDestroy(ev.first);
      hipEventDestroy(ev.second);
    }
    kv.second.pending.clear();
  }
}

static int validate(const pu_config *c) {
  if (!c) return fail(PU_ERR_ARG, "null config");
  if (c->nx < 1 || c->ny < 1 || c->nz < 1) return fail(PU_ERR_ARG, "grid sizes must be positive");
  if (!(c->xlen > 0 && c->ylen > 0 && c->zlen > 0)) return fail(PU_ERR_ARG, "domain lengths must be positive");
  if (c->ord != 3 && c->ord != 5 && c->ord != 7 && c->ord != 9 && c->ord != 11)
    return fail(PU_ERR_UNSUPPORTED, "ord must be 3, 5, 7, 9 or 11");
  if (c->num_tracers < 1 || c->num_tracers > PU_MAX_TRACERS)
    return fail(PU_ERR_ARG, "num_tracers must be in [1, PU_MAX_TRACERS] (water_vapor is required)");
  if (c->idWV < 0 || c->idWV >= c->num_tracers) return fail(PU_ERR_ARG, "idWV must index a tracer");
  if (c->idTKE >= c->num_tracers) return fail(PU_ERR_ARG, "idTKE out of range");
  if (c->bc_x != PU_BC_PERIODIC && c->bc_x != PU_BC_PRECURSOR) return fail(PU_ERR_ARG, "invalid bc_x");
  if (c->bc_y != PU_BC_PERIODIC && c->bc_y != PU_BC_PRECURSOR) return fail(PU_ERR_ARG, "invalid bc_y");
  if (c->bc_z != PU_BC_PERIODIC && c->bc_z != PU_BC_SOLID_WALL)
    return fail(PU_ERR_ARG, "ERROR: Specified invalid bc_z in coupler options");
  if (c->precision != PU_FP64 && c->precision != PU_FP32) return fail(PU_ERR_ARG, "invalid precision");
  int hs = (c->ord - 1) / 2;
  // a line segment's neighbour cell must come from the periodic halo, not from a second wrap
  if (c->bc_x == PU_BC_PERIODIC && c->nx < hs + 1) return fail(PU_ERR_UNSUPPORTED, "nx smaller than the stencil halo");
  if (c->bc_y == PU_BC_PERIODIC && c->ny < hs + 1 && c->ny > 1) return fail(PU_ERR_UNSUPPORTED, "ny smaller than the stencil halo");
  if (c->nz < hs + 1) return fail(PU_ERR_UNSUPPORTED, "nz smaller than the stencil halo");
  return PU_OK;
}

template <class P> static int alloc(P **p, size_t bytes) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, bytes < 256 ? 256 : bytes);
  if (e != hipSuccess) return fail(PU_ERR_HIP, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
  *p = (P *)q;
  return PU_OK;
}

static void free_ctx(pu_ctx *c) {
  void *ptrs[] = {c->d_hy, c->d_imm_flag, c->d_any10, c->d_tau, c->d_relax, c->d_imm_les, c->d_imm_les_val,
                  c->d_U, c->d_Us, c->d_Ulo, c->d_Uslo, c->d_tend, c->d_prim, c->d_les, c->d_red,
                  c->d_sgs_tau, c->d_sgs_flags, c->d_sgs_valid, c->d_sgs_fmul, c->d_sfc_rough, c->d_sfc_temp,
                  c->d_col};
  for (void *p : ptrs) if (p) hipFree(p);
  if (c->t0) hipEventDestroy(c->t0);
  if (c->t1) hipEventDestroy(c->t1);
  for (auto &kv : c->timers)
    for (auto &ev : kv.second.pending) { hipEventDestroy(ev.first); hipEventDestroy(ev.second); }
}

}  // namespace pu

using namespace pu;

extern "C" {

const char *pu_last_error(void) { return g_last_error.c_str(); }

const char *pu_build_info(void) {
  return "libporturb: gfx950 HIP kernels; WENO orders 3,5,7,9,11; precisions FP64 and FP32(+compensated rho, rho*theta)";
}

int pu_create(const pu_config *cfg, void *stream, pu_ctx **out) {
  if (!out) return fail(PU_ERR_ARG, "null output pointer");
  *out = nullptr;
  int rc = validate(cfg);
  if (rc) return rc;
  pu_ctx *c = new pu_ctx();
  c->cfg = *cfg;
  c->stream = (hipStream_t)stream;
  if (hipGetDevice(&c->device) != hipSuccess) { delete c; return fail(PU_ERR_HIP, "hipGetDevice failed"); }
  if (getenv("PORTURB_SEGMENT_SWEEPS")) {
    const char *dirs = getenv("PORTURB_SEGMENT_DIRS");      // optional subset, e.g. "z" or "xy"
    for (int d = 0; d < 3; d++) c->segment_sweeps[d] = !dirs || strchr(dirs, "xyz"[d]) != nullptr;
  }
  Grid &g = c->g;
  g.nx = cfg->nx; g.ny = cfg->ny; g.nz = cfg->nz;
  g.hs = (cfg->ord - 1) / 2;
  g.hp = prim_halo(g.hs);
  g.ox = PRIM_OX;
  g.nxh = (g.ox + g.nx + g.hp + PRIM_ROW_ALIGN - 1) / PRIM_ROW_ALIGN * PRIM_ROW_ALIGN;
  g.nyh = g.ny + 2 * g.hp; g.nzh = g.nz + 2 * g.hp;
  c->nv = NUM_STATE + cfg->num_tracers;
  c->nf = c->nv + 1;
  size_t n3 = (size_t)g.n3(), nh = (size_t)g.nh();
  size_t es = cfg->precision == PU_FP32 ? sizeof(float) : sizeof(double);
  size_t nl = (size_t)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
  int nq = 0;
  for (int tr = 0; tr < cfg->num_tracers; tr++) nq += cfg->tracer_diffuse[tr] ? 1 : 0;
  if ((rc = alloc(&c->d_hy, sizeof(double) * (3 * (g.nz + 2 * g.hs) + 3 * (g.nz + 1)))) ||
      (rc = alloc(&c->d_imm_flag, nh)) || (rc = alloc(&c->d_any10, n3)) || (rc = alloc(&c->d_tau, n3 * sizeof(int))) ||
      (rc = alloc(&c->d_relax, n3 * sizeof(double))) || (rc = alloc(&c->d_imm_les_val, nl * sizeof(double))) || (rc = alloc(&c->d_imm_les, nl)) ||
      (rc = alloc(&c->d_U, es * c->nv * n3)) || (rc = alloc(&c->d_Us, es * c->nv * n3)) ||
      (rc = alloc(&c->d_tend, es * c->nv * n3)) || (rc = alloc(&c->d_prim, es * c->nf * nh)) ||
      (rc = alloc(&c->d_red, sizeof(double) * (size_t)std::max(4096, 3 * g.nz + 16)))) {
    free_ctx(c); delete c; return rc;
  }
  if (cfg->precision == PU_FP32) {
    if ((rc = alloc(&c->d_Ulo, sizeof(float) * 2 * n3)) || (rc = alloc(&c->d_Uslo, sizeof(float) * 2 * n3))) {
      free_ctx(c); delete c; return rc;
    }
  }
  if (cfg->idTKE >= 0) {
    if ((rc = alloc(&c->d_les, sizeof(double) * (6 + nq) * nl))) { free_ctx(c); delete c; return rc; }
  }
  // the primitive halo corners are never written; keep them defined
  hipMemset(c->d_prim, 0, es * c->nf * nh);
  hipEventCreate(&c->t0);
  hipEventCreate(&c->t1);
  *out = c;
  return PU_OK;
}

int pu_destroy(pu_ctx *ctx) {
  if (!ctx) return PU_OK;
  hipStreamSynchronize(ctx->stream);
  free_ctx(ctx);
  delete ctx;
  return PU_OK;
}

static int check_state(const pu_ctx *ctx, const pu_coupler_state *cs, bool need_imm) {
  if (!ctx) return fail(PU_ERR_ARG, "null context");
  if (!cs) return fail(PU_ERR_ARG, "null coupler state");
  if (!cs->density_dry || !cs->uvel || !cs->vvel || !cs->wvel || !cs->temp)
    return fail(PU_ERR_ARG, "coupler state is missing density_dry/uvel/vvel/wvel/temp");
  for (int tr = 0; tr < ctx->cfg.num_tracers; tr++)
    if (!cs->tracers[tr]) return fail(PU_ERR_ARG, "coupler state is missing tracer " + std::to_string(tr));
  if (need_imm && !cs->immersed_proportion) return fail(PU_ERR_ARG, "coupler state is missing immersed_proportion");
  return PU_OK;
}

int pu_dycore_init(pu_ctx *ctx, const pu_coupler_state *cs) {
  int rc = check_state(ctx, cs, true);
  if (rc) return rc;
  rc = dycore_init_device(ctx, cs);
  if (rc) return rc;
  ctx->dycore_ready = true;
  ctx->sgs_valid_ready = false;
  return PU_OK;
}

double pu_dycore_compute_time_step(const pu_ctx *ctx) {
  if (!ctx) return 0;
  const pu_config &c = ctx->cfg;
  double dx = c.xlen / c.nx, dy = c.ylen / c.ny, dz = c.zlen / c.nz;
  double maxwave = 350 + 100;
  return c.cfl * std::fmin(std::fmin(dx, dy), dz) / maxwave;
}

int pu_dycore_time_step(pu_ctx *ctx, pu_coupler_state *cs, double dt_phys) {
  int rc = check_state(ctx, cs, false);
  if (rc) return rc;
  if (!ctx->dycore_ready) return fail(PU_ERR_STATE, "pu_dycore_init has not been called");
  if (!(dt_phys > 0)) return fail(PU_ERR_ARG, "dt_phys must be positive");
  hipStream_t s = ctx->stream;
  double dt_dyn = pu_dycore_compute_time_step(ctx);
  int ncycles = (int)std::ceil(dt_phys / dt_dyn);
  dt_dyn = dt_phys / ncycles;
  for (int cyc = 0; cyc < ncycles; cyc++) {
    for (int stage = 0; stage < 3; stage++) {
      const void *U = stage == 0 ? ctx->d_U : ctx->d_Us;
      const void *Ulo = stage == 0 ? ctx->d_Ulo : ctx->d_Uslo;
      // the first stage converts the coupler state (with the relaxation that opens the first
      // SSPRK3 step) inside its prep
      if (cyc == 0 && stage == 0) rc = launch_c2d_prep(ctx, cs, s);
      else rc = launch_prep(ctx, U, Ulo, s);
      if (rc) return rc;
      // the final relaxation of a step is followed by the opening relaxation of the next one
      int nrelax = (stage == 2 && cyc + 1 < ncycles) ? 2 : 1;
      // z, y sweeps accumulate tendencies; the x sweep finishes them and combines the stage
      if ((rc = launch_sweeps(ctx, s, stage, dt_dyn, nrelax))) return rc;
    }
  }
  if ((rc = launch_convert_d2c(ctx, cs, s))) return rc;
  if (ctx->kernel_timing) collect_timers(ctx);
  return PU_OK;
}

int pu_dycore_tendencies(pu_ctx *ctx, const pu_coupler_state *cs, double *tend) {
  int rc = check_state(ctx, cs, false);
  if (rc) return rc;
  if (!ctx->dycore_ready) return fail(PU_ERR_STATE, "pu_dycore_init has not been called");
  if (!tend) return fail(PU_ERR_ARG, "null tend");
  hipStream_t s = ctx->stream;
  if ((rc = launch_convert_c2d(ctx, cs, 0, s))) return rc;
  if ((rc = launch_prep(ctx, ctx->d_U, ctx->d_Ulo, s))) return rc;
  if ((rc = launch_sweeps(ctx, s))) return rc;
  if ((rc = launch_tend_export(ctx, tend, s))) return rc;
  PU_HIP_CHECK(hipStreamSynchronize(s));
  return PU_OK;
}

int pu_dycore_flat_census(pu_ctx *ctx, const pu_coupler_state *cs, long long *counts) {
  int rc = check_state(ctx, cs, false);
  if (rc) return rc;
  if (!ctx->dycore_ready) return fail(PU_ERR_STATE, "pu_dycore_init has not been called");
  if (!counts) return fail(PU_ERR_ARG, "null counts");
  hipStream_t s = ctx->stream;
  if ((rc = launch_c2d_prep(ctx, cs, s))) return rc;
  PU_HIP_CHECK(hipMemsetAsync(ctx->d_red, 0, 9 * sizeof(unsigned long long), s));
  if ((rc = launch_census(ctx, s, (unsigned long long *)ctx->d_red))) return rc;
  PU_HIP_CHECK(hipMemcpyAsync(counts, ctx->d_red, 9 * sizeof(long long), hipMemcpyDeviceToHost, s));
  PU_HIP_CHECK(hipStreamSynchronize(s));
  return PU_OK;
}

int pu_dycore_hydrostatics(pu_ctx *ctx, double *dc, double *tc, double *pc, double *de, double *te, double *pe) {
  if (!ctx || !ctx->dycore_ready) return fail(PU_ERR_STATE, "dycore not initialised");
  double *outc[3] = {dc, tc, pc}, *oute[3] = {de, te, pe};
  for (int w = 0; w < 3; w++) {
    if (outc[w]) std::memcpy(outc[w], ctx->hy_cells[w].data(), ctx->hy_cells[w].size() * sizeof(double));
    if (oute[w]) std::memcpy(oute[w], ctx->hy_edges[w].data(), ctx->hy_edges[w].size() * sizeof(double));
  }
  return PU_OK;
}

int pu_dycore_immersed_masks(pu_ctx *ctx, unsigned char *any10, int *tau) {
  if (!ctx || !ctx->dycore_ready) return fail(PU_ERR_STATE, "dycore not initialised");
  size_t n3 = (size_t)ctx->g.n3();
  PU_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (any10) PU_HIP_CHECK(hipMemcpy(any10, ctx->d_any10, n3, hipMemcpyDeviceToHost));
  if (tau) PU_HIP_CHECK(hipMemcpy(tau, ctx->d_tau, n3 * sizeof(int), hipMemcpyDeviceToHost));
  return PU_OK;
}

int pu_les_closure_init(pu_ctx *ctx, const pu_coupler_state *cs) {
  int rc = check_state(ctx, cs, true);
  if (rc) return rc;
  rc = les_init_device(ctx, cs);
  if (rc) return rc;
  ctx->les_ready = true;
  ctx->sgs_valid_ready = false;
  return PU_OK;
}

int pu_les_closure_set_immersed_halos(pu_ctx *ctx, const double *halos) {
  if (!ctx || !halos) return fail(PU_ERR_ARG, "pu_les_closure_set_immersed_halos: null argument");
  int rc = les_set_immersed_halos_device(ctx, halos);
  if (rc) return rc;
  ctx->les_ready = true;
  ctx->sgs_valid_ready = false;          // the learned closure's mask follows the immersed cells
  return PU_OK;
}

int pu_les_closure_apply(pu_ctx *ctx, pu_coupler_state *cs, double dt_phys) {
  int rc = check_state(ctx, cs, false);
  if (rc) return rc;
  if (!ctx->dycore_ready && !ctx->les_ready)
    return fail(PU_ERR_STATE, "pu_les_closure_init or pu_dycore_init has not been called (immersed halos)");
  if (!(dt_phys > 0)) return fail(PU_ERR_ARG, "dt_phys must be positive");
  rc = les_apply_device(ctx, cs, dt_phys);
  if (rc) return rc;
  if (ctx->kernel_timing) collect_timers(ctx);
  return PU_OK;
}

int pu_diagnostics(pu_ctx *ctx, const pu_coupler_state *cs, double *mass, double *max_wind, double *max_abs_w) {
  int rc = check_state(ctx, cs, false);
  if (rc) return rc;
  return diagnostics_device(ctx, cs, mass, max_wind, max_abs_w);
}

int pu_alloc_device(size_t bytes, void **ptr) {
  if (!ptr) return fail(PU_ERR_ARG, "null pointer");
  return alloc(ptr, bytes);
}
int pu_free_device(void *ptr) { PU_HIP_CHECK(hipFree(ptr)); return PU_OK; }
int pu_memcpy_h2d(void *dst, const void *src, size_t bytes) { PU_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); return PU_OK; }
int pu_memcpy_d2h(void *dst, const void *src, size_t bytes) { PU_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); return PU_OK; }
int pu_memset_device(void *ptr, int value, size_t bytes) { PU_HIP_CHECK(hipMemset(ptr, value, bytes)); return PU_OK; }
int pu_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  if ((!dst || !src) && bytes) return fail(PU_ERR_ARG, "null pointer");
  return copy_device(dst, src, bytes);
}
int pu_fill_device(double *ptr, double value, size_t count) {
  if (!ptr && count) return fail(PU_ERR_ARG, "null pointer");
  return fill_device(ptr, value, count);
}
int pu_synchronize(pu_ctx *ctx) { PU_HIP_CHECK(hipStreamSynchronize(ctx ? ctx->stream : nullptr)); return PU_OK; }

int pu_case_init(const pu_config *cfg, int case_id, unsigned long long seed, pu_coupler_state *cs, double *imm) {
  int rc = validate(cfg);
  if (rc) return rc;
  if (case_id != PU_CASE_BUBBLE && case_id != PU_CASE_ABL_URBAN && case_id != PU_CASE_WINDFARM)
    return fail(PU_ERR_ARG, "unknown case id");
  if (!cs || !cs->density_dry || !cs->uvel || !cs->vvel || !cs->wvel || !cs->temp) return fail(PU_ERR_ARG, "bad state");
  for (int tr = 0; tr < cfg->num_tracers; tr++) if (!cs->tracers[tr]) return fail(PU_ERR_ARG, "missing tracer");
  return case_init_device(cfg, case_id, seed, cs, imm, nullptr);
}

int pu_timer_start(pu_ctx *ctx) {
  if (!ctx) return fail(PU_ERR_ARG, "null context");
  PU_HIP_CHECK(hipEventRecord(ctx->t0, ctx->stream));
  return PU_OK;
}
int pu_timer_stop(pu_ctx *ctx, double *ms) {
  if (!ctx) return fail(PU_ERR_ARG, "null context");
  PU_HIP_CHECK(hipEventRecord(ctx->t1, ctx->stream));
  PU_HIP_CHECK(hipEventSynchronize(ctx->t1));
  float f = 0;
  PU_HIP_CHECK(hipEventElapsedTime(&f, ctx->t0, ctx->t1));
  if (ms) *ms = f;
  return PU_OK;
}
int pu_kernel_timing(pu_ctx *ctx, int enable) {
  if (!ctx) return fail(PU_ERR_ARG, "null context");
  collect_timers(ctx);
  ctx->kernel_timing = enable != 0;
  if (enable) ctx->timers.clear();
  return PU_OK;
}
int pu_kernel_time(pu_ctx *ctx, const char *name, double *total_ms, long long *count) {
  if (!ctx || !name) return fail(PU_ERR_ARG, "null argument");
  collect_timers(ctx);
  auto it = ctx->timers.find(name);
  if (total_ms) *total_ms = it == ctx->timers.end() ? 0 : it->second.total_ms;
  if (count) *count = it == ctx->timers.end() ? 0 : it->second.count;
  return PU_OK;
}

}  // extern "C"
