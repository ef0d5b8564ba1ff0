// f2d_runtime.hip -- multigrid hierarchy, models and the extern "C" ABI
// (include/fluid2d_amd.h) of libf2dgpu.so.
//
// The cycle logic restates core/gmg/hierarchy.py (Vcycle, Fcycle, solve,
// twoVcycle) and the model logic core/operators.py, core/timescheme.py,
// core/euler.py and core/boussinesq.py; every arithmetic step is a kernel
// from f2d_ops.hip on device-resident buffers.  The only host round trips
// are the residual norms that decide the solve loop (hierarchy.py:169) and
// the diagnostics that feed cst[3] (fluid2d.py:397).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/fluid2d_amd.h"
#include "f2d_ops.h"

using namespace f2d;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(F2D_E_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #expr,             \
                  hipGetErrorString(e_));                                            \
  } while (0)

#define LAUNCHCHK()                                                                  \
  do {                                                                               \
    hipError_t e_ = hipGetLastError();                                               \
    if (e_ != hipSuccess)                                                            \
      return fail(F2D_E_HIP, "%s:%d launch: %s", __FILE__, __LINE__,                \
                  hipGetErrorString(e_));                                            \
  } while (0)

#define TRY(expr)              \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != F2D_OK) return rc_; \
  } while (0)

static hipStream_t S(f2d_stream_t s) { return (hipStream_t)s; }

template <typename T>
static int dalloc(T** p, size_t cnt) {
  *p = nullptr;
  if (cnt == 0) return F2D_OK;
  hipError_t e = hipMalloc((void**)p, cnt * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(F2D_E_NOMEM, "hipMalloc(%zu bytes): %s", cnt * sizeof(T), hipGetErrorString(e));
  }
  return F2D_OK;
}
template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// ---------------------------------------------------------------------------
// multigrid hierarchy
// ---------------------------------------------------------------------------
struct Level {
  int m = 0, n = 0;
  int8_t* msk = nullptr;
  double* A = nullptr;  // 5 planes
  double *x = nullptr, *b = nullptr, *r = nullptr, *t = nullptr;
  size_t size() const { return (size_t)m * n; }
};

struct f2d_gmg {
  int nlev = 0;
  double omega = 8. / 9.;
  int npre = 1, npost = 1, ndeepest = 16;
  Level L[24];
  double* partial = nullptr;  // reduction partials
  double* scal = nullptr;     // device scalars
  double* hscal = nullptr;    // pinned host mirror
  hipStream_t s = nullptr;
};

static int gmg_free(f2d_gmg* g) {
  if (!g) return F2D_OK;
  for (int l = 0; l < g->nlev; l++) {
    Level& L = g->L[l];
    dfree(L.msk);
    dfree(L.A);
    dfree(L.x);
    dfree(L.b);
    dfree(L.r);
    dfree(L.t);
  }
  dfree(g->partial);
  dfree(g->scal);
  if (g->hscal) (void)hipHostFree(g->hscal);
  delete g;
  return F2D_OK;
}

// Level sizes: core/gmg/level.py:62-115 with one subdomain.
static int gmg_levels(int nx, int ny, std::vector<std::pair<int, int>>& out) {
  out.clear();
  int n = nx, m = ny;
  for (int l = 0;; l++) {
    if (l > 0) {
      if (n % 2 || m % 2)
        return fail(F2D_E_ARG, "grid %dx%d: level %d size %dx%d is odd (need powers of two)", nx,
                    ny, l - 1, n, m);
      n /= 2;
      m /= 2;
    }
    if (n < NH || m < NH)
      return fail(F2D_E_ARG, "grid %dx%d: level %d interior %dx%d below the halo width", nx, ny, l,
                  n, m);
    out.push_back({m + 2 * NH, n + 2 * NH});
    if (n <= 4 || m <= 4) break;
    if (l > 20) return fail(F2D_E_ARG, "too many levels");
  }
  return F2D_OK;
}

// core/gmg/hierarchy.py:38-84 and core/gmg/level.py:227-335, on the device.
static int gmg_build(int nx, int ny, double dx, double dy, double hydroepsilon, double omega,
                     const int8_t* mskp_host, f2d_gmg** out) {
  std::vector<std::pair<int, int>> lv;
  TRY(gmg_levels(nx, ny, lv));
  f2d_gmg* g = new f2d_gmg();
  g->nlev = (int)lv.size();
  g->omega = omega;
  int rc = F2D_OK;
  auto bail = [&](int code) {
    gmg_free(g);
    return code;
  };
  for (int l = 0; l < g->nlev; l++) {
    Level& L = g->L[l];
    L.m = lv[l].first;
    L.n = lv[l].second;
    const size_t sz = L.size();
    if ((rc = dalloc(&L.msk, sz)) || (rc = dalloc(&L.A, 5 * sz)) || (rc = dalloc(&L.x, sz)) ||
        (rc = dalloc(&L.b, sz)) || (rc = dalloc(&L.r, sz)) || (rc = dalloc(&L.t, sz)))
      return bail(rc);
    if (hipMemset(L.x, 0, sz * 8) || hipMemset(L.b, 0, sz * 8) || hipMemset(L.r, 0, sz * 8) ||
        hipMemset(L.t, 0, sz * 8))
      return bail(fail(F2D_E_HIP, "hipMemset"));
  }
  if ((rc = dalloc(&g->partial, (size_t)kRedMaxBlocks * 8)) || (rc = dalloc(&g->scal, 64)))
    return bail(rc);
  if (hipHostMalloc((void**)&g->hscal, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return bail(fail(F2D_E_NOMEM, "hipHostMalloc"));
  hipStream_t s = g->s;
  // level 0 mask and 9-point (or 5-point) operator
  std::vector<double*> A9(g->nlev, nullptr);
  for (int l = 0; l < g->nlev; l++)
    if ((rc = dalloc(&A9[l], 9 * g->L[l].size()))) {
      for (auto p : A9) dfree(p);
      return bail(rc);
    }
  auto cleanup9 = [&]() {
    for (auto& p : A9) dfree(p);
  };
  {
    Level& L = g->L[0];
    if (hipMemcpy(L.msk, mskp_host, L.size(), hipMemcpyHostToDevice) != hipSuccess) {
      cleanup9();
      return bail(fail(F2D_E_HIP, "upload mask"));
    }
    St9 st;
    if (dx == dy && hydroepsilon == 1.) {
      const double a = -6. / 2, b = 1. / 2, c = 0.5 / 2;
      const double w[9] = {c, b, c, b, a, b, c, b, c};
      memcpy(st.w, w, sizeof w);
    } else {
      const double bx = dy / dx * hydroepsilon, by = dx / dy, a = -2 * (bx + by);
      const double w[9] = {0., by, 0., bx, a, bx, 0., by, 0.};
      memcpy(st.w, w, sizeof w);
    }
    launch_afinest(L.msk, A9[0], st, 1. / (dx * dy), L.m, L.n, s);
  }
  for (int l = 1; l < g->nlev; l++) {
    Level &F = g->L[l - 1], &C = g->L[l];
    launch_coarse_mask(F.msk, C.msk, C.m, C.n, F.n, s);
    launch_coarsen(A9[l - 1], F.msk, C.msk, A9[l], F.m, F.n, C.m, C.n, s);
  }
  // keep the 5 lower planes (ninetofive, core/gmg/level.py:332-335)
  for (int l = 0; l < g->nlev; l++)
    if (hipMemcpyAsync(g->L[l].A, A9[l], 5 * g->L[l].size() * 8, hipMemcpyDeviceToDevice, s) !=
        hipSuccess) {
      cleanup9();
      return bail(fail(F2D_E_HIP, "copy A"));
    }
  hipError_t e = hipStreamSynchronize(s);
  cleanup9();
  if (e != hipSuccess) return bail(fail(F2D_E_HIP, "gmg setup: %s", hipGetErrorString(e)));
  *out = g;
  return F2D_OK;
}

// core/gmg/level.py:340-368 (smooth nite times, fill after each)
static void g_smooth(f2d_gmg* g, int l, double* x, const double* b, int nite) {
  Level& L = g->L[l];
  for (int k = 0; k < nite; k++) launch_smooth(L.msk, L.A, x, b, g->omega, L.t, L.m, L.n, g->s);
}
static void g_residual(f2d_gmg* g, int l, const double* x, const double* b, double* r) {
  Level& L = g->L[l];
  launch_residual(L.msk, L.A, x, b, r, L.m, L.n, g->s);
}
static void g_finetocoarse(f2d_gmg* g, int l, const double* xf, double* xc) {
  Level &F = g->L[l], &C = g->L[l + 1];
  launch_restrict(C.msk, xf, xc, C.m, C.n, F.n, g->s);
}
static void g_coarsetofine(f2d_gmg* g, int l, const double* xc, double* xf, bool acc) {
  Level &F = g->L[l], &C = g->L[l + 1];
  launch_interpolate(F.msk, C.msk, xc, xf, C.m, C.n, F.n, acc, g->s);
}

// core/gmg/hierarchy.py:98-126.  r[lev] = I x[lev+1]; x[lev] += r[lev] is
// fused (the interpolated r is not read again before being overwritten).
static void g_vcycle(f2d_gmg* g, int lev1) {
  const int deep = g->nlev - 1;
  for (int l = lev1; l < deep; l++) {
    Level& L = g->L[l];
    if (l > lev1) launch_fill_value(L.x, 0., L.size(), g->s);
    g_smooth(g, l, L.x, L.b, g->npre);
    g_residual(g, l, L.x, L.b, L.r);
    g_finetocoarse(g, l, L.r, g->L[l + 1].b);
  }
  Level& D = g->L[deep];
  launch_fill_value(D.x, 0., D.size(), g->s);
  g_smooth(g, deep, D.x, D.b, g->ndeepest);
  for (int l = deep - 1; l >= lev1; l--) {
    Level& L = g->L[l];
    g_coarsetofine(g, l, g->L[l + 1].x, L.x, true);
    g_smooth(g, l, L.x, L.b, g->npost);
  }
}

// core/gmg/hierarchy.py:131-151
static void g_fcycle(f2d_gmg* g, int lev1) {
  const int deep = g->nlev - 1;
  for (int l = lev1; l < deep; l++) g_finetocoarse(g, l, g->L[l].b, g->L[l + 1].b);
  Level& D = g->L[deep];
  launch_fill_value(D.x, 0., D.size(), g->s);
  g_smooth(g, deep, D.x, D.b, g->ndeepest);
  for (int l = deep - 1; l >= lev1; l--) {
    g_coarsetofine(g, l, g->L[l + 1].x, g->L[l].x, false);
    g_vcycle(g, l);
  }
}

static int g_norm(f2d_gmg* g, int l, const double* x, double* out_host) {
  Level& L = g->L[l];
  launch_norm2(L.msk, x, L.m, L.n, g->partial, g->scal, g->s);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(g->hscal, g->scal, 8, hipMemcpyDeviceToHost, g->s));
  HIPCHK(hipStreamSynchronize(g->s));
  *out_host = sqrt(g->hscal[0] / 1.);
  return F2D_OK;
}

// core/gmg/hierarchy.py:154-192
static int g_solve(f2d_gmg* g, double* x, const double* b, int maxite, double tol, int* nite_out,
                   double* res_out) {
  Level& L0 = g->L[0];
  g_residual(g, 0, x, b, L0.b);
  double normeb, nb0;
  TRY(g_norm(g, 0, b, &normeb));
  *nite_out = 0;
  *res_out = 0.;
  if (!(normeb > 0)) return F2D_OK;
  TRY(g_norm(g, 0, L0.b, &nb0));
  double res0 = nb0 / normeb, res = res0;
  int nite = 0, ndiv = 0;
  while (nite < maxite && res0 > tol) {
    g_fcycle(g, 0);
    launch_add_inplace(x, (const double*)L0.x, L0.size(), g->s);
    g_residual(g, 0, x, b, L0.b);
    double nb;
    TRY(g_norm(g, 0, L0.b, &nb));
    res = nb / normeb;
    const double conv = res0 / res;
    res0 = res;
    nite++;
    if (conv < 1) ndiv++;
    if (ndiv > 4) {
      *nite_out = nite;
      *res_out = res;
      return fail(F2D_E_DIVERGED, "multigrid solve is not converging (res %g)", res);
    }
  }
  *nite_out = nite;
  *res_out = res;
  return F2D_OK;
}

// core/gmg/hierarchy.py:207-218 (the residual at the top of each cycle is
// overwritten by Vcycle before it is read, so it is not recomputed)
static void g_twovcycle(f2d_gmg* g, double* x, const double* b) {
  Level& L0 = g->L[0];
  launch_copy(L0.x, (const double*)x, L0.size(), g->s);
  launch_copy(L0.b, b, L0.size(), g->s);
  for (int k = 0; k < 2; k++) g_vcycle(g, 0);
  launch_copy(x, (const double*)L0.x, L0.size(), g->s);
}

// ---------------------------------------------------------------------------
// models
// ---------------------------------------------------------------------------
enum { I_VORT = 0, I_PSI = 1, I_U = 2, I_V = 3, I_SRC = 4, I_BUOY = 4, I_BANOM = 5 };

struct f2d_model {
  f2d_model_params p;
  int m = 0, n = 0, nvar = 0;
  size_t fsz = 0, ssz = 0;  // field, state sizes
  double dx = 0, dy = 0, area = 0;
  double cst[5];
  int kstage = 0, kforcing = 0;
  double dtcoef = 1.;
  int8_t *msk = nullptr, *mskp = nullptr;
  double *state = nullptr, *k[3] = {nullptr, nullptr, nullptr}, *xs = nullptr, *work = nullptr;
  double *xr = nullptr, *yr = nullptr;
  double* partial = nullptr;
  double* dscal = nullptr;
  double* hscal = nullptr;
  f2d_gmg* gmg = nullptr;
  hipStream_t s = nullptr;
  double diags[8];
  int last_nite = 0;
  double last_res = 0.;
  // flux diagnostics (core/fluxes.py): extended state, input copy, forward
  // result, stage state, persistent tendencies; allocated on first use
  double *fx_x = nullptr, *fx_xe = nullptr, *fx_xw = nullptr, *fx_xs = nullptr;
  double* fx_k[3] = {nullptr, nullptr, nullptr};
  // no-slip walls (core/operators.py:157-185, :245-290) and islands
  // (core/island.py:21-43)
  bool noslip = false, enforce = false, isisland = false;
  int8_t *msknoslip = nullptr, *mskbc = nullptr;
  double *xr0 = nullptr, *yr0 = nullptr, *source_b = nullptr, *rhsp = nullptr, *ipsi = nullptr;
  double bcarea = 0., x2bc = 0., y2bc = 0.;
  double* F(double* base, int k) const { return base + (size_t)k * fsz; }
};

static void model_free(f2d_model* M) {
  if (!M) return;
  dfree(M->msk);
  dfree(M->mskp);
  dfree(M->state);
  for (auto& q : M->k) dfree(q);
  dfree(M->xs);
  dfree(M->work);
  dfree(M->xr);
  dfree(M->yr);
  dfree(M->partial);
  dfree(M->dscal);
  dfree(M->fx_x);
  dfree(M->fx_xe);
  dfree(M->fx_xw);
  dfree(M->fx_xs);
  for (auto& q : M->fx_k) dfree(q);
  dfree(M->msknoslip);
  dfree(M->mskbc);
  dfree(M->xr0);
  dfree(M->yr0);
  dfree(M->source_b);
  dfree(M->rhsp);
  dfree(M->ipsi);
  if (M->hscal) (void)hipHostFree(M->hscal);
  gmg_free(M->gmg);
  delete M;
}

// core/operators.py:214-236: advection of each tracer, then fillhalo
static int rhs_adv(f2d_model* M, const double* x, double* dxdt) {
  const int ntr = M->p.model == 0 ? 1 : 2;
  const int tr[2] = {I_VORT, I_BUOY};
  for (int t = 0; t < ntr; t++) {
    double* y = M->F(dxdt, tr[t]);
    const double* q = x + (size_t)tr[t] * M->fsz;
    const double* u = x + (size_t)I_U * M->fsz;
    const double* v = x + (size_t)I_V * M->fsz;
    if (M->p.order % 2 == 0)
      launch_adv_centered<double>(M->msk, q, y, u, v, nullptr, nullptr, M->cst, M->p.order, M->m,
                                  M->n, M->s);
    else
      launch_adv_upwind<double>(M->msk, q, y, u, v, nullptr, nullptr, M->cst,
                                M->p.flux_splitting, M->p.order, M->m, M->n, M->s);
    launch_fillhalo(y, M->m, M->n, M->s);
  }
  LAUNCHCHK();
  return F2D_OK;
}

// core/operators.py:292-301
static void rhs_diffusion(f2d_model* M, const double* x, double* dxdt, double coef) {
  const int ntr = M->p.model == 0 ? 1 : 2;
  const int tr[2] = {I_VORT, I_BUOY};
  const double kd[2] = {M->p.kdiff_vort, M->p.kdiff_buoy};
  for (int t = 0; t < ntr; t++)
    launch_add_diffusion(M->msk, x + (size_t)tr[t] * M->fsz, M->dx, coef * kd[t],
                         M->F(dxdt, tr[t]), M->m, M->n, M->s);
}

// core/operators.py:303-314
static void rhs_torque(f2d_model* M, const double* x, double* dxdt) {
  launch_torque(M->msk, x + (size_t)I_BUOY * M->fsz, M->dx, M->p.gravity, M->F(dxdt, I_VORT),
                M->m, M->n, M->s);
}

// core/operators.py:421-498
// island: the islands' rhs correction and boundary values (core/euler.py
// passes island=isisland for the state's inversions)
static int invert_vorticity(f2d_model* M, double* x, bool full, bool island = false) {
  double* psi = M->F(x, I_PSI);
  island = island && M->isisland;
  launch_celltocorner(M->F(x, I_VORT), M->work, M->m, M->n, M->s);
  if (island) launch_sub_inplace(M->work, M->rhsp, M->fsz, M->s);
  if (!full) {
    g_twovcycle(M->gmg, psi, M->work);
  } else {
    int nite = 0;
    double res = 0.;
    int rc = g_solve(M->gmg, psi, M->work, 4, 1e-11, &nite, &res);
    M->last_nite = nite;
    M->last_res = res;
    TRY(rc);
    if (M->p.perio) {  // psi -= domain_integration(psi) / area
      launch_domain_sum(psi, M->m, M->n, M->area, M->partial, M->dscal, M->s);
      launch_sub_dev_scalar(psi, (const double*)M->dscal, M->fsz, M->s);
    }
  }
  launch_mask_mul(psi, M->mskp, M->fsz, M->s);
  if (island) launch_add_inplace(psi, (const double*)M->ipsi, M->fsz, M->s);
  launch_orthogradient(M->msk, (const double*)psi, M->dx, M->dy, M->F(x, I_U), M->F(x, I_V), M->m,
                       M->n, M->s);
  LAUNCHCHK();
  return F2D_OK;
}

// Operators.rhs_noslip (core/operators.py:245-290) into `source`, then
// vorticity -= source, then the fast inversion with islands: add_noslip of
// core/euler.py:171-179 and core/boussinesq.py:116-118.  The reference
// first writes cornertocell(psi) into the work array that the source term
// then overwrites; the cells it leaves are halo cells, which the final
// periodic fill rewrites, so that pass has no effect and is not run.
static int add_noslip(f2d_model* M, double* x, double* src) {
  const int m = M->m, n = M->n;
  launch_noslip_source(M->msknoslip, (const double*)M->F(x, I_PSI), src, M->dx, M->dy, m, n, M->s);
  launch_domain_sum(src, m, n, M->bcarea, M->partial, M->dscal, M->s);  // meansource
  launch_sub_mean_mask(src, M->mskbc, M->dscal, M->fsz, M->s);
  if (M->enforce) {
    launch_dot2(M->msk, src, M->xr0, M->yr0, m, n, M->partial, M->dscal + 2, M->s);
    launch_noslip_momentum(src, M->mskbc, M->xr0, M->yr0, M->dscal + 2, M->x2bc, M->y2bc,
                           M->fsz, M->s);
  }
  launch_fillhalo(src, m, n, M->s);
  launch_sub_inplace(M->F(x, I_VORT), src, M->fsz, M->s);
  LAUNCHCHK();
  return invert_vorticity(M, x, false, true);
}

// core/euler.py:149-169, core/boussinesq.py:102-114
static int dynamics(f2d_model* M, const double* x, double* dxdt) {
  TRY(rhs_adv(M, x, dxdt));
  if (M->p.model == 0) {
    if (M->kstage == M->kforcing) {
      if (M->p.diffusion) {
        rhs_diffusion(M, x, dxdt, 1.);
        return invert_vorticity(M, dxdt, false);
      }
      return F2D_OK;
    }
    return invert_vorticity(M, dxdt, false);
  }
  rhs_torque(M, x, dxdt);
  if (M->kstage == M->kforcing && M->p.diffusion) rhs_diffusion(M, x, dxdt, M->dtcoef);
  return invert_vorticity(M, dxdt, false);
}

// core/timescheme.py:167-180 (RK3_SSP), :152-164 (RK3) over cnt values of
// x with stage state xs, tendencies k[0..2] and right-hand side rhs
using RhsFn = int (*)(f2d_model*, const double*, double*);
static int rk_forward(f2d_model* M, double* x, double* xs, double* const* k, size_t cnt,
                      double dt, RhsFn rhs) {
  double *k0 = k[0], *k1 = k[1], *k2 = k[2];
  if (M->p.timestepping == 0) {
    M->kstage = 0;
    TRY(rhs(M, x, k0));
    launch_rk(0, xs, x, k0, k1, k2, dt, cnt, M->s);
    M->kstage = 1;
    TRY(rhs(M, xs, k1));
    launch_rk(1, xs, x, k0, k1, k2, 0.25 * dt, cnt, M->s);
    M->kstage = 2;
    TRY(rhs(M, xs, k2));
    launch_rk(2, x, x, k0, k1, k2, dt / 6., cnt, M->s);
  } else {
    M->kstage = 0;
    TRY(rhs(M, x, k0));
    launch_rk(0, xs, x, k0, k1, k2, dt / 3., cnt, M->s);
    M->kstage = 1;
    TRY(rhs(M, xs, k1));
    launch_rk(3, xs, x, k0, k1, k2, 0.5 * dt, cnt, M->s);
    M->kstage = 2;
    TRY(rhs(M, xs, k2));
    launch_rk(4, x, x, k0, k1, k2, dt, cnt, M->s);
  }
  LAUNCHCHK();
  return F2D_OK;
}

static int forward(f2d_model* M, double dt) {
  return rk_forward(M, M->state, M->xs, M->k, M->ssz, dt, dynamics);
}

// ---------------------------------------------------------------------------
// Flux diagnostics, core/fluxes.py:20-213 (class Fluxes).  Extended state:
// the model variables, the cell-centred velocities uc, vc, then an (x, y)
// edge-flux plane pair per advected quantity: vorticity [buoyancy] uc vc.
// ---------------------------------------------------------------------------
static int fx_tracers(const f2d_model* M, int* tr) {
  int k = 0;
  tr[k++] = I_VORT;
  if (M->p.model != 0) tr[k++] = I_BUOY;
  tr[k++] = M->nvar;      // uc
  tr[k++] = M->nvar + 1;  // vc
  return k;
}

// Fluxes.advection + rhs_adv (core/fluxes.py:172-213): advection with the
// edge-flux outputs (core/fortran_fluxes.f90), fillhalo of the three
// results, the torque for Boussinesq, the fast inversion
static int fx_advection(f2d_model* M, const double* x, double* dxdt) {
  int tr[4];
  const int nt = fx_tracers(M, tr), nvs = M->nvar + 2;
  const double* u = x + (size_t)I_U * M->fsz;
  const double* v = x + (size_t)I_V * M->fsz;
  for (int t = 0; t < nt; t++) {
    double* y = M->F(dxdt, tr[t]);
    double* xf = M->F(dxdt, nvs + 2 * t);
    double* yf = M->F(dxdt, nvs + 2 * t + 1);
    const double* q = x + (size_t)tr[t] * M->fsz;
    if (M->p.order % 2 == 0)
      launch_adv_centered<double>(M->msk, q, y, u, v, xf, yf, M->cst, M->p.order, M->m, M->n,
                                  M->s);
    else
      launch_adv_upwind<double>(M->msk, q, y, u, v, xf, yf, M->cst, M->p.flux_splitting,
                                M->p.order, M->m, M->n, M->s);
    launch_fillhalo(y, M->m, M->n, M->s);
    launch_fillhalo(xf, M->m, M->n, M->s);
    launch_fillhalo(yf, M->m, M->n, M->s);
  }
  if (M->p.model != 0) rhs_torque(M, x, dxdt);
  LAUNCHCHK();
  return invert_vorticity(M, dxdt, false);
}

// Fluxes.diag_fluxes (core/fluxes.py:100-170): one step dt forward from x,
// one step -dt from the time-reversed x (u, v, psi, vorticity, uc, vc
// negated); per flux plane, with a the forward and b the backward result
// and cff = 0.5/dt: flx[k] = cff*(a + sign*b) ('rev'), flx[nflx + k] =
// cff*(a - sign*b) ('irr'), sign = -1 for vorticity, uc, vc, +1 for
// buoyancy.  The tendencies persist between calls (psi initial guesses).
static int diag_fluxes(f2d_model* M, const double* x, double dt, double* flx) {
  int tr[4];
  const int nt = fx_tracers(M, tr), nvs = M->nvar + 2, nflx = 2 * nt, ne = nvs + nflx;
  const size_t sz = M->fsz, cnt = (size_t)ne * sz;
  if (!M->fx_x) {
    int rc;
    if ((rc = dalloc(&M->fx_x, cnt)) || (rc = dalloc(&M->fx_xe, (size_t)nvs * sz)) ||
        (rc = dalloc(&M->fx_xw, cnt)) || (rc = dalloc(&M->fx_xs, cnt)) ||
        (rc = dalloc(&M->fx_k[0], cnt)) || (rc = dalloc(&M->fx_k[1], cnt)) ||
        (rc = dalloc(&M->fx_k[2], cnt)))
      return rc;
    for (double* q : {M->fx_x, M->fx_xw, M->fx_xs, M->fx_k[0], M->fx_k[1], M->fx_k[2]})
      HIPCHK(hipMemsetAsync(q, 0, cnt * sizeof(double), M->s));
    HIPCHK(hipMemsetAsync(M->fx_xe, 0, (size_t)nvs * sz * sizeof(double), M->s));
  }
  double *xe = M->fx_xe, *X = M->fx_x;
  HIPCHK(hipMemcpyAsync(xe, x, (size_t)M->nvar * sz * sizeof(double), hipMemcpyDeviceToDevice,
                        M->s));
  double* uc = M->F(xe, M->nvar);
  double* vc = M->F(xe, M->nvar + 1);
  launch_roll_mean(x + (size_t)I_U * sz, uc, 1, M->m, M->n, M->s);
  launch_fillhalo(uc, M->m, M->n, M->s);
  launch_roll_mean(x + (size_t)I_V * sz, vc, 0, M->m, M->n, M->s);
  launch_fillhalo(vc, M->m, M->n, M->s);
  LAUNCHCHK();
  auto load = [&](bool reverse) -> int {
    HIPCHK(hipMemcpyAsync(X, xe, (size_t)nvs * sz * sizeof(double), hipMemcpyDeviceToDevice,
                          M->s));
    if (reverse)
      for (int a : {(int)I_U, (int)I_V, (int)I_PSI, (int)I_VORT, M->nvar, M->nvar + 1})
        launch_negate(M->F(X, a), sz, M->s);
    launch_fill_value(M->F(X, nvs), 0., (size_t)nflx * sz, M->s);
    LAUNCHCHK();
    return F2D_OK;
  };
  TRY(load(false));
  TRY(rk_forward(M, X, M->fx_xs, M->fx_k, cnt, dt, fx_advection));
  HIPCHK(hipMemcpyAsync(M->fx_xw, X, cnt * sizeof(double), hipMemcpyDeviceToDevice, M->s));
  TRY(load(true));
  TRY(rk_forward(M, X, M->fx_xs, M->fx_k, cnt, -dt, fx_advection));
  const double cff = 0.5 / dt;
  for (int k = 0; k < nflx; k++) {
    const double sign = (k < 2 || k >= nflx - 4) ? -1. : 1.;
    launch_flux_split(M->F(M->fx_xw, nvs + k), M->F(X, nvs + k), flx + (size_t)k * sz,
                      flx + (size_t)(nflx + k) * sz, sign, cff, sz, M->s);
  }
  LAUNCHCHK();
  return F2D_OK;
}

static int model_step(f2d_model* M, double dt) {
  TRY(forward(M, dt));
  if (M->noslip) {  // core/euler.py:108-112, core/boussinesq.py:92-94
    double* src = M->p.model == 0 ? M->F(M->state, I_SRC) : M->source_b;
    TRY(add_noslip(M, M->state, src));
    if (M->p.model == 0) launch_div_scalar(src, dt, M->fsz, M->s);
  }
  if (M->p.model == 0) return invert_vorticity(M, M->state, true, true);
  // banom = buoyancy - bref, bref = the all-zero buoyancy at construction
  launch_copy(M->F(M->state, I_BANOM), (const double*)M->F(M->state, I_BUOY), M->fsz, M->s);
  LAUNCHCHK();
  return F2D_OK;
}

static int model_diag(f2d_model* M, double* out) {
  launch_diag(M->msk, M->state, M->p.model, M->xr, M->yr, M->m, M->n, M->partial, M->dscal + 8,
              M->s);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(M->hscal, M->dscal + 8, 8 * sizeof(double), hipMemcpyDeviceToHost, M->s));
  HIPCHK(hipStreamSynchronize(M->s));
  for (int k = 0; k < 8; k++) out[k] = M->hscal[k];
  if (M->p.model == 1) out[4] = -M->p.gravity * out[4];
  memcpy(M->diags, out, sizeof M->diags);
  return F2D_OK;
}

// ---------------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------------
extern "C" {

const char* f2d_last_error(void) { return g_err.c_str(); }
int f2d_version(void) { return 100; }

int f2d_device_count(int* count) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *count = 0;
    return fail(F2D_E_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
  }
  *count = c;
  return F2D_OK;
}

int f2d_malloc(void** ptr, size_t bytes) {
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) return fail(F2D_E_NOMEM, "hipMalloc: %s", hipGetErrorString(e));
  return F2D_OK;
}
int f2d_free(void* ptr) {
  HIPCHK(hipFree(ptr));
  return F2D_OK;
}
int f2d_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return F2D_OK;
}
int f2d_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return F2D_OK;
}
int f2d_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return F2D_OK;
}
int f2d_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return F2D_OK;
}

static int check_field(int nh, int m, int n) {
  if (nh != NH) return fail(F2D_E_ARG, "nh must be %d (got %d)", NH, nh);
  if (m < 3 * NH || n < 3 * NH) return fail(F2D_E_ARG, "field %dx%d too small", m, n);
  return F2D_OK;
}

int f2d_adv_upwind_f64(const int8_t* msk, const double* x, double* y, const double* u,
                       const double* v, double* xflx, double* yflx, const double* cst, int nh,
                       int method, int order, int m, int n, f2d_stream_t stream) {
  TRY(check_field(nh, m, n));
  if (method < 0 || method > 2) return fail(F2D_E_ARG, "method must be 0, 1 or 2");
  if (order != 1 && order != 3 && order != 5) return fail(F2D_E_ARG, "upwind order 1, 3 or 5");
  if ((xflx == nullptr) != (yflx == nullptr)) return fail(F2D_E_ARG, "xflx and yflx together");
  launch_adv_upwind<double>(msk, x, y, u, v, xflx, yflx, cst, method, order, m, n, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

int f2d_adv_centered_f64(const int8_t* msk, const double* x, double* y, const double* u,
                         const double* v, double* xflx, double* yflx, const double* cst, int nh,
                         int order, int m, int n, f2d_stream_t stream) {
  TRY(check_field(nh, m, n));
  if (order != 2 && order != 4 && order != 6) return fail(F2D_E_ARG, "centred order 2, 4 or 6");
  if ((xflx == nullptr) != (yflx == nullptr)) return fail(F2D_E_ARG, "xflx and yflx together");
  launch_adv_centered<double>(msk, x, y, u, v, xflx, yflx, cst, order, m, n, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

int f2d_celltocorner_f64(const double* xr, double* xp, int m, int n, f2d_stream_t stream) {
  if (m < 2 || n < 2) return fail(F2D_E_ARG, "bad shape");
  launch_celltocorner(xr, xp, m, n, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

int f2d_orthogradient_f64(const int8_t* msk, const double* psi, double dx, double dy, int nh,
                          double* u, double* v, int m, int n, f2d_stream_t stream) {
  if (m < 3 || n < 3) return fail(F2D_E_ARG, "bad shape");
  launch_orthogradient(msk, psi, dx, dy, u, v, m, n, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

// add_diffusion then fillhalo (the reference always fills after it)
int f2d_add_diffusion_f64(const int8_t* msk, const double* trac, double dx, int nh, double kdiff,
                          double* dtrac, int m, int n, f2d_stream_t stream) {
  TRY(check_field(nh, m, n));
  launch_add_diffusion(msk, trac, dx, kdiff, dtrac, m, n, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

// dw *= msk; add_torque; fillhalo (core/operators.py:311-313)
int f2d_add_torque_f64(const int8_t* msk, const double* buoy, double dx, int nh, double gravity,
                       double* domega, int m, int n, f2d_stream_t stream) {
  TRY(check_field(nh, m, n));
  launch_torque(msk, buoy, dx, gravity, domega, m, n, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

int f2d_fillhalo_f64(double* x, int nh, int m, int n, f2d_stream_t stream) {
  TRY(check_field(nh, m, n));
  launch_fillhalo(x, m, n, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

int f2d_smooth_f64(const int8_t* msk, const double* A, double* x, const double* b, double omega,
                   double* work, int m, int n, f2d_stream_t stream) {
  TRY(check_field(NH, m, n));
  launch_smooth(msk, A, x, b, omega, work, m, n, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

int f2d_residual_f64(const int8_t* msk, const double* A, const double* x, const double* b,
                     double* r, int m, int n, f2d_stream_t stream) {
  TRY(check_field(NH, m, n));
  launch_residual(msk, A, x, b, r, m, n, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

static int check_pair(int nh, int m2, int n2, int m1, int n1) {
  TRY(check_field(nh, m2, n2));
  if (m1 != 2 * (m2 - 2 * NH) + 2 * NH || n1 != 2 * (n2 - 2 * NH) + 2 * NH)
    return fail(F2D_E_ARG, "fine %dx%d is not twice coarse %dx%d", m1, n1, m2, n2);
  return F2D_OK;
}

int f2d_restrict_f64(const int8_t* msk2, const double* x1, int nh, double* x2, int m2, int n2,
                     int m1, int n1, f2d_stream_t stream) {
  TRY(check_pair(nh, m2, n2, m1, n1));
  launch_restrict(msk2, x1, x2, m2, n2, n1, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

int f2d_interpolate_f64(const int8_t* msk1, const int8_t* msk2, const double* x2, int nh,
                        double* x1, int m2, int n2, int m1, int n1, int accumulate,
                        f2d_stream_t stream) {
  TRY(check_pair(nh, m2, n2, m1, n1));
  launch_interpolate(msk1, msk2, x2, x1, m2, n2, n1, accumulate != 0, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

int f2d_coarsenmatrix_f64(const double* Af, const int8_t* msk1, const int8_t* msk2, int nh,
                          double* Ac, int m1, int n1, int m2, int n2, f2d_stream_t stream) {
  TRY(check_pair(nh, m2, n2, m1, n1));
  launch_coarsen(Af, msk1, msk2, Ac, m1, n1, m2, n2, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

// scratch for the stand-alone reductions (one per thread of the caller)
static int red_scratch(double** partial) {
  static thread_local double* p = nullptr;
  if (!p) TRY(dalloc(&p, (size_t)kRedMaxBlocks * 8));
  *partial = p;
  return F2D_OK;
}

int f2d_norm2_f64(const int8_t* msk, const double* x, int nh, int m, int n, double* out,
                  f2d_stream_t stream) {
  TRY(check_field(nh, m, n));
  double* partial;
  TRY(red_scratch(&partial));
  launch_norm2(msk, x, m, n, partial, out, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

int f2d_kemaxu_f64(const int8_t* msk, const double* u, const double* v, int nh, int m, int n,
                   double* out, f2d_stream_t stream) {
  TRY(check_field(nh, m, n));
  double* partial;
  TRY(red_scratch(&partial));
  launch_kemaxu(msk, u, v, m, n, partial, out, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

int f2d_dotprod_f64(const int8_t* msk, const double* x, const double* y, int nh, int m, int n,
                    double* out, f2d_stream_t stream) {
  TRY(check_field(nh, m, n));
  double* partial;
  TRY(red_scratch(&partial));
  launch_msum(0, msk, x, y, m, n, partial, out, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

int f2d_sum_f64(const int8_t* msk, const double* x, int nh, int m, int n, double* out,
                f2d_stream_t stream) {
  TRY(check_field(nh, m, n));
  double* partial;
  TRY(red_scratch(&partial));
  launch_msum(1, msk, x, nullptr, m, n, partial, out, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

int f2d_sumandnorm_f64(const int8_t* msk, const double* x, int nh, int m, int n, double* out,
                       f2d_stream_t stream) {
  TRY(check_field(nh, m, n));
  double* partial;
  TRY(red_scratch(&partial));
  launch_msum(2, msk, x, nullptr, m, n, partial, out, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

int f2d_domain_sum_f64(const double* x, int nh, int m, int n, double* out, f2d_stream_t stream) {
  TRY(check_field(nh, m, n));
  double* partial;
  TRY(red_scratch(&partial));
  const size_t tot = (size_t)(m - 2 * NH) * (n - 2 * NH);
  if ((tot + 8191) / 8192 > (size_t)kRedMaxBlocks * 8)
    return fail(F2D_E_ARG, "field too large for the NumPy-order sum");
  launch_domain_sum(x, m, n, 0., partial, out, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

// ---- allocation-free NumPy-order sum ----
// One partial per 8192-element chunk of the interior lives in a file-scope
// device array: at most kNoAllocChunks chunks (64 KiB per GPU, held until
// the process ends), i.e. interiors of up to 8192 * 8192 cells.  Every chunk
// partial is written before the fold reads it, so nothing needs clearing
// between calls.  All calls on one GPU share the array: calls that overlap
// on that GPU (other streams, other host threads) race.
constexpr int kNoAllocChunks = kRedMaxBlocks * 8;
static __device__ __align__(256) double scratch[kNoAllocChunks];

// Device address of `scratch` on the current device, looked up once per
// device (the name itself is a host-side stand-in, not a device address).
static std::mutex g_scratch_mu;
static double* g_scratch_dev[64] = {};

static int scratch_address(double** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(F2D_E_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(F2D_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch_dev[dev] = (double*)p;
  }
  *out = g_scratch_dev[dev];
  return F2D_OK;
}

int f2d_domain_sum_f64_noalloc(const double* x, int nh, int m, int n, double* out,
                               f2d_stream_t stream) {
  TRY(check_field(nh, m, n));
  const size_t tot = (size_t)(m - 2 * NH) * (n - 2 * NH);
  const size_t nch = (tot + 8191) / 8192;
  if (nch > (size_t)kNoAllocChunks)
    return fail(F2D_E_ARG,
                "f2d_domain_sum_f64_noalloc: %zu interior cells need %zu chunk partials; the "
                "allocation-free sum holds at most %d (%lld cells)",
                tot, nch, kNoAllocChunks, (long long)kNoAllocChunks * 8192);
  double* partial = nullptr;
  TRY(scratch_address(&partial));
  launch_domain_sum(x, m, n, 0., partial, out, S(stream));
  LAUNCHCHK();
  return F2D_OK;
}

// ---- gmg ----
int f2d_gmg_create(int nx, int ny, int nh, double dx, double dy, double hydroepsilon, double omega,
                   const int8_t* mskp_host, f2d_gmg** out) {
  if (nh != NH) return fail(F2D_E_ARG, "nh must be 3");
  return gmg_build(nx, ny, dx, dy, hydroepsilon, omega, mskp_host, out);
}
int f2d_gmg_destroy(f2d_gmg* g) { return gmg_free(g); }
int f2d_gmg_set_helmholtz(f2d_gmg* g, double Rd) {
  if (!(Rd > 0.) || !isfinite(Rd)) return fail(F2D_E_ARG, "Rd must be positive and finite");
  const double shift = 1. / (Rd * Rd);
  for (int l = 0; l < g->nlev; l++)
    launch_helmholtz_shift(g->L[l].A + 4 * g->L[l].size(), g->L[l].size(), shift, g->s);
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(g->s));
  return F2D_OK;
}
int f2d_gmg_nlevels(const f2d_gmg* g, int* nlev) {
  *nlev = g->nlev;
  return F2D_OK;
}
int f2d_gmg_level(const f2d_gmg* g, int lev, int which, void** dev_ptr, int* m, int* n) {
  if (lev < 0 || lev >= g->nlev) return fail(F2D_E_ARG, "level %d out of range", lev);
  const Level& L = g->L[lev];
  void* p = nullptr;
  switch (which) {
    case 0: p = L.msk; break;
    case 1: p = L.A; break;
    case 2: p = L.x; break;
    case 3: p = L.b; break;
    case 4: p = L.r; break;
    default: return fail(F2D_E_ARG, "which must be 0..4");
  }
  *dev_ptr = p;
  *m = L.m;
  *n = L.n;
  return F2D_OK;
}
int f2d_gmg_vcycle(f2d_gmg* g, int lev1) {
  if (lev1 < 0 || lev1 >= g->nlev) return fail(F2D_E_ARG, "level out of range");
  g_vcycle(g, lev1);
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(g->s));
  return F2D_OK;
}
int f2d_gmg_fcycle(f2d_gmg* g, int lev1) {
  if (lev1 < 0 || lev1 >= g->nlev) return fail(F2D_E_ARG, "level out of range");
  g_fcycle(g, lev1);
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(g->s));
  return F2D_OK;
}
int f2d_gmg_twovcycle(f2d_gmg* g, double* x, const double* b) {
  g_twovcycle(g, x, b);
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(g->s));
  return F2D_OK;
}
int f2d_gmg_solve(f2d_gmg* g, double* x, const double* b, int maxite, double tol, int* nite,
                  double* res) {
  TRY(g_solve(g, x, b, maxite, tol, nite, res));
  HIPCHK(hipStreamSynchronize(g->s));
  return F2D_OK;
}

// ---- model ----
int f2d_model_create(const f2d_model_params* p, const int8_t* msk_host, f2d_model** out) {
  if (p->model != 0 && p->model != 1) return fail(F2D_E_ARG, "model must be 0 or 1");
  if (p->order < 1 || p->order > 6) return fail(F2D_E_ARG, "order must be 1..6");
  if (p->flux_splitting < 0 || p->flux_splitting > 1)
    return fail(F2D_E_ARG, "flux_splitting must be 0 (minmax) or 1 (parabolic)");
  if (p->timestepping != 0 && p->timestepping != 1)
    return fail(F2D_E_ARG, "timestepping must be 0 (RK3_SSP) or 1 (RK3)");
  f2d_model* M = new f2d_model();
  M->p = *p;
  M->m = p->ny + 2 * NH;
  M->n = p->nx + 2 * NH;
  M->nvar = p->model == 0 ? 5 : 6;
  M->fsz = (size_t)M->m * M->n;
  M->ssz = M->fsz * M->nvar;
  M->dx = p->Lx / p->nx;
  M->dy = p->Ly / p->ny;
  M->kforcing = p->timestepping == 0 ? 2 : 0;
  M->dtcoef = p->timestepping == 0 ? 1.5 : 1.;
  for (auto& d : M->diags) d = 0.;
  int rc;
  auto bail = [&](int code) {
    model_free(M);
    return code;
  };
  if ((rc = dalloc(&M->msk, M->fsz)) || (rc = dalloc(&M->mskp, M->fsz)) ||
      (rc = dalloc(&M->state, M->ssz)) || (rc = dalloc(&M->k[0], M->ssz)) ||
      (rc = dalloc(&M->k[1], M->ssz)) || (rc = dalloc(&M->k[2], M->ssz)) ||
      (rc = dalloc(&M->xs, M->ssz)) || (rc = dalloc(&M->work, M->fsz)) ||
      (rc = dalloc(&M->xr, M->fsz)) || (rc = dalloc(&M->yr, M->fsz)) ||
      (rc = dalloc(&M->partial, (size_t)kRedMaxBlocks * 8)) || (rc = dalloc(&M->dscal, 64)))
    return bail(rc);
  if (hipHostMalloc((void**)&M->hscal, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return bail(fail(F2D_E_NOMEM, "hipHostMalloc"));
  if (hipMemset(M->state, 0, M->ssz * 8) || hipMemset(M->k[0], 0, M->ssz * 8) ||
      hipMemset(M->k[1], 0, M->ssz * 8) || hipMemset(M->k[2], 0, M->ssz * 8) ||
      hipMemset(M->xs, 0, M->ssz * 8) || hipMemset(M->work, 0, M->fsz * 8) ||
      hipMemset(M->xr, 0, M->fsz * 8) || hipMemset(M->yr, 0, M->fsz * 8))
    return bail(fail(F2D_E_HIP, "hipMemset"));
  if (hipMemcpy(M->msk, msk_host, M->fsz, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(F2D_E_HIP, "upload mask"));
  // area = np.sum(msk[nh:-nh, nh:-nh]) (exact integer sum)
  long long a = 0;
  for (int j = NH; j < M->m - NH; j++)
    for (int i = NH; i < M->n - NH; i++) a += msk_host[(size_t)j * M->n + i];
  M->area = (double)a;
  // corner mask (core/operators.py:59-67); work keeps its zero last row/column
  launch_corner_mask(M->msk, M->mskp, M->m, M->n, M->s);
  if (hipStreamSynchronize(M->s) != hipSuccess) return bail(fail(F2D_E_HIP, "corner mask"));
  std::vector<int8_t> mskp(M->fsz);
  if (hipMemcpy(mskp.data(), M->mskp, M->fsz, hipMemcpyDeviceToHost) != hipSuccess)
    return bail(fail(F2D_E_HIP, "download corner mask"));
  if ((rc = gmg_build(p->nx, p->ny, M->dx, M->dy, p->hydroepsilon, 8. / 9., mskp.data(),
                      &M->gmg)))
    return bail(rc);
  M->cst[0] = M->dx;
  M->cst[1] = M->dy;
  M->cst[2] = 0.05;
  M->cst[3] = 0.;
  M->cst[4] = p->order % 2 == 0 ? 0. : p->aparab;
  *out = M;
  return F2D_OK;
}

int f2d_model_destroy(f2d_model* M) {
  model_free(M);
  return F2D_OK;
}
int f2d_model_state(f2d_model* M, double** state_dev, int* nvar, int* m, int* n) {
  *state_dev = M->state;
  *nvar = M->nvar;
  *m = M->m;
  *n = M->n;
  return F2D_OK;
}
int f2d_model_tendency(f2d_model* M, int k, double** dev) {
  if (k < 0 || k > 2) return fail(F2D_E_ARG, "k must be 0..2");
  *dev = M->k[k];
  return F2D_OK;
}
int f2d_model_corner_mask(f2d_model* M, int8_t** dev) {
  *dev = M->mskp;
  return F2D_OK;
}
int f2d_model_gmg(f2d_model* M, f2d_gmg** g) {
  *g = M->gmg;
  return F2D_OK;
}
int f2d_model_set_psi_from_vorticity(f2d_model* M) {
  TRY(invert_vorticity(M, M->state, true, true));
  HIPCHK(hipStreamSynchronize(M->s));
  return F2D_OK;
}
int f2d_model_invert_vorticity(f2d_model* M, double* x_dev, int flags) {
  TRY(invert_vorticity(M, x_dev, (flags & 1) != 0, (flags & 2) != 0));
  HIPCHK(hipStreamSynchronize(M->s));
  return F2D_OK;
}
int f2d_model_set_noslip(f2d_model* M, const int8_t* msknoslip_host, const int8_t* mskbc_host,
                         const double* xr0_host, const double* yr0_host, double bcarea,
                         double x2bc, double y2bc, int enforce_momentum) {
  if (!M || !msknoslip_host || !mskbc_host || !xr0_host || !yr0_host)
    return fail(F2D_E_ARG, "f2d_model_set_noslip: null argument");
  if (!(bcarea > 0.))
    return fail(F2D_E_ARG, "f2d_model_set_noslip: no boundary cells (bcarea = %g)", bcarea);
  const size_t sz = M->fsz;
  if (!M->msknoslip) {
    int rc;
    if ((rc = dalloc(&M->msknoslip, sz)) || (rc = dalloc(&M->mskbc, sz)) ||
        (rc = dalloc(&M->xr0, sz)) || (rc = dalloc(&M->yr0, sz)) || (rc = dalloc(&M->source_b, sz)))
      return rc;
    HIPCHK(hipMemset(M->source_b, 0, sz * sizeof(double)));
  }
  HIPCHK(hipMemcpy(M->msknoslip, msknoslip_host, sz, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(M->mskbc, mskbc_host, sz, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(M->xr0, xr0_host, sz * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(M->yr0, yr0_host, sz * sizeof(double), hipMemcpyHostToDevice));
  M->bcarea = bcarea;
  M->x2bc = x2bc;
  M->y2bc = y2bc;
  M->enforce = enforce_momentum != 0;
  M->noslip = true;
  return F2D_OK;
}
int f2d_model_set_islands(f2d_model* M, const double* rhsp_host, const double* psi_host) {
  if (!M || !rhsp_host || !psi_host) return fail(F2D_E_ARG, "f2d_model_set_islands: null argument");
  const size_t sz = M->fsz;
  if (!M->rhsp) {
    int rc;
    if ((rc = dalloc(&M->rhsp, sz)) || (rc = dalloc(&M->ipsi, sz))) return rc;
  }
  HIPCHK(hipMemcpy(M->rhsp, rhsp_host, sz * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(M->ipsi, psi_host, sz * sizeof(double), hipMemcpyHostToDevice));
  M->isisland = true;
  return F2D_OK;
}
int f2d_model_rhs_adv(f2d_model* M, const double* x_dev, double* dxdt_dev) {
  TRY(rhs_adv(M, x_dev, dxdt_dev));
  HIPCHK(hipStreamSynchronize(M->s));
  return F2D_OK;
}
int f2d_model_set_umax(f2d_model* M, double umax) {
  M->cst[3] = umax;
  return F2D_OK;
}
int f2d_model_step(f2d_model* M, double dt) {
  TRY(model_step(M, dt));
  HIPCHK(hipStreamSynchronize(M->s));
  return F2D_OK;
}
int f2d_model_diagnostics(f2d_model* M, const double* xr_host, const double* yr_host,
                          double* out) {
  if (xr_host) HIPCHK(hipMemcpy(M->xr, xr_host, M->fsz * 8, hipMemcpyHostToDevice));
  if (yr_host) HIPCHK(hipMemcpy(M->yr, yr_host, M->fsz * 8, hipMemcpyHostToDevice));
  return model_diag(M, out);
}
int f2d_model_diag_fluxes(f2d_model* M, const double* x_dev, double dt, double* flx_dev) {
  if (!M || !x_dev || !flx_dev) return fail(F2D_E_ARG, "f2d_model_diag_fluxes: null argument");
  if (!(dt != 0.)) return fail(F2D_E_ARG, "f2d_model_diag_fluxes: dt must be nonzero");
  TRY(diag_fluxes(M, x_dev, dt, flx_dev));
  HIPCHK(hipStreamSynchronize(M->s));
  return F2D_OK;
}
int f2d_model_run(f2d_model* M, int nsteps, double dt, int* nite_log) {
  double out[8];
  for (int k = 0; k < nsteps; k++) {
    M->cst[3] = M->diags[0];
    TRY(model_step(M, dt));
    if (nite_log) nite_log[k] = M->last_nite;
    TRY(model_diag(M, out));
  }
  return F2D_OK;
}
int f2d_model_last_solve(f2d_model* M, int* nite, double* res) {
  *nite = M->last_nite;
  *res = M->last_res;
  return F2D_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// multigrid set-up shared with the batched engine (f2d_engine.hip): builds a
// hierarchy on the device and hands out its level masks and operators.
// ---------------------------------------------------------------------------
int f2d_gmg_build_for_engine(int nx, int ny, double dx, double dy, double hydroepsilon,
                             const int8_t* mskp_host, void** gmg_out) {
  f2d_gmg* g = nullptr;
  int rc = gmg_build(nx, ny, dx, dy, hydroepsilon, 8. / 9., mskp_host, &g);
  *gmg_out = g;
  return rc;
}
int f2d_gmg_level_raw(void* gv, int lev, int8_t** msk, double** A, int* m, int* n) {
  f2d_gmg* g = (f2d_gmg*)gv;
  *msk = g->L[lev].msk;
  *A = g->L[lev].A;
  *m = g->L[lev].m;
  *n = g->L[lev].n;
  return F2D_OK;
}
void f2d_gmg_free_raw(void* g) { gmg_free((f2d_gmg*)g); }
