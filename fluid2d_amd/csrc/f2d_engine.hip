// f2d_engine.hip -- batched engine kernels, launch program and hipGraph replay.
//
// Kernels (each a fused form of reference routines; FP64 operation order is
// kept, so results are bit-identical to the unfused path and the oracle):
//   k_rhs        stage state formed on the fly (core/timescheme.py:167-180)
//                + advection (core/fortran_advection.f90) [+ y*=msk + torque,
//                core/fortran_operators.f90:330-381] [+ diffusion, :125-156]
//                + periodic images (the fills of core/operators.py)
//   k_c2c        celltocorner (core/fortran_operators.f90:44-64)
//   k_smooth     smoothtwicewithA + fill (core/gmg/fortran_multigrid.f90:2-127,
//                core/gmg/level.py:340-368) in one launch through an LDS tile;
//                input x, zero, x + I(xc) (hierarchy.py:123-126, fused
//                prolongation and correction) or I(xc) (hierarchy.py:145-151)
//   k_resid_restrict  residual + fill + full weighting + coarse fill
//                (hierarchy.py:105-108), optional residual store and norms
//   k_coarse     all levels below the fine part of one instance in one
//                workgroup, operands in LDS (V-cycle or F-cycle tail)
//   k_ortho      computeorthogradient (core/fortran_operators.f90:2-39)
//   k_combine    final RK combination (core/timescheme.py:180)
//   k_update, k_fold, k_npsum, k_finalize_psi, k_diag: solve control,
//                norms, zero-mean psi and diagnostics, all on the device.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/fluid2d_amd.h"
#include "f2d_dev.h"
#include "f2d_engine.h"

namespace f2d {
namespace eng {

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
// both flags read at once, no branch between the loads
__device__ __forceinline__ bool skip(const DynD* dyn, int inst, int pred) {
  const int run = dyn[inst].run, act = dyn[inst].active;
  return (run == 0) | ((pred != 0) & (act == 0));
}
__device__ __forceinline__ bool idle(const DynD* dyn, int inst) { return !dyn[inst].run; }

// Task of this workgroup.  Workgroups are dealt to the 8 XCDs round-robin
// (block b runs on XCD b mod 8); each XCD takes a contiguous run of the
// task list as the builder made it (tasks ordered by instance, level and
// tile row), which keeps neighbouring tiles -- and the same tiles in
// consecutive launches -- on one XCD and its L2.  The device copy of a
// tile launch's task list is stored in dispatch order (xcd_task below, at
// upload), so workgroup b reads task b: no grid size is read (a memory
// round trip before the task read), and a task's place in the builder's
// list (partial-sum slot) is its field slot.
constexpr int NXCD = 8;
__device__ __forceinline__ int task_index() { return blockIdx.x; }
// builder's task index of workgroup b of an n-workgroup launch
__host__ __device__ inline int xcd_task(int b, int n) {
  const int q = n / NXCD, r = n % NXCD, x = b % NXCD, i = b / NXCD;
  return x * q + (x < r ? x : r) + i;
}

// AND of a predicate over the 256 threads of a (64, 4) workgroup: one vote
// per wavefront, combined through LDS.  (The runtime's block-wide AND reads
// the block size from the dispatch arguments in the middle of the kernel, a
// memory round trip on the critical path.)  Once per kernel.
__device__ __forceinline__ int block_all4(bool p) {
  __shared__ int wf[4];
  const int w = __all(p);
  if (threadIdx.x == 0) wf[threadIdx.y] = w;
  __syncthreads();
  return wf[0] & wf[1] & wf[2] & wf[3];
}

// Step gate (first launch of a step): an instance runs this step while its
// step budget lasts; every kernel of the step skips instances that do not.
__global__ void k_gate(DynD* dyn, int ninst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  DynD& d = dyn[i];
  d.run = d.steps_left != 0;
  if (d.steps_left > 0) d.steps_left -= 1;
  if (d.run) d.steps_done += 1;
}

// stage state (core/timescheme.py:167-180, :152-164), NumPy order
// stage 0: x; RK3_SSP 1: x + dt*k0; 2: x + (0.25 dt)(k0 + k1);
// RK3 1: x + (dt/3) k0; 2: x + (0.5 dt) k1
// The state is FP64; tendencies are T (double, or float in mixed mode) and
// are widened, so in FP64 mode this is exactly the reference's expression.
template <typename T>
struct Stage {
  const double* x;
  const T *k0, *k1;
  double c;
  int kind;  // 0 x, 1 x + c*k0, 2 x + c*(k0+k1), 3 x + c*k1
  __device__ __forceinline__ double operator()(size_t q) const {
    switch (kind) {
      case 0: return x[q];
      case 1: return x[q] + c * (double)k0[q];
      case 2: return x[q] + c * ((double)k0[q] + (double)k1[q]);
      default: return x[q] + c * (double)k1[q];
    }
  }
  // the same with the kind fixed at compile time (no branch between loads)
  // and a 32-bit element offset into the instance's planes
  template <int KIND>
  __device__ __forceinline__ double at(uint32_t q) const {
    if constexpr (KIND == 0) return ld32(x, q);
    if constexpr (KIND == 1) return ld32(x, q) + c * (double)ld32(k0, q);
    if constexpr (KIND == 2) return ld32(x, q) + c * ((double)ld32(k0, q) + (double)ld32(k1, q));
    if constexpr (KIND == 3) return ld32(x, q) + c * (double)ld32(k1, q);
    return 0.;
  }
};

// ---------------------------------------------------------------------------
// k_rhs: tendencies of the tracers for one RK stage.  The stage state of the
// tracer(s) and of (u, v) is formed once per cell (FP64) into LDS tiles
// (3-cell halo for the tracer, 1 for the velocities); each edge flux of the
// tile is then computed once into LDS, and each cell takes the divergence,
// torque and diffusion.
//
// FP64 (T = double): the reference arithmetic, operation for operation.
// Mixed (T = float): the flux arithmetic runs in FP32 on differences
// q_k - q_L taken in FP64 against the tracer q_L of the edge's left (lower)
// cell, and q_L times the flux of a unit field is added in FP64.  The flux
// is linear in q, so the sum is the same flux; rounding then scales with the
// local variation of q instead of with q itself (a plain FP32 evaluation
// loses about L/dx digits to the cancellation in f_e - f_w), and the
// divergence is taken in FP64.  Torque and diffusion difference in FP64 and
// multiply in FP32.
// ---------------------------------------------------------------------------
constexpr int QW = TX + 6, QH = TY + 6;

// Load phase of a 256-thread tile: element q = tid + 256 k for k < IT,
// every load issued before the first store (load(q) sees q clamped to
// [0, N), store runs for q < N only).
template <int IT, typename LD, typename ST>
__device__ __forceinline__ void tile_load(int tid, int N, LD load, ST store) {
  decltype(load(0)) v[IT];
#pragma unroll
  for (int k = 0; k < IT; k++) v[k] = load(min(tid + 256 * k, N - 1));
#pragma unroll
  for (int k = 0; k < IT; k++)
    if (tid + 256 * k < N) store(tid + 256 * k, v[k]);
}

struct TrVal {
  double q0, q1;
  int8_t m;
};

// NTR: tracers advected (1 Euler, 2 Boussinesq); each launch holds instances
// of one model
template <typename T, int NTR>
__global__ void __launch_bounds__(256) k_rhs(const Task* __restrict__ tasks,
                                             const InstD* __restrict__ I, double* P, T* Q,
                                             const int8_t* __restrict__ B,
                                             const DynD* __restrict__ dyn, int stage, int out_k) {
  __shared__ double qs[NTR][QH * QW];
  __shared__ double us[TY * (TX + 1)];
  __shared__ double vs[(TY + 1) * TX];
  __shared__ int8_t ms[QH * QW];
  const int ti = task_index();
  const Task tk = tasks[ti];
  if (idle(dyn, tk.inst)) return;
  const InstD& in = I[tk.inst];
  const int m = in.m, n = in.n;
  const size_t F = (size_t)m * n;
  const DynD& d = dyn[tk.inst];
  const double dt = d.dt;
  Stage<T> S;
  S.x = P + in.state;
  S.k0 = Q + in.k[0];
  S.k1 = Q + in.k[1];
  if (stage == 0) {
    S.kind = 0;
    S.c = 0.;
  } else if (in.tscheme == 0) {
    S.kind = stage == 1 ? 1 : 2;
    S.c = stage == 1 ? dt : 0.25 * dt;
  } else {
    S.kind = stage == 1 ? 1 : 3;
    S.c = stage == 1 ? dt / 3. : 0.5 * dt;
  }
  const int8_t* msk = B + in.msk;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x, nt = blockDim.x * blockDim.y;
  const int y0 = tk.y0, x0 = tk.x0;
  // tiles: tracer rows [y0-3, y0+TY+2] x cols [x0-3, x0+TX+2]; each tile's
  // loads are issued together (compile-time trip counts, clamped in-bounds
  // addresses), then stored
  (void)nt;
  auto load_tiles = [&](auto kind) {
    constexpr int K = decltype(kind)::value;
    tile_load<(QH * QW + 255) / 256>(tid, QH * QW, [&](int q) {
      const int r = q / QW, c = q - r * QW;
      const int j = y0 - 3 + r, i = x0 - 3 + c;
      const bool ok = j < m && i < n;
      const uint32_t e = ok ? at32(j, i, n) : 0u;
      TrVal v;
      v.m = ld32(msk, e);
      v.q0 = S.template at<K>(e);
      v.q1 = NTR > 1 ? S.template at<K>((uint32_t)(4 * F) + e) : 0.;
      if (!ok) {
        v.m = 0;
        v.q0 = v.q1 = 0.;
      }
      return v;
    }, [&](int q, const TrVal& v) {
      ms[q] = v.m;
      qs[0][q] = v.q0;
      if constexpr (NTR > 1) qs[NTR - 1][q] = v.q1;
    });
    tile_load<(TY * (TX + 1) + 255) / 256>(tid, TY * (TX + 1), [&](int q) {
      const int r = q / (TX + 1), c = q - r * (TX + 1);  // u on rows y0.., cols x0-1..
      const int j = y0 + r, i = x0 - 1 + c;
      const bool ok = j < m && i < n;
      const double v = S.template at<K>((uint32_t)(2 * F) + (ok ? at32(j, i, n) : 0u));
      return ok ? v : 0.;
    }, [&](int q, double v) { us[q] = v; });
    tile_load<((TY + 1) * TX + 255) / 256>(tid, (TY + 1) * TX, [&](int q) {
      const int r = q / TX, c = q - r * TX;  // v on rows y0-1.., cols x0..
      const int j = y0 - 1 + r, i = x0 + c;
      const bool ok = j < m && i < n;
      const double v = S.template at<K>((uint32_t)(3 * F) + (ok ? at32(j, i, n) : 0u));
      return ok ? v : 0.;
    }, [&](int q, double v) { vs[q] = v; });
  };
  auto by_kind = [&]() {
    switch (S.kind) {
      case 0: load_tiles(std::integral_constant<int, 0>()); break;
      case 1: load_tiles(std::integral_constant<int, 1>()); break;
      case 2: load_tiles(std::integral_constant<int, 2>()); break;
      default: load_tiles(std::integral_constant<int, 3>()); break;
    }
  };
  by_kind();
  __syncthreads();
  // advection parameters (core/fortran_advection.f90:46-60), cst[3] = umax
  UpParams<double> p;
  const double umax = d.umax, logcosh = 0.05;
  p.zdx = 1. / in.dx;
  p.zdy = 1. / in.dy;
  {
    const double a = logcosh / umax;
    p.a = a;
    p.KK = (umax - log(cosh(umax * a)) / a) * 0.5;
    const double u1 = in.aparab * umax;
    p.u1 = u1;
    p.aa = 1. / (2. * u1);
    p.bb = u1 * 0.5;
  }
  p.method = in.method;
  p.order = in.order;
  UpParams<T> pt;
  pt.zdx = T(p.zdx);
  pt.zdy = T(p.zdy);
  pt.a = T(p.a);
  pt.KK = T(p.KK);
  pt.u1 = T(p.u1);
  pt.aa = T(p.aa);
  pt.bb = T(p.bb);
  pt.method = p.method;
  pt.order = p.order;
  const bool upwind = in.order % 2 == 1;
  const bool kforce = stage == (in.tscheme == 0 ? 2 : 0);
  T* out = Q + in.k[out_k];
  constexpr bool MIXED = sizeof(T) == 4;
  // Edge fluxes, each computed once: fx[r][e] on the x-edge between tile
  // cells (r, e-1) and (r, e), fy[e][c] on the y-edge between (e-1, c) and
  // (e, c).  FP64: the reference's flux arithmetic.  Mixed: the flux is
  // linear in q, F(q) = F(q - q_L) + q_L F(1) with q_L the tracer of the
  // edge's left (lower) cell: the first term in FP32 on differences taken
  // in FP64, the unit-field flux F(1) and the sum in FP64, so the rounding
  // scales with the local variation of q.
  // Each edge's flux overwrites the edge velocity it was computed from (the
  // same thread reads and writes the same slot; with two tracers the thread
  // keeps its edges' velocities in registers for the second), so the tile
  // needs no flux arrays of its own: four workgroups per CU with one tracer,
  // three with two.
  double(*fx)[TX + 1] = (double(*)[TX + 1])us;
  double(*fy)[TX] = (double(*)[TX])vs;
  constexpr int KU = (TY * (TX + 1) + 255) / 256, KV = ((TY + 1) * TX + 255) / 256;
  double uvel[NTR > 1 ? KU : 1], vvel[NTR > 1 ? KV : 1];
  (void)uvel;
  (void)vvel;
  auto edge_flux = [&](const double* qt, int lq, ptrdiff_t st, double vel) -> double {
    const double ql = MIXED ? qt[lq] : 0.;
    auto qv = [qt, lq, st, ql](int k) { return T(qt[lq + k * st] - ql); };
    T f;
    if (upwind)
      f = up_edge_flux_q(ms + lq, st, T(vel), pt, qv);
    else
      f = ce_edge_flux_q(ms + lq, st, T(vel), in.order, true, qv);
    if (!MIXED) return (double)f;
    auto one = [](int) { return 1.; };
    const double u1 = upwind ? up_edge_flux_q(ms + lq, st, vel, p, one)
                             : ce_edge_flux_q(ms + lq, st, vel, in.order, true, one);
    return (double)f + ql * u1;
  };
#pragma unroll
  for (int t = 0; t < NTR; t++) {
    const double* qt = qs[t];
    if (t > 0) __syncthreads();  // the previous tracer's divergence read fx, fy
    if constexpr (NTR == 1) {
      for (int q = tid; q < TY * (TX + 1); q += 256) {
        const int r = q / (TX + 1), e = q - r * (TX + 1);
        fx[r][e] = edge_flux(qt, (r + 3) * QW + e + 2, 1, us[q]);
      }
      for (int q = tid; q < (TY + 1) * TX; q += 256) {
        const int e = q / TX, c = q - e * TX;
        fy[e][c] = edge_flux(qt, (e + 2) * QW + c + 3, (ptrdiff_t)QW, vs[q]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < KU; k++) {
        const int q = tid + 256 * k;
        if (q < TY * (TX + 1)) {
          const int r = q / (TX + 1), e = q - r * (TX + 1);
          if (t == 0) uvel[k] = us[q];
          fx[r][e] = edge_flux(qt, (r + 3) * QW + e + 2, 1, uvel[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < KV; k++) {
        const int q = tid + 256 * k;
        if (q < (TY + 1) * TX) {
          const int e = q / TX, c = q - e * TX;
          if (t == 0) vvel[k] = vs[q];
          fy[e][c] = edge_flux(qt, (e + 2) * QW + c + 3, (ptrdiff_t)QW, vvel[k]);
        }
      }
    }
    __syncthreads();
    for (int r = threadIdx.y; r < TY; r += blockDim.y) {
      const int J = y0 + r, Ic = x0 + threadIdx.x;
      if (J >= m - NH || Ic >= n - NH) continue;
      const int lq = (r + 3) * QW + threadIdx.x + 3;  // cell in the tracer/mask tile
      const int c = threadIdx.x;
      T val;
      if (MIXED)
        val = T(-p.zdx * (fx[r][c + 1] - fx[r][c]) - p.zdy * (fy[r + 1][c] - fy[r][c]));
      else
        val = T(-pt.zdx * T(fx[r][c + 1] - fx[r][c]) - pt.zdy * T(fy[r + 1][c] - fy[r][c]));
      const int8_t mc = ms[lq];
      if constexpr (NTR > 1) {
        if (t == 0) {  // Boussinesq: y *= msk; add_torque
          val = val * T(mc);
          int ml = mc + ms[lq - 1], mr = ms[lq + 1] + mc;
          ml = ml < 1 ? 1 : ml;
          mr = mr < 1 ? 1 : mr;
          if (ml + mr == 4)
            val = val + T(qs[NTR - 1][lq + 1] - qs[NTR - 1][lq - 1]) * T(0.5 * in.gravity / in.dx) *
                            T(mc);
        }
      }
      if (in.diffusion && kforce && mc == 1) {  // add_diffusion
        const double coefd =
            (in.model == 0 ? 1. : (in.tscheme == 0 ? 1.5 : 1.)) * (t == 0 ? in.kd0 : in.kd1);
        const T coef = T(coefd / (in.dx * in.dx));
        const double tc = qt[lq];
        T sm = T(ms[lq - 1]) * T(qt[lq - 1] - tc);
        sm = sm + T(ms[lq + 1]) * T(qt[lq + 1] - tc);
        sm = sm + T(ms[lq - QW]) * T(qt[lq - QW] - tc);
        sm = sm + T(ms[lq + QW]) * T(qt[lq + QW] - tc);
        val = val + coef * sm;
      }
      store_img_mask32<T>(out + (t == 0 ? 0 : 4 * F), nullptr, J, Ic, val, m, n);
    }
  }
}

// ---------------------------------------------------------------------------
// k_c2c: dst = celltocorner(src) on [0, m-2] x [0, n-2]
// ---------------------------------------------------------------------------
// T: source type, D: destination type (an FP32 source summed in FP64 is exact)
template <typename T, typename D>
__global__ void __launch_bounds__(256) k_c2c(const Task* __restrict__ tasks,
                                             const InstD* __restrict__ I, const T* SP, D* DP,
                                             const DynD* __restrict__ dyn, int pred) {
  const int ti = task_index();
  const Task tk = tasks[ti];
  if (skip(dyn, tk.inst, pred)) return;
  const InstD& in = I[tk.inst];
  const int m = in.m, n = in.n;
  const T* src = SP + tk.xin;
  D* dst = DP + tk.xout;
  // task flag 1: minus the islands' rhs correction (core/operators.py:444-446)
  const double* rp = (tk.flags & 1) ? I[tk.inst].rhsp : nullptr;
  for (int r = threadIdx.y; r < TY; r += blockDim.y) {
    const int j = tk.y0 + r, i = tk.x0 + threadIdx.x;
    if (j >= m - 1 || i >= n - 1) continue;
    const size_t c = at(j, i, n);
    D v = D(0.25) * (D(src[c]) + D(src[c + 1]) + D(src[c + n]) + D(src[c + n + 1]));
    if (rp) v = v - D(rp[c]);
    dst[c] = v;
  }
}

// ---------------------------------------------------------------------------
// k_smooth: two damped-Jacobi sweeps + fill, LDS tile.
// MODE 0: x = xin; 1: x = 0; 2: x = xin + I(xc); 3: x = I(xc).
// FINAL: each stored copy multiplied by mskp (psi * mskp).
// ---------------------------------------------------------------------------
// One tile (task ti of `tasks`) of a (64, 4) workgroup; the kernel k_smooth
// below runs one tile per workgroup, the persistent kernel k_persist any
// number of them one after another.
// Phase timestamps of the 64 x 16 smoother tiles (profiling build only,
// make stamps): thread 0 of each of the first 2048 workgroups of a launch
// records the real-time clock (100 MHz) at the start, after the tile loads
// are issued and the step gate read, after the tile is in LDS, after the
// prolongation, after each sweep; slot 7 holds the compute unit and the
// tile.  Each launch overwrites the previous one's records.
#ifdef F2D_PHASE_STAMPS
__device__ unsigned long long smooth_stamps[2048 * 8];
#define F2D_STAMP(k)                                                               \
  do {                                                                             \
    if (TYV == 16 && threadIdx.x == 0 && threadIdx.y == 0 && blockIdx.x < 2048)     \
      smooth_stamps[blockIdx.x * 8 + (k)] = (unsigned long long)wall_clock64();      \
  } while (0)
#define F2D_STAMP_SYNC(k) \
  do {                    \
    __syncthreads();      \
    F2D_STAMP(k);         \
  } while (0)
#else
#define F2D_STAMP(k) \
  do {               \
  } while (0)
#define F2D_STAMP_SYNC(k) \
  do {                    \
  } while (0)
#endif

// ADD (task flag 4): the stored value is also added to the FP64 plane at
// task offset xcc (in the P pool; flag 8: in the T pool of an FP64 batch)
// at every position stored -- x0 += x (hierarchy.py:171) folded into the
// last smoothing of a refinement V-cycle (no tile of the launch reads that
// plane); a separate instantiation, so the other smoothings carry none of
// its code.
template <typename T, int TYV, bool ADD = false>
__device__ __forceinline__ void smooth_tile(const Task* __restrict__ tasks, const int ti,
                                            const InstD* __restrict__ I, T* Q, double* P,
                                            const int8_t* __restrict__ B,
                                            const DynD* __restrict__ dyn, int pred) {
  // Block (64, 4).  Each phase first issues the global loads of all its
  // iterations (compile-time trip counts, clamped in-bounds addresses, no
  // branch between the loads), then computes: one memory round trip per
  // phase instead of one per iteration.
  constexpr int NT = 256;
  constexpr int XW = TX + 4, XH = TYV + 4, JW = TX + 2, JH = TYV + 2;
  constexpr int N1 = XH * XW, IT1 = (N1 + NT - 1) / NT;
  constexpr int N2 = JH * JW, IT2 = (N2 + NT - 1) / NT;
  constexpr int IT3 = TYV / 4;
  __shared__ T xs[XH][XW];
  __shared__ T j1[JH][JW];
  __shared__ T bs[JH][JW];         // right-hand side on the first sweep's region
  __shared__ uint8_t sbs[XH][XW];  // stencil bits (finest level)
  F2D_STAMP(0);
  const Task tk = tasks[ti];
  // the skip test is read and taken after the tile loads are issued (the
  // task holds the addressing, so the data loads follow the task read)
  const int MODE = tk.mode;
#ifdef F2D_PHASE_STAMPS
  if (TYV == 16 && threadIdx.x == 0 && threadIdx.y == 0 && blockIdx.x < 2048)
    smooth_stamps[blockIdx.x * 8 + 7] =
        ((unsigned long long)__smid() << 32) | (unsigned)(MODE | (tk.y0 << 4) | (tk.x0 << 16));
#endif
  const int final_mask = tk.flags & 1;
  const InstD& in = I[tk.inst];
  const int m = tk.m, n = tk.n;
  const size_t PL = (size_t)m * n;
  const int8_t* msk = B + tk.msk;
  const T* A = Q + tk.A;
  const T* b = Q + tk.b;
  const T* xin = Q + tk.xin;
  const int tid = threadIdx.y * 64 + threadIdx.x;
  const int y0 = tk.y0, x0 = tk.x0;
  const bool fly = tk.sb >= 0;
  const uint8_t* sb = (const uint8_t*)(B + (fly ? tk.sb : 0));
  const bool rdx = MODE == 0 || MODE == 2, rdi = MODE == 2 || MODE == 3;
  struct {
    int m, n;
  } C = {tk.m2, tk.n2};
  const int8_t* msk2 = B + tk.msk2;
  const T* xc = Q + tk.xc;
  int full_tile = 0;
  // One round trip: every global load of the tile is issued before the
  // first is used -- the right-hand side of both sweeps, x and the stencil
  // bits on the first sweep's stencil region, and for a prolongation the
  // fine mask there and the coarse tile under it.
  //
  // Two forms of the same computation.  The general form maps thread t to
  // the flat region positions t, t + 256, ... and clips them to the array.
  // The unclipped form serves tiles whose x region lies inside the array
  // (every tile of a grid that the tile size divides): thread (x, y) holds
  // column x of the rows y, y + 4, ..., the four (x region) or two
  // (first-sweep region) columns past the 64th go to the first threads of
  // the block, and every offset is the thread's base plus a multiple of
  // four rows -- no division, bounds test or select per element.
  constexpr int KX = XH / 4, KB = (JH + 3) / 4;
  static_assert(XH % 4 == 0 && KB <= IT2 && KX + 1 <= IT1, "unclipped form registers");
  const bool lean = y0 + TYV + 1 <= m - 2 && x0 + TX + 1 <= n - 2;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const uint32_t s4 = 4u * (uint32_t)n;
  const uint32_t ox = at32(y0 - 2 + ty, x0 - 2 + tx, n);  // x region (ty, tx)
  const uint32_t ob = ox + (uint32_t)n + 1u;              // first-sweep region (ty, tx)
  const bool xe_on = tid < 4 * XH, be_on = tid < 2 * JH;
  const int xer = xe_on ? tid >> 2 : 0, xec = TX + (tid & 3);
  const int ber = be_on ? tid >> 1 : 0, bec = TX + (tid & 1);
  const uint32_t oxe = at32(y0 - 2 + xer, x0 - 2 + xec, n);
  const uint32_t obe = at32(y0 - 1 + ber, x0 - 1 + bec, n);
  uint32_t cc2[IT2];
  bool ok2[IT2];
  T bv[IT2];
  T be = T(0);
  // the plane the stored values are added to (task flag 4), and its values
  // at the tile's cells in the unclipped form, loaded with the tile
  const bool addp = ADD && (tk.flags & 4) != 0;
  double* pw = nullptr;
  if constexpr (sizeof(T) == 8) {
    pw = addp ? ((tk.flags & 8) ? (double*)Q : P) + tk.xcc : nullptr;
  } else {
    pw = addp ? P + tk.xcc : nullptr;
  }
  double pv[IT3];
  if (lean) {
#pragma unroll
    for (int k = 0; k < KB; k++) bv[k] = ld32(b, ty + 4 * k < JH ? ob + k * s4 : ob);
    be = ld32(b, obe);
    if (addp) {
#pragma unroll
      for (int k = 0; k < IT3; k++) pv[k] = ld32(pw, ob + (uint32_t)n + 1u + k * s4);
    }
  } else {
#pragma unroll
    for (int it = 0; it < IT2; it++) {
      const int q = tid + it * NT;
      const int rr = q / JW, cc = q - rr * JW;
      const int j = y0 - 1 + rr, i = x0 - 1 + cc;
      ok2[it] = q < N2 && j <= m - NH && i <= n - NH;
      cc2[it] = ok2[it] ? at32(j, i, n) : 0u;
    }
#pragma unroll
    for (int it = 0; it < IT2; it++) bv[it] = ld32(b, cc2[it]);
  }
  const int Ic = x0 + threadIdx.x;
  // x region: rows [y0-2, y0+TY+1] x cols [x0-2, x0+TX+1], clipped to
  // [1, m-2]; the coarse points its prolongation reads are rows
  // [y0>>1, +CH) x cols [x0>>1, +CW) (clamped to the coarse array)
  constexpr int CH = TYV / 2 + 4, CW = TX / 2 + 4, NC = CH * CW, ITC = (NC + NT - 1) / NT;
  __shared__ T cxs[CH][CW];
  __shared__ int8_t cms[CH][CW];
  __shared__ int8_t ms1[XH][XW];  // fine mask of the prolongation
  const int cy0 = y0 >> 1, cx0 = x0 >> 1;
  bool ok1[IT1];
  T xv[IT1];
  uint8_t bt[IT1];
  int8_t m1v[IT1];
  T cxv[ITC];
  int8_t cmv[ITC];
  // the fine mask of the prolongation: bit 4 of the stencil bits where
  // they exist (the level's 0/1 mask), else loaded
  if (lean) {
#pragma unroll
    for (int k = 0; k <= KX; k++) {  // rows ty + 4k, then the extra columns
      const uint32_t c = k < KX ? ox + k * s4 : oxe;
      xv[k] = rdx ? ld32(xin, c) : T(0);
      bt[k] = fly ? ld32(sb, c) : (uint8_t)0;
      m1v[k] = (rdi && !fly) ? ld32(msk, c) : (int8_t)0;
    }
  } else {
#pragma unroll
    for (int it = 0; it < IT1; it++) {
      const int q = tid + it * NT;
      const int rr = q / XW, cc = q - rr * XW;
      const int j = y0 - 2 + rr, i = x0 - 2 + cc;
      ok1[it] = q < N1 && j <= m - 2 && i <= n - 2;
      const uint32_t c = ok1[it] ? at32(j, i, n) : at32(1, 1, n);
      xv[it] = rdx ? ld32(xin, c) : T(0);
      bt[it] = fly ? ld32(sb, c) : (uint8_t)0;
      m1v[it] = (rdi && !fly) ? ld32(msk, c) : (int8_t)0;
    }
  }
  if (rdi) {
#pragma unroll
    for (int it = 0; it < ITC; it++) {
      const int q = min(tid + it * NT, NC - 1);
      const int rr = q / CW, cc = q - rr * CW;
      const uint32_t c = at32(min(cy0 + rr, C.m - 1), min(cx0 + cc, C.n - 1), C.n);
      cxv[it] = ld32(xc, c);
      cmv[it] = ld32(msk2, c);
    }
  }
  if (skip(dyn, tk.inst, pred)) return;
  F2D_STAMP(1);
  bool full = true;
  if (lean) {
#pragma unroll
    for (int k = 0; k < KB; k++)
      if (ty + 4 * k < JH) bs[ty + 4 * k][tx] = bv[k];
    if (be_on) bs[ber][bec] = be;
#pragma unroll
    for (int k = 0; k <= KX; k++) {
      if (k == KX && !xe_on) continue;
      const int rr = k < KX ? ty + 4 * k : xer, cc = k < KX ? tx : xec;
      xs[rr][cc] = xv[k];
      sbs[rr][cc] = bt[k];
      ms1[rr][cc] = fly ? (int8_t)((bt[k] >> 4) & 1) : m1v[k];
      full = full && bt[k] == 31;
    }
  } else {
#pragma unroll
    for (int it = 0; it < IT2; it++) {
      const int q = tid + it * NT;
      if (q < N2) {
        const int rr = q / JW, cc = q - rr * JW;
        bs[rr][cc] = ok2[it] ? bv[it] : T(0);
      }
    }
#pragma unroll
    for (int it = 0; it < IT1; it++) {
      const int q = tid + it * NT;
      if (q < N1) {
        const int rr = q / XW, cc = q - rr * XW;
        xs[rr][cc] = ok1[it] ? xv[it] : T(0);
        sbs[rr][cc] = ok1[it] ? bt[it] : (uint8_t)0;
        ms1[rr][cc] = fly ? (int8_t)((bt[it] >> 4) & 1) : m1v[it];
        full = full && ok1[it] && bt[it] == 31;
      }
    }
  }
  if (rdi) {
#pragma unroll
    for (int it = 0; it < ITC; it++) {
      const int q = tid + it * NT;
      if (q < NC) {
        const int rr = q / CW, cc = q - rr * CW;
        cxs[rr][cc] = cxv[it];
        cms[rr][cc] = cmv[it];
      }
    }
    __syncthreads();
    // prolongation (ip_eval: interp_at's expressions) of the coarse tile,
    // added to x (MODE 2) or replacing it (MODE 3).  Work goes by coarse
    // cell: coarse (J2, I2) and its right/lower neighbours give the four
    // fine points (2 J2 - 2 + dj, 2 I2 - 2 + di), whose parities are
    // (dj, di), so the operands are read once for four points.
    constexpr int CRH = XH / 2 + 1, CRW = XW / 2 + 1, NCR = CRH * CRW;
#pragma unroll
    for (int it = 0; it < (NCR + NT - 1) / NT; it++) {
      const int q = tid + it * NT;
      if (q >= NCR) continue;
      const int a = q / CRW, bc = q - a * CRW;
      IpOps<T> o;
      o.a = cxs[a][bc];
      o.b = cxs[a][bc + 1];
      o.c = cxs[a + 1][bc];
      o.d = cxs[a + 1][bc + 1];
      o.sa = cms[a][bc];
      o.sb = cms[a][bc + 1];
      o.sc = cms[a + 1][bc];
      o.sd = cms[a + 1][bc + 1];
#pragma unroll
      for (int dj = 0; dj < 2; dj++)
#pragma unroll
        for (int di = 0; di < 2; di++) {
          const int j = 2 * (cy0 + a) - 2 + dj, i = 2 * (cx0 + bc) - 2 + di;
          const int rr = j - (y0 - 2), cc = i - (x0 - 2);
          if (rr < 0 || rr >= XH || cc < 0 || cc >= XW || j > m - 2 || i > n - 2) continue;
          o.m1 = ms1[rr][cc];
          o.par = di | (dj << 1);
          const T r = ip_eval<T>(o);
          xs[rr][cc] = MODE == 2 ? xs[rr][cc] + r : r;
        }
    }
  }
  // every cell of the tile has the full 9-point stencil: constant weights
  F2D_STAMP_SYNC(2);
  full_tile = block_all4(fly && full);
  F2D_STAMP(3);
  const T c1 = T(in.omega), c2 = T(1. - in.omega);
  T* xout = Q + tk.xout;
  const int8_t* mp = final_mask ? B + in.mskp : nullptr;
  T w[5];
  for (int k = 0; k < 5; k++) w[k] = T(in.w0[k]);
  // the plane += the value stored at interior (J, I) and at each image
  auto add_img = [&](int J, int I, T v) {
    each_img(J, I, m, n, [&](int j, int i) {
      const uint32_t c = at32(j, i, n);
      const T sv = mp ? v * T(ld32(mp, c)) : v;
      st32(pw, c, ld32(pw, c) + (double)sv);
    });
  };
  // Jac(c, r, q, b, X): damped-Jacobi value of the cell at global c, tile
  // (r, q) of the x/bits tiles, right-hand side b, neighbour accessor X
  auto sweeps = [&](auto Jac) {
    // phase 2: first sweep on rows [y0-1, y0+TY] x cols [x0-1, x0+TX],
    // clipped to [.., m-NH]
    {
#pragma unroll
      for (int it = 0; it < IT2; it++) {
        const int q = tid + it * NT;
        if (q < N2) {
          const int rr = q / JW, cc = q - rr * JW;
          T v = T(0);
          if (ok2[it]) {
            const size_t c = cc2[it];
            v = Jac(c, rr + 1, cc + 1, bv[it],
                    [&](int dj, int di) { return xs[rr + 1 + dj][cc + 1 + di]; });
          }
          j1[rr][cc] = v;
        }
      }
    }
    __syncthreads();
    // phase 3: second sweep on the tile interior, stored with its images
#pragma unroll
    for (int k = 0; k < IT3; k++) {
      const int rr = threadIdx.y + 4 * k, cc = threadIdx.x;
      const int J = y0 + rr;
      if (J >= m - NH || Ic >= n - NH) continue;
      const size_t c = at32(J, Ic, n);
      const T v = Jac(c, rr + 2, cc + 2, bs[rr + 1][cc + 1],
                      [&](int dj, int di) { return j1[rr + 1 + dj][cc + 1 + di]; });
      store_img_mask32<T>(xout, mp, J, Ic, v, m, n);
      if (addp) add_img(J, Ic, v);
    }
  };
  // the same two sweeps in the unclipped form; stores without images when
  // no cell of the tile has one
  const bool noimg =
      y0 >= 2 * NH && y0 + TYV <= m - 2 * NH && x0 >= 2 * NH && x0 + TX <= n - 2 * NH;
  auto sweeps_lean = [&](auto Jac) {
#pragma unroll
    for (int k = 0; k < KB; k++) {
      const int rr = ty + 4 * k;
      if (rr < JH)
        j1[rr][tx] = Jac(ob + k * s4, rr + 1, tx + 1, bv[k],
                         [&](int dj, int di) { return xs[rr + 1 + dj][tx + 1 + di]; });
    }
    if (be_on)
      j1[ber][bec] = Jac(obe, ber + 1, bec + 1, be,
                         [&](int dj, int di) { return xs[ber + 1 + dj][bec + 1 + di]; });
    __syncthreads();
    F2D_STAMP(4);
#pragma unroll
    for (int k = 0; k < IT3; k++) {
      const int rr = ty + 4 * k;
      const uint32_t c = ob + (uint32_t)n + 1u + k * s4;  // (y0 + rr, x0 + tx)
      const T v = Jac(c, rr + 2, tx + 2, bs[rr + 1][tx + 1],
                      [&](int dj, int di) { return j1[rr + 1 + dj][tx + 1 + di]; });
      if (noimg) {
        const T sv = mp ? v * T(ld32(mp, c)) : v;
        st32(xout, c, sv);
        if (addp) st32(pw, c, pv[k] + (double)sv);
      } else {
        store_img_mask32<T>(xout, mp, y0 + rr, Ic, v, m, n);
        if (addp) add_img(y0 + rr, Ic, v);
      }
    }
    F2D_STAMP_SYNC(5);
  };
  auto run = [&](auto Jac) {
    if (lean)
      sweeps_lean(Jac);
    else
      sweeps(Jac);
  };
  if (full_tile) {
    // A_k = w_k everywhere, the mask is 1, and c1 / |A_5| is one value
    const T c3 = c1 / fabs(w[4]);
    run([&](size_t, int, int, T bc, auto X) {
      return jacobi_c3_2d<T>(bc, c3, c2, [&](int k, int, int) { return w[k]; }, X);
    });
  } else if (fly) {
    // A_5 = w_5 wherever the mask (bit 4) is set, and the value is
    // discarded where it is not: one c1 / |A_5| serves the tile
    const T c3 = c1 / fabs(w[4]);
    run([&](size_t, int r, int q, T bc, auto X) {
      const T v = jacobi_c3_2d<T>(bc, c3, c2, ABitsTile2<T, XW>{&sbs[r][q], w}, X);
      return bit_select<T>(sbs[r][q], 4, v);
    });
  } else {
    run([&](size_t c, int, int, T bc, auto X) {
      const APlanes<T> Ap{A, PL, c};
      return jacobi_2d<T>(msk[c], bc, c1, c2,
                          [&](int k, int dj, int di) { return Ap(k, (ptrdiff_t)dj * n + di); },
                          X);
    });
  }
}

template <typename T, int TYV, bool ADD = false>
__global__ void __launch_bounds__(256) k_smooth(const Task* __restrict__ tasks,
                                                const InstD* __restrict__ I, T* Q, double* P,
                                                const int8_t* __restrict__ B,
                                                const DynD* __restrict__ dyn, int pred) {
  smooth_tile<T, TYV, ADD>(tasks, task_index(), I, Q, P, B, dyn, pred);
}

// ---------------------------------------------------------------------------
// k_resid_restrict: r = b - A x on the fine points a coarse tile restricts
// (halo points through their periodic images), then the full weighting onto
// the coarse interior + images.
// flags: 1 store r on the owned fine cells (+ images) into rout (tk.xc),
//        2 restrict b itself instead of the residual (F-cycle descent),
//        4 norm partial of r, 8 norm partial of b (owned interior, msk != 0),
//        16 rhs an R plane at tk.xout, 32 the FP64 level-0 operator A0d,
//        64 rhs = celltocorner of the T plane at tk.xout, 128 x in the T pool
// ---------------------------------------------------------------------------
template <typename T, typename R, int TY2V>
__device__ __forceinline__ void rr_tile(const Task* __restrict__ tasks, const int ti,
                                        const InstD* __restrict__ I, const R* RP, T* Q,
                                        const int8_t* __restrict__ B,
                                        const DynD* __restrict__ dyn, int pred, int flags,
                                        double* partial) {
  constexpr int RY = 2 * TY2V + 1, RX = 2 * TX2 + 1;
  __shared__ R rs[RY][RX];
  __shared__ double red[2][8];
  const Task tk = tasks[ti];
  const int tid = threadIdx.y * blockDim.x + threadIdx.x, nt = blockDim.x * blockDim.y;
  // the skip test is read and taken after the tile loads are issued (the
  // task holds the addressing, so the data loads follow the task read)
  const InstD& in = I[tk.inst];
  const int m1 = tk.m, n1 = tk.n, m2 = tk.m2, n2 = tk.n2;
  const size_t PL = (size_t)m1 * n1;
  const int8_t* msk1 = B + tk.msk;
  const int8_t* msk2 = B + tk.msk2;
  // operands in RP: the level's A, x and b, or (flag 32) the FP64 level-0
  // operator with an explicit rhs buffer (flag 16) -- the outer residual
  const R* A = RP + ((flags & 32) ? in.A0d : tk.A);
  // flag 128 (FP64 pool type only): x is a plane of the T pool
  const R* x = RP + tk.xin;
  if constexpr (std::is_same<T, R>::value)
    if (flags & 128) x = Q + tk.xin;
  const R* bp = (flags & 64) ? x : RP + ((flags & 16) ? tk.xout : tk.b);
  // flag 64: the rhs is celltocorner of the T plane at tk.xout
  // (core/fortran_operators.f90:44-64, evaluated as k_c2c does)
  const T* cq = Q + tk.xout;
  auto bat = [&](uint32_t c) -> R {
    if (flags & 64)
      return R(0.25) * (R(ld32(cq, c)) + R(ld32(cq, c + 1)) + R(ld32(cq, c + n1)) +
                        R(ld32(cq, c + n1 + 1)));
    return ld32(bp, c);
  };
  const int fy0 = 2 * tk.y0 - 3, fx0 = 2 * tk.x0 - 3;
  R w[5];
  for (int k = 0; k < 5; k++) w[k] = R(in.w0[k]);
  const bool fly = tk.sb >= 0;
  const uint8_t* sb = (const uint8_t*)(B + (fly ? tk.sb : 0));
  double nr = 0., nb = 0.;
  // x (and the stencil bits) of the tile's region in one round trip into
  // LDS (rows/columns clamped to the array); points off the periodic halo
  // take their operands there, the few halo points the restriction needs
  // (their values are their interior images') read them around the image.
  // Both forms evaluate the same expressions in the same order.
  constexpr int NT = 256, XRW = RX + 2, XRH = RY + 2, NX = XRW * XRH;
  constexpr int ITX = (NX + NT - 1) / NT, ITP = (RY * RX + NT - 1) / NT;
  __shared__ R xsr[XRH][XRW];
  __shared__ uint8_t sbr[XRH][XRW];
  int full = 0;
  R xv[ITX];
  uint8_t bv[ITX];
  if (!(flags & 2)) {
#pragma unroll
    for (int it = 0; it < ITX; it++) {
      const int q = min(tid + it * NT, NX - 1);
      const int rr = q / XRW, cc = q - rr * XRW;
      const uint32_t c = at32(min(fy0 - 1 + rr, m1 - 1), min(fx0 - 1 + cc, n1 - 1), n1);
      xv[it] = ld32(x, c);
      bv[it] = fly ? ld32(sb, c) : (uint8_t)0;
    }
  }
  // the coarse mask of the points restricted, then the rhs and mask of the
  // points evaluated, issued with the x loads
  constexpr int NR = TY2V * TX2, ITR = (NR + NT - 1) / NT;
  int8_t m2v[ITR];
#pragma unroll
  for (int it = 0; it < ITR; it++) {
    const int q = min(tid + it * NT, NR - 1);
    const int ry = q / TX2, rx = q - ry * TX2;
    m2v[it] = ld32(msk2, at32(min(tk.y0 + ry, m2 - 1), min(tk.x0 + rx, n2 - 1), n2));
  }
  R bq[ITP];
  int8_t mq[ITP];
#pragma unroll
  for (int it = 0; it < ITP; it++) {
    const int q = min(tid + it * NT, RY * RX - 1);
    const int rr = q / RX, cc = q - rr * RX;
    const int jf = min(fy0 + rr, m1 - NH), if_ = min(fx0 + cc, n1 - NH);
    const uint32_t c = at32(img(jf, m1), img(if_, n1), n1);
    bq[it] = bat(c);
    mq[it] = ld32(msk1, c);
  }
  if (skip(dyn, tk.inst, pred)) {
    if ((flags & 12) && tid == 0) {
      partial[2 * tk.slot] = 0.;
      partial[2 * tk.slot + 1] = 0.;
    }
    return;
  }
  if (!(flags & 2)) {
    bool allfull = true;
#pragma unroll
    for (int it = 0; it < ITX; it++) {
      const int q = tid + it * NT;
      if (q < NX) {
        const int rr = q / XRW, cc = q - rr * XRW;
        xsr[rr][cc] = xv[it];
        sbr[rr][cc] = bv[it];
        allfull = allfull && bv[it] == 31;
      }
    }
    // the whole region has the full stencil: constant weights
    full = block_all4(fly && allfull);
  }
#pragma unroll
  for (int it = 0; it < ITP; it++) {
    const int q = tid + it * NT;
    if (q >= RY * RX) continue;
    const int rr = q / RX, cc = q - rr * RX;
    const int jf = fy0 + rr, if_ = fx0 + cc;
    if (jf > m1 - NH || if_ > n1 - NH) continue;  // beyond the last needed point
    const bool halo = jf >= m1 - NH || if_ >= n1 - NH;
    R v;
    if (flags & 2) {
      v = bq[it];
    } else if (!halo) {
      const size_t c = at(jf, if_, n1);
      auto xl = [&](int dj, int di) { return xsr[rr + 1 + dj][cc + 1 + di]; };
      if (full) {
        v = residual_2d<R>(mq[it], bq[it], [&](int k, int, int) { return w[k]; }, xl);
      } else if (fly) {
        v = residual_2d<R>(mq[it], bq[it], ABitsTile2<R, XRW>{&sbr[rr + 1][cc + 1], w}, xl);
      } else {
        const APlanes<R> Ap{A, PL, c};
        v = residual_2d<R>(mq[it], bq[it],
                           [&](int k, int dj, int di) { return Ap(k, (ptrdiff_t)dj * n1 + di); },
                           xl);
      }
    } else {
      const size_t c = at(img(jf, m1), img(if_, n1), n1);
      if (fly)
        v = residual_f<R>(mq[it], n1, bq[it], ABits<R>{sb, w, c}, [&](int dd) { return x[c + dd]; });
      else
        v = residual_x<R>(mq[it], A, PL, c, n1, bq[it], [&](int dd) { return x[c + dd]; });
    }
    rs[rr][cc] = v;
    // owned fine cells: local rows/cols [0, 2*T-1] off the halo, under the
    // coarse interior
    if (rr < 2 * TY2V && cc < 2 * TX2 && !halo && (tk.y0 + rr / 2) < m2 - NH &&
        (tk.x0 + cc / 2) < n2 - NH) {
      if (flags & 1) store_img_mask32<T>(Q + tk.xc, nullptr, jf, if_, T(v), m1, n1);
      if (mq[it] != 0) {
        nr += (double)v * (double)v;
        const double bb = (double)bq[it];
        nb += bb * bb;
      }
    }
  }
  if (flags & 12) {
    for (int o = 32; o > 0; o >>= 1) {
      nr += __shfl_xor(nr, o, 64);
      nb += __shfl_xor(nb, o, 64);
    }
    if ((tid & 63) == 0) {
      red[0][tid >> 6] = nr;
      red[1][tid >> 6] = nb;
    }
  }
  __syncthreads();
  if ((flags & 12) && tid == 0) {
    double a = 0., c = 0.;
    for (int w = 0; w < nt / 64; w++) {
      a += red[0][w];
      c += red[1][w];
    }
    partial[2 * tk.slot] = a;
    partial[2 * tk.slot + 1] = c;
  }
  T* b2 = Q + in.L[tk.lev + 1].b;
#pragma unroll
  for (int it = 0; it < ITR; it++) {
    const int q = tid + it * NT;
    if (q >= NR) continue;
    const int ry = q / TX2, rx = q - ry * TX2;
    const int J2 = tk.y0 + ry, I2 = tk.x0 + rx;
    if (J2 >= m2 - NH || I2 >= n2 - NH) continue;
    R v = R(0);
    if (m2v[it] != 0) {
      const int cy = 2 * ry + 1, cx = 2 * rx + 1;
      v = restrict_r<R>([&](int dj, int di) { return rs[cy + dj][cx + di]; });
    }
    store_img_mask32<T>(b2, nullptr, J2, I2, T(v), m2, n2);
  }
}

template <typename T, typename R, int TY2V>
__global__ void __launch_bounds__(256) k_resid_restrict(const Task* __restrict__ tasks,
                                                        const InstD* __restrict__ I,
                                                        const R* RP, T* Q,
                                                        const int8_t* __restrict__ B,
                                                        const DynD* __restrict__ dyn, int pred,
                                                        int flags, double* partial) {
  rr_tile<T, R, TY2V>(tasks, task_index(), I, RP, Q, B, dyn, pred, flags, partial);
}

// ---------------------------------------------------------------------------
// k_restrict_b2: the F-cycle's descent of b (core/gmg/hierarchy.py:131-151;
// restrict, core/gmg/fortran_multigrid.f90:501-546) over two levels in one
// tile program.  Task mode 2: b of level lev restricted onto lev+1 over the
// region the tile needs (the cells it owns stored with their images, as
// the one-level restriction stores them), then onto the tile's cells of
// lev+2 (its b at xc); mode 1: one level, lev onto the tile's cells of
// lev+1.  Below the finest level a halo value is its periodic image's, so
// a lev+1 value on the halo is evaluated at its image.  Every value is
// the restriction of k_resid_restrict with flag 2, operands in the same
// order, rounded to T before the next level uses it: the results equal two
// launches of that kernel bit for bit.
// ---------------------------------------------------------------------------
template <typename T, int TYO, int NTH>
__global__ void __launch_bounds__(NTH) k_restrict_b2(const Task* __restrict__ tasks,
                                                     const InstD* __restrict__ I, T* Q,
                                                     const int8_t* __restrict__ B,
                                                     const DynD* __restrict__ dyn, int pred) {
  constexpr int MY = 2 * TYO + 1, MX = 2 * TX2 + 1;  // region below the output level
  constexpr int FY = 2 * MY + 1, FX = 2 * MX + 1;    // region of level lev (mode 2)
  constexpr int NF = FY * FX, ITF = (NF + NTH - 1) / NTH;
  constexpr int NM = MY * MX, ITM = (NM + NTH - 1) / NTH;
  constexpr int NO = TYO * TX2, ITO = (NO + NTH - 1) / NTH;
  __shared__ T fs[FY][FX];
  __shared__ T ms[MY][MX];
  const Task tk = tasks[task_index()];
  const int tid = threadIdx.x;
  const bool two = tk.mode == 2;
  const T* bf = Q + tk.b;
  const int m = tk.m, n = tk.n;     // level lev
  const int mm = tk.m2, nm = tk.n2;  // level lev+1
  const int mo = two ? tk.m3 : mm, no = two ? tk.n3 : nm;
  const int8_t* mko = B + (two ? tk.msk3 : tk.msk2);
  const int8_t* mkm = B + tk.msk2;
  // region origins: the output level's tile (y0, x0) reads the level below
  // from (2 y0 - 3, 2 x0 - 3); in mode 2 that level is lev+1, which reads
  // lev from twice that again
  const int my0 = 2 * tk.y0 - 3, mx0 = 2 * tk.x0 - 3;
  const int fy0 = 2 * my0 - 3, fx0 = 2 * mx0 - 3;
  // loads issued before the step gate is read: b of the region below the
  // output (mode 1) or of level lev's region with the lev+1 mask (mode 2),
  // then the output mask; rows and columns past the last needed point are
  // clamped and never used.  Level lev's region is read at the images of
  // its points: the lev+1 halo row mm - NH restricts lev rows m - 3 .. m - 1,
  // whose images NH .. NH + 2 are the rows its image (row NH) restricts, so
  // the halo values come out of the region as their images' values.
  T fv[ITF];
  int8_t mmv[ITM];
  T mv[ITM];
  if (two) {
#pragma unroll
    for (int it = 0; it < ITF; it++) {
      const int q = min(tid + it * NTH, NF - 1);
      const int r = q / FX, s = q - r * FX;
      fv[it] = ld32(bf, at32(img(min(fy0 + r, m - 1), m), img(min(fx0 + s, n - 1), n), n));
    }
#pragma unroll
    for (int it = 0; it < ITM; it++) {
      const int q = min(tid + it * NTH, NM - 1);
      const int r = q / MX, s = q - r * MX;
      mmv[it] = ld32(mkm, at32(img(min(my0 + r, mm - NH), mm), img(min(mx0 + s, nm - NH), nm), nm));
    }
  } else {
#pragma unroll
    for (int it = 0; it < ITM; it++) {
      const int q = min(tid + it * NTH, NM - 1);
      const int r = q / MX, s = q - r * MX;
      mv[it] = ld32(bf, at32(img(min(my0 + r, m - NH), m), img(min(mx0 + s, n - NH), n), n));
    }
  }
  int8_t mov[ITO];
#pragma unroll
  for (int it = 0; it < ITO; it++) {
    const int q = min(tid + it * NTH, NO - 1);
    const int ry = q / TX2, rx = q - ry * TX2;
    mov[it] = ld32(mko, at32(min(tk.y0 + ry, mo - 1), min(tk.x0 + rx, no - 1), no));
  }
  if (skip(dyn, tk.inst, pred)) return;
  if (two) {
#pragma unroll
    for (int it = 0; it < ITF; it++) {
      const int q = tid + it * NTH;
      if (q < NF) fs[q / FX][q % FX] = fv[it];
    }
    __syncthreads();
    T* bm = Q + tk.b2;
#pragma unroll
    for (int it = 0; it < ITM; it++) {
      const int q = tid + it * NTH;
      if (q >= NM) continue;
      const int r = q / MX, s = q - r * MX;
      const int J = my0 + r, Ii = mx0 + s;
      if (J > mm - NH || Ii > nm - NH) continue;  // beyond the last needed point
      const bool halo = J == mm - NH || Ii == nm - NH;
      T v = T(0);
      if (mmv[it] != 0)
        v = restrict_r<T>([&](int dj, int di) { return fs[2 * r + 1 + dj][2 * s + 1 + di]; });
      ms[r][s] = v;
      // owned: under the output tile's interior cells, off the halo
      if (r < 2 * TYO && s < 2 * TX2 && !halo && tk.y0 + r / 2 < mo - NH &&
          tk.x0 + s / 2 < no - NH)
        store_img_mask32<T>(bm, nullptr, J, Ii, v, mm, nm);
    }
  } else {
#pragma unroll
    for (int it = 0; it < ITM; it++) {
      const int q = tid + it * NTH;
      if (q < NM) ms[q / MX][q % MX] = mv[it];
    }
  }
  __syncthreads();
  T* bo = Q + (two ? tk.xc : tk.b2);
#pragma unroll
  for (int it = 0; it < ITO; it++) {
    const int q = tid + it * NTH;
    if (q >= NO) continue;
    const int ry = q / TX2, rx = q - ry * TX2;
    const int J = tk.y0 + ry, Ii = tk.x0 + rx;
    if (J >= mo - NH || Ii >= no - NH) continue;
    T v = T(0);
    if (mov[it] != 0)
      v = restrict_r<T>([&](int dj, int di) { return ms[2 * ry + 1 + dj][2 * rx + 1 + di]; });
    store_img_mask32<T>(bo, nullptr, J, Ii, v, mo, no);
  }
}

// ---------------------------------------------------------------------------
// k_smooth_rr: the down-leg of a V-cycle on a level below the finest entered
// with x = 0 (hierarchy.py:103-108 for lev > lev1: smooth, residual,
// restrict) or, at the start level of an F-cycle's V-cycle, with
// x = I(x of the coarser level) (hierarchy.py:145-151), in one launch
// instead of two.  One task is one coarse tile
// (TY2V x TX2 cells of level lev+1); its workgroup smooths the fine region
// the tile's full weighting reads (the 2 TY2V + 1 by 2 TX2 + 1 residual
// points, plus one ring for the residual's stencil and one for the second
// sweep's), forms the residuals, restricts them onto the tile (+ images) and
// stores the smoothed x on the fine cells the tile owns (+ images).  The
// regions of neighbouring tiles overlap and are recomputed with the same
// operands in the same order.
//
// Positions are taken at their periodic images.  On every level below the
// finest, the right-hand side, the operator and the mask are stored with
// periodic images in their halo (the restriction, the coarsening and the
// coarse-mask set-up all fill it), so a value evaluated at a halo position
// from halo operands equals the value at its image: the first sweep's halo
// rows, which smoothtwicewithA evaluates from the halo, and the second
// sweep's and the residual's halo values, which the fills copy from the
// images.  The per-cell arithmetic is jacobi_2d's and residual_2d's, as in
// k_smooth (mode 1, explicit operator) and k_resid_restrict.
// ---------------------------------------------------------------------------
// Phase timestamps of k_smooth_rr (profiling build only, make stamps):
// thread 0 of each of the first 256 workgroups of the last launch on each
// level (1..7) records the real-time clock (100 MHz) at the start, after the
// task read, after the loads are issued and the step gate read, after each
// sweep, after the residuals and at the end; slot 7 holds the compute unit.
#ifdef F2D_PHASE_STAMPS
__device__ unsigned long long rr_stamps[8][256][8];
#define F2D_RSTAMP(lev, k)                                                  \
  do {                                                                     \
    if (threadIdx.x == 0 && threadIdx.y == 0 && blockIdx.x < 256 && (lev) < 8) \
      rr_stamps[(lev)][blockIdx.x][(k)] = (unsigned long long)wall_clock64();  \
  } while (0)
#else
#define F2D_RSTAMP(lev, k) \
  do {                     \
  } while (0)
#endif

__device__ __forceinline__ int wrap_img(int g, int m) {
  const int M = m - 2 * NH;
  while (g < NH) g += M;
  while (g >= m - NH) g -= M;
  return g;
}

// X0: 1 the level is entered with x = 0 (a V-cycle's descent), 3 with
// x = I(x of the coarser level) (an F-cycle's V-cycle from this level,
// hierarchy.py:145-151); the prolongation is evaluated at the images too (the
// coarse x and both masks are stored with their images).
// STAGE: the five operator planes on the region (plus the row and the
// columns the stencil reaches past it) are loaded with the tile, at their
// images, and the sweeps and residuals read them from LDS: one memory round
// trip for the launch instead of one per phase (the small launches of the
// mid levels, whose time is those round trips).  NTH: threads per
// workgroup, (64, NTH / 64); 1024 for launches of at most two workgroups
// per CU, where each phase is then one pass over its region instead of
// several.
template <typename T, int TY2V, int X0, int STAGE = 0, int NTH = 256>
__global__ void __launch_bounds__(NTH) k_smooth_rr(const Task* __restrict__ tasks,
                                                   const InstD* __restrict__ I, T* Q,
                                                   const int8_t* __restrict__ B,
                                                   const DynD* __restrict__ dyn, int pred) {
  static_assert(X0 == 1 || (X0 == 3 && TY2V <= 4), "x0 form");
  constexpr int NT = NTH;
  constexpr int RY = 2 * TY2V + 1, RX = 2 * TX2 + 1;  // residual points
  constexpr int H1 = RY + 2, W1 = RX + 2;             // second sweep
  constexpr int H2 = RY + 4, W2 = RX + 4;             // first sweep, rhs, mask
  constexpr int H3 = X0 == 3 ? RY + 6 : 1, W3 = X0 == 3 ? RX + 6 : 1;  // x0 (prolongation)
  constexpr int NR2 = H2 * W2, IT2 = (NR2 + NT - 1) / NT;
  __shared__ T x0s[H3][W3];
  __shared__ T x1s[H2][W2];
  __shared__ T bs2[H2][W2];
  __shared__ T x2s[H1][W1];
  __shared__ T rs[RY][RX];
  __shared__ int8_t ms[H2][W2];
  __shared__ uint32_t cs[STAGE ? 1 : H2][STAGE ? 1 : W2];  // image cell index of each region position
  // staged operator: rows [gy, gy + H2], columns [gx - 1, gx + W2]
  constexpr int HA = H2 + 1, WA = W2 + 2, NA = HA * WA, ITA = STAGE ? (NA + NT - 1) / NT : 1;
  __shared__ T As[STAGE ? 5 : 1][STAGE ? HA : 1][STAGE ? WA : 1];
  const int ti = task_index();
#ifdef F2D_PHASE_STAMPS
  const unsigned long long st0 = wall_clock64();
#endif
  const Task tk = tasks[ti];
#ifdef F2D_PHASE_STAMPS
  F2D_RSTAMP(tk.lev, 1);
  if (threadIdx.x == 0 && threadIdx.y == 0 && blockIdx.x < 256 && tk.lev < 8) {
    rr_stamps[tk.lev][blockIdx.x][0] = st0;
    rr_stamps[tk.lev][blockIdx.x][7] = ((unsigned long long)__smid() << 32) | gridDim.x;
  }
#endif
  const InstD& in = I[tk.inst];
  const int m1 = tk.m, n1 = tk.n, m2 = tk.m2, n2 = tk.n2;
  const size_t PL = (size_t)m1 * n1;
  const int8_t* msk = B + tk.msk;
  const int8_t* msk2 = B + tk.msk2;
  const T* A = Q + tk.A;
  const T* b = Q + tk.b;
  const int tid = threadIdx.y * 64 + threadIdx.x;
  const int fy0 = 2 * tk.y0 - 3, fx0 = 2 * tk.x0 - 3;  // first residual point
  const int gy = fy0 - 2, gx = fx0 - 2;                // first-sweep region origin
  // rhs and mask of the first sweep's region, every load issued before use
  T bv[IT2];
  int8_t mv[IT2];
  uint32_t cv[IT2];
#pragma unroll
  for (int it = 0; it < IT2; it++) {
    const int q = min(tid + it * NT, NR2 - 1);
    const int rr = q / W2, cc = q - rr * W2;
    cv[it] = at32(wrap_img(gy + rr, m1), wrap_img(gx + cc, n1), n1);
    bv[it] = ld32(b, cv[it]);
    mv[it] = ld32(msk, cv[it]);
  }
  T av[5][ITA];
  if constexpr (STAGE != 0) {
#pragma unroll
    for (int it = 0; it < ITA; it++) {
      const int q = min(tid + it * NT, NA - 1);
      const int rr = q / WA, cc = q - rr * WA;
      const uint32_t e = at32(wrap_img(gy + rr, m1), wrap_img(gx - 1 + cc, n1), n1);
#pragma unroll
      for (int k = 0; k < 5; k++) av[k][it] = ld32(A, (uint32_t)k * (uint32_t)PL + e);
    }
  }
  // X0 3: x = I(xc) on the first sweep's region plus one ring.  Fine row
  // gy - 1 + rr (gy odd) lies over coarse row cy + rr / 2, so the coarse x
  // and mask under the region and the fine mask on it are loaded at their
  // images in the same round trip, and ip_eval (interp_at's expressions)
  // runs out of LDS
  constexpr int CH3 = (H3 - 1) / 2 + 2, CW3 = (W3 - 1) / 2 + 2, NC3 = CH3 * CW3;
  constexpr int NM3 = H3 * W3;
  constexpr int ITC3 = X0 == 3 ? (NC3 + NT - 1) / NT : 1, ITM3 = X0 == 3 ? (NM3 + NT - 1) / NT : 1;
  __shared__ T cx3[CH3][CW3];
  __shared__ int8_t cm3[CH3][CW3];
  __shared__ int8_t m13[H3][W3];
  const int cy = (gy + 1) >> 1, cxo = (gx + 1) >> 1;
  T c3v[ITC3];
  int8_t s3v[ITC3], m3v[ITM3];
  if constexpr (X0 == 3) {
    const T* xc = Q + tk.xc;
#pragma unroll
    for (int it = 0; it < ITC3; it++) {
      const int q = min(tid + it * NT, NC3 - 1);
      const int a = q / CW3, bb = q - a * CW3;
      const uint32_t c = at32(wrap_img(cy + a, m2), wrap_img(cxo + bb, n2), n2);
      c3v[it] = ld32(xc, c);
      s3v[it] = ld32(msk2, c);
    }
#pragma unroll
    for (int it = 0; it < ITM3; it++) {
      const int q = min(tid + it * NT, NM3 - 1);
      const int rr = q / W3, cc = q - rr * W3;
      m3v[it] = ld32(msk, at32(wrap_img(gy - 1 + rr, m1), wrap_img(gx - 1 + cc, n1), n1));
    }
  }
  if (skip(dyn, tk.inst, pred)) return;
  F2D_RSTAMP(tk.lev, 2);
#pragma unroll
  for (int it = 0; it < IT2; it++) {
    const int q = tid + it * NT;
    if (q < NR2) {
      const int rr = q / W2, cc = q - rr * W2;
      bs2[rr][cc] = bv[it];
      ms[rr][cc] = mv[it];
      if constexpr (STAGE == 0) cs[rr][cc] = cv[it];
    }
  }
  if constexpr (STAGE != 0) {
#pragma unroll
    for (int it = 0; it < ITA; it++) {
      const int q = tid + it * NT;
      if (q < NA) {
        const int rr = q / WA, cc = q - rr * WA;
#pragma unroll
        for (int k = 0; k < 5; k++) As[k][rr][cc] = av[k][it];
      }
    }
    if constexpr (X0 == 1) __syncthreads();  // (X0 3 synchronises below)
  }
  if constexpr (X0 == 3) {
#pragma unroll
    for (int it = 0; it < ITC3; it++) {
      const int q = tid + it * NT;
      if (q < NC3) {
        const int a = q / CW3, bb = q - a * CW3;
        cx3[a][bb] = c3v[it];
        cm3[a][bb] = s3v[it];
      }
    }
#pragma unroll
    for (int it = 0; it < ITM3; it++) {
      const int q = tid + it * NT;
      if (q < NM3) {
        const int rr = q / W3, cc = q - rr * W3;
        m13[rr][cc] = m3v[it];
      }
    }
    __syncthreads();
    for (int q = tid; q < NM3; q += NT) {
      const int rr = q / W3, cc = q - rr * W3;
      const int a = rr >> 1, bb = cc >> 1;
      IpOps<T> o;
      o.a = cx3[a][bb];
      o.b = cx3[a][bb + 1];
      o.c = cx3[a + 1][bb];
      o.d = cx3[a + 1][bb + 1];
      o.sa = cm3[a][bb];
      o.sb = cm3[a][bb + 1];
      o.sc = cm3[a + 1][bb];
      o.sd = cm3[a + 1][bb + 1];
      o.m1 = m13[rr][cc];
      o.par = ((gx - 1 + cc) & 1) | (((gy - 1 + rr) & 1) << 1);
      x0s[rr][cc] = ip_eval<T>(o);
    }
    __syncthreads();
  }
  const T c1 = T(in.omega), c2 = T(1. - in.omega);
  auto Aat = [&](uint32_t c) {
    const APlanes<T> Ap{A, PL, c};
    return [=](int k, int dj, int di) { return Ap(k, (ptrdiff_t)dj * n1 + di); };
  };
  // the operator of region position (rr, cc): from LDS when staged, else at
  // the position's image in the level's planes
  auto Aop = [&](int rr, int cc, uint32_t e) {
    if constexpr (STAGE != 0) {
      (void)e;
      return [&, rr, cc](int k, int dj, int di) { return As[k][rr + dj][cc + 1 + di]; };
    } else {
      (void)rr;
      (void)cc;
      return Aat(e);
    }
  };
  // first sweep on the whole region (rhs from registers), from x = 0 or
  // from the prolongation
#pragma unroll
  for (int it = 0; it < IT2; it++) {
    const int q = tid + it * NT;
    if (q < NR2) {
      const int rr = q / W2, cc = q - rr * W2;
      if constexpr (X0 == 3)
        x1s[rr][cc] = jacobi_2d<T>(mv[it], bv[it], c1, c2, Aop(rr, cc, cv[it]),
                                   [&](int dj, int di) { return x0s[rr + 1 + dj][cc + 1 + di]; });
      else
        x1s[rr][cc] = jacobi_2d<T>(mv[it], bv[it], c1, c2, Aop(rr, cc, cv[it]),
                                   [](int, int) { return T(0); });
    }
  }
  __syncthreads();
  F2D_RSTAMP(tk.lev, 3);
  // second sweep on the region less one ring
  constexpr int NR1 = H1 * W1;
  for (int q = tid; q < NR1; q += NT) {
    const int r1 = q / W1, c1i = q - r1 * W1;
    const int rr = r1 + 1, cc = c1i + 1;
    x2s[r1][c1i] = jacobi_2d<T>(ms[rr][cc], bs2[rr][cc], c1, c2,
                                Aop(rr, cc, STAGE ? 0u : cs[STAGE ? 0 : rr][STAGE ? 0 : cc]),
                                [&](int dj, int di) { return x1s[rr + dj][cc + di]; });
  }
  __syncthreads();
  F2D_RSTAMP(tk.lev, 4);
  // residuals at the restriction's points; the smoothed x of the owned
  // fine cells (local rows/cols [0, 2*T) off the halo, under the coarse
  // interior) stored with their images
  T* xout = Q + tk.xout;
  constexpr int NRR = RY * RX;
  for (int q = tid; q < NRR; q += NT) {
    const int r0 = q / RX, c0 = q - r0 * RX;
    const int rr = r0 + 2, cc = c0 + 2;
    rs[r0][c0] = residual_2d<T>(ms[rr][cc], bs2[rr][cc],
                                Aop(rr, cc, STAGE ? 0u : cs[STAGE ? 0 : rr][STAGE ? 0 : cc]),
                                [&](int dj, int di) { return x2s[r0 + 1 + dj][c0 + 1 + di]; });
    const int jf = fy0 + r0, if_ = fx0 + c0;
    if (r0 < 2 * TY2V && c0 < 2 * TX2 && jf < m1 - NH && if_ < n1 - NH &&
        (tk.y0 + r0 / 2) < m2 - NH && (tk.x0 + c0 / 2) < n2 - NH)
      store_img_mask32<T>(xout, nullptr, jf, if_, x2s[r0 + 1][c0 + 1], m1, n1);
  }
  __syncthreads();
  F2D_RSTAMP(tk.lev, 5);
  // full weighting onto the coarse tile (+ images)
  T* b2 = Q + in.L[tk.lev + 1].b;
  constexpr int NC = TY2V * TX2;
  for (int q = tid; q < NC; q += NT) {
    const int ry = q / TX2, rx = q - ry * TX2;
    const int J2 = tk.y0 + ry, I2 = tk.x0 + rx;
    if (J2 >= m2 - NH || I2 >= n2 - NH) continue;
    T v = T(0);
    if (ld32(msk2, at32(J2, I2, n2)) != 0) {
      const int cy = 2 * ry + 1, cx = 2 * rx + 1;
      v = restrict_r<T>([&](int dj, int di) { return rs[cy + dj][cx + di]; });
    }
    store_img_mask32<T>(b2, nullptr, J2, I2, v, m2, n2);
  }
#ifdef F2D_PHASE_STAMPS
  __syncthreads();
  F2D_RSTAMP(tk.lev, 6);
#endif
}

// ---------------------------------------------------------------------------
// k_up2: two consecutive up-leg smoothings of a V-cycle (hierarchy.py:
// 123-126 on levels lc = lev + 1 and lev, both below the finest) in one
// launch.  A task is a tile of level lev; its workgroup computes the
// coarse level's x + I(x of lc + 1) and two sweeps on the coarse region its
// own prolongation reads (plus the rings the sweeps need), stores the
// coarse cells its tile lies over (+ images), then the fine level's
// x + I(coarse x) and two sweeps on the tile (+ images).  Neighbouring
// tiles recompute the overlap of their coarse regions with the same
// operands in the same order.  As in k_smooth_rr, every position is taken
// at its periodic image (below the finest level every array is stored with
// its images), so the values equal the launch-per-level ones bit for bit.
// Tile origins are odd (NH plus multiples of the even tile sizes): fine
// region row rr lies over coarse region row (rr + 1) / 2, and so on down.
// ---------------------------------------------------------------------------
// NTH: threads per workgroup, (64, NTH / 64); its launches are small (at
// most 256 tiles), so 1024 threads make each phase one pass over its region.
template <typename T, int TYV, int NTH = 1024>
__global__ void __launch_bounds__(NTH) k_up2(const Task* __restrict__ tasks,
                                             const InstD* __restrict__ I, T* Q,
                                             const int8_t* __restrict__ B,
                                             const DynD* __restrict__ dyn, int pred) {
  static_assert(TYV == 4 || TYV == 8 || TYV == 16, "tile rows");
  constexpr int NT = NTH;
  constexpr int XH = TYV + 4, XW = TX + 4, JH = TYV + 2, JW = TX + 2;  // fine x0, first sweep
  constexpr int CH = TYV / 2 + 4, CW = TX / 2 + 4;  // coarse x the prolongation reads
  constexpr int C1H = CH + 2, C1W = CW + 2;          // coarse first sweep
  constexpr int C2H = CH + 4, C2W = CW + 4;          // coarse x0
  constexpr int DH = C2H / 2 + 2, DW = C2W / 2 + 2;  // next coarser x under the coarse x0
  constexpr int ND = DH * DW, NC2 = C2H * C2W, NC1 = C1H * C1W, NX = XH * XW, NJ = JH * JW;
  constexpr int ITD = (ND + NT - 1) / NT, ITC2 = (NC2 + NT - 1) / NT, ITC1 = (NC1 + NT - 1) / NT;
  constexpr int ITX = (NX + NT - 1) / NT, ITJ = (NJ + NT - 1) / NT;
  __shared__ T xd[DH][DW];
  __shared__ int8_t md[DH][DW];
  __shared__ T x0c[C2H][C2W];
  __shared__ int8_t mc[C2H][C2W];
  __shared__ T bc[C1H][C1W];
  __shared__ uint32_t ic[C1H][C1W];  // image cell index (coarse level)
  __shared__ T j1c[C1H][C1W];
  __shared__ T cx[CH][CW];
  __shared__ T xs[XH][XW];
  __shared__ int8_t mf[XH][XW];
  __shared__ T bs[JH][JW];
  __shared__ uint32_t jf[JH][JW];  // image cell index (fine level)
  __shared__ T j1[JH][JW];
  const int ti = task_index();
  const Task tk = tasks[ti];
  const InstD& in = I[tk.inst];
  const int tid = threadIdx.y * 64 + threadIdx.x;
  const int m = tk.m, n = tk.n, m2 = tk.m2, n2 = tk.n2, m3 = tk.m3, n3 = tk.n3;
  const int y0 = tk.y0, x0 = tk.x0;
  const int cy0 = y0 >> 1, cx0 = x0 >> 1;  // coarse region origin
  const int dy0 = cy0 >> 1, dx0 = cx0 >> 1;  // next coarser region origin (of coarse row cy0 - 2)
  const int8_t* mskf = B + tk.msk;
  const int8_t* mskc = B + tk.msk2;
  const int8_t* mskd = B + tk.msk3;
  // every load of the tile issued before the first is used
  T vd[ITD], vc2[ITC2], vb1[ITC1], vx[ITX], vj[ITJ];
  int8_t wd[ITD], wc2[ITC2], wx[ITX];
  uint32_t uc1[ITC1], uj[ITJ];
#pragma unroll
  for (int it = 0; it < ITD; it++) {
    const int q = min(tid + it * NT, ND - 1), a = q / DW, c = q - a * DW;
    const uint32_t e = at32(wrap_img(dy0 + a, m3), wrap_img(dx0 + c, n3), n3);
    vd[it] = ld32(Q + tk.xcc, e);
    wd[it] = ld32(mskd, e);
  }
#pragma unroll
  for (int it = 0; it < ITC2; it++) {
    const int q = min(tid + it * NT, NC2 - 1), a = q / C2W, c = q - a * C2W;
    const uint32_t e = at32(wrap_img(cy0 - 2 + a, m2), wrap_img(cx0 - 2 + c, n2), n2);
    vc2[it] = ld32(Q + tk.xin2, e);
    wc2[it] = ld32(mskc, e);
  }
#pragma unroll
  for (int it = 0; it < ITC1; it++) {
    const int q = min(tid + it * NT, NC1 - 1), a = q / C1W, c = q - a * C1W;
    uc1[it] = at32(wrap_img(cy0 - 1 + a, m2), wrap_img(cx0 - 1 + c, n2), n2);
    vb1[it] = ld32(Q + tk.b2, uc1[it]);
  }
#pragma unroll
  for (int it = 0; it < ITX; it++) {
    const int q = min(tid + it * NT, NX - 1), a = q / XW, c = q - a * XW;
    const uint32_t e = at32(wrap_img(y0 - 2 + a, m), wrap_img(x0 - 2 + c, n), n);
    vx[it] = ld32(Q + tk.xin, e);
    wx[it] = ld32(mskf, e);
  }
#pragma unroll
  for (int it = 0; it < ITJ; it++) {
    const int q = min(tid + it * NT, NJ - 1), a = q / JW, c = q - a * JW;
    uj[it] = at32(wrap_img(y0 - 1 + a, m), wrap_img(x0 - 1 + c, n), n);
    vj[it] = ld32(Q + tk.b, uj[it]);
  }
  if (skip(dyn, tk.inst, pred)) return;
#pragma unroll
  for (int it = 0; it < ITD; it++) {
    const int q = tid + it * NT, a = q / DW, c = q - a * DW;
    if (q < ND) {
      xd[a][c] = vd[it];
      md[a][c] = wd[it];
    }
  }
#pragma unroll
  for (int it = 0; it < ITC2; it++) {
    const int q = tid + it * NT, a = q / C2W, c = q - a * C2W;
    if (q < NC2) mc[a][c] = wc2[it];
  }
#pragma unroll
  for (int it = 0; it < ITC1; it++) {
    const int q = tid + it * NT, a = q / C1W, c = q - a * C1W;
    if (q < NC1) {
      bc[a][c] = vb1[it];
      ic[a][c] = uc1[it];
    }
  }
#pragma unroll
  for (int it = 0; it < ITX; it++) {
    const int q = tid + it * NT, a = q / XW, c = q - a * XW;
    if (q < NX) mf[a][c] = wx[it];
  }
#pragma unroll
  for (int it = 0; it < ITJ; it++) {
    const int q = tid + it * NT, a = q / JW, c = q - a * JW;
    if (q < NJ) {
      bs[a][c] = vj[it];
      jf[a][c] = uj[it];
    }
  }
  __syncthreads();
  // prolongation at region position (rr, cc) from the coarser tile X with
  // its mask M held at offset mo (ip_eval: interp_at's expressions)
  auto ip = [&](auto& X, auto& M, int mo, int8_t m1, int rr, int cc) {
    const int a = (rr + 1) >> 1, b = (cc + 1) >> 1;
    IpOps<T> o;
    o.a = X[a][b];
    o.b = X[a][b + 1];
    o.c = X[a + 1][b];
    o.d = X[a + 1][b + 1];
    o.sa = M[a + mo][b + mo];
    o.sb = M[a + mo][b + 1 + mo];
    o.sc = M[a + 1 + mo][b + mo];
    o.sd = M[a + 1 + mo][b + 1 + mo];
    o.m1 = m1;
    o.par = ((cc + 1) & 1) | (((rr + 1) & 1) << 1);
    return ip_eval<T>(o);
  };
  // coarse x0 = x + I(x of the next coarser level)
#pragma unroll
  for (int it = 0; it < ITC2; it++) {
    const int q = tid + it * NT, a = q / C2W, c = q - a * C2W;
    if (q < NC2) x0c[a][c] = vc2[it] + ip(xd, md, 0, mc[a][c], a, c);
  }
  __syncthreads();
  const T c1 = T(in.omega), c2 = T(1. - in.omega);
  const T* Ac = Q + tk.A2;
  const T* Af = Q + tk.A;
  const size_t PLc = (size_t)m2 * n2, PLf = (size_t)m * n;
  auto Aat = [&](const T* A, size_t PL, int nn, uint32_t e) {
    const APlanes<T> Ap{A, PL, e};
    return [=](int k, int dj, int di) { return Ap(k, (ptrdiff_t)dj * nn + di); };
  };
  // coarse first sweep (region less one ring of x0), second sweep on the
  // prolongation's region; the cells the tile lies over are stored
  for (int q = tid; q < NC1; q += NT) {
    const int a = q / C1W, c = q - a * C1W;
    j1c[a][c] = jacobi_2d<T>(mc[a + 1][c + 1], bc[a][c], c1, c2, Aat(Ac, PLc, n2, ic[a][c]),
                             [&](int dj, int di) { return x0c[a + 1 + dj][c + 1 + di]; });
  }
  __syncthreads();
  T* xoc = Q + tk.xc;
  for (int q = tid; q < CH * CW; q += NT) {
    const int a = q / CW, c = q - a * CW;
    const T v = jacobi_2d<T>(mc[a + 2][c + 2], bc[a + 1][c + 1], c1, c2,
                             Aat(Ac, PLc, n2, ic[a + 1][c + 1]),
                             [&](int dj, int di) { return j1c[a + 1 + dj][c + 1 + di]; });
    cx[a][c] = v;
    const int J = cy0 + a, Ic = cx0 + c;
    if (a >= 2 && a < 2 + TYV / 2 && c >= 2 && c < 2 + TX2 && J < m2 - NH && Ic < n2 - NH)
      store_img_mask32<T>(xoc, nullptr, J, Ic, v, m2, n2);
  }
  __syncthreads();
  // fine x0 = x + I(coarse x)
#pragma unroll
  for (int it = 0; it < ITX; it++) {
    const int q = tid + it * NT, a = q / XW, c = q - a * XW;
    if (q < NX) xs[a][c] = vx[it] + ip(cx, mc, 2, mf[a][c], a, c);  // cx = mc's region + 2
  }
  __syncthreads();
  for (int q = tid; q < NJ; q += NT) {
    const int a = q / JW, c = q - a * JW;
    j1[a][c] = jacobi_2d<T>(mf[a + 1][c + 1], bs[a][c], c1, c2, Aat(Af, PLf, n, jf[a][c]),
                            [&](int dj, int di) { return xs[a + 1 + dj][c + 1 + di]; });
  }
  __syncthreads();
  T* xo = Q + tk.xout;
  for (int q = tid; q < TYV * TX; q += NT) {
    const int a = q / TX, c = q - a * TX;
    const int J = y0 + a, Ic = x0 + c;
    if (J >= m - NH || Ic >= n - NH) continue;
    const T v = jacobi_2d<T>(mf[a + 2][c + 2], bs[a + 1][c + 1], c1, c2,
                             Aat(Af, PLf, n, jf[a + 1][c + 1]),
                             [&](int dj, int di) { return j1[a + 1 + dj][c + 1 + di]; });
    store_img_mask32<T>(xo, nullptr, J, Ic, v, m, n);
  }
}

// ---------------------------------------------------------------------------
// k_persist: a run of consecutive smoother and residual/restriction phases
// of the multigrid levels below the finest (the tile lists of what would
// otherwise be one launch each), executed by one launch of G co-resident
// workgroups with a device-wide barrier between phases.  Each phase is the
// same tile function over the same tasks as its own launch would run, so the
// results are bit for bit those of the launch-per-phase program.
//
// Barrier: one slot per persistent launch of the step program.  Workgroups
// add to `arrive` and wait until it reaches (phase + 1) * G; the last
// workgroup to leave the kernel re-arms the slot (every workgroup has then
// left every wait).  Every wait is bounded: after BAR_SPIN polls a
// workgroup gives up, raises the batch's fault flag (the host reports it as
// an error) and the launch ends, so no configuration can hang the device.
// ---------------------------------------------------------------------------
struct BarSlot {
  unsigned arrive, leave;
  unsigned pad[30];  // one slot per 128-byte line
};
struct PPhase {
  int kind;   // 0 smoother, 1 residual/restriction
  int toff, ntask, pred, flags;
};
constexpr int BAR_SPIN = 1 << 18;

__device__ __forceinline__ bool grid_sync(BarSlot* bs, int* fault, unsigned target) {
  __shared__ int ok_sh;
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    __hip_atomic_fetch_add(&bs->arrive, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    int ok = 1;
    for (int n = 0;; n++) {
      if (__hip_atomic_load(&bs->arrive, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >= target)
        break;
      if (n > BAR_SPIN || __hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    ok_sh = ok;
  }
  __syncthreads();
  return ok_sh != 0;
}

template <typename T, int TYV>
__global__ void __launch_bounds__(256) k_persist(const PPhase* __restrict__ ph, int nph,
                                                 const Task* __restrict__ tasks,
                                                 const InstD* __restrict__ I, T* Q,
                                                 const int8_t* __restrict__ B,
                                                 const DynD* __restrict__ dyn, BarSlot* bar,
                                                 int* fault) {
  const unsigned G = gridDim.x;
  bool ok = true;
  for (int p = 0; p < nph && ok; p++) {
    const PPhase f = ph[p];
    for (int t = blockIdx.x; t < f.ntask; t += G) {
      if (f.kind == 0)
        smooth_tile<T, TYV>(tasks + f.toff, t, I, Q, nullptr, B, dyn, f.pred);
      else
        rr_tile<T, T, TYV / 2>(tasks + f.toff, t, I, (const T*)Q, Q, B, dyn, f.pred, f.flags,
                               nullptr);
      __syncthreads();  // the next tile reuses the LDS tiles
    }
    if (p + 1 < nph) ok = grid_sync(bar, fault, (unsigned)(p + 1) * G);
  }
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const unsigned e =
        __hip_atomic_fetch_add(&bar->leave, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (e + 1 == G) {
      __hip_atomic_store(&bar->arrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&bar->leave, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// ---------------------------------------------------------------------------
// k_coarse: every level from nfine to the coarsest of one instance, in LDS.
// mode 0: V-cycle tail with x(nfine) = 0 (hierarchy.py:98-126 from a level
// above); mode 1: F-cycle tail (hierarchy.py:131-151 from nfine).
// Input b(nfine) from global, output x(nfine) (filled) to tk.xout.
// ---------------------------------------------------------------------------
// LDS (address space 3) pointers: accesses compile to ds_* instructions
// even when the pointers themselves are read from the LDS level table
#define LDSP __attribute__((address_space(3)))
// T: arithmetic type (FP64); TA: storage type of the operator (the pool's,
// FP32 in mixed mode, widened exactly on use)
template <typename T, typename TA = T>
struct CL {
  int m, n;
  LDSP int8_t* msk;
  LDSP TA* A;
  LDSP T *x, *b, *t;  // the residual reuses t (never live at the same time)
};

// Phase stamps of the coarse kernel (profiling build only, make stamps):
// thread 0 of workgroup 0 records the shader clock and a phase code at the
// end of each level visit of the last launch of each mode (stamps[mode]:
// clock, code pairs; code 100 + k: the real-time clock, 100 MHz).
#ifdef F2D_PHASE_STAMPS
__device__ unsigned long long coarse_stamps[2][256];
#define F2D_CSTAMP(sy, mode, code)                                                     \
  do {                                                                                \
    if ((sy).tid == 0 && blockIdx.x == 0 && (sy).ck < 127) {                            \
      coarse_stamps[(mode)][2 * (sy).ck] = (unsigned long long)clock64();                \
      coarse_stamps[(mode)][2 * (sy).ck + 1] = (unsigned long long)(code);               \
      (sy).ck++;                                                                      \
    }                                                                                 \
  } while (0)
#else
#define F2D_CSTAMP(sy, mode, code) \
  do {                             \
  } while (0)
#endif

// Phases over a level: levels of at most SMALL cells run on wavefront 0
// alone and end with a wavefront-level barrier (LDS operations of one
// wavefront execute in order); larger levels use the whole workgroup and a
// workgroup barrier.  `prev_small` (uniform) inserts the workgroup barrier
// needed when a large phase follows small ones.
struct Sync {
  int tid, nt;
  int small_cells;  // levels of at most this many cells run on wavefront 0
  bool prev_small;
  int ck;           // next phase stamp (profiling build)
  __device__ __forceinline__ bool begin(int cells) {
    const bool small = cells <= small_cells;
    if (!small && prev_small) __syncthreads();
    prev_small = small;
    return small ? tid < 64 : true;
  }
  __device__ __forceinline__ int stride(int cells) const { return cells <= small_cells ? 64 : nt; }
  __device__ __forceinline__ void end(int cells) {
    if (cells <= small_cells) {
      if (tid < 64) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
    } else {
      __syncthreads();
    }
  }
};

// q / w for the small q, w of the coarse levels (q < 2^16, w < 2^10):
// (q + 0.5) / w is at least 0.5 / w from an integer and the float product
// is within 1e-4 of it, so the truncation is exact
__device__ __forceinline__ int qdiv(int q, int w) {
  return (int)(((float)q + 0.5f) * __builtin_amdgcn_rcpf((float)w));
}

// zero_in: the smoother's input x is zero (a level entered from above);
// the zero is read, not stored, with the arithmetic of a stored zero
template <typename T, typename TA>
__device__ __forceinline__ void c_smooth(const CL<T, TA>& L, T c1, T c2, Sync& sy, bool zero_in = false) {
  const int m = L.m, n = L.n, cells = m * n;
  const size_t P = (size_t)m * n;
  const int wj = m - 2 * NH + 2, wi = n - 2 * NH + 2;
  if (sy.begin(cells)) {
    const int st = sy.stride(cells);
    for (int q = sy.tid; q < wj * wi; q += st) {
      const int qj = qdiv(q, wi);
      const int j = NH - 1 + qj, i = NH - 1 + (q - qj * wi);
      const size_t c = at(j, i, n);
      auto Af = [&](int k, ptrdiff_t d) { return T(L.A[k * P + c + d]); };
      L.t[c] = zero_in ? jacobi_f<T>(L.msk[c], n, L.b[c], c1, c2, Af, [&](int) { return T(0); })
                       : jacobi_f<T>(L.msk[c], n, L.b[c], c1, c2, Af,
                                     [&](int dd) { return L.x[c + dd]; });
    }
  }
  sy.end(cells);
  const int N = n - 2 * NH, M = m - 2 * NH;
  if (sy.begin(cells)) {
    const int st = sy.stride(cells);
    for (int q = sy.tid; q < M * N; q += st) {
      const int qj = qdiv(q, N);
      const int J = NH + qj, I2 = NH + (q - qj * N);
      const size_t c = at(J, I2, n);
      auto Af = [&](int k, ptrdiff_t d) { return T(L.A[k * P + c + d]); };
      const T v = jacobi_f<T>(L.msk[c], n, L.b[c], c1, c2, Af, [&](int dd) { return L.t[c + dd]; });
      store_img_mask<T>(L.x, nullptr, J, I2, v, m, n);
    }
  }
  sy.end(cells);
}
template <typename T, typename TA>
__device__ __forceinline__ void c_resid_restrict(const CL<T, TA>& F, const CL<T, TA>& C,
                                                 bool restrict_b, Sync& sy) {
  const int m = F.m, n = F.n, cells = m * n;
  const size_t P = (size_t)m * n;
  const int N = n - 2 * NH, M = m - 2 * NH;
  if (!restrict_b) {
    if (sy.begin(cells))
      for (int q = sy.tid; q < M * N; q += sy.stride(cells)) {
        const int qj = qdiv(q, N);
        const int J = NH + qj, I2 = NH + (q - qj * N);
        const size_t c = at(J, I2, n);
        const T v = residual_f<T>(F.msk[c], n, F.b[c],
                                  [&](int k, ptrdiff_t d) { return T(F.A[k * P + c + d]); },
                                  [&](int dd) { return F.x[c + dd]; });
        store_img_mask<T>(F.t, nullptr, J, I2, v, m, n);
      }
    sy.end(cells);
  }
  LDSP T* src = restrict_b ? F.b : F.t;
  const int N2 = C.n - 2 * NH, M2 = C.m - 2 * NH, ccells = C.m * C.n;
  if (sy.begin(ccells))
    for (int q = sy.tid; q < M2 * N2; q += sy.stride(ccells)) {
      const int qj = qdiv(q, N2);
      const int J2 = NH + qj, I2 = NH + (q - qj * N2);
      T v = T(0);
      if (C.msk[at(J2, I2, C.n)] != 0) {
        const size_t c = at(2 * J2 + 1 - NH, 2 * I2 + 1 - NH, n);
        v = restrict_r<T>([&](int dj, int di) { return src[c + (ptrdiff_t)dj * n + di]; });
      }
      store_img_mask<T>(C.b, nullptr, J2, I2, v, C.m, C.n);
    }
  sy.end(ccells);
}
template <typename T, typename TA>
__device__ __forceinline__ void c_interp(const CL<T, TA>& F, const CL<T, TA>& C, bool acc, Sync& sy) {
  const int cnt = F.m * F.n;
  if (sy.begin(cnt))
    for (int q = sy.tid; q < cnt; q += sy.stride(cnt)) {
      const int j = qdiv(q, F.n), i = q - j * F.n;
      const T r = interp_at<T>(F.msk, F.n, C.msk, C.x, C.n, j, i);
      F.x[q] = acc ? F.x[q] + r : r;
    }
  sy.end(cnt);
}

// A cell's operator row in registers: the eight couplings offdiag_2d reads
// and the centre A_5 (residual_2d); the arguments are literals after
// inlining
template <typename T>
struct ARow {
  T a0, a1, a2, a3, a4, a3e, a2s, a1s, a0s;
  __device__ __forceinline__ T operator()(int k, int dj, int di) const {
    if (dj == 0 && di == 0)
      return k == 0 ? a0 : (k == 1 ? a1 : (k == 2 ? a2 : (k == 3 ? a3 : a4)));
    return k == 3 ? a3e : (k == 2 ? a2s : (k == 1 ? a1s : a0s));
  }
};

// An interior cell held by its owning thread: position, operator row,
// c1 / |A_5|, rhs and mask (read once per level visit)
template <typename T>
struct OwnCell {
  int j, i;
  int c;
  bool own;
  ARow<T> a;
  T c3, bc;
  int8_t mk;
};

// Level visits where each thread (of the workgroup, or of wavefront 0 when
// the level is small) owns at most OWNK interior cells keep those cells'
// operands in registers: a sweep, and the residual, read only x or t from
// LDS.  The arithmetic is jacobi_f's and residual_f's, operand for operand
// (c1 / |A_5| is the same division, done once).  The first sweep runs on
// the interior only and stores t with its periodic images: below the finest
// level the operator, rhs, mask and x hold periodic images in their halo,
// so the first sweep's value at a halo position (which smoothtwicewithA
// evaluates there from the halo) is the value at its image.
constexpr int OWNK = 1;

// `iters` smoothings (two damped-Jacobi sweeps each; the first reads x = 0
// when zero_in), then with `resid` the residual b - A x on the interior into
// t with its images (the first half of c_resid_restrict).  Returns false,
// doing nothing, when the level has more cells than the owners hold.
template <typename T, typename TA>
__device__ __forceinline__ bool c_smooth_own(const CL<T, TA>& L, T c1, T c2, Sync& sy, bool zero_in, int iters,
                             bool resid) {
  const int m = L.m, n = L.n, cells = m * n;
  const int wi = n - 2 * NH, w = (m - 2 * NH) * wi;
  // one wavefront when the interior fits it or the level is small (LDS
  // operations of a wavefront execute in order: a wave-level barrier
  // separates the phases)
  const bool wave = w <= 64 || cells <= sy.small_cells;
  const int st = wave ? 64 : sy.nt;
  if (w > OWNK * st) return false;
  if (!wave && sy.prev_small) __syncthreads();
  sy.prev_small = wave;
  auto sync = [&]() {
    if (wave) {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    } else {
      __syncthreads();
    }
  };
  const int P = cells;
  const bool active = !wave || sy.tid < 64;
  OwnCell<T> o[OWNK];
#pragma unroll
  for (int k = 0; k < OWNK; k++) {
    const int q = sy.tid + k * st;
    o[k].own = active && q < w;
    const int qj = o[k].own ? qdiv(q, wi) : 0;
    o[k].j = NH + qj;
    o[k].i = NH + (o[k].own ? q - qj * wi : 0);
    o[k].c = o[k].j * n + o[k].i;
    if (o[k].own) {
      const int c = o[k].c;
      o[k].a.a0 = T(L.A[c]);
      o[k].a.a1 = T(L.A[P + c]);
      o[k].a.a2 = T(L.A[2 * P + c]);
      o[k].a.a3 = T(L.A[3 * P + c]);
      o[k].a.a4 = T(L.A[4 * P + c]);
      o[k].a.a3e = T(L.A[3 * P + c + 1]);
      o[k].a.a2s = T(L.A[2 * P + c + n - 1]);
      o[k].a.a1s = T(L.A[P + c + n]);
      o[k].a.a0s = T(L.A[c + n + 1]);
      o[k].bc = L.b[c];
      o[k].mk = L.msk[c];
    }
  }
#pragma unroll
  for (int k = 0; k < OWNK; k++)
    if (o[k].own) o[k].c3 = c1 / fabs(o[k].a.a4);
  LDSP T* x = L.x;
  LDSP T* t = L.t;
  for (int it = 0; it < iters; it++) {
    const bool z = zero_in && it == 0;
#pragma unroll
    for (int k = 0; k < OWNK; k++)
      if (o[k].own) {
        const int c = o[k].c;
        T v;
        if (z)
          v = jacobi_c3_2d<T>(o[k].bc, o[k].c3, c2, o[k].a, [&](int, int) { return T(0); });
        else
          v = jacobi_c3_2d<T>(o[k].bc, o[k].c3, c2, o[k].a,
                              [&](int dj, int di) { return x[c + dj * n + di]; });
        store_img_mask<T>(t, nullptr, o[k].j, o[k].i, o[k].mk == 0 ? T(0) : v, m, n);
      }
    sync();
#pragma unroll
    for (int k = 0; k < OWNK; k++)
      if (o[k].own) {
        const int c = o[k].c;
        const T v = jacobi_c3_2d<T>(o[k].bc, o[k].c3, c2, o[k].a,
                                    [&](int dj, int di) { return t[c + dj * n + di]; });
        store_img_mask<T>(x, nullptr, o[k].j, o[k].i, o[k].mk == 0 ? T(0) : v, m, n);
      }
    sync();
  }
  if (resid) {
#pragma unroll
    for (int k = 0; k < OWNK; k++)
      if (o[k].own) {
        const int c = o[k].c;
        const T v = residual_2d<T>(o[k].mk, o[k].bc, o[k].a,
                                   [&](int dj, int di) { return x[c + dj * n + di]; });
        store_img_mask<T>(t, nullptr, o[k].j, o[k].i, v, m, n);
      }
    sync();
  }
  return true;
}

// the second half of c_resid_restrict: full weighting of `src` (level F)
// onto the coarse interior of C (+ images)
template <typename T, typename TA>
__device__ __forceinline__ void c_restrict(const CL<T, TA>& F, const CL<T, TA>& C, LDSP T* src, Sync& sy) {
  const int n = F.n;
  const int N2 = C.n - 2 * NH, M2 = C.m - 2 * NH, ccells = C.m * C.n;
  if (sy.begin(ccells))
    for (int q = sy.tid; q < M2 * N2; q += sy.stride(ccells)) {
      const int qj = qdiv(q, N2);
      const int J2 = NH + qj, I2 = NH + (q - qj * N2);
      T v = T(0);
      if (C.msk[at(J2, I2, C.n)] != 0) {
        const size_t c = at(2 * J2 + 1 - NH, 2 * I2 + 1 - NH, n);
        v = restrict_r<T>([&](int dj, int di) { return src[c + (ptrdiff_t)dj * n + di]; });
      }
      store_img_mask<T>(C.b, nullptr, J2, I2, v, C.m, C.n);
    }
  sy.end(ccells);
}

// a smoothing of level L (one smoothtwice), owned or generic
template <typename T, typename TA>
__device__ __forceinline__ void c_smooth_level(const CL<T, TA>& L, T c1, T c2, Sync& sy, bool zero_in) {
  if (!c_smooth_own(L, c1, c2, sy, zero_in, 1, false)) c_smooth(L, c1, c2, sy, zero_in);
}

// a double from another lane of the wavefront (two 32-bit permutes)
__device__ __forceinline__ double lane_get(double v, int src_lane) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(src_lane << 2, (int)u);
  const int hi = __builtin_amdgcn_ds_bpermute(src_lane << 2, (int)(u >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// The coarsest-level solve on one wavefront with x and t in registers:
// lane q holds cell q of the first sweep's region (interior + one ring) with
// its operator row, c1 / |A_5|, rhs and mask; a sweep fetches its nine
// neighbour values from other lanes.  x lives on the interior lanes (x at a
// halo position is x at its periodic image, which is what the LDS copy of
// the level holds there), t on all lanes of the region.  Same arithmetic as
// c_smooth_own, sweep for sweep; x (interior + images) is written to LDS at
// the end, t is not needed afterwards.  Levels whose region exceeds a
// wavefront, or with an interior narrower than 2, return false.
template <typename TA>
__device__ __forceinline__ bool c_coarsest_wave(const CL<double, TA>& L, double c1, double c2, Sync& sy,
                                int iters) {
  typedef double T;
  const int m = L.m, n = L.n, M = m - 2 * NH, N = n - 2 * NH;
  const int wj = M + 2, wi = N + 2, w = wj * wi;
  if (w > 64 || M < 2 || N < 2) return false;
  sy.prev_small = true;  // a workgroup-wide phase after this one synchronises first
  if (sy.tid >= 64) return true;
  const int lane = sy.tid, P = m * n;
  const bool own = lane < w;
  const int qj = own ? qdiv(lane, wi) : 0;
  const int j = NH - 1 + qj, i = NH - 1 + (own ? lane - qj * wi : 0);
  const int c = j * n + i;
  const bool inner = own && j >= NH && j < m - NH && i >= NH && i < n - NH;
  ARow<T> a = {};
  T c3 = T(0), bc = T(0);
  int8_t mk = 0;
  if (own) {
    a.a0 = T(L.A[c]);
    a.a1 = T(L.A[P + c]);
    a.a2 = T(L.A[2 * P + c]);
    a.a3 = T(L.A[3 * P + c]);
    a.a4 = T(L.A[4 * P + c]);
    a.a3e = T(L.A[3 * P + c + 1]);
    a.a2s = T(L.A[2 * P + c + n - 1]);
    a.a1s = T(L.A[P + c + n]);
    a.a0s = T(L.A[c + n + 1]);
    c3 = c1 / fabs(a.a4);
    bc = L.b[c];
    mk = L.msk[c];
  }
  // source lanes of the nine neighbours: for x, the image of the position;
  // for t, the position itself (inner cells' neighbours lie in the region)
  int sx[9], st[9];
#pragma unroll
  for (int dj = -1; dj <= 1; dj++)
#pragma unroll
    for (int di = -1; di <= 1; di++) {
      const int k = (dj + 1) * 3 + di + 1;
      sx[k] = own ? (img(j + dj, m) - (NH - 1)) * wi + img(i + di, n) - (NH - 1) : 0;
      st[k] = inner ? (j + dj - (NH - 1)) * wi + i + di - (NH - 1) : 0;
    }
  T xv = T(0), tv = T(0);
  for (int it = 0; it < iters; it++) {
    T v;
    if (it == 0) {
      v = jacobi_c3_2d<T>(bc, c3, c2, a, [&](int, int) { return T(0); });
    } else {
      T xn[9];
#pragma unroll
      for (int k = 0; k < 9; k++) xn[k] = lane_get(xv, sx[k]);
      v = jacobi_c3_2d<T>(bc, c3, c2, a, [&](int dj, int di) { return xn[(dj + 1) * 3 + di + 1]; });
    }
    tv = (own && mk != 0) ? v : T(0);
    T tn[9];
#pragma unroll
    for (int k = 0; k < 9; k++) tn[k] = lane_get(tv, st[k]);
    const T v2 = jacobi_c3_2d<T>(bc, c3, c2, a, [&](int dj, int di) { return tn[(dj + 1) * 3 + di + 1]; });
    xv = inner ? (mk == 0 ? T(0) : v2) : T(0);
  }
  if (inner) store_img_mask<T>(L.x, nullptr, j, i, xv, m, n);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  return true;
}

// a double from lane (lane moved by the DPP control CTRL) of the same row of
// 16 lanes: two 32-bit DPP moves
template <int CTRL>
__device__ __forceinline__ double dpp_get(double v) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)u, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// DPP controls on the 4 x 4 torus held by a row of 16 lanes, lane 4 r + c:
// the value of cell (r - 1, c) (row rotated right by 4 lanes), (r + 1, c)
// (by 12), (r, c - 1) and (r, c + 1) (quad permutes [3,0,1,2], [1,2,3,0])
constexpr int DPP_N = 0x124, DPP_S = 0x12C, DPP_W = 0x93, DPP_E = 0x39;

// The coarsest-level solve of a 4 x 4 interior on a row of 16 lanes, lane
// 4 r + c holding interior cell (r, c) with its operator row, c1 / |A_5|,
// rhs and mask.  Below the finest level every array holds periodic images
// in its halo (operator, rhs, mask, and x as stored), so the first sweep's
// value at a halo position is the value at its image: both sweeps are
// sweeps on the 4 x 4 torus, and the nine neighbour values of a cell are
// DPP lane moves within the row (a row rotation, then a quad permute for
// the diagonals) instead of LDS permutes.  Same arithmetic as
// c_coarsest_wave, sweep for sweep.  Each of the four rows of wavefront 0
// computes the same torus; row 0 stores x (interior + images).
template <typename TA>
__device__ __forceinline__ bool c_coarsest_torus4(const CL<double, TA>& L, double c1, double c2,
                                                  Sync& sy, int iters) {
  typedef double T;
  const int m = L.m, n = L.n;
  if (m - 2 * NH != 4 || n - 2 * NH != 4) return false;
  sy.prev_small = true;  // a workgroup-wide phase after this one synchronises first
  if (sy.tid >= 64) return true;
  const int lane = sy.tid & 15, P = m * n;
  const int j = NH + (lane >> 2), i = NH + (lane & 3), c = j * n + i;
  ARow<T> a;
  a.a0 = T(L.A[c]);
  a.a1 = T(L.A[P + c]);
  a.a2 = T(L.A[2 * P + c]);
  a.a3 = T(L.A[3 * P + c]);
  a.a4 = T(L.A[4 * P + c]);
  a.a3e = T(L.A[3 * P + c + 1]);
  a.a2s = T(L.A[2 * P + c + n - 1]);
  a.a1s = T(L.A[P + c + n]);
  a.a0s = T(L.A[c + n + 1]);
  const T c3 = c1 / fabs(a.a4);
  const T bc = L.b[c];
  const int8_t mk = L.msk[c];
  // the nine torus neighbours of v, o[(dj + 1) * 3 + di + 1] = v(r + dj, c + di)
  auto nbrs = [](T v, T* o) {
    const T up = dpp_get<DPP_N>(v), dn = dpp_get<DPP_S>(v);
    o[0] = dpp_get<DPP_W>(up);
    o[1] = up;
    o[2] = dpp_get<DPP_E>(up);
    o[3] = dpp_get<DPP_W>(v);
    o[4] = v;
    o[5] = dpp_get<DPP_E>(v);
    o[6] = dpp_get<DPP_W>(dn);
    o[7] = dn;
    o[8] = dpp_get<DPP_E>(dn);
  };
  T xv = T(0);
  for (int it = 0; it < iters; it++) {
    T v;
    if (it == 0) {
      v = jacobi_c3_2d<T>(bc, c3, c2, a, [&](int, int) { return T(0); });
    } else {
      T xn[9];
      nbrs(xv, xn);
      v = jacobi_c3_2d<T>(bc, c3, c2, a, [&](int dj, int di) { return xn[(dj + 1) * 3 + di + 1]; });
    }
    const T tv = mk != 0 ? v : T(0);
    T tn[9];
    nbrs(tv, tn);
    const T v2 = jacobi_c3_2d<T>(bc, c3, c2, a, [&](int dj, int di) { return tn[(dj + 1) * 3 + di + 1]; });
    xv = mk == 0 ? T(0) : v2;
  }
  if (sy.tid < 16) store_img_mask<T>(L.x, nullptr, j, i, xv, m, n);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  return true;
}

// the coarsest-level solve: 16 smoothings of a level entered with x = 0
// (hierarchy.py:117-118 / 144-145)
template <typename T, typename TA>
__device__ __forceinline__ void c_coarsest(const CL<T, TA>& L, T c1, T c2, Sync& sy) {
  if (c_coarsest_torus4(L, c1, c2, sy, 16)) return;
  if (c_coarsest_wave(L, c1, c2, sy, 16)) return;
  if (!c_smooth_own(L, c1, c2, sy, true, 16, false))
    for (int k = 0; k < 16; k++) c_smooth(L, c1, c2, sy, k == 0);
}

// V-cycle from level lev1 down to level deep of the coarse part
// (hierarchy.py:98-126)
template <typename T, typename TA>
__device__ __forceinline__ void c_vcycle(const CL<T, TA>* Ls, int lev1, int deep, bool zero_first,
                                         T c1, T c2, Sync& sy, int mode) {
  for (int l = lev1; l < deep; l++) {
    const bool z = l > lev1 || zero_first;
    if (c_smooth_own(Ls[l], c1, c2, sy, z, 1, true)) {
      c_restrict(Ls[l], Ls[l + 1], Ls[l].t, sy);
    } else {
      c_smooth(Ls[l], c1, c2, sy, z);
      c_resid_restrict(Ls[l], Ls[l + 1], false, sy);
    }
    F2D_CSTAMP(sy, mode, 10 + l);
  }
  c_coarsest(Ls[deep], c1, c2, sy);
  F2D_CSTAMP(sy, mode, 20);
  for (int l = deep - 1; l >= lev1; l--) {
    c_interp(Ls[l], Ls[l + 1], true, sy);
    c_smooth_level(Ls[l], c1, c2, sy, false);
    F2D_CSTAMP(sy, mode, 30 + l);
  }
  (void)mode;
}

constexpr int COARSE_THREADS = 1024;  // one workgroup per instance holds the CU's LDS anyway
constexpr int MAXCL = 6;              // most levels the coarse kernel takes

// Operands come from the T pool and are solved in FP64 in LDS (the
// coarsest-grid solve stays FP64 in mixed mode).  One block per instance.
template <typename T>
__global__ void __launch_bounds__(COARSE_THREADS) k_coarse(const Task* __restrict__ tasks,
                                                const InstD* __restrict__ I, T* Q,
                                                const int8_t* __restrict__ B,
                                                const DynD* __restrict__ dyn, int pred, int mode,
                                                int small_cells, int inst_first) {
  typedef double C;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  LDSP unsigned char* lds = (LDSP unsigned char*)lds_raw;
  const int ti = task_index();
  const Task& tk = tasks[ti];
  // the instance: first + task index where the launch's instances are
  // consecutive (inst_first >= 0), so that the descriptor read, and the
  // operator transfers addressed by it, need not wait for the task read
  const int inst = inst_first >= 0 ? inst_first + ti : tk.inst;
  const InstD& in = I[inst];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int k0 = in.nfine, deep = in.nlev - 1, nl = deep - k0 + 1;
  __shared__ CL<C, T> Ls[MAXLEV];
  Sync sy;
  sy.tid = tid;
  sy.nt = nt;
  sy.small_cells = small_cells;
  sy.prev_small = false;
  sy.ck = 0;
#ifdef F2D_PHASE_STAMPS
  if (tid == 0 && blockIdx.x == 0) {
    coarse_stamps[mode][254] = wall_clock64();
    coarse_stamps[mode][255] = 100 + nl;
  }
#endif
  F2D_CSTAMP(sy, mode, 0);
  // LDS: per level A (5 planes, pool type), x, b, t (FP64), 16-byte
  // aligned; then the masks.  Every thread computes the layout (thread 0
  // publishes it for the phases) and issues its share of the operator and
  // mask DMA (global_load_lds, 16 bytes per lane, 1 KB per wavefront; the
  // pools carry slack past their last array for the rounded-up tail) and
  // of the rhs loads before the step gate is read, so the gate's round trip
  // overlaps the transfers.
  auto r16 = [](unsigned v) { return (v + 15u) & ~15u; };
  {
    const int lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
    auto dma = [&](LDSP void* dst, const void* src, int bytes) {
      for (int c = wv; c * 1024 < bytes; c += nw) {
        const int off = c * 1024 + lane * 16;
        if (off < bytes)
          __builtin_amdgcn_global_load_lds((const void*)((const char*)src + off),
                                           (LDSP void*)((LDSP char*)dst + c * 1024), 16, 0, 0);
      }
    };
    unsigned off = 0;
    LDSP C* b0 = nullptr;
    for (int l = 0; l < nl; l++) {
      const LevD& G = in.L[k0 + l];
      const unsigned sz = (unsigned)(G.m * G.n);
      LDSP T* A = (LDSP T*)(lds + off);
      off += r16(sizeof(T) * 5 * sz);
      LDSP C* x = (LDSP C*)(lds + off);
      off += sizeof(C) * sz;
      LDSP C* bl = (LDSP C*)(lds + off);
      off += sizeof(C) * sz;
      LDSP C* t = (LDSP C*)(lds + off);
      off += sizeof(C) * sz;
      if (l == 0) b0 = bl;
      if (tid == 0) {
        Ls[l].m = G.m;
        Ls[l].n = G.n;
        Ls[l].A = A;
        Ls[l].x = x;
        Ls[l].b = bl;
        Ls[l].t = t;
      }
      dma(A, Q + G.A, (int)(5 * sz * sizeof(T)));
      for (unsigned q = tid; q < sz; q += nt) t[q] = C(0);
    }
    for (int l = 0; l < nl; l++) {
      const LevD& G = in.L[k0 + l];
      LDSP int8_t* mk = (LDSP int8_t*)(lds + off);
      if (tid == 0) Ls[l].msk = mk;
      off += r16(G.m * G.n);
      dma(mk, B + G.msk, G.m * G.n);
    }
    // the top level's rhs, widened to FP64: its first 8 * nt values are
    // loaded with the transfers, before the gate is read
    const LevD& G0 = in.L[k0];
    const int sz0 = G0.m * G0.n;
    const T* bg = Q + G0.b;
    T v0[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v0[k] = tid + k * nt < sz0 ? bg[tid + k * nt] : T(0);
    if (skip(dyn, inst, pred)) {
      __builtin_amdgcn_s_waitcnt(0);  // no transfer into LDS outlives the block
      return;
    }
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (tid + k * nt < sz0) b0[tid + k * nt] = (C)v0[k];
    for (int base = tid + 8 * nt; base < sz0; base += 8 * nt) {
      T v[8];
#pragma unroll
      for (int k = 0; k < 8; k++) v[k] = bg[min(base + k * nt, sz0 - 1)];
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (base + k * nt < sz0) b0[base + k * nt] = (C)v[k];
    }
    __builtin_amdgcn_s_waitcnt(0);
  }
  __syncthreads();
  F2D_CSTAMP(sy, mode, 1);
  const C c1 = C(in.omega), c2 = C(1. - in.omega);
  if (mode == 0) {
    c_vcycle(Ls, 0, nl - 1, true, c1, c2, sy, 0);
  } else {
    for (int l = 0; l < nl - 1; l++) {
      c_resid_restrict(Ls[l], Ls[l + 1], true, sy);
      F2D_CSTAMP(sy, 1, 40 + l);
    }
    c_coarsest(Ls[nl - 1], c1, c2, sy);
    F2D_CSTAMP(sy, 1, 20);
    for (int l = nl - 2; l >= 0; l--) {
      c_interp(Ls[l], Ls[l + 1], false, sy);
      F2D_CSTAMP(sy, 1, 50 + l);
      c_vcycle(Ls, l, nl - 1, false, c1, c2, sy, 1);
    }
  }
  __syncthreads();
  T* xg = Q + tk.xout;
  const int sz = Ls[0].m * Ls[0].n;
  for (int q = tid; q < sz; q += nt) xg[q] = T(Ls[0].x[q]);
  F2D_CSTAMP(sy, mode, 2);
#ifdef F2D_PHASE_STAMPS
  if (tid == 0 && blockIdx.x == 0) {
    coarse_stamps[mode][252] = wall_clock64();
    coarse_stamps[mode][253] = 102;
  }
#endif
}

// ---------------------------------------------------------------------------
// k_ortho: (u,v) from psi on [1, m-2] (computeorthogradient)
// ---------------------------------------------------------------------------
template <typename T, typename S>
__global__ void __launch_bounds__(256) k_ortho(const Task* __restrict__ tasks,
                                               const InstD* __restrict__ I, T* P, const S* SP,
                                               const int8_t* __restrict__ B,
                                               const DynD* __restrict__ dyn) {
  // P: the pool holding the state-shaped array (FP64 state or T tendency),
  // u and v written at +2F, +3F of tk.xout; psi read at SP + tk.xin (its
  // own plane, FP64 for the refined tendencies of the mixed mode)
  const int ti = task_index();
  const Task tk = tasks[ti];
  if (idle(dyn, tk.inst)) return;
  const InstD& in = I[tk.inst];
  const int m = in.m, n = in.n;
  const size_t F = (size_t)m * n;
  const int8_t* msk = B + in.msk;
  T* base = P + tk.xout;
  const S* psi = SP + tk.xin;
  const S zdx = S(1. / in.dx), zdy = S(1. / in.dy);
  for (int r = threadIdx.y; r < TY; r += blockDim.y) {
    const int j = tk.y0 + r, i = tk.x0 + threadIdx.x;
    if (j < 1 || i < 1 || j >= m - 1 || i >= n - 1) continue;
    const size_t c = at(j, i, n);
    base[2 * F + c] = (msk[c] + msk[c + 1] == 2) ? T(zdy * (psi[c - n] - psi[c])) : T(0);
    base[3 * F + c] = (msk[c] + msk[c + n] == 2) ? T(zdx * (psi[c] - psi[c - 1])) : T(0);
  }
}

// ---------------------------------------------------------------------------
// flat (element-wise) kernels: task.y0 = first element, task.x0 = count
// ---------------------------------------------------------------------------
constexpr int FLAT = 4096;

// final RK combination of the fields in fmask (bit k: field k)
template <typename T>
__global__ void __launch_bounds__(256) k_combine(const Task* __restrict__ tasks,
                                                 const InstD* __restrict__ I, double* P, T* Q,
                                                 const DynD* __restrict__ dyn) {
  const int ti = task_index();
  const Task tk = tasks[ti];
  if (idle(dyn, tk.inst)) return;
  const InstD& in = I[tk.inst];
  const size_t F = (size_t)in.m * in.n;
  const double dt = dyn[tk.inst].dt;
  double* x = P + in.state;
  const T *k0 = Q + in.k[0], *k1 = Q + in.k[1], *k2 = Q + in.k[2];
  const int field = tk.lev;
  const size_t base = (size_t)field * F;
  // psi tendencies of the mixed mode are FP64 planes of their own
  const bool p64 = field == 1 && in.kp[0] >= 0;
  const double *p0 = P + in.kp[0], *p1 = P + in.kp[1], *p2 = P + in.kp[2];
  for (int q = tk.y0 + threadIdx.x; q < tk.y0 + tk.x0; q += blockDim.x) {
    const size_t e = base + q;
    const double a0 = p64 ? p0[q] : (double)k0[e], a1 = p64 ? p1[q] : (double)k1[e],
                 a2 = p64 ? p2[q] : (double)k2[e];
    if (in.tscheme == 0) {
      const double c6 = dt / 6.;
      x[e] = x[e] + c6 * ((a0 + a1) + 4. * a2);
    } else {
      x[e] = x[e] + dt * a2;
    }
    if (in.model == 1 && field == 4) x[5 * F + q] = x[e] - 0.;  // banom = b - bref(=0)
  }
}

// psi += x0 over the whole array (hierarchy.py:171)
template <typename T>
__global__ void __launch_bounds__(256) k_update(const Task* __restrict__ tasks,
                                                const InstD* __restrict__ I, double* P, T* Q,
                                                const DynD* __restrict__ dyn, int pred) {
  const int ti = task_index();
  const Task tk = tasks[ti];
  if (skip(dyn, tk.inst, pred)) return;
  double* x = P + tk.xout;
  const T* y = Q + tk.xin;
  for (int q = tk.y0 + threadIdx.x; q < tk.y0 + tk.x0; q += blockDim.x)
    x[q] = x[q] + (double)y[q];
}

// psi = (psi - psim) * mskp (core/operators.py:474-481), psim used if perio.
// Task flag 2: the plane at tk.xout instead of the state's psi, masked only
// (the fast path, core/operators.py:483); flag 1: that plane is in PQ (the
// T pool of an FP64 batch).
__global__ void __launch_bounds__(256) k_finalize_psi(const Task* __restrict__ tasks,
                                                      const InstD* __restrict__ I, double* P,
                                                      double* PQ, const int8_t* __restrict__ B,
                                                      const DynD* __restrict__ dyn) {
  const int ti = task_index();
  const Task tk = tasks[ti];
  if (idle(dyn, tk.inst)) return;
  const InstD& in = I[tk.inst];
  const size_t F = (size_t)in.m * in.n;
  const bool plane = tk.flags & 2;
  double* psi = plane ? ((tk.flags & 1) ? PQ : P) + tk.xout : P + in.state + F;
  const bool demean = !plane && in.perio;
  // the islands' boundary psi after the mask (core/operators.py:483-489):
  // the state's inversions (flag 4 on an explicit plane)
  const bool isl = in.island && (!plane || (tk.flags & 4));
  const int8_t* mp = B + in.mskp;
  const double psim = dyn[tk.inst].psim;
  for (int q = tk.y0 + threadIdx.x; q < tk.y0 + tk.x0; q += blockDim.x) {
    double v = psi[q];
    if (demean) v = v - psim;
    v = v * (double)mp[q];
    if (isl) v = v + in.ipsi[q];
    psi[q] = v;
  }
}

// ---------------------------------------------------------------------------
// no-slip walls (core/operators.py:245-290) and momentum enforcement
// ---------------------------------------------------------------------------
// source = the no-slip wall source of the interior cells of the tile
// (noslip_source_at, from the state's psi after the time step)
__global__ void __launch_bounds__(256) k_ns_source(const Task* __restrict__ tasks,
                                                   const InstD* __restrict__ I, const double* P,
                                                   const DynD* __restrict__ dyn) {
  const int ti = task_index();
  const Task tk = tasks[ti];
  if (idle(dyn, tk.inst)) return;
  const InstD& in = I[tk.inst];
  const int m = in.m, n = in.n;
  const double* psi = P + in.state + (size_t)m * n;
  const double cff = 1. / (2 * in.dx * in.dy);  // core/fortran_operators.f90:240
  for (int r = threadIdx.y; r < TY; r += blockDim.y) {
    const int J = tk.y0 + r, Ic = tk.x0 + threadIdx.x;
    if (J >= m - NH || Ic >= n - NH) continue;
    in.nsrc[(size_t)J * n + Ic] = noslip_source_at(in.msknoslip, psi, J, Ic, n, cff);
  }
}

// Momentum part of the no-slip source (enforce_momentum): source -=
// meansource * mskbc on the interior, and per-tile partials of
// {sum source*xr0, sum source*yr0} over msk != 0 (computedotprod,
// core/fortran_diag.f90:3-30, summed per tile and folded)
__global__ void __launch_bounds__(256) k_ns_dot(const Task* __restrict__ tasks,
                                                const InstD* __restrict__ I,
                                                const int8_t* __restrict__ B, double* partial,
                                                const DynD* __restrict__ dyn) {
  __shared__ double red[2][4];
  const int ti = task_index();
  const Task tk = tasks[ti];
  if (idle(dyn, tk.inst)) return;
  const InstD& in = I[tk.inst];
  const int m = in.m, n = in.n;
  const int8_t* msk = B + in.msk;
  const double srcm = dyn[tk.inst].srcm;
  double a[2] = {0., 0.};
  for (int r = threadIdx.y; r < TY; r += blockDim.y) {
    const int J = tk.y0 + r, Ic = tk.x0 + threadIdx.x;
    if (J >= m - NH || Ic >= n - NH) continue;
    const size_t c = (size_t)J * n + Ic;
    const double v = in.nsrc[c] - srcm * (double)in.mskbc[c];
    in.nsrc[c] = v;
    if (msk[c] != 0) {
      a[0] += v * in.xr0[c];
      a[1] += v * in.yr0[c];
    }
  }
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int k = 0; k < 2; k++) {
    double x = a[k];
    for (int o = 32; o > 0; o >>= 1) x = x + __shfl_xor(x, o, 64);
    if ((tid & 63) == 0) red[k][tid >> 6] = x;
  }
  __syncthreads();
  if (tid < 2) {
    double x = red[tid][0];
    for (int q = 1; q < (int)(blockDim.x * blockDim.y) / 64; q++) x = x + red[tid][q];
    partial[(size_t)tk.slot * 2 + tid] = x;
  }
}

// The rest of rhs_noslip on the interior cells of the tile: source -=
// meansource * mskbc (or, after k_ns_dot, source -= (px xr0 + py yr0) mskbc
// with px = sum/x2bc, py = sum/y2bc), the periodic fill, then vorticity -=
// source on the cell and its images (core/operators.py:274-290)
__global__ void __launch_bounds__(256) k_ns_fix(const Task* __restrict__ tasks,
                                                const InstD* __restrict__ I, double* P,
                                                const DynD* __restrict__ dyn) {
  const int ti = task_index();
  const Task tk = tasks[ti];
  if (idle(dyn, tk.inst)) return;
  const InstD& in = I[tk.inst];
  const int m = in.m, n = in.n;
  double* w = P + in.state;
  const DynD& d = dyn[tk.inst];
  for (int r = threadIdx.y; r < TY; r += blockDim.y) {
    const int J = tk.y0 + r, Ic = tk.x0 + threadIdx.x;
    if (J >= m - NH || Ic >= n - NH) continue;
    const size_t c = (size_t)J * n + Ic;
    double v = in.nsrc[c];
    if (in.enforce) {
      const double px = d.nsdot[0] / in.x2bc, py = d.nsdot[1] / in.y2bc;
      v = v - (px * in.xr0[c] + py * in.yr0[c]) * (double)in.mskbc[c];
    } else {
      v = v - d.srcm * (double)in.mskbc[c];
    }
    each_img(J, Ic, m, n, [&](int j, int i) {
      const size_t e = (size_t)j * n + i;
      in.nsrc[e] = v;
      w[e] = w[e] - v;
    });
  }
}

// Flat element ranges (task y0 = first element, x0 = count) of one plane:
// flag 1: the no-slip source divided by the step's dt (core/euler.py:112);
// flag 2: the loop's momentum enforcement, vorticity -= xr0 px + yr0 py
// (core/fluid2d.py:352-358, px and py host-written);
// flag 4 / 8: copy of an FP64 plane into the T pool of an FP64 batch (P at
// xin -> PQ at xout) / back (PQ at xin -> P at xout, plus the islands'
// boundary psi when the task's mode is 1: psi += island psi,
// core/operators.py:483-489)
__global__ void __launch_bounds__(256) k_flat_ops(const Task* __restrict__ tasks,
                                                  const InstD* __restrict__ I, double* P,
                                                  double* PQ, const DynD* __restrict__ dyn) {
  const int ti = task_index();
  const Task tk = tasks[ti];
  if (idle(dyn, tk.inst)) return;
  const InstD& in = I[tk.inst];
  const DynD& d = dyn[tk.inst];
  const int f = tk.flags;
  for (int q = tk.y0 + threadIdx.x; q < tk.y0 + tk.x0; q += blockDim.x) {
    if (f & 1) {
      in.nsrc[q] = in.nsrc[q] / d.dt;
    } else if (f & 2) {
      double* w = P + in.state;
      w[q] = w[q] - (in.xr0[q] * d.mom[0] + in.yr0[q] * d.mom[1]);
    } else if (f & 4) {
      PQ[tk.xout + q] = P[tk.xin + q];
    } else {
      double v = PQ[tk.xin + q];
      if (tk.mode == 1) v = v + in.ipsi[q];
      P[tk.xout + q] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
// NumPy pairwise sum of interior elements [s, s+cnt) (flattened C order)
template <typename T>
__device__ double np_pairwise_flat(const T* x, int n, int N, size_t start, size_t cnt) {
  struct Fr {
    size_t s, c;
    double left;
    int stage;
  };
  Fr st[16];
  int sp = 0;
  st[0] = {start, cnt, 0., 0};
  double ret = 0.;
  auto val = [&](size_t q) { return (double)x[at(NH + (int)(q / N), NH + (int)(q % N), n)]; };
  while (sp >= 0) {
    Fr& f = st[sp];
    if (f.c <= 128) {
      double r;
      if (f.c < 8) {
        r = -0.0;
        for (size_t k = 0; k < f.c; k++) r += val(f.s + k);
      } else {
        double a[8];
        for (int k = 0; k < 8; k++) a[k] = val(f.s + k);
        size_t k;
        for (k = 8; k < f.c - (f.c % 8); k += 8)
          for (int q = 0; q < 8; q++) a[q] += val(f.s + k + q);
        r = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
        for (; k < f.c; k++) r += val(f.s + k);
      }
      ret = r;
      sp--;
      while (sp >= 0) {
        Fr& pp = st[sp];
        if (pp.stage == 1) {
          pp.left = ret;
          pp.stage = 2;
          size_t h = pp.c / 2;
          h -= h % 8;
          st[++sp] = {pp.s + h, pp.c - h, 0., 0};
          break;
        } else {
          ret = pp.left + ret;
          sp--;
        }
      }
      continue;
    }
    size_t h = f.c / 2;
    h -= h % 8;
    f.stage = 1;
    st[sp + 1] = {f.s, h, 0., 0};
    sp++;
  }
  return ret;
}

// One wavefront per 8192-element chunk of psi's interior (task.y0 = first
// chunk, task.x0 = chunk count, one block per task with 64 threads per
// chunk).  When the chunk length is 128 * 2^k -- every chunk of a
// power-of-two grid -- NumPy's recursion splits it into equal halves down to
// 128-element leaves, so lane L sums leaf L with NumPy's eight-accumulator
// loop and the leaves are combined as the same balanced tree.  Other
// lengths run NumPy's recursion on lane 0.
__global__ void __launch_bounds__(256) k_npsum(const Task* __restrict__ tasks,
                                               const InstD* __restrict__ I, double* P,
                                               double* partial, const int* __restrict__ slot,
                                               const DynD* __restrict__ dyn) {
  const int ti = task_index();
  const Task tk = tasks[ti];
  if (idle(dyn, tk.inst)) return;
  const InstD& in = I[tk.inst];
  // task flag 1: the no-slip source (meansource, core/operators.py:274);
  // else the state's psi of a periodic instance (core/operators.py:474-477)
  const bool src = tk.flags & 1;
  if (!src && !in.perio) return;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m = in.m, n = in.n, N = n - 2 * NH;
  const size_t tot = (size_t)(m - 2 * NH) * N;
  const double* psi = src ? in.nsrc : P + in.state + (size_t)m * n;
  for (int ch = wv; ch < tk.x0; ch += blockDim.x >> 6) {
    const size_t s0 = (size_t)(tk.y0 + ch) * 8192;
    const size_t cnt = tot - s0 < 8192 ? tot - s0 : 8192;
    const size_t leaves = cnt / 128;
    const bool balanced = cnt >= 128 && cnt % 128 == 0 && (leaves & (leaves - 1)) == 0;
    double r = 0.;
    if (!balanced) {
      if (lane == 0) r = np_pairwise_flat<double>(psi, n, N, s0, cnt);
    } else {
      if ((size_t)lane < leaves) {
        const size_t b0 = s0 + (size_t)lane * 128;
        auto val = [&](size_t q) { return psi[at(NH + (int)(q / N), NH + (int)(q % N), n)]; };
        double a[8];
        for (int k = 0; k < 8; k++) a[k] = val(b0 + k);
        for (int k = 8; k < 128; k += 8)
          for (int q = 0; q < 8; q++) a[q] += val(b0 + k + q);
        r = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
      }
      // balanced tree: at width w, lane L (L % 2w == 0) adds lane L + w
      for (size_t w = 1; w < leaves; w <<= 1) {
        const double o = __shfl_down(r, (unsigned)w, 64);
        if ((lane % (2 * w)) == 0) r = r + o;
      }
    }
    if (lane == 0) partial[slot[tk.slot] + ch] = r;
  }
}

// diagnostics partials: {maxu, ke, sum w, sum w^2, d4, d5, d6, d7}
__global__ void __launch_bounds__(256) k_diag(const Task* __restrict__ tasks,
                                              const InstD* __restrict__ I, const double* P,
                                              const int8_t* __restrict__ B, double* partial,
                                              const DynD* __restrict__ dyn) {
  typedef double T;
  __shared__ double red[8][4];
  const int ti = task_index();
  const Task tk = tasks[ti];
  if (idle(dyn, tk.inst)) return;
  const InstD& in = I[tk.inst];
  const int m = in.m, n = in.n;
  const size_t F = (size_t)m * n;
  const int8_t* msk = B + in.msk;
  const T* s = P + in.state;
  const T *w = s, *psi = s + F, *u = s + 2 * F, *v = s + 3 * F, *f4 = s + 4 * F;
  const double* xr = P + in.xr;
  const double* yr = P + in.yr;
  double a[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  for (int r = threadIdx.y; r < TY; r += blockDim.y) {
    const int J = tk.y0 + r, Ic = tk.x0 + threadIdx.x;
    if (J >= m - NH || Ic >= n - NH) continue;
    const size_t c = at(J, Ic, n);
    const int mk = msk[c];
    if (mk == 1) {
      const double zu = (double)u[c] * u[c] + (double)u[c - 1] * u[c - 1];
      const double zv = (double)v[c] * v[c] + (double)v[c - n] * v[c - n];
      a[0] = fmax(a[0], fabs((double)u[c] + u[c - 1]) + fabs((double)v[c] + v[c - n]));
      a[1] += zu + zv;
      a[2] += w[c];
      a[3] += (double)w[c] * w[c];
    }
    if (in.model == 0) {
      if (mk != 0) {
        a[4] += (double)w[c] * xr[c];
        a[5] += (double)w[c] * yr[c];
        a[6] += psi[c];
        a[7] += (double)w[c] * f4[c];
      }
    } else {
      if (mk != 0) a[4] += (double)f4[c] * yr[c];
      if (mk == 1) {
        a[5] += f4[c];
        a[6] += (double)f4[c] * f4[c];
      }
    }
  }
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int k = 0; k < 8; k++) {
    double x = a[k];
    for (int o = 32; o > 0; o >>= 1) {
      const double y = __shfl_xor(x, o, 64);
      x = k == 0 ? fmax(x, y) : x + y;
    }
    if ((tid & 63) == 0) red[k][tid >> 6] = x;
  }
  __syncthreads();
  if (tid < 8) {
    double x = red[tid][0];
    for (int q = 1; q < (int)(blockDim.x * blockDim.y) / 64; q++)
      x = tid == 0 ? fmax(x, red[tid][q]) : x + red[tid][q];
    partial[(size_t)tk.slot * 8 + tid] = x;
  }
}

// Per-instance fold of partial sums plus solve control.  256 threads; thread
// t folds slots t, t+256, ... in order, then a fixed tree combines the
// threads, so the result is the same on every run.
// what: 0 initial norms (slot pairs {r, b}), 1 iteration norms,
//       2 diagnostics (8 per slot), 3 NumPy chunks -> psim (sequential,
//       as NumPy adds its chunk sums in order), 4 NumPy chunks of the
//       no-slip source -> srcm, 5 no-slip momentum sums (pairs)
struct FoldSpan {
  int first, count;  // partial slots of this instance
};
// One block per instance of the window [inst0, inst0 + gridDim.x); spans
// holds one entry per instance of the window.
__global__ void __launch_bounds__(256) k_fold(const FoldSpan* __restrict__ spans,
                                              const InstD* __restrict__ I, DynD* dyn,
                                              const double* __restrict__ partial, int what,
                                              int maxite, double tol, int pred, int adapt_dt,
                                              double cfl, double dtmax, int inst0) {
  __shared__ double sh[8][256];
  const int inst = inst0 + blockIdx.x, tid = threadIdx.x;
  DynD& d = dyn[inst];
  const FoldSpan sp = spans[blockIdx.x];
  if (sp.count == 0 || !d.run) return;
  if (what == 3 || what == 4) {
    if (tid != 0 || (what == 3 && !I[inst].perio)) return;
    double acc = 0.;
    for (int k = 0; k < sp.count; k++) acc = acc + partial[sp.first + k];
    // domain_integration(z) = np.sum(interior) * 1. (core/grid.py:132-140)
    if (what == 3)
      d.psim = acc * 1. / I[inst].area;
    else
      d.srcm = acc * 1. / I[inst].bcarea;
    return;
  }
  if (what < 2 && pred && !d.active) return;
  const int nv = what == 2 ? 8 : 2;
  double a[8];
  for (int k = 0; k < nv; k++) a[k] = 0.;
  for (int q = tid; q < sp.count; q += 256)
    for (int k = 0; k < nv; k++) {
      const double y = partial[(size_t)(sp.first + q) * nv + k];
      a[k] = (what == 2 && k == 0) ? fmax(a[k], y) : a[k] + y;
    }
  for (int k = 0; k < nv; k++) sh[k][tid] = a[k];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h)
      for (int k = 0; k < nv; k++)
        sh[k][tid] = (what == 2 && k == 0) ? fmax(sh[k][tid], sh[k][tid + h])
                                           : sh[k][tid] + sh[k][tid + h];
    __syncthreads();
  }
  if (tid != 0) return;
  for (int k = 0; k < nv; k++) a[k] = sh[k][0];
  if (what == 5) {
    d.nsdot[0] = a[0];
    d.nsdot[1] = a[1];
    return;
  }
  if (what == 2) {
    a[0] *= 0.5;
    a[1] *= 0.25;
    if (I[inst].model == 1) a[4] = -I[inst].gravity * a[4];
    for (int k = 0; k < 8; k++) d.diag[k] = a[k];
    d.umax = a[0];  // cst[3] of the next step (core/fluid2d.py:397)
    if (adapt_dt && a[0] != 0.) {  // core/fluid2d.py:378-391
      double dt = cfl * fmin(I[inst].dx, I[inst].dy) / a[0];
      if (dt > dtmax) dt = dtmax;
      d.dt = dt;
    }
    return;
  }
  const double nr = a[0], nb = a[1];
  if (what == 0) {  // core/gmg/hierarchy.py:159-169
    d.normeb = sqrt(nb / 1.);
    d.nite = 0;
    d.ndiv = 0;
    d.diverged = 0;
    if (!(d.normeb > 0)) {
      d.res = 0.;
      d.res0 = 0.;
      d.active = 0;
      return;
    }
    d.res0 = sqrt(nr / 1.) / d.normeb;
    d.res = d.res0;
    d.active = (maxite > 0 && d.res0 > tol) ? 1 : 0;
    return;
  }
  // what == 1: core/gmg/hierarchy.py:172-188
  const double res = sqrt(nr / 1.) / d.normeb;
  const double conv = d.res0 / res;
  d.res0 = res;
  d.res = res;
  d.nite += 1;
  if (conv < 1) d.ndiv += 1;
  if (d.ndiv > 4) d.diverged = 1;
  d.active = (d.nite < maxite && d.res0 > tol && !d.diverged) ? 1 : 0;
}

}  // namespace eng
}  // namespace f2d

// ===========================================================================
// host side: batch layout, launch program, graph
// ===========================================================================
extern int f2d_gmg_build_for_engine(int nx, int ny, double dx, double dy, double hydroepsilon,
                                    const int8_t* mskp_host, void** gmg_out);
extern int f2d_gmg_level_raw(void* g, int lev, int8_t** msk, double** A, int* m, int* n);
extern void f2d_gmg_free_raw(void* g);

namespace f2d {
namespace eng {

static thread_local std::string e_err;
static int efail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static int efail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  e_err = buf;
  return code;
}
#define ECHK(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return efail(F2D_E_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #expr,              \
                   hipGetErrorString(e_));                                             \
  } while (0)
#define ETRY(expr)                   \
  do {                               \
    int rc_ = (expr);                \
    if (rc_ != F2D_OK) return rc_;   \
  } while (0)

enum Kind {
  L_RHS,
  L_C2C,
  L_SMOOTH,
  L_RR,
  L_COARSE,
  L_ORTHO,
  L_COMBINE,
  L_UPDATE,
  L_FINPSI,
  L_NPSUM,
  L_DIAG,
  L_FOLD,
  L_GATE,
  L_PERSIST,
  L_SMOOTH_RR,
  L_NSSRC,
  L_NSDOT,
  L_NSFIX,
  L_FLATOPS,
  L_UP2,
  L_RB2
};

struct Launch {
  int kind = 0;
  int toff = 0, ntask = 0;   // task table slice
  int pred = 0, flags = 0;   // predication, variant flags
  int stage = 0, outk = 0;   // RK stage, tendency slot (k_rhs)
  int mode = 0;              // coarse mode / fold what
  int span_off = 0;          // fold spans (one entry per instance of the window)
  int inst0 = 0, ninst = 0;  // fold: instance window [inst0, inst0 + ninst)
  int ph_off = 0, nph = 0;   // persistent launch: phase table slice
  int bar = 0;               // persistent launch: barrier slot
  int slot_off = 0;          // npsum slots
  size_t shmem = 0;          // k_coarse dynamic LDS
  int ty = TY;               // fine tile rows of k_smooth (k_resid_restrict: ty / 2 coarse rows)
  int maxite = -1;           // k_fold stopping test (< 0: the batch's full-solve values)
  double tol = 0.;
};

struct HostInst {
  InstD d;
  f2d_model_params p;
  long long x0alt = 0;
  int cur[MAXLEV];
  long long buf[MAXLEV][2];  // x buffers per level (level 0 set per phase)
  size_t coarse_lds = 0;
  std::vector<int8_t> hmsk;  // cell mask (host copy): tile classification
};

struct Batch {
  int ninst = 0;
  int prec = 0;  // 0: FP64 throughout; 1: FP32 tendencies and V-cycles
  std::vector<HostInst> H;
  double* P = nullptr;  // FP64 pool
  void* Q = nullptr;    // T pool (double or float)
  int8_t* B = nullptr;
  size_t nP = 0, nQ = 0, nB = 0;
  InstD* dI = nullptr;
  DynD* dD = nullptr;
  DynD* hD = nullptr;  // pinned mirror
  Task* dT = nullptr;
  FoldSpan* dS = nullptr;
  int* dSlot = nullptr;
  double* partial = nullptr;
  size_t npartial = 0;
  std::vector<Task> tasks;
  std::vector<FoldSpan> spans;
  std::vector<int> slots;
  std::vector<Launch> prog_init, prog_step, prog_diag;
  // per instance: its own initial inversion + diagnostics (a slot re-used
  // for a new run while the other instances keep stepping)
  std::vector<std::vector<Launch>> prog_one;
  // per instance: the loop's momentum enforcement (empty unless enabled)
  std::vector<std::vector<Launch>> prog_mom;
  // device arrays of the no-slip walls, islands and momentum enforcement
  std::vector<void*> extra;
  hipStream_t s = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool use_graph = true;
  int adapt = 0;
  double cfl = 0.8, dtmax = 5.0;
  int maxite = 4;
  double tol = 1e-11;
  // fast (per-stage) inversions: twoVcycle (fast_maxite == 0, the
  // reference) or gmg.solve(fast_maxite, fast_tol) -- the alternative
  // core/operators.py:452-455 keeps commented out; in mixed precision its
  // F-cycles run in FP32 between FP64 residuals and corrections (iterative
  // refinement)
  int fast_maxite = 0;
  double fast_tol = 1e-10;
  // down-legs below the finest level entered from zero in one launch
  // (k_smooth_rr) instead of a smoothing and a residual/restriction launch
  bool fuse_down = true;
  // the same for the start level of an F-cycle's V-cycle below the finest
  // (entered from the prolongation)
  bool fuse_ip = true;
  // two consecutive up-leg smoothings below the finest level in one launch
  // (k_up2)
  bool fuse_up = true;
  size_t max_shmem = 0;
  // persistent multigrid launches (tuning.persist workgroups per CU, 0 off)
  int persist = 0;
  std::vector<PPhase> phases;
  PPhase* dPh = nullptr;
  BarSlot* bar = nullptr;
  int nbar = 0;
  int* fault = nullptr;  // device flag: a persistent barrier gave up
  int min_tiles = 0;           // fine-phase tile choice (Builder::pick_ty)
  int coarse_wave_cells = 256; // coarse levels of at most this many cells run on one wavefront
  size_t coarse_budget = 0;    // LDS bytes for k_coarse (0: device maximum)
};

// launches of the fused down-leg of at most this many tiles run the staged
// 1024-thread form (two workgroups per CU: 512 tiles of 64 x 8 on the 512^2
// level take 0.8 % less per 1024^2 step than 256 tiles of 64 x 16)
constexpr double RR_STAGE_TILES_B = 512.;
// the smallest fused down-leg launches (at most this many tiles of 4 rows:
// the 128^2 and 64^2 levels at 1024^2) take tiles of 2 fine rows, twice the
// workgroups, each with half the cells (-0.6 % per 1024^2 step)
constexpr double RR_SMALL_TILES = 64.;
// two-level restrictions of b: launches of at most RB2_TILES tiles of 8
// output rows take tiles of RB2_TY_SMALL rows
constexpr int RB2_TILES = 256, RB2_TY_SMALL = 4;

// ---- task helpers ----------------------------------------------------------
static Task mk_task(int inst, int lev, int y0, int x0) {
  Task t;
  memset(&t, 0, sizeof t);
  t.inst = inst;
  t.lev = lev;
  t.y0 = y0;
  t.x0 = x0;
  return t;
}

// tiles of ty x TX over rows [r0, r1) x cols [c0, c1)
// Tiles row by row; each row's columns start at a rotation of 5 per row.
// Workgroups go to the compute units of an XCD round-robin, so without the
// rotation a unit would get tiles of one column only, and the units holding
// the (slower) boundary columns would finish last.
static void tiles(std::vector<Task>& out, const Task& proto, int r0, int r1, int c0, int c1,
                  int ty = TY) {
  const int nc = (c1 - c0 + TX - 1) / TX;
  int row = 0;
  for (int y = r0; y < r1; y += ty, row++)
    for (int k = 0; k < nc; k++) {
      Task t = proto;
      t.y0 = y;
      t.x0 = c0 + ((k + 5 * row) % nc) * TX;
      out.push_back(t);
    }
}

struct Builder {
  Batch& b;
  std::vector<Launch>* prog;
  explicit Builder(Batch& bb) : b(bb), prog(nullptr) {}

  Launch& add(int kind, const std::vector<Task>& ts) {
    Launch L;
    L.kind = kind;
    L.toff = (int)b.tasks.size();
    L.ntask = (int)ts.size();
    b.tasks.insert(b.tasks.end(), ts.begin(), ts.end());
    prog->push_back(L);
    return prog->back();
  }
  // spans of a reduction launch (tasks grouped by instance, contiguous)
  // instance window of the program being built: folds cover [lo, hi)
  int lo = 0, hi = 0;
  int spans_of(const std::vector<Task>& ts) {
    const int off = (int)b.spans.size();
    std::vector<FoldSpan> sp(hi - lo, FoldSpan{0, 0});
    for (int k = 0; k < (int)ts.size(); k++) {
      FoldSpan& f = sp[ts[k].inst - lo];
      if (f.count == 0) f.first = k;
      f.count++;
    }
    b.spans.insert(b.spans.end(), sp.begin(), sp.end());
    return off;
  }
  void fold(int what, int span_off, int pred, int maxite = -1, double tol = 0.) {
    Launch L;
    L.kind = L_FOLD;
    L.mode = what;
    L.span_off = span_off;
    L.inst0 = lo;
    L.ninst = hi - lo;
    L.pred = pred;
    L.maxite = maxite;
    L.tol = tol;
    prog->push_back(L);
  }

  const LevD& lev(int i, int l) { return b.H[i].d.L[l]; }

  // Fine-phase tile height: the tallest of 16, 8, 4 rows that still gives
  // at least min_tiles workgroups over the cells of the phase (all
  // instances at bottom-aligned level dd); min_tiles <= 0: always 16.
  int pick_ty_cells(double cells) const {
    if (b.min_tiles <= 0) return TY;
    for (int ty = TY; ty > 4; ty /= 2)
      if (cells / ((double)TX * ty) >= (double)b.min_tiles) return ty;
    return 4;
  }
  // tile rows of a fused down-leg phase: 2 on the smallest phases, else ty,
  // or the lowest of 8 and `most` (rows) that brings the phase to at most
  // RR_STAGE_TILES_B tiles
  int rr_tile_rows(const std::vector<int>& who, const std::vector<int>& start, int dd, int ty,
                   int most) const {
    double cells = 0.;
    for (int i : who) {
      const int D = b.H[i].d.nfine - 1 - start[i];
      if (dd > D) continue;
      const LevD& L = b.H[i].d.L[b.H[i].d.nfine - 1 - dd];
      cells += (double)(L.m - 2 * NH) * (L.n - 2 * NH);
    }
    if (cells / ((double)TX * 4) <= RR_SMALL_TILES) return 2;
    if (cells / ((double)TX * ty) <= RR_STAGE_TILES_B) return ty;
    for (int t = 2 * ty; t <= most; t *= 2)
      if (cells / ((double)TX * t) <= RR_STAGE_TILES_B) return t;
    return ty;
  }
  int pick_ty(const std::vector<int>& who, const std::vector<int>& start, int dd) const {
    double cells = 0.;
    for (int i : who) {
      const int D = b.H[i].d.nfine - 1 - start[i];
      if (dd > D) continue;
      const LevD& L = b.H[i].d.L[b.H[i].d.nfine - 1 - dd];
      cells += (double)(L.m - 2 * NH) * (L.n - 2 * NH);
    }
    return pick_ty_cells(cells);
  }

  // A finest-level tile whose stencil region may reach a solid cell takes
  // the smoother's per-cell stencil-bit path, about twice the arithmetic of
  // a tile with the constant stencil (judged from the cell mask with a
  // 3-cell margin).
  bool heavy_tile(const Task& t, int ty, bool rr) const {
    if (t.lev != 0) return false;
    const HostInst& h = b.H[t.inst];
    const int m = h.d.m, n = h.d.n;
    // the fine rows and columns the task reads (a residual/restriction task
    // is placed in coarse coordinates, ty coarse rows of TX2 columns)
    const int j0 = rr ? 2 * t.y0 - 4 : t.y0 - 3, j1 = rr ? 2 * (t.y0 + ty) + 2 : t.y0 + ty + 3;
    const int i0 = rr ? 2 * t.x0 - 4 : t.x0 - 3, i1 = rr ? 2 * (t.x0 + TX2) + 2 : t.x0 + TX + 3;
    for (int j = std::max(0, j0); j < std::min(m, j1); j++)
      for (int i = std::max(0, i0); i < std::min(n, i1); i++)
        if (h.hmsk[(size_t)j * n + i] == 0) return true;
    return false;
  }
  // Reorder each XCD's share of a launch's tasks (task_index gives XCD x a
  // contiguous share; its first workgroups land on distinct CUs) so that
  // its heavy tiles come first: no CU then holds two of them while others
  // hold none.  The shares keep their tiles, so neighbouring tiles still
  // meet in one XCD's L2.  Single-instance launches without a reduction
  // over the task order only.
  void balance(std::vector<Task>& ts, int ty, bool rr = false) const {
    const int n = (int)ts.size();
    if (n < 4 * NXCD) return;
    for (const Task& t : ts)
      if (t.inst != ts[0].inst) return;
    const int q = n / NXCD, r = n % NXCD;
    for (int x = 0; x < NXCD; x++) {
      const int a = x * q + std::min(x, r), e = a + q + (x < r ? 1 : 0);
      std::stable_partition(ts.begin() + a, ts.begin() + e,
                            [&](const Task& t) { return heavy_tile(t, ty, rr); });
    }
  }
  // smoother on level l of instance i: input mode, output the other buffer
  // per instance: the FP64 plane the finest level's up-leg smoothing adds
  // its result to (x0 += x folded into it), -1 for none; in the T pool of an
  // FP64 batch when upd_q
  std::vector<long long> upd_plane;
  bool upd_q = false;
  // interior cells of instance i's finest level
  long long h_cells(int i) const {
    const LevD& L = b.H[i].d.L[0];
    return (long long)(L.m - 2 * NH) * (L.n - 2 * NH);
  }
  Task smooth_task(int i, int l, int mode, int final_mask, std::vector<Task>& out, int ty) {
    HostInst& h = b.H[i];
    Task t = mk_task(i, l, 0, 0);
    t.mode = mode;
    t.flags = final_mask;
    t.xin = h.buf[l][h.cur[l]];
    t.xout = h.buf[l][h.cur[l] ^ 1];
    if (mode == 2 || mode == 3) t.xc = h.buf[l + 1][h.cur[l + 1]];
    if (l == 0 && mode == 2 && !upd_plane.empty() && upd_plane[i] >= 0) {
      t.flags |= 4 | (upd_q ? 8 : 0);
      t.xcc = upd_plane[i];
    }
    const LevD& L = lev(i, l);
    tiles(out, t, NH, L.m - NH, NH, L.n - NH, ty);
    h.cur[l] ^= 1;
    return t;
  }
  // residual (or b) of level l restricted onto l+1
  void rr_tasks(int i, int l, long long xin, long long rhs, long long rstore,
                std::vector<Task>& out, int ty2 = TY2) {
    const LevD& C = lev(i, l + 1);
    Task t = mk_task(i, l, 0, 0);
    t.xin = xin;
    t.xout = rhs;
    t.xc = rstore;
    // rows of coarse tiles, columns rotated per row as in tiles()
    const int nc = (C.n - 2 * NH + TX2 - 1) / TX2;
    int row = 0;
    for (int y = NH; y < C.m - NH; y += ty2, row++)
      for (int k = 0; k < nc; k++) {
        Task u = t;
        u.y0 = y;
        u.x0 = NH + ((k + 5 * row) % nc) * TX2;
        out.push_back(u);
      }
  }

  // b of level l restricted onto l + mode (k_restrict_b2): tiles of
  // ty x TX2 over the interior of level l + mode, columns rotated per row
  // as in tiles()
  void rb_tasks(int i, int l, int mode, std::vector<Task>& out, int ty) {
    const LevD& C = lev(i, l + mode);
    Task t = mk_task(i, l, 0, 0);
    t.mode = mode;
    t.xc = mode == 2 ? C.b : 0;  // b of level l + 2
    const int nc = (C.n - 2 * NH + TX2 - 1) / TX2;
    int row = 0;
    for (int y = NH; y < C.m - NH; y += ty, row++)
      for (int k = 0; k < nc; k++) {
        Task u = t;
        u.y0 = y;
        u.x0 = NH + ((k + 5 * row) % nc) * TX2;
        out.push_back(u);
      }
  }

  // V-cycles of the instances in `who`; instance i starts at top-down level
  // start[i] (its bottom-aligned index D_i = nfine_i - 1 - start[i]) with
  // smoother mode m0 there; levels below are zeroed; the coarse part runs in
  // k_coarse.  final_top: mask the last smooth of the start level by mskp.
  void vcycle(const std::vector<int>& who, const std::vector<int>& start, int m0, int pred,
              bool final_top) {
    int Dmax = -1;
    for (int i : who) Dmax = std::max(Dmax, b.H[i].d.nfine - 1 - start[i]);
    for (int dd = Dmax; dd >= 0; dd--) {
      const int ty = pick_ty(who, start, dd);
      // the fused down-leg takes taller tiles where that brings its launch
      // to at most RR_STAGE_TILES tiles (at most two 1024-thread workgroups
      // per CU, operator staged in LDS); x0 = I(xc) tiles stay at most 8 rows
      const int tyr1 = rr_tile_rows(who, start, dd, ty, 16), tyr3 = rr_tile_rows(who, start, dd, ty, 8);
      std::vector<Task> ts, rs, fs, f3;
      for (int i : who) {
        const int D = b.H[i].d.nfine - 1 - start[i];
        if (dd > D) continue;
        const int l = b.H[i].d.nfine - 1 - dd;
        const int mode = dd == D ? m0 : 1;
        HostInst& h = b.H[i];
        if (l >= 1 && b.fuse_down && (mode == 1 || (mode == 3 && ty <= 8 && b.fuse_ip))) {
          // below the finest level, entered with x = 0 or x = I(xc):
          // smoothing, residual and restriction in one launch (k_smooth_rr)
          const long long xo = h.buf[l][h.cur[l] ^ 1];  // the smoothed x (task xout)
          const long long xc = h.buf[l + 1][h.cur[l + 1]];
          h.cur[l] ^= 1;
          std::vector<Task>& out = mode == 3 ? f3 : fs;
          const size_t n0 = out.size();
          rr_tasks(i, l, 0, xo, mode == 3 ? xc : -1, out, (mode == 3 ? tyr3 : tyr1) / 2);
          for (size_t k = n0; k < out.size(); k++) out[k].mode = mode;
          continue;
        }
        smooth_task(i, l, mode, 0, ts, ty);
        rr_tasks(i, l, h.buf[l][h.cur[l]], -1, -1, rs, ty / 2);
      }
      if (!ts.empty()) {
        balance(ts, ty);
        balance(rs, ty / 2, true);
        Launch& Ls = add(L_SMOOTH, ts);
        Ls.pred = pred;
        Ls.ty = ty;
        Launch& Lr = add(L_RR, rs);
        Lr.pred = pred;
        Lr.ty = ty;
      }
      for (std::vector<Task>* f : {&fs, &f3}) {
        if (f->empty()) continue;
        Launch& Lf = add(L_SMOOTH_RR, *f);
        Lf.pred = pred;
        Lf.ty = f == &f3 ? tyr3 : tyr1;
        Lf.mode = f == &f3 ? 3 : 1;
      }
    }
    coarse(who, 0, pred);
    // up-leg: levels below the finest two at a time (k_up2, one 1024-thread
    // workgroup per tile) where the instance has both, the finer one is not
    // the masked start level and the finer level's launch is small (at most
    // UP2_TILES tiles, taking tiles of up to 16 rows for that): there a pair
    // costs less than its two latency-bound launches (64^2 + 128^2, and
    // 256^2 + 512^2 on 64 x 16 tiles: -0.5 % per 1024^2 step); larger
    // launches would not be resident in one wave
    constexpr double UP2_TILES = 256.;
    std::vector<int> paired(b.ninst, -1);
    for (int dd = 0; dd <= Dmax; dd++) {
      const int ty = pick_ty(who, start, dd);
      int tyf = dd < Dmax ? pick_ty(who, start, dd + 1) : 0;
      if (tyf) {
        double cells = 0.;
        for (int i : who) {
          const int D = b.H[i].d.nfine - 1 - start[i];
          if (dd + 1 > D) continue;
          const LevD& L = b.H[i].d.L[b.H[i].d.nfine - 2 - dd];
          cells += (double)(L.m - 2 * NH) * (L.n - 2 * NH);
        }
        // taller tiles where those bring the finer level to UP2_TILES
        while (tyf < 16 && cells / ((double)TX * tyf) > UP2_TILES) tyf *= 2;
        if (cells / ((double)TX * tyf) > UP2_TILES) tyf = 0;
      }
      std::vector<Task> ts, ps;
      for (int i : who) {
        const int D = b.H[i].d.nfine - 1 - start[i];
        if (dd > D || paired[i] == dd) continue;
        const int l = b.H[i].d.nfine - 1 - dd;
        if (b.fuse_up && dd < D && l - 1 >= 1 && !(final_top && dd + 1 == D) &&
            (tyf == 4 || tyf == 8 || tyf == 16)) {
          up2_task(i, l - 1, ps, tyf);
          paired[i] = dd + 1;
          continue;
        }
        smooth_task(i, l, 2, (final_top && dd == D) ? 1 : 0, ts, ty);
      }
      if (!ts.empty()) {
        balance(ts, ty);
        bool adds = false;
        for (const Task& t : ts) adds = adds || (t.flags & 4);
        Launch& Ls = add(L_SMOOTH, ts);
        Ls.pred = pred;
        Ls.ty = ty;
        if (adds) Ls.flags |= 16;
      }
      if (!ps.empty()) {
        Launch& Lp = add(L_UP2, ps);
        Lp.pred = pred;
        Lp.ty = tyf;
      }
    }
  }

  // up-leg smoothings of levels lf + 1 and lf of instance i in one task list
  // (k_up2; mode 2 on both, unmasked)
  void up2_task(int i, int lf, std::vector<Task>& out, int ty) {
    HostInst& h = b.H[i];
    const int lc = lf + 1;
    Task t = mk_task(i, lf, 0, 0);
    t.mode = 2;
    t.xin2 = h.buf[lc][h.cur[lc]];
    t.xc = h.buf[lc][h.cur[lc] ^ 1];  // the coarse result, the fine prolongation's source
    t.xcc = h.buf[lc + 1][h.cur[lc + 1]];
    h.cur[lc] ^= 1;
    t.xin = h.buf[lf][h.cur[lf]];
    t.xout = h.buf[lf][h.cur[lf] ^ 1];
    h.cur[lf] ^= 1;
    const LevD& L = lev(i, lf);
    tiles(out, t, NH, L.m - NH, NH, L.n - NH, ty);
  }

  void coarse(const std::vector<int>& who, int mode, int pred) {
    std::vector<Task> ts;
    size_t sh = 0;
    for (int i : who) {
      HostInst& h = b.H[i];
      const int k = h.d.nfine;
      Task t = mk_task(i, k, 0, 0);
      h.cur[k] ^= 1;
      t.xout = h.buf[k][h.cur[k]];
      ts.push_back(t);
      sh = std::max(sh, h.coarse_lds);
    }
    Launch& L = add(L_COARSE, ts);
    L.mode = mode;
    L.pred = pred;
    L.shmem = sh;
    // consecutive instances: the kernel takes instance first + task index
    L.inst0 = who.empty() ? -1 : who[0];
    for (size_t k = 0; k < who.size(); k++)
      if (who[k] != who[0] + (int)k) L.inst0 = -1;
  }

  using Off = std::function<long long(HostInst&)>;

  // pools: 0 T -> T, 1 FP64 -> FP64 (state / work), 2 T -> FP64
  // island: minus the islands' rhs correction on instances that have them
  void c2c(const std::vector<int>& who, Off src, Off dst, int pred, int pools,
           bool island = false) {
    std::vector<Task> ts;
    for (int i : who) {
      HostInst& h = b.H[i];
      Task t = mk_task(i, 0, 0, 0);
      t.xin = src(h);
      t.xout = dst(h);
      t.flags = (island && h.d.island) ? 1 : 0;
      tiles(ts, t, 0, h.d.m - 1, 0, h.d.n - 1);
    }
    Launch& L = add(L_C2C, ts);
    L.pred = pred;
    L.flags = pools;
  }

  // (u, v) at +2F, +3F of base from psi; pools: 0 T (psi at base + F),
  // 1 FP64 (psi at base + F), 2 T output from the FP64 psi plane `psi`
  void ortho(const std::vector<int>& who, Off base, int pools, Off psi = nullptr) {
    std::vector<Task> ts;
    for (int i : who) {
      HostInst& h = b.H[i];
      Task t = mk_task(i, 0, 0, 0);
      t.xout = base(h);
      t.xin = psi ? psi(h) : base(h) + (long long)h.d.m * h.d.n;
      tiles(ts, t, 1, h.d.m - 1, 1, h.d.n - 1);
    }
    add(L_ORTHO, ts).flags = pools;
  }

  // flat element ranges [0, cnt) per instance
  std::vector<Task> flat(const std::vector<int>& who, Off cnt, int field) {
    std::vector<Task> ts;
    for (int i : who) {
      HostInst& h = b.H[i];
      const long long c = cnt(h);
      for (long long q = 0; q < c; q += FLAT) {
        Task t = mk_task(i, field, (int)q, (int)std::min<long long>(FLAT, c - q));
        ts.push_back(t);
      }
    }
    return ts;
  }

  std::vector<int> all() {
    std::vector<int> v(b.ninst);
    for (int i = 0; i < b.ninst; i++) v[i] = i;
    return v;
  }

  // fast inversion of tendency slot k (core/operators.py:421-498, 'fast')
  void invert_fast(const std::vector<int>& who, int k) {
    if (who.empty()) return;
    if (b.fast_maxite > 0) return invert_fast_solve(who, k);
    if (b.prec != 0) return invert_fast_refined(who, k);
    for (int i : who) {
      HostInst& h = b.H[i];
      const size_t F = (size_t)h.d.m * h.d.n;
      h.buf[0][0] = h.d.k[k] + F;  // psi tendency: the guess and the result
      h.buf[0][1] = h.x0alt;
      h.cur[0] = 0;
    }
    c2c(who, [k](HostInst& h) { return h.d.k[k]; }, [](HostInst& h) { return h.d.L[0].b; }, 0,
        false);
    std::vector<int> st(b.ninst, 0);
    vcycle(who, st, 0, 0, false);
    vcycle(who, st, 0, 0, true);
    for (int i : who)
      if (b.H[i].cur[0] != 0) fprintf(stderr, "engine: fast inversion parity\n");
    ortho(who, [k](HostInst& h) { return h.d.k[k]; }, 0);
  }

  // Mixed mode: the same two V-cycles (hierarchy.py:207-218) in correction
  // form.  V-cycles are affine, V(x0, b) = x0 + V(0, b - A x0), so the FP64
  // guess x0 (the previous psi tendency, an FP64 plane) is kept, the FP64
  // defect r0 = b - A x0 is the FP32 V-cycles' right-hand side, and the
  // correction is added in FP64.  u, v tendencies come from the FP64 psi
  // tendency, so their differences do not cancel in FP32.
  void invert_fast_refined(const std::vector<int>& who, int k) {
    std::vector<Task> rs;
    for (int i : who) {
      HostInst& h = b.H[i];
      h.buf[0][0] = h.d.x0[0];
      h.buf[0][1] = h.d.x0[1];
      h.cur[0] = 0;
      rr_tasks(i, 0, h.d.kp[k], h.d.k[k], h.d.L[0].b, rs);
    }
    // FP64 defect b - A x0 with b = celltocorner(omega tendency) formed on
    // the fly, stored (T) as the level-0 rhs and restricted
    add(L_RR, rs).flags = 1 | 32 | 64;
    std::vector<int> st(b.ninst, 0);
    vcycle(who, st, 1, 0, false);
    vcycle(who, st, 0, 0, true);
    std::vector<Task> up;
    for (int i : who) {
      HostInst& h = b.H[i];
      const size_t F = (size_t)h.d.m * h.d.n;
      for (size_t q = 0; q < F; q += FLAT) {
        Task t = mk_task(i, 0, (int)q, (int)std::min<size_t>(FLAT, F - q));
        t.xout = h.d.kp[k];
        t.xin = h.buf[0][h.cur[0]];
        up.push_back(t);
      }
    }
    add(L_UPDATE, up);
    ortho(who, [k](HostInst& h) { return h.d.k[k]; }, 2, [k](HostInst& h) { return h.d.kp[k]; });
  }

  // Defect-correction solve, core/gmg/hierarchy.py:154-192 (solve): the
  // FP64 plane psi(h) (pool P, or the T pool of an FP64 batch when psi_q)
  // against the right-hand side rhs(h): rhs_flag 16 an FP64 plane (pool P),
  // 64 celltocorner of a T plane formed on the fly.  The residual and the
  // correction are FP64; the F-cycles run in the pool type T.  The stopping
  // test (maxite, tol) is evaluated on the device.
  void solve_defect(const std::vector<int>& who, Off psi, bool psi_q, Off rhs, int rhs_flag,
                    int maxite, double tol) {
    for (int i : who) {
      HostInst& h = b.H[i];
      h.buf[0][0] = h.d.x0[0];
      h.buf[0][1] = h.d.x0[1];
      h.cur[0] = 0;
    }
    auto resid_norm = [&](int what, int pred) {
      std::vector<Task> rs;
      for (int i : who) {
        HostInst& h = b.H[i];
        rr_tasks(i, 0, psi(h), rhs(h), h.d.L[0].b, rs);
      }
      const int sp = spans_of(rs);
      Launch& L = add(L_RR, rs);
      L.flags = 1 | 4 | 8 | 32 | rhs_flag | (psi_q ? 128 : 0);  // FP64 outer residual
      L.pred = pred;
      fold(what, sp, pred, maxite, tol);
    };
    resid_norm(0, 0);
    for (int it = 0; it < maxite; it++) {
      // F-cycle: b restricted down the fine levels (b0 -> b1 came with the
      // residual), coarse tail, then V-cycles from each fine level upward.
      // Level l = nfine - 1 - dd of an instance goes to l + 1 at
      // bottom-aligned step dd; steps dd and dd - 1 share a k_restrict_b2
      // launch (one-level launches where the program's mid-level phases
      // are merged into persistent launches)
      int Dmax = 0;
      for (int i : who) Dmax = std::max(Dmax, b.H[i].d.nfine - 1);
      for (int dd = Dmax - 1; dd >= 0; dd -= (b.persist > 0 ? 1 : 2)) {
        if (b.persist > 0) {
          double cells = 0.;
          for (int i : who) {
            const int l = b.H[i].d.nfine - 1 - dd;
            if (l >= 1) cells += (double)(b.H[i].d.L[l].m - 2 * NH) * (b.H[i].d.L[l].n - 2 * NH);
          }
          const int ty = pick_ty_cells(cells);
          std::vector<Task> rs;
          for (int i : who) {
            const int l = b.H[i].d.nfine - 1 - dd;
            if (l >= 1) rr_tasks(i, l, 0, -1, -1, rs, ty / 2);
          }
          if (!rs.empty()) {
            Launch& L = add(L_RR, rs);
            L.flags = 2;
            L.pred = 1;
            L.ty = ty;
          }
          continue;
        }
        // output tiles of 8 rows, or of RB2_TY_SMALL rows where 8 would give
        // at most RB2_TILES tiles (latency-bound launches: more, shorter tiles)
        std::vector<Task> rs;
        int rty = 8;
        for (int ty : {8, RB2_TY_SMALL}) {
          rs.clear();
          rty = ty;
          for (int i : who) {
            const int l = b.H[i].d.nfine - 1 - dd;
            if (dd >= 1 && l >= 1)
              rb_tasks(i, l, 2, rs, ty);
            else if (dd >= 1 && l == 0)
              rb_tasks(i, 1, 1, rs, ty);
            else if (dd == 0 && l >= 1)
              rb_tasks(i, l, 1, rs, ty);
          }
          if (ty == 8 && (int)rs.size() > RB2_TILES) break;
        }
        if (!rs.empty()) {
          Launch& L = add(L_RB2, rs);
          L.pred = 1;
          L.ty = rty;
        }
      }
      coarse(who, 1, 1);
      // x0 += x (hierarchy.py:171) is folded into each instance's last
      // smoothing of the finest level (the up-leg of its V-cycle from level
      // 0, the last of the F-cycle's V-cycles), at every position it stores:
      // a launch less where the solve is latency-bound (at most 4 M finest
      // cells; on larger batches the smoothing's read-modify-write of the
      // plane costs what the separate pass does)
      double fine_cells = 0.;
      for (int i : who) fine_cells += (double)h_cells(i);
      const bool fold_up = fine_cells <= 4194304.;
      upd_plane.assign(b.ninst, -1);
      upd_q = psi_q;
      if (fold_up)
        for (int i : who) upd_plane[i] = psi(b.H[i]);
      for (int dd = 0; dd <= Dmax; dd++) {
        std::vector<int> sub, st(b.ninst, 0);
        for (int i : who)
          if (b.H[i].d.nfine - 1 >= dd) {
            sub.push_back(i);
            st[i] = b.H[i].d.nfine - 1 - dd;
          }
        vcycle(sub, st, 3, 1, false);
      }
      upd_plane.clear();
      if (!fold_up) {
        std::vector<Task> up;
        for (int i : who) {
          HostInst& h = b.H[i];
          const size_t F = (size_t)h.d.m * h.d.n;
          for (size_t q = 0; q < F; q += FLAT) {
            Task t = mk_task(i, 0, (int)q, (int)std::min<size_t>(FLAT, F - q));
            t.xout = psi(h);
            t.xin = h.buf[0][h.cur[0]];
            up.push_back(t);
          }
        }
        Launch& U = add(L_UPDATE, up);
        U.pred = 1;
        U.flags = psi_q ? 1 : 0;
      }
      resid_norm(1, 1);
    }
  }

  // Fast inversion of tendency slot k by gmg.solve (fast_maxite, fast_tol)
  // instead of twoVcycle: core/operators.py:452-455 (commented out in the
  // reference), then psi *= mskp and (u, v) (:483-496).  The psi tendency
  // is FP64: kp[k] (pool P) in mixed precision, the T pool's own plane in
  // an FP64 batch.
  void invert_fast_solve(const std::vector<int>& who, int k) {
    const bool fp64 = b.prec == 0;
    Off psi = [k, fp64](HostInst& h) {
      return fp64 ? h.d.k[k] + (long long)h.d.m * h.d.n : h.d.kp[k];
    };
    solve_defect(who, psi, fp64, [k](HostInst& h) { return h.d.k[k]; }, 64, b.fast_maxite,
                 b.fast_tol);
    std::vector<Task> fin;
    for (int i : who) {
      HostInst& h = b.H[i];
      const long long F = (long long)h.d.m * h.d.n;
      for (long long q = 0; q < F; q += FLAT) {
        Task t = mk_task(i, 1, (int)q, (int)std::min<long long>(FLAT, F - q));
        t.xout = psi(h);
        t.flags = 2 | (fp64 ? 1 : 0);
        fin.push_back(t);
      }
    }
    add(L_FINPSI, fin);
    if (fp64)
      ortho(who, [k](HostInst& h) { return h.d.k[k]; }, 0);
    else
      ortho(who, [k](HostInst& h) { return h.d.k[k]; }, 2, [k](HostInst& h) { return h.d.kp[k]; });
  }

  // full solve of the state's psi (core/operators.py:470-481)
  // NumPy-order sum of an interior plane per instance (k_npsum chunks, then
  // the fold `what`): flags 0 the state's psi (periodic instances), 1 the
  // no-slip source
  void npsum(const std::vector<int>& who, int flags, int what) {
    std::vector<Task> ns;
    for (int i : who) {
      HostInst& h = b.H[i];
      const long long tot = (long long)(h.d.m - 2 * NH) * (h.d.n - 2 * NH);
      const int nch = (int)((tot + 8191) / 8192);
      for (int c = 0; c < nch; c += 4) {  // one wavefront per chunk
        Task t = mk_task(i, 0, c, std::min(4, nch - c));
        t.flags = flags;
        ns.push_back(t);
      }
    }
    const int soff = (int)b.slots.size();
    std::vector<FoldSpan> sp(hi - lo, FoldSpan{0, 0});
    int slot = 0;
    for (const Task& t : ns) {
      b.slots.push_back(slot);
      if (sp[t.inst - lo].count == 0) sp[t.inst - lo].first = slot;
      sp[t.inst - lo].count += t.x0;
      slot += t.x0;
    }
    const int spoff = (int)b.spans.size();
    b.spans.insert(b.spans.end(), sp.begin(), sp.end());
    Launch& L = add(L_NPSUM, ns);
    L.slot_off = soff;
    fold(what, spoff, 0);
  }

  // full solve of the state's psi (core/operators.py:470-481), with the
  // islands' rhs correction and boundary psi where the instance has islands
  void solve_full(const std::vector<int>& who) {
    if (who.empty()) return;
    c2c(who, [](HostInst& h) { return h.d.state; }, [](HostInst& h) { return h.d.work; }, 0, 1,
        true);
    solve_defect(who, [](HostInst& h) { return h.d.state + (long long)h.d.m * h.d.n; }, false,
                 [](HostInst& h) { return h.d.work; }, 16, b.maxite, b.tol);
    // zero mean (perio: the mean's sum and fold only there) and mask, then
    // (u, v)
    std::vector<int> per;
    for (int i : who)
      if (b.H[i].d.perio) per.push_back(i);
    if (!per.empty()) npsum(per, 0, 3);
    add(L_FINPSI, flat(who, [](HostInst& h) { return (long long)h.d.m * h.d.n; }, 1));
    ortho(who, [](HostInst& h) { return h.d.state; }, 1);
  }

  // k_flat_ops tasks over the m x n plane of instance i
  void flat_ops(std::vector<Task>& out, int i, int flags, long long xin, long long xout,
                int mode) {
    const HostInst& h = b.H[i];
    const long long F = (long long)h.d.m * h.d.n;
    for (long long q = 0; q < F; q += FLAT) {
      Task t = mk_task(i, 0, (int)q, (int)std::min<long long>(FLAT, F - q));
      t.flags = flags;
      t.xin = xin;
      t.xout = xout;
      t.mode = mode;
      out.push_back(t);
    }
  }

  // Fast inversion of the state (core/operators.py:421-498, flag 'fast':
  // twoVcycle on the state's psi, psi * mskp, then (u, v)); island: with the
  // islands' rhs correction and boundary psi where the instance has them
  // (no-slip: core/euler.py:171-179; the loop's momentum enforcement runs
  // it without islands).  An FP64 batch runs the two V-cycles on T-pool
  // copies of psi and the rhs (bit-identical to the reference); a mixed
  // batch in the correction form of invert_fast_refined: FP64 defect, FP32
  // V-cycles, FP64 correction.
  void invert_state_fast(const std::vector<int>& who, bool island) {
    if (who.empty()) return;
    c2c(who, [](HostInst& h) { return h.d.state; }, [](HostInst& h) { return h.d.work; }, 0, 1,
        island);
    std::vector<int> st(b.ninst, 0);
    if (b.prec == 0) {
      std::vector<Task> cp;
      for (int i : who) {
        HostInst& h = b.H[i];
        const long long F = (long long)h.d.m * h.d.n;
        h.buf[0][0] = h.x0alt;
        h.buf[0][1] = h.d.x0[0];
        h.cur[0] = 0;
        flat_ops(cp, i, 4, h.d.state + F, h.x0alt, 0);
        flat_ops(cp, i, 4, h.d.work, h.d.L[0].b, 0);
      }
      add(L_FLATOPS, cp);
      vcycle(who, st, 0, 0, false);
      vcycle(who, st, 0, 0, true);  // the last smoothing stores psi * mskp
      std::vector<Task> back;
      for (int i : who) {
        HostInst& h = b.H[i];
        const long long F = (long long)h.d.m * h.d.n;
        flat_ops(back, i, 8, h.buf[0][h.cur[0]], h.d.state + F, (island && h.d.island) ? 1 : 0);
      }
      add(L_FLATOPS, back);
    } else {
      std::vector<Task> rs;
      for (int i : who) {
        HostInst& h = b.H[i];
        h.buf[0][0] = h.d.x0[0];
        h.buf[0][1] = h.d.x0[1];
        h.cur[0] = 0;
        rr_tasks(i, 0, h.d.state + (long long)h.d.m * h.d.n, h.d.work, h.d.L[0].b, rs);
      }
      add(L_RR, rs).flags = 1 | 16 | 32;  // FP64 defect, stored (T) and restricted
      vcycle(who, st, 1, 0, false);
      vcycle(who, st, 0, 0, false);
      std::vector<Task> up, fin;
      for (int i : who) {
        HostInst& h = b.H[i];
        const long long F = (long long)h.d.m * h.d.n;
        for (long long q = 0; q < F; q += FLAT) {
          Task t = mk_task(i, 1, (int)q, (int)std::min<long long>(FLAT, F - q));
          t.xout = h.d.state + F;
          t.xin = h.buf[0][h.cur[0]];
          up.push_back(t);
          t.flags = 2 | ((island && h.d.island) ? 4 : 0);
          fin.push_back(t);
        }
      }
      add(L_UPDATE, up);
      add(L_FINPSI, fin);
    }
    ortho(who, [](HostInst& h) { return h.d.state; }, 1);
  }

  // No-slip walls after the time step (core/euler.py:108-112,
  // core/boussinesq.py:92-94; Operators.rhs_noslip, core/operators.py:
  // 245-290): the wall source, minus its mean over the boundary cells (and,
  // with enforce_momentum, its momentum), filled and subtracted from the
  // vorticity, then the fast inversion with islands; Euler keeps source/dt.
  void noslip(const std::vector<int>& who) {
    if (who.empty()) return;
    std::vector<Task> ts, en;
    for (int i : who) {
      HostInst& h = b.H[i];
      tiles(ts, mk_task(i, 0, 0, 0), NH, h.d.m - NH, NH, h.d.n - NH);
      if (h.d.enforce) tiles(en, mk_task(i, 0, 0, 0), NH, h.d.m - NH, NH, h.d.n - NH);
    }
    add(L_NSSRC, ts);
    npsum(who, 1, 4);
    if (!en.empty()) {
      const int sp = spans_of(en);
      add(L_NSDOT, en);
      fold(5, sp, 0);
    }
    add(L_NSFIX, ts);
    invert_state_fast(who, true);
    std::vector<Task> dv;
    for (int i : who)
      if (b.H[i].d.model == 0) flat_ops(dv, i, 1, 0, 0, 0);
    if (!dv.empty()) add(L_FLATOPS, dv);
  }

  // The loop's momentum enforcement of instance i (core/fluid2d.py:352-358):
  // vorticity -= xr0 px + yr0 py (px, py host-written), fast inversion.
  void momentum(int i) {
    std::vector<Task> mt;
    flat_ops(mt, i, 2, 0, 0, 0);
    add(L_FLATOPS, mt);
    invert_state_fast({i}, false);
  }

  void diag() {
    std::vector<Task> ts;
    for (int i = lo; i < hi; i++) {
      HostInst& h = b.H[i];
      tiles(ts, mk_task(i, 0, 0, 0), NH, h.d.m - NH, NH, h.d.n - NH);
    }
    const int sp = spans_of(ts);
    add(L_DIAG, ts);
    fold(2, sp, 0);
  }

  void rk_step() {
    std::vector<int> who = all(), eul;
    const bool ssp = b.H[0].d.tscheme == 0;
    const int kforcing = ssp ? 2 : 0;
    for (int stage = 0; stage < 3; stage++) {
      // one advection launch per model (tracer count)
      for (int model = 0; model < 2; model++) {
        std::vector<Task> ts;
        for (int i : who) {
          HostInst& h = b.H[i];
          if (h.d.model == model) tiles(ts, mk_task(i, 0, 0, 0), NH, h.d.m - NH, NH, h.d.n - NH);
        }
        if (ts.empty()) continue;
        Launch& L = add(L_RHS, ts);
        L.stage = stage;
        L.outk = stage;
        L.flags = model + 1;
      }
      // Euler inverts except at stage kforcing (unless diffusion);
      // Boussinesq inverts at every stage (core/euler.py:156-169,
      // core/boussinesq.py:114)
      std::vector<int> inv;
      for (int i : who) {
        const InstD& d = b.H[i].d;
        if (d.model == 1 || stage != kforcing || d.diffusion) inv.push_back(i);
      }
      invert_fast(inv, stage);
    }
    // final combination: Euler omega, psi; Boussinesq all fields (+ banom)
    std::vector<Task> cs;
    for (int i : who) {
      HostInst& h = b.H[i];
      const int nf = h.d.model == 0 ? 2 : 5;
      for (int f = 0; f < nf; f++) {
        auto t = flat({i}, [](HostInst& hh) { return (long long)hh.d.m * hh.d.n; }, f);
        cs.insert(cs.end(), t.begin(), t.end());
      }
      if (h.d.model == 0) eul.push_back(i);
    }
    add(L_COMBINE, cs);
    std::vector<int> ns;
    for (int i : who)
      if (b.H[i].d.noslip) ns.push_back(i);
    noslip(ns);
    solve_full(eul);
    diag();
  }
};

// ---- execution ------------------------------------------------------------
// Optional timing of one launch kind (eager runs only): every matching
// launch is issued PROF_REP times back to back between one pair of events
// on the batch stream, so the events' own latency is spread over the
// repeats.  The launches read buffers they do not write (smoother
// ping-pong buffers, residual into its own buffer, tendencies from stage
// states), so a repeat leaves the state as one launch does -- except a
// smoothing that also adds its result into the solution (launch flag 16,
// x0 += x): it is never timed, since a repeat would add it again.
constexpr int PROF_REP = 8;
struct ProfSel {
  int kind = -1, level = -1;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  std::vector<const Launch*> which;
};

#define F2D_LAUNCH(K, G, B, S, STREAM, ...) hipLaunchKernelGGL(K, G, B, S, STREAM, __VA_ARGS__)

template <typename T, typename R>
static void launch_rr(int ty, unsigned g, hipStream_t s, const Task* TK, const InstD* dI,
                      const R* RP, T* Q, const int8_t* B, const DynD* dD, int pred, int flags,
                      double* partial) {
  const dim3 B2(64, 4);
  if (ty == 4)
    F2D_LAUNCH((k_resid_restrict<T, R, 2>), dim3(g), B2, 0, s, TK, dI, RP, Q, B, dD, pred,
                       flags, partial);
  else if (ty == 8)
    F2D_LAUNCH((k_resid_restrict<T, R, 4>), dim3(g), B2, 0, s, TK, dI, RP, Q, B, dD, pred,
                       flags, partial);
  else
    F2D_LAUNCH((k_resid_restrict<T, R, TY2>), dim3(g), B2, 0, s, TK, dI, RP, Q, B, dD,
                       pred, flags, partial);
}

constexpr unsigned RR_STAGE_TILES = 512;

template <typename T>
static int run_prog_t(Batch& b, const std::vector<Launch>& prog, ProfSel* prof = nullptr) {
  const dim3 B2(64, 4), B1(256);
  T* Q = (T*)b.Q;
  double* P = b.P;
  // the T pool as FP64 planes (FP64 batches only; null in mixed precision)
  double* PQ = std::is_same<T, double>::value ? (double*)b.Q : nullptr;
  for (const Launch& L : prog) {
    const Task* TK = b.dT + L.toff;
    const unsigned g = (unsigned)L.ntask;
    if (L.kind != L_FOLD && L.kind != L_GATE && g == 0) continue;
    bool timed = false;
    if (prof && L.kind == prof->kind && !(L.kind == L_SMOOTH && (L.flags & 16)) &&
        (prof->level < 0 || (L.ntask > 0 && b.tasks[L.toff].lev == prof->level))) {
      hipEvent_t e0, e1;
      ECHK(hipEventCreate(&e0));
      ECHK(hipEventCreate(&e1));
      ECHK(hipEventRecord(e0, b.s));
      prof->ev.push_back({e0, e1});
      prof->which.push_back(&L);
      timed = true;
    }
    const int nrep = timed ? PROF_REP : 1;
    for (int rep = 0; rep < nrep; rep++) switch (L.kind) {
      case L_RHS:
        if (L.flags == 2)
          F2D_LAUNCH((k_rhs<T, 2>), dim3(g), B2, 0, b.s, TK, b.dI, P, Q, b.B, b.dD, L.stage,
                     L.outk);
        else
          F2D_LAUNCH((k_rhs<T, 1>), dim3(g), B2, 0, b.s, TK, b.dI, P, Q, b.B, b.dD, L.stage,
                     L.outk);
        break;
      case L_C2C:
        if (L.flags == 1)
          F2D_LAUNCH((k_c2c<double, double>), dim3(g), B2, 0, b.s, TK, b.dI, P, P, b.dD,
                             L.pred);
        else if (L.flags == 2)
          F2D_LAUNCH((k_c2c<T, double>), dim3(g), B2, 0, b.s, TK, b.dI, Q, P, b.dD,
                             L.pred);
        else
          F2D_LAUNCH((k_c2c<T, T>), dim3(g), B2, 0, b.s, TK, b.dI, Q, Q, b.dD, L.pred);
        break;
      case L_SMOOTH:
        // launch flag 16: tasks that add their result to a plane (k_smooth ADD)
        if ((L.flags & 16) && L.ty == 4)
          F2D_LAUNCH((k_smooth<T, 4, true>), dim3(g), B2, 0, b.s, TK, b.dI, Q, P, b.B, b.dD,
                     L.pred);
        else if ((L.flags & 16) && L.ty == 8)
          F2D_LAUNCH((k_smooth<T, 8, true>), dim3(g), B2, 0, b.s, TK, b.dI, Q, P, b.B, b.dD,
                     L.pred);
        else if (L.flags & 16)
          F2D_LAUNCH((k_smooth<T, TY, true>), dim3(g), B2, 0, b.s, TK, b.dI, Q, P, b.B, b.dD,
                     L.pred);
        else if (L.ty == 4)
          F2D_LAUNCH((k_smooth<T, 4>), dim3(g), B2, 0, b.s, TK, b.dI, Q, P, b.B, b.dD, L.pred);
        else if (L.ty == 8)
          F2D_LAUNCH((k_smooth<T, 8>), dim3(g), B2, 0, b.s, TK, b.dI, Q, P, b.B, b.dD, L.pred);
        else
          F2D_LAUNCH((k_smooth<T, TY>), dim3(g), B2, 0, b.s, TK, b.dI, Q, P, b.B, b.dD,
                             L.pred);
        break;
      case L_RR:
        if (L.flags & 32)
          launch_rr<T, double>(L.ty, g, b.s, TK, b.dI, P, Q, b.B, b.dD, L.pred, L.flags, b.partial);
        else
          launch_rr<T, T>(L.ty, g, b.s, TK, b.dI, (const T*)Q, Q, b.B, b.dD, L.pred, L.flags,
                          b.partial);
        break;
      case L_COARSE:
        F2D_LAUNCH(k_coarse<T>, dim3(g), dim3(COARSE_THREADS), L.shmem, b.s, TK, b.dI, Q, b.B,
                           b.dD, L.pred, L.mode, b.coarse_wave_cells, L.inst0);
        break;
      case L_ORTHO:
        if (L.flags == 1)
          F2D_LAUNCH((k_ortho<double, double>), dim3(g), B2, 0, b.s, TK, b.dI, P, P, b.B,
                             b.dD);
        else if (L.flags == 2)
          F2D_LAUNCH((k_ortho<T, double>), dim3(g), B2, 0, b.s, TK, b.dI, Q, P, b.B,
                             b.dD);
        else
          F2D_LAUNCH((k_ortho<T, T>), dim3(g), B2, 0, b.s, TK, b.dI, Q, Q, b.B, b.dD);
        break;
      case L_COMBINE:
        F2D_LAUNCH(k_combine<T>, dim3(g), B1, 0, b.s, TK, b.dI, P, Q, b.dD);
        break;
      case L_UPDATE:
        // flag 1: the corrected plane is in the T pool (FP64 batches only)
        F2D_LAUNCH(k_update<T>, dim3(g), B1, 0, b.s, TK, b.dI, (L.flags & 1) ? PQ : P, Q, b.dD,
                   L.pred);
        break;
      case L_FINPSI:
        F2D_LAUNCH(k_finalize_psi, dim3(g), B1, 0, b.s, TK, b.dI, P, PQ, b.B, b.dD);
        break;
      case L_NPSUM:
        F2D_LAUNCH(k_npsum, dim3(g), B1, 0, b.s, TK, b.dI, P, b.partial,
                           b.dSlot + L.slot_off, b.dD);
        break;
      case L_DIAG:
        F2D_LAUNCH(k_diag, dim3(g), B2, 0, b.s, TK, b.dI, P, b.B, b.partial, b.dD);
        break;
      case L_GATE:
        F2D_LAUNCH(k_gate, dim3((b.ninst + 255) / 256), B1, 0, b.s, b.dD, b.ninst);
        break;
      case L_SMOOTH_RR: {
        // launches of at most RR_STAGE_TILES workgroups (two per CU or less)
        // run 1024-thread workgroups and stage the operator in LDS
        const bool st = g <= RR_STAGE_TILES;
        const dim3 B16(64, 16);
        if (L.mode == 3) {  // x0 = I(xc): tiles of at most 4 coarse rows
          if (L.ty == 2)
            F2D_LAUNCH((k_smooth_rr<T, 1, 3, 1, 1024>), dim3(g), B16, 0, b.s, TK, b.dI, Q, b.B,
                       b.dD, L.pred);
          else if (L.ty == 4 && st)
            F2D_LAUNCH((k_smooth_rr<T, 2, 3, 1, 1024>), dim3(g), B16, 0, b.s, TK, b.dI, Q, b.B,
                       b.dD, L.pred);
          else if (L.ty == 4)
            F2D_LAUNCH((k_smooth_rr<T, 2, 3>), dim3(g), B2, 0, b.s, TK, b.dI, Q, b.B, b.dD,
                       L.pred);
          else if (st)
            F2D_LAUNCH((k_smooth_rr<T, 4, 3, 1, 1024>), dim3(g), B16, 0, b.s, TK, b.dI, Q, b.B,
                       b.dD, L.pred);
          else
            F2D_LAUNCH((k_smooth_rr<T, 4, 3>), dim3(g), B2, 0, b.s, TK, b.dI, Q, b.B, b.dD,
                       L.pred);
        } else if (L.ty == 2) {
          F2D_LAUNCH((k_smooth_rr<T, 1, 1, 1, 1024>), dim3(g), B16, 0, b.s, TK, b.dI, Q, b.B,
                     b.dD, L.pred);
        } else if (L.ty == 4) {
          if (st)
            F2D_LAUNCH((k_smooth_rr<T, 2, 1, 1, 1024>), dim3(g), B16, 0, b.s, TK, b.dI, Q, b.B,
                       b.dD, L.pred);
          else
            F2D_LAUNCH((k_smooth_rr<T, 2, 1>), dim3(g), B2, 0, b.s, TK, b.dI, Q, b.B, b.dD,
                       L.pred);
        } else if (L.ty == 8) {
          if (st)
            F2D_LAUNCH((k_smooth_rr<T, 4, 1, 1, 1024>), dim3(g), B16, 0, b.s, TK, b.dI, Q, b.B,
                       b.dD, L.pred);
          else
            F2D_LAUNCH((k_smooth_rr<T, 4, 1>), dim3(g), B2, 0, b.s, TK, b.dI, Q, b.B, b.dD,
                       L.pred);
        } else {
          if (st)
            F2D_LAUNCH((k_smooth_rr<T, TY2, 1, 1, 1024>), dim3(g), B16, 0, b.s, TK, b.dI, Q, b.B,
                       b.dD, L.pred);
          else
            F2D_LAUNCH((k_smooth_rr<T, TY2, 1>), dim3(g), B2, 0, b.s, TK, b.dI, Q, b.B, b.dD,
                       L.pred);
        }
        break;
      }
      case L_PERSIST: {
        const PPhase* PH = b.dPh + L.ph_off;
        BarSlot* BS = b.bar + L.bar;
        if (L.ty == 4)
          F2D_LAUNCH((k_persist<T, 4>), dim3(g), B2, 0, b.s, PH, L.nph, b.dT, b.dI, Q, b.B, b.dD,
                     BS, b.fault);
        else if (L.ty == 8)
          F2D_LAUNCH((k_persist<T, 8>), dim3(g), B2, 0, b.s, PH, L.nph, b.dT, b.dI, Q, b.B, b.dD,
                     BS, b.fault);
        else
          F2D_LAUNCH((k_persist<T, TY>), dim3(g), B2, 0, b.s, PH, L.nph, b.dT, b.dI, Q, b.B,
                     b.dD, BS, b.fault);
        break;
      }
      case L_RB2:
        if (L.ty == 8)
          F2D_LAUNCH((k_restrict_b2<T, 8, 1024>), dim3(g), dim3(1024), 0, b.s, TK, b.dI, Q, b.B,
                     b.dD, L.pred);
        else
          F2D_LAUNCH((k_restrict_b2<T, RB2_TY_SMALL, 1024>), dim3(g), dim3(1024), 0, b.s, TK,
                     b.dI, Q, b.B, b.dD, L.pred);
        break;
      case L_UP2:
        if (L.ty == 4)
          F2D_LAUNCH((k_up2<T, 4>), dim3(g), dim3(64, 16), 0, b.s, TK, b.dI, Q, b.B, b.dD, L.pred);
        else if (L.ty == 8)
          F2D_LAUNCH((k_up2<T, 8>), dim3(g), dim3(64, 16), 0, b.s, TK, b.dI, Q, b.B, b.dD, L.pred);
        else
          F2D_LAUNCH((k_up2<T, 16>), dim3(g), dim3(64, 16), 0, b.s, TK, b.dI, Q, b.B, b.dD, L.pred);
        break;
      case L_NSSRC:
        F2D_LAUNCH(k_ns_source, dim3(g), B2, 0, b.s, TK, b.dI, P, b.dD);
        break;
      case L_NSDOT:
        F2D_LAUNCH(k_ns_dot, dim3(g), B2, 0, b.s, TK, b.dI, b.B, b.partial, b.dD);
        break;
      case L_NSFIX:
        F2D_LAUNCH(k_ns_fix, dim3(g), B2, 0, b.s, TK, b.dI, P, b.dD);
        break;
      case L_FLATOPS:
        F2D_LAUNCH(k_flat_ops, dim3(g), B1, 0, b.s, TK, b.dI, P, PQ, b.dD);
        break;
      case L_FOLD:
        F2D_LAUNCH(k_fold, dim3(L.ninst), B1, 0, b.s, b.dS + L.span_off, b.dI, b.dD,
                           b.partial, L.mode, L.maxite < 0 ? b.maxite : L.maxite,
                           L.maxite < 0 ? b.tol : L.tol, L.pred, b.adapt, b.cfl, b.dtmax,
                           L.inst0);
        break;
    }
    if (timed) ECHK(hipEventRecord(prof->ev.back().second, b.s));
  }
  ECHK(hipGetLastError());
  return F2D_OK;
}

static int run_prog(Batch& b, const std::vector<Launch>& prog, ProfSel* prof = nullptr) {
  return b.prec == 0 ? run_prog_t<double>(b, prog, prof) : run_prog_t<float>(b, prog, prof);
}

// Bytes a launch moves: algorithmic (each operand read once, results
// written once, operator matrix-free, masks ignored) and as implemented
// (operator planes or stencil bytes and masks included).
static void launch_bytes(const Batch& b, const Launch& L, double* alg, double* impl) {
  const double tz = b.prec == 0 ? 8. : 4.;
  double a = 0., im = 0.;
  if (L.kind == L_PERSIST) {  // the sum over its phases
    for (int p = 0; p < L.nph; p++) {
      const PPhase& f = b.phases[L.ph_off + p];
      Launch q;
      q.kind = f.kind == 0 ? L_SMOOTH : L_RR;
      q.toff = f.toff;
      q.ntask = f.ntask;
      q.flags = f.flags;
      q.ty = L.ty;
      double qa = 0., qi = 0.;
      launch_bytes(b, q, &qa, &qi);
      a += qa;
      im += qi;
    }
    *alg = a;
    *impl = im;
    return;
  }
  for (int k = 0; k < L.ntask; k++) {
    const Task& t = b.tasks[L.toff + k];
    const InstD& d = b.H[t.inst].d;
    const LevD& lv = d.L[t.lev];
    if (L.kind == L_SMOOTH) {
      const int cj = std::min(L.ty, lv.m - NH - t.y0), ci = std::min(TX, lv.n - NH - t.x0);
      const double cells = (double)std::max(cj, 0) * std::max(ci, 0);
      const bool rd_x = t.mode == 0 || t.mode == 2;
      const double wrd = (rd_x ? 1. : 0.) + 1.;  // x (if read), b
      a += cells * (wrd + 1.) * tz;               // + x written
      const bool fly = t.lev == 0 && d.a0fly;
      im += cells * ((wrd + 1.) * tz + 1. + (fly ? 1. : 5. * tz) + (t.mode >= 2 ? 0.25 * tz : 0.));
    } else if (L.kind == L_UP2) {
      // two smoothings in mode 2: x and b read, x written, on the tile's
      // fine cells and on the coarse cells it lies over (+ the next coarser
      // x under those)
      const int cj = std::min(L.ty, lv.m - NH - t.y0), ci = std::min(TX, lv.n - NH - t.x0);
      const double cells = (double)std::max(cj, 0) * std::max(ci, 0);
      const double cc = 0.25 * cells;
      a += (cells + cc) * 3. * tz + 0.25 * cc * tz;
      im += (cells + cc) * (3. * tz + 1. + 5. * tz) * 1.3 + 0.25 * cc * (tz + 1.);
    } else if (L.kind == L_SMOOTH_RR) {
      // rhs read and smoothed x written on the tile's fine cells, the
      // restricted residual written on its coarse cells
      const LevD& lc = d.L[t.lev + 1];
      const int cj = std::min(L.ty / 2, lc.m - NH - t.y0), ci = std::min(TX2, lc.n - NH - t.x0);
      const double cc = (double)std::max(cj, 0) * std::max(ci, 0);
      a += 4. * cc * 2. * tz + cc * tz;
      im += 4. * cc * (2. * tz + 1. + 5. * tz) * 1.3 + cc * (tz + 1.);
    } else if (L.kind == L_RR) {
      const LevD& lc = d.L[t.lev + 1];
      const int cj = std::min(L.ty / 2, lc.m - NH - t.y0), ci = std::min(TX2, lc.n - NH - t.x0);
      const double cc = (double)std::max(cj, 0) * std::max(ci, 0);
      const double fine = 4. * cc;
      const double rz = (L.flags & 32) ? 8. : tz;
      const double rd = (L.flags & 2) ? 1. : 2.;  // b (and x)
      a += fine * rd * rz + cc * tz + ((L.flags & 1) ? fine * tz : 0.);
      const bool fly = t.lev == 0 && d.a0fly;
      im += fine * (rd * rz + 1. + ((L.flags & 2) ? 0. : (fly ? 1. : 5. * rz))) + cc * (tz + 1.) +
            ((L.flags & 1) ? fine * tz : 0.);
    } else if (L.kind == L_RHS) {
      // stage state (x + c k: 1, 2 or 3 reads per field) of the tracer(s)
      // and (u, v), one tendency written per tracer
      const int cj = std::min(TY, d.m - NH - t.y0), ci = std::min(TX, d.n - NH - t.x0);
      const double cells = (double)std::max(cj, 0) * std::max(ci, 0);
      const int ntr = d.model == 0 ? 1 : 2;
      const int kr = L.stage == 0 ? 0 : ((d.tscheme == 0 && L.stage == 2) ? 2 : 1);
      a += cells * ((ntr + 2) * (8. + kr * tz) + ntr * tz);
      im += cells * ((ntr + 2) * (8. + kr * tz) + ntr * tz + 1.);
    } else if (L.kind == L_COARSE) {
      // rhs in, solution out at the first coarse level; the rest stays in LDS
      const double cells = (double)lv.m * lv.n;
      a += cells * 2. * tz;
      double opb = 0.;
      for (int l = t.lev; l < d.nlev; l++) opb += (double)d.L[l].m * d.L[l].n * (5. * tz + 1.);
      im += cells * 2. * tz + opb;
    }
  }
  *alg = a;
  *impl = im;
}

}  // namespace eng
}  // namespace f2d
namespace f2d {
namespace eng {

// Destroys the step graph once the stream has drained: an executable graph
// must not be destroyed while a launch of it may still be in flight.
static void drop_graph(Batch& b) {
  if (!b.gexec && !b.graph) return;
  if (b.s) (void)hipStreamSynchronize(b.s);
  if (b.gexec) (void)hipGraphExecDestroy(b.gexec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  b.gexec = nullptr;
  b.graph = nullptr;
}

// After the batch's stream has drained: did a persistent launch's barrier
// give up?  (The flag is cleared; the results of that step are invalid.)
static int check_fault(Batch& b) {
  if (b.nbar == 0 || !b.fault) return F2D_OK;
  int f = 0;
  ECHK(hipMemcpy(&f, b.fault, sizeof(int), hipMemcpyDeviceToHost));
  if (!f) return F2D_OK;
  ECHK(hipMemset(b.fault, 0, sizeof(int)));
  return efail(F2D_E_HIP, "a persistent multigrid barrier gave up waiting: not all of the "
               "launch's workgroups could be resident (another persistent batch running on "
               "the device?); the results since the last check are invalid");
}

static void batch_free(Batch* b) {
  if (!b) return;
  // everything enqueued on the batch's stream reads the buffers freed below
  if (b->s) (void)hipStreamSynchronize(b->s);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->P) (void)hipFree(b->P);
  if (b->Q) (void)hipFree(b->Q);
  if (b->B) (void)hipFree(b->B);
  if (b->dI) (void)hipFree(b->dI);
  if (b->dD) (void)hipFree(b->dD);
  if (b->hD) (void)hipHostFree(b->hD);
  if (b->dT) (void)hipFree(b->dT);
  if (b->dS) (void)hipFree(b->dS);
  if (b->dSlot) (void)hipFree(b->dSlot);
  if (b->partial) (void)hipFree(b->partial);
  if (b->dPh) (void)hipFree(b->dPh);
  if (b->bar) (void)hipFree(b->bar);
  if (b->fault) (void)hipFree(b->fault);
  for (void* p : b->extra) (void)hipFree(p);
  if (b->s) (void)hipStreamDestroy(b->s);
  delete b;
}

// Workgroups of one persistent launch of tile height ty: at most what can be
// resident at once (occupancy of the kernel times the compute units), at
// most `persist` per compute unit, at most the largest phase's tile count.
template <typename T>
static int persist_grid(int ty, int per_cu, int max_tasks, int* grid) {
  int dev = 0, ncu = 0, occ = 0;
  ECHK(hipGetDevice(&dev));
  ECHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const void* fn = ty == 4 ? (const void*)k_persist<T, 4>
                           : (ty == 8 ? (const void*)k_persist<T, 8> : (const void*)k_persist<T, TY>);
  ECHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 256, 0));
  if (occ < 1) return efail(F2D_E_HIP, "persistent kernel cannot be resident");
  *grid = std::max(1, std::min({max_tasks, std::min(occ, per_cu) * ncu}));
  return F2D_OK;
}

// Runs of consecutive smoother and residual/restriction launches (the
// residual only in its plain forms: restrict r or b, no norms, no stored
// residual) whose tasks all lie below the finest level and share a tile
// height become one persistent launch (k_persist), phases in order.
static int merge_persistent(Batch& b, std::vector<Launch>& prog) {
  if (b.persist <= 0) return F2D_OK;
  auto mergeable = [&](const Launch& L) {
    if (L.ntask <= 0) return false;
    if (L.kind == L_RR) {
      if ((L.flags & ~2) != 0) return false;
    } else if (L.kind != L_SMOOTH) {
      return false;
    }
    for (int k = 0; k < L.ntask; k++)
      if (b.tasks[L.toff + k].lev == 0) return false;
    return true;
  };
  std::vector<Launch> out;
  size_t i = 0;
  while (i < prog.size()) {
    size_t j = i;
    if (mergeable(prog[i]))
      while (j < prog.size() && mergeable(prog[j]) && prog[j].ty == prog[i].ty) j++;
    if (j - i < 2) {
      out.push_back(prog[i]);
      i++;
      continue;
    }
    Launch P;
    P.kind = L_PERSIST;
    P.ty = prog[i].ty;
    P.ph_off = (int)b.phases.size();
    P.nph = (int)(j - i);
    P.bar = b.nbar++;
    int maxt = 0;
    for (size_t k = i; k < j; k++) {
      const Launch& L = prog[k];
      b.phases.push_back(PPhase{L.kind == L_SMOOTH ? 0 : 1, L.toff, L.ntask, L.pred, L.flags});
      maxt = std::max(maxt, L.ntask);
    }
    int g = 0;
    ETRY(b.prec == 0 ? persist_grid<double>(P.ty, b.persist, maxt, &g)
                     : persist_grid<float>(P.ty, b.persist, maxt, &g));
    P.ntask = g;  // the grid
    out.push_back(P);
    i = j;
  }
  prog.swap(out);
  return F2D_OK;
}

static int build_programs(Batch& b) {
  b.tasks.clear();
  b.spans.clear();
  b.slots.clear();
  b.prog_init.clear();
  b.prog_step.clear();
  Builder bd(b);
  bd.lo = 0;
  bd.hi = b.ninst;
  bd.prog = &b.prog_init;
  // initial inversion of every instance's state + diagnostics.  Euler and
  // Boussinesq both start from set_psi_from_vorticity (a full solve).
  bd.solve_full(bd.all());
  bd.diag();
  // diagnostics of the current states alone (after a host write)
  b.prog_diag.clear();
  bd.prog = &b.prog_diag;
  bd.diag();
  bd.prog = &b.prog_step;
  {
    Launch g;
    g.kind = L_GATE;
    b.prog_step.push_back(g);
  }
  bd.rk_step();
  // one instance's initial inversion and diagnostics, per instance
  b.prog_one.assign(b.ninst, {});
  b.prog_mom.assign(b.ninst, {});
  for (int i = 0; i < b.ninst; i++) {
    bd.lo = i;
    bd.hi = i + 1;
    bd.prog = &b.prog_one[i];
    bd.solve_full({i});
    bd.diag();
    if (b.H[i].d.momentum) {
      bd.prog = &b.prog_mom[i];
      bd.momentum(i);
      bd.diag();
    }
  }
  bd.lo = 0;
  bd.hi = b.ninst;
  // persistent launches over runs of mid-level multigrid phases
  b.phases.clear();
  b.nbar = 0;
  ETRY(merge_persistent(b, b.prog_init));
  ETRY(merge_persistent(b, b.prog_step));
  ETRY(merge_persistent(b, b.prog_diag));
  for (auto& po : b.prog_one) ETRY(merge_persistent(b, po));
  for (auto& po : b.prog_mom) ETRY(merge_persistent(b, po));
  size_t np = 0;
  std::vector<const std::vector<Launch>*> progs = {&b.prog_init, &b.prog_step, &b.prog_diag};
  for (const auto& po : b.prog_one) progs.push_back(&po);
  for (const auto& po : b.prog_mom) progs.push_back(&po);
  for (auto* pr : progs)
    for (const Launch& L : *pr)
      if (L.kind == L_RR || L.kind == L_DIAG || L.kind == L_NSDOT)
        np = std::max(np, (size_t)L.ntask * 8);
  np = std::max(np, (size_t)b.slots.size() * 256 + 256);
  if (b.partial) (void)hipFree(b.partial);
  if (b.dT) (void)hipFree(b.dT);
  if (b.dS) (void)hipFree(b.dS);
  if (b.dSlot) (void)hipFree(b.dSlot);
  ECHK(hipMalloc(&b.partial, np * sizeof(double)));
  ECHK(hipMalloc(&b.dT, std::max<size_t>(1, b.tasks.size()) * sizeof(Task)));
  ECHK(hipMalloc(&b.dS, std::max<size_t>(1, b.spans.size()) * sizeof(FoldSpan)));
  ECHK(hipMalloc(&b.dSlot, std::max<size_t>(1, b.slots.size()) * sizeof(int)));
  for (Task& t : b.tasks) {
    const InstD& d = b.H[t.inst].d;
    const LevD& L = d.L[t.lev];
    t.m = L.m;
    t.n = L.n;
    t.b = L.b;
    t.A = L.A;
    t.msk = L.msk;
    const bool cl = t.lev + 1 < d.nlev;
    t.m2 = cl ? d.L[t.lev + 1].m : 0;
    t.n2 = cl ? d.L[t.lev + 1].n : 0;
    t.msk2 = cl ? d.L[t.lev + 1].msk : 0;
    t.sb = (t.lev == 0 && d.a0fly) ? d.sb0 : -1;
    const bool cl2 = t.lev + 2 < d.nlev;
    t.b2 = cl ? d.L[t.lev + 1].b : 0;
    t.A2 = cl ? d.L[t.lev + 1].A : 0;
    t.msk3 = cl2 ? d.L[t.lev + 2].msk : 0;
    t.m3 = cl2 ? d.L[t.lev + 2].m : 0;
    t.n3 = cl2 ? d.L[t.lev + 2].n : 0;
  }
  // the device copy: each launch's task list in dispatch order (workgroup k
  // of a tile launch takes the builder's task xcd_task(k, n); the coarse
  // kernel's workgroup k takes task k, its instance first + k), each task's
  // slot its index in the builder's list.  Persistent phases keep the
  // builder's order (the persistent kernel walks them itself).
  std::vector<Task> dev(b.tasks);
  for (size_t k = 0; k < dev.size(); k++) dev[k].slot = 0;
  std::vector<char> placed(b.tasks.size(), 0);
  for (auto* pr : progs)
    for (const Launch& L : *pr) {
      if (L.ntask <= 0 || L.kind == L_FOLD || L.kind == L_GATE || L.kind == L_PERSIST) continue;
      if (placed[L.toff]) continue;
      for (int k = 0; k < L.ntask; k++) {
        const int src = L.kind == L_COARSE ? k : xcd_task(k, L.ntask);
        dev[L.toff + k] = b.tasks[L.toff + src];
        dev[L.toff + k].slot = src;
        placed[L.toff + k] = 1;
      }
    }
  ECHK(hipMemcpy(b.dT, dev.data(), dev.size() * sizeof(Task), hipMemcpyHostToDevice));
  if (b.dPh) (void)hipFree(b.dPh);
  if (b.bar) (void)hipFree(b.bar);
  b.dPh = nullptr;
  b.bar = nullptr;
  if (!b.phases.empty()) {
    ECHK(hipMalloc(&b.dPh, b.phases.size() * sizeof(PPhase)));
    ECHK(hipMemcpy(b.dPh, b.phases.data(), b.phases.size() * sizeof(PPhase),
                   hipMemcpyHostToDevice));
    ECHK(hipMalloc(&b.bar, (size_t)b.nbar * sizeof(BarSlot)));
    ECHK(hipMemset(b.bar, 0, (size_t)b.nbar * sizeof(BarSlot)));
  }
  if (!b.fault) {
    ECHK(hipMalloc(&b.fault, sizeof(int)));
    ECHK(hipMemset(b.fault, 0, sizeof(int)));
  }
  ECHK(hipMemcpy(b.dS, b.spans.data(), b.spans.size() * sizeof(FoldSpan), hipMemcpyHostToDevice));
  if (!b.slots.empty())
    ECHK(hipMemcpy(b.dSlot, b.slots.data(), b.slots.size() * sizeof(int), hipMemcpyHostToDevice));
  return F2D_OK;
}

// LDS of the coarse kernel from level k0 down: per level the operator (5
// planes of the pool type, tsz bytes) and x, b, t in FP64, then the masks,
// each region rounded to 16 bytes
static size_t coarse_bytes(const InstD& d, int k0, size_t tsz) {
  auto r16 = [](size_t v) { return (v + 15) & ~(size_t)15; };
  size_t sz = 0;
  for (int l = k0; l < d.nlev; l++) {
    const size_t c = (size_t)d.L[l].m * d.L[l].n;
    sz += r16(5 * c * tsz) + 3 * c * sizeof(double) + r16(c);
  }
  return sz;
}

// LDS bytes the coarse kernel may use: the requested budget (0: all), at
// most the device's per-workgroup maximum less the static level table
static size_t lds_budget(int requested) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  int maxsh = 65536;
  (void)hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  const size_t avail = (size_t)maxsh > 4096 ? (size_t)maxsh - 4096 : (size_t)maxsh;
  return requested > 0 ? std::min((size_t)requested, avail) : avail;
}

// first level of the coarse kernel: the largest set of coarse levels whose
// operands fit the budget (at least the coarsest, at most all but level 0)
static int first_coarse_level(const InstD& d, size_t budget, size_t tsz, int* k0) {
  int k = d.nlev - 1;
  while (k > 1 && coarse_bytes(d, k - 1, tsz) <= budget && d.nlev - (k - 1) <= MAXCL) k--;
  if (coarse_bytes(d, k, tsz) > budget)
    return efail(F2D_E_ARG, "coarse levels need %zu bytes of LDS, budget %zu",
                 coarse_bytes(d, k, tsz), budget);
  *k0 = k;
  return F2D_OK;
}

// The kernels address each array of one instance through 32-bit byte
// offsets from a uniform base (at32/ld32/st32, APlanes, Stage::at): every
// per-instance array must stay below 4 GiB.  The largest are the state
// (nvar planes, FP64), a tendency (nvar planes of the pool type), the FP64
// level-0 operator A0d (5 planes) and the level operators (5 planes of the
// pool type, the finest the largest).  Host-only: no device call.
static int extent_check(const f2d_model_params& p, size_t tsz) {
  const double lim = 4294967296.0;  // 2^32 bytes
  const double F = (double)(p.ny + 2 * NH) * (double)(p.nx + 2 * NH);
  const double nvar = p.model == 0 ? 5. : 6.;
  struct {
    const char* what;
    double bytes;
  } ext[] = {{"state", nvar * F * 8.}, {"tendency", nvar * F * (double)tsz},
             {"FP64 level-0 operator", 5. * F * 8.}, {"level-0 operator", 5. * F * (double)tsz}};
  for (const auto& e : ext)
    if (e.bytes >= lim)
      return efail(F2D_E_ARG, "grid %dx%d: the %s takes %.0f bytes; each array of one instance "
                   "must stay below 2^32 bytes", p.nx, p.ny, e.what, e.bytes);
  return F2D_OK;
}

// apply a tuning: coarse split per instance, tile policy, programs rebuilt
static int apply_tuning(Batch& b, const f2d_batch_tuning* t) {
  const size_t budget = lds_budget(t ? t->coarse_lds_budget : 0);
  std::vector<int> k0(b.ninst);
  const size_t tsz = b.prec == 0 ? sizeof(double) : sizeof(float);
  for (int i = 0; i < b.ninst; i++) ETRY(first_coarse_level(b.H[i].d, budget, tsz, &k0[i]));
  size_t maxlds = 0;
  for (int i = 0; i < b.ninst; i++) {
    b.H[i].d.nfine = k0[i];
    b.H[i].coarse_lds = coarse_bytes(b.H[i].d, k0[i], tsz);
    maxlds = std::max(maxlds, b.H[i].coarse_lds);
  }
  b.min_tiles = t ? std::max(0, t->min_tiles) : 0;
  b.persist = t ? std::max(0, t->persist) : 0;
  b.coarse_wave_cells = (t && t->coarse_wave_cells >= 0) ? t->coarse_wave_cells : 256;
  b.coarse_budget = budget;
  if (maxlds > 65536) {
    const void* fn = b.prec == 0 ? (const void*)k_coarse<double> : (const void*)k_coarse<float>;
    ECHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)maxlds));
  }
  std::vector<InstD> di(b.ninst);
  for (int i = 0; i < b.ninst; i++) di[i] = b.H[i].d;
  drop_graph(b);
  ECHK(hipMemcpy(b.dI, di.data(), b.ninst * sizeof(InstD), hipMemcpyHostToDevice));
  return build_programs(b);
}

}  // namespace eng
}  // namespace f2d

using namespace f2d;
using namespace f2d::eng;

struct f2d_batch : public Batch {};

extern "C" {

const char* f2d_batch_last_error(void) { return e_err.c_str(); }

int f2d_batch_create(int ninst, const f2d_model_params* params, const int8_t* const* msks,
                     int precision, const f2d_batch_tuning* tuning, f2d_batch** out) {
  if (ninst < 1) return efail(F2D_E_ARG, "ninst must be >= 1");
  if (precision != 0 && precision != 1) return efail(F2D_E_ARG, "precision must be 0 or 1");
  // host-only checks first: nothing below runs for a rejected batch
  for (int i = 0; i < ninst; i++) {
    if (params[i].nx < 1 || params[i].ny < 1)
      return efail(F2D_E_ARG, "instance %d: grid %dx%d", i, params[i].nx, params[i].ny);
    const int rc = extent_check(params[i], precision == 0 ? sizeof(double) : sizeof(float));
    if (rc) return efail(rc, "instance %d: %s", i, std::string(e_err).c_str());
  }
  f2d_batch* b = new f2d_batch();
  b->prec = precision;
  const size_t tsz = precision == 0 ? sizeof(double) : sizeof(float);
  auto bail = [&](int rc) {
    batch_free(b);
    return rc;
  };
  b->ninst = ninst;
  b->H.resize(ninst);
  // static LDS of k_coarse (level table) comes on top of the dynamic part
  const size_t budget = lds_budget(tuning ? tuning->coarse_lds_budget : 0);
  const int tscheme0 = params[0].timestepping;
  // layout
  size_t nP = 0, nQ = 0, nB = 0;
  auto takeP = [&](size_t cnt) {
    const size_t o = nP;
    nP += (cnt + 31) & ~(size_t)31;
    return (long long)o;
  };
  auto takeQ = [&](size_t cnt) {
    const size_t o = nQ;
    nQ += (cnt + 63) & ~(size_t)63;
    return (long long)o;
  };
  auto takeB = [&](size_t cnt) {
    const size_t o = nB;
    nB += (cnt + 255) & ~(size_t)255;
    return (long long)o;
  };
  for (int i = 0; i < ninst; i++) {
    const f2d_model_params& p = params[i];
    HostInst& h = b->H[i];
    h.p = p;
    h.hmsk.assign(msks[i], msks[i] + (size_t)(p.ny + 2 * NH) * (p.nx + 2 * NH));
    InstD& d = h.d;
    memset(&d, 0, sizeof d);
    if (p.model != 0 && p.model != 1) return bail(efail(F2D_E_ARG, "model"));
    if (p.timestepping != tscheme0)
      return bail(efail(F2D_E_ARG, "one time scheme per batch"));
    if (p.order < 1 || p.order > 6 || p.flux_splitting < 0 || p.flux_splitting > 1)
      return bail(efail(F2D_E_ARG, "order/flux_splitting"));
    d.model = p.model;
    d.m = p.ny + 2 * NH;
    d.n = p.nx + 2 * NH;
    d.order = p.order;
    d.method = p.flux_splitting;
    d.perio = p.perio;
    d.diffusion = p.diffusion;
    d.tscheme = p.timestepping;
    d.nvar = p.model == 0 ? 5 : 6;
    d.dx = p.Lx / p.nx;
    d.dy = p.Ly / p.ny;
    d.gravity = p.gravity;
    d.kd0 = p.kdiff_vort;
    d.kd1 = p.kdiff_buoy;
    d.aparab = p.order % 2 == 0 ? 0. : p.aparab;
    d.omega = 8. / 9.;
    // levels (core/gmg/level.py:62-115)
    int n = p.nx, m = p.ny, nl = 0;
    for (;;) {
      if (nl > 0) {
        if (n % 2 || m % 2) return bail(efail(F2D_E_ARG, "grid %dx%d not halvable", p.nx, p.ny));
        n /= 2;
        m /= 2;
      }
      if (n < NH || m < NH) return bail(efail(F2D_E_ARG, "grid too small"));
      if (nl >= MAXLEV) return bail(efail(F2D_E_ARG, "too many levels"));
      d.L[nl].m = m + 2 * NH;
      d.L[nl].n = n + 2 * NH;
      nl++;
      if (n <= 4 || m <= 4) break;
    }
    if (nl < 2) return bail(efail(F2D_E_ARG, "grid must have at least two levels"));
    d.nlev = nl;
    int k0 = 0;
    if (int rc = first_coarse_level(d, budget, tsz, &k0)) return bail(rc);
    d.nfine = k0;
    h.coarse_lds = coarse_bytes(d, k0, tsz);
    const size_t F = (size_t)d.m * d.n;
    d.state = takeP(d.nvar * F);
    d.work = takeP(F);
    d.xr = takeP(F);
    d.yr = takeP(F);
    d.A0d = takeP(5 * F);
    for (int k = 0; k < 3; k++) d.k[k] = takeQ(d.nvar * F);
    for (int k = 0; k < 3; k++) d.kp[k] = precision != 0 ? takeP(F) : -1;
    d.x0[0] = takeQ(F);
    d.x0[1] = takeQ(F);
    h.x0alt = takeQ(F);
    d.msk = takeB(F);
    d.mskp = takeB(F);
    d.sb0 = takeB(F);
    for (int l = 0; l < nl; l++) {
      LevD& L = d.L[l];
      const size_t S = (size_t)L.m * L.n;
      L.A = takeQ(5 * S);
      L.b = takeQ(S);
      if (l > 0) {
        L.x[0] = takeQ(S);
        L.x[1] = takeQ(S);
      }
      L.msk = l == 0 ? d.mskp : takeB(S);
      h.buf[l][0] = L.x[0];
      h.buf[l][1] = L.x[1];
      h.cur[l] = 0;
    }
    long long a = 0;
    for (int j = NH; j < d.m - NH; j++)
      for (int q = NH; q < d.n - NH; q++) a += msks[i][(size_t)j * d.n + q];
    d.area = (double)a;
  }
  b->nP = nP;
  b->nQ = nQ;
  b->nB = nB;
  if (hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) != hipSuccess)
    return bail(efail(F2D_E_HIP, "stream"));
  // 256 bytes of slack after the T and byte pools: the coarse kernel's
  // 16-byte LDS loads round the last array's tail up
  nQ += 256 / tsz;
  nB += 256;
  if (hipMalloc(&b->P, nP * sizeof(double)) != hipSuccess ||
      hipMalloc(&b->Q, nQ * tsz) != hipSuccess || hipMalloc(&b->B, nB) != hipSuccess)
    return bail(efail(F2D_E_NOMEM, "pools (%zu doubles, %zu reals, %zu bytes)", nP, nQ, nB));
  if (hipMemset(b->P, 0, nP * sizeof(double)) != hipSuccess ||
      hipMemset(b->Q, 0, nQ * tsz) != hipSuccess || hipMemset(b->B, 0, nB) != hipSuccess)
    return bail(efail(F2D_E_HIP, "memset pools"));
  // masks, operators: level hierarchy built by the multigrid set-up kernels
  for (int i = 0; i < ninst; i++) {
    HostInst& h = b->H[i];
    InstD& d = h.d;
    const f2d_model_params& p = h.p;
    const size_t F = (size_t)d.m * d.n;
    if (hipMemcpy(b->B + d.msk, msks[i], F, hipMemcpyHostToDevice) != hipSuccess)
      return bail(efail(F2D_E_HIP, "mask upload"));
    // corner mask on the host (celltocorner of the float mask == 1)
    std::vector<int8_t> mp(F, 0);
    const int8_t* mk = msks[i];
    for (int j = 0; j < d.m - 1; j++)
      for (int q = 0; q < d.n - 1; q++) {
        const size_t c = (size_t)j * d.n + q;
        const double s4 = 0.25 * ((double)mk[c] + (double)mk[c + 1] + (double)mk[c + d.n] +
                                  (double)mk[c + d.n + 1]);
        mp[c] = s4 == 1.0 ? 1 : 0;
      }
    if (hipMemcpy(b->B + d.mskp, mp.data(), F, hipMemcpyHostToDevice) != hipSuccess)
      return bail(efail(F2D_E_HIP, "mskp upload"));
    void* g = nullptr;
    int rc = f2d_gmg_build_for_engine(p.nx, p.ny, d.dx, d.dy, p.hydroepsilon, mp.data(), &g);
    if (rc) return bail(efail(rc, "multigrid set-up failed"));
    for (int l = 0; l < d.nlev; l++) {
      int8_t* gm;
      double* gA;
      int gmm, gnn;
      f2d_gmg_level_raw(g, l, &gm, &gA, &gmm, &gnn);
      const size_t S = (size_t)gmm * gnn;
      if (gmm != d.L[l].m || gnn != d.L[l].n) {
        f2d_gmg_free_raw(g);
        return bail(efail(F2D_E_ARG, "level shape mismatch"));
      }
      bool ok = true;
      if (b->prec == 0) {
        ok = hipMemcpy((double*)b->Q + d.L[l].A, gA, 5 * S * sizeof(double),
                       hipMemcpyDeviceToDevice) == hipSuccess;
      } else {  // operators rounded to FP32 for the FP32 V-cycles
        std::vector<double> hA(5 * S);
        std::vector<float> fA(5 * S);
        ok = hipMemcpy(hA.data(), gA, 5 * S * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
        for (size_t q = 0; q < 5 * S; q++) fA[q] = (float)hA[q];
        ok = ok && hipMemcpy((float*)b->Q + d.L[l].A, fA.data(), 5 * S * sizeof(float),
                             hipMemcpyHostToDevice) == hipSuccess;
      }
      if (l == 0) {
        ok = ok && hipMemcpy(b->P + d.A0d, gA, 5 * S * sizeof(double),
                             hipMemcpyDeviceToDevice) == hipSuccess;
        // stencil bits of the finest operator, verified value by value
        std::vector<double> hA(5 * S);
        ok = ok && hipMemcpy(hA.data(), gA, 5 * S * sizeof(double), hipMemcpyDeviceToHost) ==
                       hipSuccess;
        // bits 0-3: coupling A_k nonzero (all nonzero values of a plane
        // equal); bit 4: the cell mask.  The centre coefficient is read
        // only where the mask is set (Jacobi and residual are 0 elsewhere),
        // and must equal one nonzero value there; the halo fill leaves
        // nonzero centres under a zero mask, which are never read.
        std::vector<uint8_t> bits(S, 0);
        bool fly = ok;
        for (int k = 0; k < 5 && fly; k++) {
          double wk = 0.;
          for (size_t q = 0; q < S; q++) {
            const double a = hA[k * S + q];
            if (k == 4) {
              if (mp[q] == 0) continue;
              if (a == 0.) {
                fly = false;
                break;
              }
            }
            if (a == 0.) continue;
            if (wk == 0.) wk = a;
            if (a != wk) {
              fly = false;
              break;
            }
            bits[q] |= (uint8_t)(1u << k);
          }
          d.w0[k] = wk;
        }
        d.a0fly = fly ? 1 : 0;
        if (fly)
          ok = ok && hipMemcpy(b->B + d.sb0, bits.data(), S, hipMemcpyHostToDevice) == hipSuccess;
      }
      if (l > 0) ok = ok && hipMemcpy(b->B + d.L[l].msk, gm, S, hipMemcpyDeviceToDevice) == hipSuccess;
      if (!ok) {
        f2d_gmg_free_raw(g);
        return bail(efail(F2D_E_HIP, "level copy"));
      }
    }
    f2d_gmg_free_raw(g);
    // coordinates for the diagnostics (core/grid.py:49-52)
    std::vector<double> xr(F), yr(F);
    for (int j = 0; j < d.m; j++)
      for (int q = 0; q < d.n; q++) {
        xr[(size_t)j * d.n + q] = ((double)q + 0.5 - NH) * d.dx;
        yr[(size_t)j * d.n + q] = ((double)j + 0.5 - NH) * d.dy;
      }
    if (hipMemcpy(b->P + d.xr, xr.data(), F * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->P + d.yr, yr.data(), F * 8, hipMemcpyHostToDevice) != hipSuccess)
      return bail(efail(F2D_E_HIP, "coords"));
  }
  std::vector<InstD> di(ninst);
  for (int i = 0; i < ninst; i++) di[i] = b->H[i].d;
  if (hipMalloc(&b->dI, ninst * sizeof(InstD)) != hipSuccess ||
      hipMemcpy(b->dI, di.data(), ninst * sizeof(InstD), hipMemcpyHostToDevice) != hipSuccess)
    return bail(efail(F2D_E_HIP, "descriptors"));
  if (hipMalloc(&b->dD, ninst * sizeof(DynD)) != hipSuccess ||
      hipMemset(b->dD, 0, ninst * sizeof(DynD)) != hipSuccess)
    return bail(efail(F2D_E_HIP, "dyn"));
  if (hipHostMalloc((void**)&b->hD, ninst * sizeof(DynD), hipHostMallocDefault) != hipSuccess)
    return bail(efail(F2D_E_NOMEM, "pinned dyn"));
  memset(b->hD, 0, ninst * sizeof(DynD));
  for (int i = 0; i < ninst; i++) {
    b->hD[i].run = 1;
    b->hD[i].steps_left = -1;
  }
  if (hipMemcpy(b->dD, b->hD, ninst * sizeof(DynD), hipMemcpyHostToDevice) != hipSuccess)
    return bail(efail(F2D_E_HIP, "dyn init"));
  int rc = apply_tuning(*b, tuning);
  if (rc) return bail(rc);
  *out = b;
  return F2D_OK;
}

int f2d_batch_retune(f2d_batch* b, const f2d_batch_tuning* tuning) {
  ECHK(hipStreamSynchronize(b->s));
  return apply_tuning(*b, tuning);
}

int f2d_batch_tuning_info(f2d_batch* b, f2d_batch_tuning* out) {
  out->coarse_lds_budget = (int)b->coarse_budget;
  out->min_tiles = b->min_tiles;
  out->coarse_wave_cells = b->coarse_wave_cells;
  out->persist = b->persist;
  return F2D_OK;
}

int f2d_batch_destroy(f2d_batch* b) {
  batch_free(b);
  return F2D_OK;
}

int f2d_batch_info(f2d_batch* b, int inst, int* nlev, int* nfine, long long* launches_per_step) {
  if (inst < 0 || inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  *nlev = b->H[inst].d.nlev;
  *nfine = b->H[inst].d.nfine;
  *launches_per_step = (long long)b->prog_step.size();
  return F2D_OK;
}

int f2d_batch_set_field(f2d_batch* b, int inst, int field, const double* host) {
  if (inst < 0 || inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  const InstD& d = b->H[inst].d;
  if (field < 0 || field >= d.nvar) return efail(F2D_E_ARG, "field out of range");
  const size_t F = (size_t)d.m * d.n;
  // the batch stream is non-blocking: steps still in flight read the state
  ECHK(hipStreamSynchronize(b->s));
  ECHK(hipMemcpyAsync(b->P + d.state + field * F, host, F * 8, hipMemcpyHostToDevice, b->s));
  ECHK(hipStreamSynchronize(b->s));
  return F2D_OK;
}

int f2d_batch_get_state(f2d_batch* b, int inst, double* host) {
  if (inst < 0 || inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  const InstD& d = b->H[inst].d;
  ECHK(hipStreamSynchronize(b->s));
  ETRY(check_fault(*b));
  ECHK(hipMemcpy(host, b->P + d.state, (size_t)d.nvar * d.m * d.n * 8, hipMemcpyDeviceToHost));
  return F2D_OK;
}

int f2d_batch_state_ptr(f2d_batch* b, int inst, double** dev) {
  if (inst < 0 || inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  *dev = b->P + b->H[inst].d.state;
  return F2D_OK;
}

static int sync_dyn_to_device(f2d_batch* b) {
  ECHK(hipMemcpyAsync(b->dD, b->hD, b->ninst * sizeof(DynD), hipMemcpyHostToDevice, b->s));
  ECHK(hipStreamSynchronize(b->s));
  return F2D_OK;
}
static int sync_dyn_to_host(f2d_batch* b) {
  ECHK(hipMemcpyAsync(b->hD, b->dD, b->ninst * sizeof(DynD), hipMemcpyDeviceToHost, b->s));
  ECHK(hipStreamSynchronize(b->s));
  return F2D_OK;
}

int f2d_batch_set_dt(f2d_batch* b, int inst, double dt) {
  if (inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  // only the dt fields are written (the rest of the device record is the
  // engine's), through the pinned mirror
  const int i0 = inst < 0 ? 0 : inst, i1 = inst < 0 ? b->ninst : inst + 1;
  ECHK(hipStreamSynchronize(b->s));
  for (int i = i0; i < i1; i++) {
    b->hD[i].dt = dt;
    ECHK(hipMemcpyAsync(&b->dD[i].dt, &b->hD[i].dt, sizeof(double), hipMemcpyHostToDevice, b->s));
  }
  ECHK(hipStreamSynchronize(b->s));
  return F2D_OK;
}

int f2d_batch_set_steps(f2d_batch* b, int inst, int nsteps) {
  ETRY(sync_dyn_to_host(b));
  for (int i = 0; i < b->ninst; i++)
    if (inst < 0 || i == inst) {
      b->hD[i].steps_left = nsteps;
      b->hD[i].run = 1;
    }
  return sync_dyn_to_device(b);
}

int f2d_batch_progress(f2d_batch* b, int inst, int* steps_left, long long* steps_done) {
  if (inst < 0 || inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  ETRY(sync_dyn_to_host(b));
  *steps_left = b->hD[inst].steps_left;
  *steps_done = b->hD[inst].steps_done;
  return F2D_OK;
}

int f2d_batch_set_adaptive(f2d_batch* b, int on, double cfl, double dtmax) {
  // the flags are kernel arguments of the captured step: recapture
  drop_graph(*b);
  b->adapt = on;
  b->cfl = cfl;
  b->dtmax = dtmax;
  return F2D_OK;
}

int f2d_batch_set_fast_solve(f2d_batch* b, int maxite, double tol) {
  if (maxite < 0 || maxite > 64) return efail(F2D_E_ARG, "maxite must be in 0..64 (got %d)", maxite);
  if (maxite > 0 && !(tol >= 0.)) return efail(F2D_E_ARG, "tol must be >= 0");
  // the step's launch program changes: rebuild it (and recapture the graph)
  ECHK(hipStreamSynchronize(b->s));
  drop_graph(*b);
  b->fast_maxite = maxite;
  b->fast_tol = tol;
  return build_programs(*b);
}

}  // extern "C"

// The launch programs again after an instance's options changed.
static int reprogram(Batch& b) {
  ECHK(hipStreamSynchronize(b.s));
  drop_graph(b);
  std::vector<InstD> di(b.ninst);
  for (int i = 0; i < b.ninst; i++) di[i] = b.H[i].d;
  ECHK(hipMemcpy(b.dI, di.data(), b.ninst * sizeof(InstD), hipMemcpyHostToDevice));
  return build_programs(b);
}
// a device copy of cnt host values, owned by the batch
template <typename T>
static int upload(Batch& b, const T* host, size_t cnt, T** dev) {
  void* p = nullptr;
  ECHK(hipMalloc(&p, cnt * sizeof(T)));
  b.extra.push_back(p);
  ECHK(hipMemcpy(p, host, cnt * sizeof(T), hipMemcpyHostToDevice));
  *dev = (T*)p;
  return F2D_OK;
}

extern "C" {

int f2d_batch_set_noslip(f2d_batch* b, int inst, const int8_t* msknoslip, const int8_t* mskbc,
                         const double* xr0, const double* yr0, double bcarea, double x2bc,
                         double y2bc, int enforce) {
  if (inst < 0 || inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  if (!msknoslip || !mskbc || !xr0 || !yr0)
    return efail(F2D_E_ARG, "f2d_batch_set_noslip: null argument");
  if (!(bcarea > 0.)) return efail(F2D_E_ARG, "no boundary cells (bcarea = %g)", bcarea);
  if (enforce && !(x2bc > 0. && y2bc > 0.))
    return efail(F2D_E_ARG, "enforce_momentum needs x2bc, y2bc > 0 (got %g, %g)", x2bc, y2bc);
  InstD& d = b->H[inst].d;
  if (d.noslip) return efail(F2D_E_ARG, "instance %d: no-slip walls already set", inst);
  const size_t F = (size_t)d.m * d.n;
  ECHK(hipStreamSynchronize(b->s));
  ETRY(upload(*b, msknoslip, F, const_cast<int8_t**>(&d.msknoslip)));
  ETRY(upload(*b, mskbc, F, const_cast<int8_t**>(&d.mskbc)));
  if (!d.xr0) {
    ETRY(upload(*b, xr0, F, const_cast<double**>(&d.xr0)));
    ETRY(upload(*b, yr0, F, const_cast<double**>(&d.yr0)));
  }
  if (d.model == 0) {
    d.nsrc = b->P + d.state + 4 * F;  // the Euler state's source field
  } else {
    void* p = nullptr;
    ECHK(hipMalloc(&p, F * sizeof(double)));
    b->extra.push_back(p);
    ECHK(hipMemset(p, 0, F * sizeof(double)));
    d.nsrc = (double*)p;
  }
  d.noslip = 1;
  d.enforce = enforce ? 1 : 0;
  d.bcarea = bcarea;
  d.x2bc = x2bc;
  d.y2bc = y2bc;
  return reprogram(*b);
}

#ifdef F2D_PHASE_STAMPS
// the phase timestamps of the last 64 x 16 smoother launch (2048 x 8)
int f2d_debug_smooth_stamps(unsigned long long* host) {
  ECHK(hipMemcpyFromSymbol(host, HIP_SYMBOL(f2d::eng::smooth_stamps),
                           sizeof(unsigned long long) * 2048 * 8));
  return F2D_OK;
}
// the phase stamps of k_smooth_rr (8 levels x 256 workgroups x 8)
int f2d_debug_rr_stamps(unsigned long long* host) {
  ECHK(hipMemcpyFromSymbol(host, HIP_SYMBOL(f2d::eng::rr_stamps),
                           sizeof(unsigned long long) * 8 * 256 * 8));
  return F2D_OK;
}
// the phase stamps of the last coarse-kernel launch of each mode (2 x 256)
int f2d_debug_coarse_stamps(unsigned long long* host) {
  ECHK(hipMemcpyFromSymbol(host, HIP_SYMBOL(f2d::eng::coarse_stamps),
                           sizeof(unsigned long long) * 2 * 256));
  return F2D_OK;
}
#endif

int f2d_batch_corner_mask(f2d_batch* b, int inst, int8_t* host) {
  if (inst < 0 || inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  const InstD& d = b->H[inst].d;
  ECHK(hipStreamSynchronize(b->s));
  ECHK(hipMemcpy(host, b->B + d.mskp, (size_t)d.m * d.n, hipMemcpyDeviceToHost));
  return F2D_OK;
}

int f2d_batch_set_islands(f2d_batch* b, int inst, const double* rhsp, const double* psi) {
  if (inst < 0 || inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  if (!rhsp || !psi) return efail(F2D_E_ARG, "f2d_batch_set_islands: null argument");
  InstD& d = b->H[inst].d;
  if (d.island) return efail(F2D_E_ARG, "instance %d: islands already set", inst);
  const size_t F = (size_t)d.m * d.n;
  ECHK(hipStreamSynchronize(b->s));
  ETRY(upload(*b, rhsp, F, const_cast<double**>(&d.rhsp)));
  ETRY(upload(*b, psi, F, const_cast<double**>(&d.ipsi)));
  d.island = 1;
  return reprogram(*b);
}

int f2d_batch_set_momentum(f2d_batch* b, int inst, const double* xr0, const double* yr0) {
  if (inst < 0 || inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  if (!xr0 || !yr0) return efail(F2D_E_ARG, "f2d_batch_set_momentum: null argument");
  InstD& d = b->H[inst].d;
  if (d.momentum) return F2D_OK;
  const size_t F = (size_t)d.m * d.n;
  ECHK(hipStreamSynchronize(b->s));
  if (!d.xr0) {
    ETRY(upload(*b, xr0, F, const_cast<double**>(&d.xr0)));
    ETRY(upload(*b, yr0, F, const_cast<double**>(&d.yr0)));
  }
  d.momentum = 1;
  return reprogram(*b);
}

int f2d_batch_enforce_momentum(f2d_batch* b, int inst, double px, double py) {
  if (inst < 0 || inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  if (!b->H[inst].d.momentum)
    return efail(F2D_E_ARG, "instance %d: momentum enforcement not enabled", inst);
  ECHK(hipStreamSynchronize(b->s));
  b->hD[inst].mom[0] = px;
  b->hD[inst].mom[1] = py;
  ECHK(hipMemcpyAsync(&b->dD[inst].mom, &b->hD[inst].mom, sizeof(double) * 2,
                      hipMemcpyHostToDevice, b->s));
  ETRY(run_prog(*b, b->prog_mom[inst]));
  ECHK(hipStreamSynchronize(b->s));
  return check_fault(*b);
}

int f2d_batch_use_graph(f2d_batch* b, int on) {
  b->use_graph = on != 0;
  return F2D_OK;
}

// A slot of the batch re-used for a new run: every per-instance buffer that
// carries values from one step (or solve) to the next is zeroed -- the state,
// the full-solve rhs, the three tendencies (their psi planes are the next
// inversions' first guesses), the FP64 psi tendencies of mixed precision, the
// level-0 correction pair and every level's rhs and iterates -- and the
// device record is reset, as at creation.  The operators, masks and
// coordinates stay.  The other instances are untouched.
int f2d_batch_reset_instance(f2d_batch* b, int inst) {
  if (inst < 0 || inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  const InstD& d = b->H[inst].d;
  const size_t F = (size_t)d.m * d.n;
  const size_t tsz = b->prec == 0 ? sizeof(double) : sizeof(float);
  char* Qb = (char*)b->Q;
  ECHK(hipStreamSynchronize(b->s));
  ECHK(hipMemsetAsync(b->P + d.state, 0, (size_t)d.nvar * F * 8, b->s));
  ECHK(hipMemsetAsync(b->P + d.work, 0, F * 8, b->s));
  for (int k = 0; k < 3; k++) {
    ECHK(hipMemsetAsync(Qb + d.k[k] * tsz, 0, (size_t)d.nvar * F * tsz, b->s));
    if (d.kp[k] >= 0) ECHK(hipMemsetAsync(b->P + d.kp[k], 0, F * 8, b->s));
  }
  ECHK(hipMemsetAsync(Qb + d.x0[0] * tsz, 0, F * tsz, b->s));
  ECHK(hipMemsetAsync(Qb + d.x0[1] * tsz, 0, F * tsz, b->s));
  ECHK(hipMemsetAsync(Qb + b->H[inst].x0alt * tsz, 0, F * tsz, b->s));
  for (int l = 0; l < d.nlev; l++) {
    const LevD& L = d.L[l];
    const size_t S = (size_t)L.m * L.n;
    ECHK(hipMemsetAsync(Qb + L.b * tsz, 0, S * tsz, b->s));
    if (l > 0)
      for (int q = 0; q < 2; q++) ECHK(hipMemsetAsync(Qb + L.x[q] * tsz, 0, S * tsz, b->s));
  }
  ECHK(hipStreamSynchronize(b->s));
  memset(&b->hD[inst], 0, sizeof(DynD));
  b->hD[inst].run = 1;
  b->hD[inst].steps_left = -1;
  ECHK(hipMemcpyAsync(&b->dD[inst], &b->hD[inst], sizeof(DynD), hipMemcpyHostToDevice, b->s));
  ECHK(hipStreamSynchronize(b->s));
  return F2D_OK;
}

// set_psi_from_vorticity of one instance (its solve and diagnostics only),
// enqueued on the batch's stream
int f2d_batch_init_instance(f2d_batch* b, int inst) {
  if (inst < 0 || inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  return run_prog(*b, b->prog_one[inst]);
}

int f2d_batch_set_psi_from_vorticity(f2d_batch* b) {
  ETRY(run_prog(*b, b->prog_init));
  ECHK(hipStreamSynchronize(b->s));
  return check_fault(*b);
}

int f2d_batch_refresh_diagnostics(f2d_batch* b) {
  ETRY(run_prog(*b, b->prog_diag));
  ECHK(hipStreamSynchronize(b->s));
  return check_fault(*b);
}

static int ensure_graph(f2d_batch* b) {
  if (b->gexec) return F2D_OK;
  ECHK(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
  int rc = run_prog(*b, b->prog_step);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->s, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return efail(F2D_E_HIP, "capture: %s", hipGetErrorString(e));
  b->graph = g;
  ECHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
  return F2D_OK;
}

int f2d_batch_step_async(f2d_batch* b, int nsteps) {
  if (b->use_graph) {
    ETRY(ensure_graph(b));
    for (int k = 0; k < nsteps; k++) ECHK(hipGraphLaunch(b->gexec, b->s));
  } else {
    for (int k = 0; k < nsteps; k++) ETRY(run_prog(*b, b->prog_step));
  }
  return F2D_OK;
}

int f2d_batch_step(f2d_batch* b, int nsteps) {
  ETRY(f2d_batch_step_async(b, nsteps));
  ECHK(hipStreamSynchronize(b->s));
  return check_fault(*b);
}

int f2d_batch_synchronize(f2d_batch* b) {
  ECHK(hipStreamSynchronize(b->s));
  return check_fault(*b);
}

int f2d_batch_query(f2d_batch* b, int* done) {
  const hipError_t e = hipStreamQuery(b->s);
  if (e == hipSuccess) {
    *done = 1;
    return check_fault(*b);
  }
  if (e == hipErrorNotReady) {
    *done = 0;
    return F2D_OK;
  }
  return efail(F2D_E_HIP, "hipStreamQuery: %s", hipGetErrorString(e));
}

int f2d_batch_profile(f2d_batch* b, int kind, int level, int nsteps, double* avg_ms,
                      long long* nlaunch, double* alg_bytes, double* impl_bytes) {
  const int kinds[7] = {L_SMOOTH, L_RR, L_RHS, L_COARSE, L_PERSIST, L_SMOOTH_RR, L_UP2};
  if (kind < 0 || kind > 6) return efail(F2D_E_ARG, "kind 0..6");
  ProfSel ps;
  ps.kind = kinds[kind];
  ps.level = level;
  ECHK(hipStreamSynchronize(b->s));
  for (int k = 0; k < nsteps; k++) {
    int rc = run_prog(*b, b->prog_step, &ps);
    if (rc) return rc;
  }
  ECHK(hipStreamSynchronize(b->s));
  double tot = 0., ab = 0., ib = 0.;
  for (size_t q = 0; q < ps.ev.size(); q++) {
    float ms = 0.f;
    ECHK(hipEventElapsedTime(&ms, ps.ev[q].first, ps.ev[q].second));
    tot += ms / PROF_REP;
    double a, im;
    launch_bytes(*b, *ps.which[q], &a, &im);
    ab += a;
    ib += im;
    (void)hipEventDestroy(ps.ev[q].first);
    (void)hipEventDestroy(ps.ev[q].second);
  }
  const double cnt = ps.ev.empty() ? 1. : (double)ps.ev.size();
  *nlaunch = (long long)ps.ev.size();
  *avg_ms = tot / cnt;
  *alg_bytes = ab / cnt;
  *impl_bytes = ib / cnt;
  return F2D_OK;
}

int f2d_batch_diagnostics(f2d_batch* b, int inst, double* out8) {
  if (inst < 0 || inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  ETRY(sync_dyn_to_host(b));
  for (int k = 0; k < NDIAG; k++) out8[k] = b->hD[inst].diag[k];
  return F2D_OK;
}

int f2d_batch_solve_info(f2d_batch* b, int inst, int* nite, double* res, int* diverged) {
  if (inst < 0 || inst >= b->ninst) return efail(F2D_E_ARG, "instance out of range");
  ETRY(sync_dyn_to_host(b));
  *nite = b->hD[inst].nite;
  *res = b->hD[inst].res;
  *diverged = b->hD[inst].diverged;
  return F2D_OK;
}

}  // extern "C"
