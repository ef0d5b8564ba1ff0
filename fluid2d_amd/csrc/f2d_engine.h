// f2d_engine.h -- batched, graph-replayable engine for the Fluid2d step.
//
// A batch holds any number of independent instances (Euler or Boussinesq,
// any power-of-two grid, any mask).  Every instance runs exactly the
// operation sequence of the reference (core/timescheme.py RK3_SSP / RK3,
// core/operators.py, core/gmg/hierarchy.py), so its result does not depend
// on what else is in the batch.  Each kernel launch covers the tiles of all
// instances that take part in that phase; a step is a fixed launch program
// with no host round trip (solve control and dt live on the device), which
// is captured once into a hipGraph and replayed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace f2d {
namespace eng {

constexpr int MAXLEV = 16;
constexpr int TX = 64, TY = 16;    // fine tile (interior cells)
constexpr int TX2 = 32, TY2 = 8;   // coarse tile of the residual/restriction
constexpr int NDIAG = 8;

struct LevD {
  int m, n;
  long long b, A;  // offsets into the real pool (A: 5 planes of m*n)
  long long x[2];  // ping-pong buffers (levels >= 1; level 0 uses per-call buffers)
  long long msk;   // offset into the byte pool
  int pad;
};

struct InstD {
  int model;   // 0 euler, 1 boussinesq
  int nlev, nfine;
  int m, n;    // finest padded dims
  int order, method, perio, diffusion, tscheme, nvar;
  double dx, dy, gravity, kd0, kd1, area, aparab, omega;
  // FP64 pool: state, work (full-solve rhs), xr, yr, A0d (FP64 level-0
  // operator of the outer residual); T pool: tendencies k, level arrays,
  // the level-0 correction pair x0 (T is double or float)
  long long state, work, xr, yr, A0d;
  long long k[3], x0[2];
  long long kp[3];                     // mixed mode: FP64 psi tendencies (P pool); else -1
  long long msk, mskp;                 // byte pool offsets
  // finest-level operator from stencil bits: A_k(cell) = w0[k] where bit k
  // of sb0[cell] is set, else 0 (verified against the explicit operator at
  // set-up; a0fly = 0 keeps the explicit planes)
  long long sb0;
  int a0fly, pad0;
  double w0[5];
  // no-slip walls (core/operators.py:157-185, :245-290), islands
  // (core/island.py:21-43) and the loop's momentum enforcement
  // (core/fluid2d.py:352-358): set after creation, device arrays of m x n
  // (nullptr when unused).  nsrc is the no-slip source plane (the Euler
  // state's field 4, a separate plane for Boussinesq).
  int noslip, enforce, island, momentum;
  double bcarea, x2bc, y2bc;
  const int8_t *msknoslip, *mskbc;
  const double *xr0, *yr0, *rhsp, *ipsi;
  double* nsrc;
  LevD L[MAXLEV];
};

// device-written per-instance values
struct DynD {
  double dt, umax;
  double normeb, res0, res;
  int nite, ndiv, active, diverged;
  double psim;
  double srcm;      // no-slip: mean source over the boundary cells
  double nsdot[2];  // no-slip momentum: sum source*xr0, sum source*yr0 (msk != 0)
  double mom[2];    // loop momentum enforcement: px / x2, py / y2 (host-written)
  double diag[NDIAG];
  int run;         // this step runs for the instance (set by the step gate)
  int steps_left;  // remaining steps (< 0: unlimited)
  long long steps_done;
};

// one tile of a 2D phase
struct Task {
  int inst, lev, y0, x0;
  long long xin, xout, xc;  // buffer offsets used by the phase
  int mode, flags;          // phase variant (smoother input form, final mask)
  // geometry of level `lev` and of the next coarser level, copied from the
  // instance's level table when the task list is uploaded, so that a tile
  // kernel can address its data right after reading its task
  int m, n, m2, n2;
  long long b, A, msk, msk2;  // level rhs, operator (T pool), mask; coarser mask (byte pool)
  long long sb;               // finest-level stencil bits (byte pool), -1 where the level has none
  // two-level up-leg tasks (k_up2): the coarser level's input x, rhs and
  // operator, the next coarser level's x, mask and size
  long long xin2, b2, A2, xcc, msk3;
  int m3, n3;
  int slot;  // the task's index in its launch's list as built (partial-sum slot)
};

}  // namespace eng
}  // namespace f2d
