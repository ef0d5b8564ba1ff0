// f2d_dev.h -- per-cell device arithmetic shared by the stand-alone kernels
// (f2d_ops.hip) and the batched engine (f2d_engine.hip).  Each function
// keeps the operation order of the reference routine it cites, so FP64
// results are bit-identical wherever it is evaluated.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "f2d_kernels.h"

namespace f2d {

// Weights of the reference are float32 literal expressions stored into
// real*8 (core/fortran_advection.f90:36-44): round in float, widen.
#define W32(num, den) ((double)((float)(num) / (float)(den)))

template <typename T>
struct UpParams {
  T zdx, zdy, u1, aa, bb, KK, a;
  int method, order;
};

template <typename T>
__device__ __forceinline__ T split_speed(T vel, const UpParams<T>& p) {
  T UU = T(0);
  if (p.method == 0) UU = fabs(vel);
  if (p.method == 1) {
    UU = fabs(vel);
    if (UU < p.u1) UU = p.aa * (vel * vel) + p.bb;
  }
  if (p.method == 2) UU = T(2) * p.KK + log(cosh(vel * p.a)) / p.a;
  return UU;
}

// Upwind flux through the edge between cell c and c+s: mask window
// mk[-2s..3s], values q[-2s..3s], edge velocity vel.
// core/fortran_advection.f90:73-146.  Generic over the value accessor Q
// (Q(k) returns the tracer at offset k*s) so stage states can be formed on
// the fly.
template <typename T, typename Q>
__device__ __forceinline__ T up_edge_flux_q(const int8_t* mk, ptrdiff_t s, T vel,
                                            const UpParams<T>& p, Q q) {
  if (mk[0] + mk[s] != 2) return T(0);
  const T d1 = T(W32(1, 30)), d2 = T(-W32(13, 60)), d3 = T(W32(47, 60)), d4 = T(W32(9, 20)),
          d5 = T(-W32(1, 20));
  const T c1 = T(-W32(1, 6)), c2 = T(W32(5, 6)), c3 = T(W32(2, 6));
  const T UU = split_speed(vel, p);
  const T up = T(0.5) * (vel + UU), um = T(0.5) * (vel - UU);
  const int m_2 = mk[-2 * s], m_1 = mk[-s], m0 = mk[0], m1 = mk[s], m2 = mk[2 * s],
            m3 = mk[3 * s];
  const int s5p = m_2 + m_1 + m0 + m1 + m2, s3p = m_1 + m0 + m1;
  const int s5m = m_1 + m0 + m1 + m2 + m3, s3m = m0 + m1 + m2;
  T qp, qm;
  if (s5p == 5 && p.order == 5)
    qp = d1 * q(-2) + d2 * q(-1) + d3 * q(0) + d4 * q(1) + d5 * q(2);
  else if (s3p == 3 && p.order >= 3)
    qp = c1 * q(-1) + c2 * q(0) + c3 * q(1);
  else
    qp = q(0);
  if (s5m == 5 && p.order == 5)
    qm = d5 * q(-1) + d4 * q(0) + d3 * q(1) + d2 * q(2) + d1 * q(3);
  else if (s3m == 3 && p.order >= 3)
    qm = c3 * q(0) + c2 * q(1) + c1 * q(2);
  else
    qm = q(1);
  return up * qp + um * qm;
}

// Centred flux, core/fortran_advection.f90:169-284 (allow6) /
// core/fortran_fluxes.f90:174-279 (no sixth order).
template <typename T, typename Q>
__device__ __forceinline__ T ce_edge_flux_q(const int8_t* mk, ptrdiff_t s, T vel, int order,
                                            bool allow6, Q q) {
  const int s2 = mk[0] + mk[s];
  if (s2 != 2) return T(0);
  const T e1 = T(W32(1, 60)), e2 = T(-W32(2, 15)), e3 = T(W32(37, 60));
  const T d1 = T(-W32(1, 12)), d2 = T(W32(7, 12));
  const T c1 = T(0.5);
  const int s4 = mk[-s] + mk[0] + mk[s] + mk[2 * s];
  const int s6 = mk[-2 * s] + s4 + mk[3 * s];
  T qp = T(0);
  if (allow6 && s6 == 6 && order == 6) {
    qp = e1 * (q(-2) + q(3)) + e2 * (q(-1) + q(2));
    qp = qp + e3 * (q(0) + q(1));
  } else if (s4 == 4 && order == 4) {
    qp = d1 * (q(-1) + q(2)) + d2 * (q(0) + q(1));
  } else if (order >= 2) {
    qp = c1 * (q(0) + q(1));
  }
  return vel * qp;
}

// Multigrid operator in two-dimensional form.  A2(k, dj, di) returns
// coefficient plane k (A_1..A_5 of the reference) at the cell offset
// (dj, di): A_k couples the cell to its lower neighbour k (SW, S, SE, W,
// centre); the upper couplings are read from the neighbours by symmetry
// (core/gmg/fortran_multigrid.f90:68-82, 342-350).  X2(dj, di) returns x
// at that offset.  Offsets are literals, so tile accessors compile to
// immediate offsets.
template <typename T, typename AF, typename X>
__device__ __forceinline__ T offdiag_2d(AF A, X x) {
  T s = A(0, 0, 0) * x(-1, -1);
  s = s + A(1, 0, 0) * x(-1, 0);
  s = s + A(2, 0, 0) * x(-1, 1);
  s = s + A(3, 0, 0) * x(0, -1);
  s = s + A(3, 0, 1) * x(0, 1);
  s = s + A(2, 1, -1) * x(1, -1);
  s = s + A(1, 1, 0) * x(1, 0);
  s = s + A(0, 1, 1) * x(1, 1);
  return s;
}

// one damped-Jacobi update (core/gmg/fortran_multigrid.f90:66-85); 0 where
// the mask is 0.  Evaluated without a branch (a masked cell's value is
// computed and discarded), so every operand load is issued at once.
template <typename T, typename AF, typename X>
__device__ __forceinline__ T jacobi_2d(int8_t mk, T bc, T c1, T c2, AF A, X x) {
  const T c3 = c1 / fabs(A(4, 0, 0));
  const T v = x(0, 0) * c2 + c3 * (offdiag_2d<T>(A, x) - bc);
  return mk == 0 ? T(0) : v;
}

// jacobi_2d for a cell known to be active, with its factor c3 = c1 / |A_5|
// computed once by the caller (the same value: the division is exact
// arithmetic on the same operands)
template <typename T, typename AF, typename X>
__device__ __forceinline__ T jacobi_c3_2d(T bc, T c3, T c2, AF A, X x) {
  return x(0, 0) * c2 + c3 * (offdiag_2d<T>(A, x) - bc);
}

// r = b - A x (core/gmg/fortran_multigrid.f90:340-357); 0 where the mask
// is 0 (evaluated without a branch, as jacobi_2d)
template <typename T, typename AF, typename X>
__device__ __forceinline__ T residual_2d(int8_t mk, T bc, AF A, X x) {
  T s = bc - A(0, 0, 0) * x(-1, -1);
  s = s - A(1, 0, 0) * x(-1, 0);
  s = s - A(2, 0, 0) * x(-1, 1);
  s = s - A(3, 0, 0) * x(0, -1);
  s = s - A(4, 0, 0) * x(0, 0);
  s = s - A(3, 0, 1) * x(0, 1);
  s = s - A(2, 1, -1) * x(1, -1);
  s = s - A(1, 1, 0) * x(1, 0);
  s = s - A(0, 1, 1) * x(1, 1);
  return mk == 0 ? T(0) : s;
}

// The same in flat form: AF(k, d) and X(d) at flat offset d = dj * n + di
// of a row-major array with row length n.
template <typename T, typename AF, typename X>
__device__ __forceinline__ T offdiag_f(int n, AF A, X x) {
  return offdiag_2d<T>([&](int k, int dj, int di) { return A(k, (ptrdiff_t)dj * n + di); },
                       [&](int dj, int di) { return x(dj * n + di); });
}
template <typename T, typename AF, typename X>
__device__ __forceinline__ T jacobi_f(int8_t mk, int n, T bc, T c1, T c2, AF A, X x) {
  return jacobi_2d<T>(mk, bc, c1, c2,
                      [&](int k, int dj, int di) { return A(k, (ptrdiff_t)dj * n + di); },
                      [&](int dj, int di) { return x(dj * n + di); });
}
template <typename T, typename AF, typename X>
__device__ __forceinline__ T jacobi_c3(int n, T bc, T c3, T c2, AF A, X x) {
  return jacobi_c3_2d<T>(bc, c3, c2,
                         [&](int k, int dj, int di) { return A(k, (ptrdiff_t)dj * n + di); },
                         [&](int dj, int di) { return x(dj * n + di); });
}
template <typename T, typename AF, typename X>
__device__ __forceinline__ T residual_f(int8_t mk, int n, T bc, AF A, X x) {
  return residual_2d<T>(mk, bc,
                        [&](int k, int dj, int di) { return A(k, (ptrdiff_t)dj * n + di); },
                        [&](int dj, int di) { return x(dj * n + di); });
}

// explicit planes of stride P at cell c
template <typename T>
struct APlanes {
  const T* A;
  size_t P, c;
  __device__ __forceinline__ T operator()(int k, ptrdiff_t d) const {
    // 32-bit element offset from the uniform base (one instance's operator)
    const uint32_t o = (uint32_t)k * (uint32_t)P + (uint32_t)c + (uint32_t)(int32_t)d;
    return *(const T*)((const char*)A + (size_t)(uint32_t)(o * (uint32_t)sizeof(T)));
  }
};
// w where bit k of `bits` is set, else +0: the value's bits masked by 0 or
// all ones (a sign-extending one-bit field extract), three instructions per
// FP64 value instead of a test and two selects
template <typename T>
__device__ __forceinline__ T bit_select(uint32_t bits, int k, T w) {
  const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int)bits, (unsigned)k, 1u);
  if constexpr (sizeof(T) == 8) {
    const uint64_t u = (uint64_t)__double_as_longlong(w);
    const uint32_t lo = (uint32_t)u & m, hi = (uint32_t)(u >> 32) & m;
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
  } else {
    return __int_as_float(__float_as_int(w) & (int)m);
  }
}

// constant weights w_k where bit k of the cell's stencil byte is set (the
// finest-level operator: stencil * coef * msk_nb * msk, core/gmg/level.py:292)
template <typename T>
struct ABits {
  const uint8_t* sb;
  const T* w;
  size_t c;
  __device__ __forceinline__ T operator()(int k, ptrdiff_t d) const {
    return bit_select<T>(sb[c + d], k, w[k]);
  }
};

// the same weights from a tile of stencil bits (two-dimensional form):
// `cell` points at the cell's byte in a tile of row length PITCH
template <typename T, int PITCH>
struct ABitsTile2 {
  const uint8_t* cell;
  const T* w;
  __device__ __forceinline__ T operator()(int k, int dj, int di) const {
    return bit_select<T>(cell[dj * PITCH + di], k, w[k]);
  }
};

template <typename T, typename X>
__device__ __forceinline__ T offdiag_x(const T* A, size_t P, size_t c, int n, X x) {
  return offdiag_f<T>(n, APlanes<T>{A, P, c}, x);
}
template <typename T, typename X>
__device__ __forceinline__ T jacobi_x(int8_t mk, const T* A, size_t P, size_t c, int n, T bc, T c1,
                                      T c2, X x) {
  return jacobi_f<T>(mk, n, bc, c1, c2, APlanes<T>{A, P, c}, x);
}
template <typename T, typename X>
__device__ __forceinline__ T residual_x(int8_t mk, const T* A, size_t P, size_t c, int n, T bc,
                                        X x) {
  return residual_f<T>(mk, n, bc, APlanes<T>{A, P, c}, x);
}

// full weighting around a fine point (core/gmg/fortran_multigrid.f90:532-534);
// R(dj, di) returns the fine residual at the offset
template <typename T, typename R>
__device__ __forceinline__ T restrict_r(R r) {
  return T(0.25) * r(0, 0) + T(0.125) * (r(0, -1) + r(0, 1) + r(-1, 0) + r(1, 0)) +
         T(0.0625) * (r(-1, -1) + r(-1, 1) + r(1, -1) + r(1, 1));
}

// Prolongation value at fine (j,i) (core/gmg/fortran_multigrid.f90:454-492,
// nh = 3): coarse (J2,I2) = ((j+2)>>1, (i+2)>>1); row/column parity picks
// the injected, edge (mask-weighted pair) or centre (mask-weighted quad)
// form.  The 1/3 weight is the float32 literal of line 439 widened.
template <typename T, typename PM1, typename PM2, typename PX>
__device__ __forceinline__ T interp_at(PM1 msk1, int n1, PM2 msk2, PX x2, int n2, int j, int i) {
  if (msk1[(size_t)j * n1 + i] <= 0) return T(0);
  const int J2 = (j + 2) >> 1, I2 = (i + 2) >> 1;
  const size_t k = (size_t)J2 * n2 + I2;
  const int pj = j & 1, pi = i & 1;
  const T a = x2[k];
  if (!pj && !pi) return a;
  // weights c2 = {0, 1, 0.5}, c3 = {0, 1, 0.5, 0.33333334f, 0.25} selected
  // without an indexed table (which would live in scratch memory)
  auto c2 = [](int s) { return s == 2 ? T(0.5) : (s == 1 ? T(1) : T(0)); };
  if (!pj) return (a + x2[k + 1]) * c2(msk2[k] + msk2[k + 1]);
  if (!pi) return (a + x2[k + n2]) * c2(msk2[k] + msk2[k + n2]);
  const int s4 = msk2[k] + msk2[k + 1] + msk2[k + n2] + msk2[k + n2 + 1];
  const T w = s4 == 4 ? T(0.25)
                      : (s4 == 3 ? T((double)0.3333333333333333333333333333f)
                                 : (s4 == 2 ? T(0.5) : (s4 == 1 ? T(1) : T(0))));
  return w * (a + x2[k + 1] + x2[k + n2] + x2[k + n2 + 1]);
}

// interp_at in two steps: IpOps holds every operand of one fine point,
// loaded up front (no load waits on the masks; all addresses in bounds for
// fine rows/columns in [1, m1-2] x [1, n1-2]); ip_eval applies the same
// expressions as interp_at.
template <typename T>
struct IpOps {
  T a, b, c, d;
  int8_t m1, sa, sb, sc, sd;
  int par;  // bit 0: column parity, bit 1: row parity
};
template <typename T>
__device__ __forceinline__ IpOps<T> ip_load(const int8_t* msk1, int n1, const int8_t* msk2,
                                            const T* x2, int n2, int j, int i) {
  IpOps<T> o;
  const int J2 = (j + 2) >> 1, I2 = (i + 2) >> 1;
  const size_t k = (size_t)J2 * n2 + I2;
  o.m1 = msk1[(size_t)j * n1 + i];
  o.a = x2[k];
  o.b = x2[k + 1];
  o.c = x2[k + n2];
  o.d = x2[k + n2 + 1];
  o.sa = msk2[k];
  o.sb = msk2[k + 1];
  o.sc = msk2[k + n2];
  o.sd = msk2[k + n2 + 1];
  o.par = (i & 1) | ((j & 1) << 1);
  return o;
}
// Evaluated without a branch: all four parity forms are computed and the
// one of this point's parity is selected (a wavefront holds every parity,
// so branches would run all four forms one after another anyway).
template <typename T>
__device__ __forceinline__ T ip_eval(const IpOps<T>& o) {
  auto c2 = [](int s) { return s == 2 ? T(0.5) : (s == 1 ? T(1) : T(0)); };
  const int pj = o.par >> 1, pi = o.par & 1;
  const T r1 = (o.a + o.b) * c2(o.sa + o.sb);
  const T r2 = (o.a + o.c) * c2(o.sa + o.sc);
  const int s4 = o.sa + o.sb + o.sc + o.sd;
  const T w = s4 == 4 ? T(0.25)
                      : (s4 == 3 ? T((double)0.3333333333333333333333333333f)
                                 : (s4 == 2 ? T(0.5) : (s4 == 1 ? T(1) : T(0))));
  const T r3 = w * (o.a + o.b + o.c + o.d);
  const T r = pj ? (pi ? r3 : r2) : (pi ? r1 : o.a);
  return o.m1 <= 0 ? T(0) : r;
}

// No-slip wall source of interior cell (j, i) (core/fortran_operators.f90:
// 221-268), cff = 1 / (2 dx dy).  The routine scatters each edge term into
// the fluid cell beside the wall; gathered here per cell in the order the
// routine's loop adds them: the cell's own x-edge and y-edge terms, then
// the term of its east neighbour's x-edge (loop index i+1) and of its north
// neighbour's y-edge (j+1).
__device__ __forceinline__ double noslip_source_at(const int8_t* msk, const double* x, int j,
                                                   int i, int n, double cff) {
  auto M = [&](int jj, int ii) { return (int)msk[(size_t)jj * n + ii]; };
  auto X = [&](int jj, int ii) { return x[(size_t)jj * n + ii]; };
  // v at the west edge of (jj, ii), u at its south edge
  auto vx = [&](int jj, int ii) {
    return (X(jj, ii) + X(jj - 1, ii) - X(jj, ii - 2) - X(jj - 1, ii - 2)) * cff;
  };
  auto uy = [&](int jj, int ii) {
    return -(X(jj, ii) + X(jj, ii - 1) - X(jj - 2, ii) - X(jj - 2, ii - 1)) * cff;
  };
  const int mc = M(j, i), me = M(j, i + 1), mn = M(j + 1, i);
  double r = 0.;
  if (mc != 0 && M(j, i - 1) + mc == 1) r = r - vx(j, i);
  if (mc != 0 && M(j - 1, i) + mc == 1) r = r + uy(j, i);
  if (me == 0 && mc + me == 1) r = r + vx(j, i + 1);
  if (mn == 0 && mc + mn == 1) r = r - uy(j + 1, i);
  return r;
}

// Calls f(j, i) at interior (J, I) and at each of its periodic halo images
// (the positions store_img_mask32 writes).
template <typename F>
__device__ __forceinline__ void each_img(int J, int I, int m, int n, F f) {
  f(J, I);
  if (J >= 2 * NH && J < m - 2 * NH && I >= 2 * NH && I < n - 2 * NH) return;
  const int M = m - 2 * NH, N = n - 2 * NH;
  const bool jp = J + M < m && J + M >= m - NH, jm = J - M >= 0 && J - M < NH;
  const bool ip = I + N < n && I + N >= n - NH, im = I - N >= 0 && I - N < NH;
  if (jp) f(J + M, I);
  if (jm) f(J - M, I);
  if (ip) f(J, I + N);
  if (im) f(J, I - N);
  if (jp && ip) f(J + M, I + N);
  if (jp && im) f(J + M, I - N);
  if (jm && ip) f(J - M, I + N);
  if (jm && im) f(J - M, I - N);
}

// Element i of a per-instance array through a 32-bit byte offset from a
// uniform base: the load or store then takes the base from scalar registers
// and one 32-bit vector offset (arrays of one instance stay below 4 GiB).
__device__ __forceinline__ uint32_t at32(int j, int i, int n) {
  return (uint32_t)j * (uint32_t)n + (uint32_t)i;
}
template <typename T>
__device__ __forceinline__ T ld32(const T* p, uint32_t i) {
  return *(const T*)((const char*)p + (size_t)(uint32_t)(i * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void st32(T* p, uint32_t i, T v) {
  *(T*)((char*)p + (size_t)(uint32_t)(i * (uint32_t)sizeof(T))) = v;
}

// periodic image index of a halo row/column
__device__ __forceinline__ int img(int j, int m) {
  const int M = m - 2 * NH;
  return j < NH ? j + M : (j >= m - NH ? j - M : j);
}

// Store at interior (J,I) and its halo images, multiplying each stored copy
// by the mask at its own location when mk != nullptr (psi * mskp of
// core/operators.py:481 applied to a filled array).  A row has an image
// above (J + M) when J < 2 NH and one below (J - M) when J >= m - 2 NH, both
// for interiors narrower than 2 NH; likewise for columns.  Straight-line
// predicated stores.
template <typename T, typename PA>
__device__ __forceinline__ void store_img_mask(PA a, const int8_t* mk, int J, int I, T v, int m,
                                               int n) {
  auto put = [&](int j, int i) {
    const size_t c = at(j, i, n);
    a[c] = mk ? v * T(mk[c]) : v;
  };
  put(J, I);
  if (J >= 2 * NH && J < m - 2 * NH && I >= 2 * NH && I < n - 2 * NH) return;  // no image
  const int M = m - 2 * NH, N = n - 2 * NH;
  const bool jp = J + M < m && J + M >= m - NH, jm = J - M >= 0 && J - M < NH;
  const bool ip = I + N < n && I + N >= n - NH, im = I - N >= 0 && I - N < NH;
  if (jp) put(J + M, I);
  if (jm) put(J - M, I);
  if (ip) put(J, I + N);
  if (im) put(J, I - N);
  if (jp && ip) put(J + M, I + N);
  if (jp && im) put(J + M, I - N);
  if (jm && ip) put(J - M, I + N);
  if (jm && im) put(J - M, I - N);
}

// store_img_mask with 32-bit offsets (see ld32)
template <typename T>
__device__ __forceinline__ void store_img_mask32(T* a, const int8_t* mk, int J, int I, T v, int m,
                                                 int n) {
  auto put = [&](int j, int i) {
    const uint32_t c = at32(j, i, n);
    st32(a, c, mk ? v * T(ld32(mk, c)) : v);
  };
  put(J, I);
  if (J >= 2 * NH && J < m - 2 * NH && I >= 2 * NH && I < n - 2 * NH) return;  // no image
  const int M = m - 2 * NH, N = n - 2 * NH;
  const bool jp = J + M < m && J + M >= m - NH, jm = J - M >= 0 && J - M < NH;
  const bool ip = I + N < n && I + N >= n - NH, im = I - N >= 0 && I - N < NH;
  if (jp) put(J + M, I);
  if (jm) put(J - M, I);
  if (ip) put(J, I + N);
  if (im) put(J, I - N);
  if (jp && ip) put(J + M, I + N);
  if (jp && im) put(J + M, I - N);
  if (jm && ip) put(J - M, I + N);
  if (jm && im) put(J - M, I - N);
}

}  // namespace f2d
