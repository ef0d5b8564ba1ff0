// f2d_ops.hip -- stencil, transfer and reduction kernels of the Fluid2d
// timestep path for CDNA4 (gfx950).
//
// One work-item per output cell; workgroups of 64 x 4 (one wavefront per
// row segment, so every load and store of a row is a coalesced 512-byte
// sweep in FP64).  These are the building blocks; the fused/LDS-tiled
// variants live in f2d_fused.hip.  Each kernel cites the reference routine
// whose arithmetic (operation order included) it reproduces.
#include "f2d_ops.h"
#include "f2d_dev.h"

namespace f2d {

// ---------------------------------------------------------------------------
// advection
// ---------------------------------------------------------------------------
// (UpParams, split_speed and the float32-rounded weights W32: f2d_dev.h)

// Flux through the edge between cell c and c+s (s = 1: east edge, s = n:
// north edge).  core/fortran_advection.f90:73-146.
template <typename T>
__device__ __forceinline__ T up_edge_flux(const int8_t* mk, const T* q, T vel, ptrdiff_t s,
                                          const UpParams<T>& p) {
  if (mk[0] + mk[s] != 2) return T(0);
  const T d1 = T(W32(1, 30)), d2 = T(-W32(13, 60)), d3 = T(W32(47, 60)), d4 = T(W32(9, 20)),
          d5 = T(-W32(1, 20));
  const T c1 = T(-W32(1, 6)), c2 = T(W32(5, 6)), c3 = T(W32(2, 6));
  const T UU = split_speed(vel, p);
  const T up = T(0.5) * (vel + UU), um = T(0.5) * (vel - UU);
  const int m_2 = mk[-2 * s], m_1 = mk[-s], m0 = mk[0], m1 = mk[s], m2 = mk[2 * s], m3 = mk[3 * s];
  const int s5p = m_2 + m_1 + m0 + m1 + m2, s3p = m_1 + m0 + m1;
  const int s5m = m_1 + m0 + m1 + m2 + m3, s3m = m0 + m1 + m2;
  T qp, qm;
  if (s5p == 5 && p.order == 5)
    qp = d1 * q[-2 * s] + d2 * q[-s] + d3 * q[0] + d4 * q[s] + d5 * q[2 * s];
  else if (s3p == 3 && p.order >= 3)
    qp = c1 * q[-s] + c2 * q[0] + c3 * q[s];
  else
    qp = q[0];
  if (s5m == 5 && p.order == 5)
    qm = d5 * q[-s] + d4 * q[0] + d3 * q[s] + d2 * q[2 * s] + d1 * q[3 * s];
  else if (s3m == 3 && p.order >= 3)
    qm = c3 * q[0] + c2 * q[s] + c1 * q[2 * s];
  else
    qm = q[s];
  return up * qp + um * qm;
}

// core/fortran_advection.f90:2-165; edge-flux stores of
// core/fortran_fluxes.f90:149-166 when xflx != nullptr.
template <typename T>
__global__ void __launch_bounds__(256) k_adv_upwind(const int8_t* __restrict__ msk,
                                                    const T* __restrict__ x, T* __restrict__ y,
                                                    const T* __restrict__ u,
                                                    const T* __restrict__ v, T* xflx, T* yflx,
                                                    UpParams<T> p, int m, int n) {
  const int I = NH + blockIdx.x * blockDim.x + threadIdx.x;
  const int J = NH + blockIdx.y * blockDim.y + threadIdx.y;
  if (I >= n - NH || J >= m - NH) return;
  const size_t c = at(J, I, n);
  const T fe = up_edge_flux(msk + c, x + c, u[c], 1, p);
  const T fw = up_edge_flux(msk + c - 1, x + c - 1, u[c - 1], 1, p);
  const T fn = up_edge_flux(msk + c, x + c, v[c], (ptrdiff_t)n, p);
  const T fs = up_edge_flux(msk + c - n, x + c - n, v[c - n], (ptrdiff_t)n, p);
  y[c] = -p.zdx * (fe - fw) - p.zdy * (fn - fs);
  if (xflx) {
    xflx[c] = fe;
    yflx[c] = fn;
    if (I == NH) xflx[c - 1] = fw;
    if (J == NH) yflx[c - n] = fs;
  }
}

// core/fortran_advection.f90:169-284 (allow6) / core/fortran_fluxes.f90:174-279
template <typename T>
__device__ __forceinline__ T ce_edge_flux(const int8_t* mk, const T* q, T vel, ptrdiff_t s,
                                          int order, bool allow6) {
  const int s2 = mk[0] + mk[s];
  if (s2 != 2) return T(0);
  const T e1 = T(W32(1, 60)), e2 = T(-W32(2, 15)), e3 = T(W32(37, 60));
  const T d1 = T(-W32(1, 12)), d2 = T(W32(7, 12));
  const T c1 = T(0.5);
  const int s4 = mk[-s] + mk[0] + mk[s] + mk[2 * s];
  const int s6 = mk[-2 * s] + s4 + mk[3 * s];
  T qp = T(0);
  if (allow6 && s6 == 6 && order == 6) {
    qp = e1 * (q[-2 * s] + q[3 * s]) + e2 * (q[-s] + q[2 * s]);
    qp = qp + e3 * (q[0] + q[s]);
  } else if (s4 == 4 && order == 4) {
    qp = d1 * (q[-s] + q[2 * s]) + d2 * (q[0] + q[s]);
  } else if (order >= 2) {
    qp = c1 * (q[0] + q[s]);
  }
  return vel * qp;
}

template <typename T>
__global__ void __launch_bounds__(256) k_adv_centered(const int8_t* __restrict__ msk,
                                                      const T* __restrict__ x, T* __restrict__ y,
                                                      const T* __restrict__ u,
                                                      const T* __restrict__ v, T* xflx, T* yflx,
                                                      T zdx, T zdy, int order, int m, int n) {
  const int I = NH + blockIdx.x * blockDim.x + threadIdx.x;
  const int J = NH + blockIdx.y * blockDim.y + threadIdx.y;
  if (I >= n - NH || J >= m - NH) return;
  const size_t c = at(J, I, n);
  const bool allow6 = xflx == nullptr;
  const T fe = ce_edge_flux(msk + c, x + c, u[c], 1, order, allow6);
  const T fw = ce_edge_flux(msk + c - 1, x + c - 1, u[c - 1], 1, order, allow6);
  const T fn = ce_edge_flux(msk + c, x + c, v[c], (ptrdiff_t)n, order, allow6);
  const T fs = ce_edge_flux(msk + c - n, x + c - n, v[c - n], (ptrdiff_t)n, order, allow6);
  y[c] = -zdx * (fe - fw) - zdy * (fn - fs);
  if (xflx) {
    xflx[c] = fe;
    yflx[c] = fn;
    if (I == NH) xflx[c - 1] = fw;
    if (J == NH) yflx[c - n] = fs;
  }
}

// ---------------------------------------------------------------------------
// halo and grid operators
// ---------------------------------------------------------------------------
// core/gmg/fortran_multigrid.f90:365-412
template <typename T>
__global__ void k_fillhalo(T* x, int m, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= n || j >= m) return;
  const int M = m - 2 * NH, N = n - 2 * NH;
  const int sj = j < NH ? j + M : (j >= m - NH ? j - M : j);
  const int si = i < NH ? i + N : (i >= n - NH ? i - N : i);
  if (sj != j || si != i) x[at(j, i, n)] = x[at(sj, si, n)];
}

// core/fortran_operators.f90:44-64 (rows/cols [0, m-2] x [0, n-2])
template <typename T>
__global__ void k_celltocorner(const T* __restrict__ xr, T* __restrict__ xp, int m, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= n - 1 || j >= m - 1) return;
  const size_t c = at(j, i, n);
  xp[c] = T(0.25) * (xr[c] + xr[c + 1] + xr[c + n] + xr[c + n + 1]);
}

// core/fortran_operators.f90:2-39 (rows/cols [1, m-2])
template <typename T>
__global__ void k_orthogradient(const int8_t* __restrict__ msk, const T* __restrict__ psi, T zdx,
                                T zdy, T* __restrict__ u, T* __restrict__ v, int m, int n) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= n - 1 || j >= m - 1) return;
  const size_t c = at(j, i, n);
  u[c] = (msk[c] + msk[c + 1] == 2) ? zdy * (psi[c - n] - psi[c]) : T(0);
  v[c] = (msk[c] + msk[c + n] == 2) ? zdx * (psi[c] - psi[c - 1]) : T(0);
}

// core/fortran_operators.f90:125-156, then fillhalo (core/operators.py:300)
template <typename T>
__global__ void k_add_diffusion(const int8_t* __restrict__ msk, const T* __restrict__ t, T coef,
                                T* __restrict__ dt, int m, int n) {
  const int I = NH + blockIdx.x * blockDim.x + threadIdx.x;
  const int J = NH + blockIdx.y * blockDim.y + threadIdx.y;
  if (I >= n - NH || J >= m - NH) return;
  const size_t c = at(J, I, n);
  T val = dt[c];
  if (msk[c] == 1) {
    const T tc = t[c];
    T s = T(msk[c - 1]) * (t[c - 1] - tc);
    s = s + T(msk[c + 1]) * (t[c + 1] - tc);
    s = s + T(msk[c - n]) * (t[c - n] - tc);
    s = s + T(msk[c + n]) * (t[c + n] - tc);
    val = val + coef * s;
  }
  store_img(dt, J, I, val, m, n);
}

// y *= msk; core/fortran_operators.f90:330-381; fillhalo
// (core/operators.py:303-314)
template <typename T>
__global__ void k_torque(const int8_t* __restrict__ msk, const T* __restrict__ b, T coef,
                         T* __restrict__ dw, int m, int n) {
  const int I = NH + blockIdx.x * blockDim.x + threadIdx.x;
  const int J = NH + blockIdx.y * blockDim.y + threadIdx.y;
  if (I >= n - NH || J >= m - NH) return;
  const size_t c = at(J, I, n);
  T val = dw[c] * T(msk[c]);
  int ml = msk[c] + msk[c - 1], mr = msk[c + 1] + msk[c];
  ml = ml < 1 ? 1 : ml;
  mr = mr < 1 ? 1 : mr;
  if (ml + mr == 4) val = val + (b[c + 1] - b[c - 1]) * coef * T(msk[c]);
  store_img(dw, J, I, val, m, n);
}

// psi = psi * mskp over the whole array (core/operators.py:481)
template <typename T>
__global__ void k_mask_mul(T* x, const int8_t* __restrict__ mk, size_t cnt) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < cnt) x[q] = x[q] * T(mk[q]);
}

// ---------------------------------------------------------------------------
// multigrid
// ---------------------------------------------------------------------------
// A is 5 planes: A_k(j,i) couples (j,i) to its lower neighbour k (SW, S,
// SE, W, centre); the upper couplings are read from the neighbours by
// symmetry (core/gmg/fortran_multigrid.f90:68-82).
template <typename T>
__device__ __forceinline__ T offdiag(const T* A, size_t P, const T* x, size_t c, int n) {
  const T* A1 = A;
  const T* A2 = A + P;
  const T* A3 = A + 2 * P;
  const T* A4 = A + 3 * P;
  T s = A1[c] * x[c - n - 1];
  s = s + A2[c] * x[c - n];
  s = s + A3[c] * x[c - n + 1];
  s = s + A4[c] * x[c - 1];
  s = s + A4[c + 1] * x[c + 1];
  s = s + A3[c + n - 1] * x[c + n - 1];
  s = s + A2[c + n] * x[c + n];
  s = s + A1[c + n + 1] * x[c + n + 1];
  return s;
}

template <typename T>
__device__ __forceinline__ T jacobi_at(const int8_t* msk, const T* A, size_t P, const T* x,
                                       const T* b, T c1, T c2, size_t c, int n) {
  if (msk[c] == 0) return T(0);
  const T c3 = c1 / fabs(A[4 * P + c]);
  return x[c] * c2 + c3 * (offdiag(A, P, x, c, n) - b[c]);
}

// First damped-Jacobi sweep of smoothtwicewithA
// (core/gmg/fortran_multigrid.f90:62-86) on rows/cols [nh-1, m-nh] -- the
// ring the second sweep reads.
template <typename T>
__global__ void __launch_bounds__(256) k_jacobi1(const int8_t* __restrict__ msk,
                                                 const T* __restrict__ A,
                                                 const T* __restrict__ x,
                                                 const T* __restrict__ b, T* __restrict__ t,
                                                 T c1, T c2, int m, int n) {
  const int i = NH - 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = NH - 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > n - NH || j > m - NH) return;
  const size_t c = at(j, i, n);
  t[c] = jacobi_at(msk, A, (size_t)m * n, x, b, c1, c2, c, n);
}

// Second sweep (core/gmg/fortran_multigrid.f90:92-118) on the interior,
// plus the halo fill of core/gmg/level.py:365.
template <typename T>
__global__ void __launch_bounds__(256) k_jacobi2(const int8_t* __restrict__ msk,
                                                 const T* __restrict__ A,
                                                 const T* __restrict__ t,
                                                 const T* __restrict__ b, T* __restrict__ x,
                                                 T c1, T c2, int m, int n) {
  const int I = NH + blockIdx.x * blockDim.x + threadIdx.x;
  const int J = NH + blockIdx.y * blockDim.y + threadIdx.y;
  if (I >= n - NH || J >= m - NH) return;
  const size_t c = at(J, I, n);
  store_img(x, J, I, jacobi_at(msk, A, (size_t)m * n, t, b, c1, c2, c, n), m, n);
}

// core/gmg/fortran_multigrid.f90:320-362 + fillhalo
template <typename T>
__global__ void __launch_bounds__(256) k_residual(const int8_t* __restrict__ msk,
                                                  const T* __restrict__ A,
                                                  const T* __restrict__ x,
                                                  const T* __restrict__ b, T* __restrict__ r,
                                                  int m, int n) {
  const int I = NH + blockIdx.x * blockDim.x + threadIdx.x;
  const int J = NH + blockIdx.y * blockDim.y + threadIdx.y;
  if (I >= n - NH || J >= m - NH) return;
  const size_t c = at(J, I, n), P = (size_t)m * n;
  T val = T(0);
  if (msk[c] != 0) {
    const T *A1 = A, *A2 = A + P, *A3 = A + 2 * P, *A4 = A + 3 * P, *A5 = A + 4 * P;
    T s = b[c] - A1[c] * x[c - n - 1];
    s = s - A2[c] * x[c - n];
    s = s - A3[c] * x[c - n + 1];
    s = s - A4[c] * x[c - 1];
    s = s - A5[c] * x[c];
    s = s - A4[c + 1] * x[c + 1];
    s = s - A3[c + n - 1] * x[c + n - 1];
    s = s - A2[c + n] * x[c + n];
    s = s - A1[c + n + 1] * x[c + n + 1];
    val = s;
  }
  store_img(r, J, I, val, m, n);
}

// core/gmg/fortran_multigrid.f90:501-546 + fillhalo on the coarse grid.
// Coarse (J2,I2) sits on fine (2*J2+1-nh, 2*I2+1-nh).
template <typename T>
__global__ void __launch_bounds__(256) k_restrict(const int8_t* __restrict__ msk2,
                                                  const T* __restrict__ x1, T* __restrict__ x2,
                                                  int m2, int n2, int n1) {
  const int I2 = NH + blockIdx.x * blockDim.x + threadIdx.x;
  const int J2 = NH + blockIdx.y * blockDim.y + threadIdx.y;
  if (I2 >= n2 - NH || J2 >= m2 - NH) return;
  T val = T(0);
  if (msk2[at(J2, I2, n2)] != 0) {
    const size_t c = at(2 * J2 + 1 - NH, 2 * I2 + 1 - NH, n1);
    val = T(0.25) * x1[c] + T(0.125) * (x1[c - 1] + x1[c + 1] + x1[c - n1] + x1[c + n1]) +
          T(0.0625) * (x1[c - n1 - 1] + x1[c - n1 + 1] + x1[c + n1 - 1] + x1[c + n1 + 1]);
  }
  store_img(x2, J2, I2, val, m2, n2);
}

// core/gmg/fortran_multigrid.f90:415-498 (nh = 3): coarse (J2,I2) for J2 in
// [1, m2-3] writes fine rows 2*J2-2, 2*J2-1 and columns 2*I2-2, 2*I2-1.
// ACC: x1 += I x2 (the x += r of core/gmg/hierarchy.py:125).
template <typename T, bool ACC>
__global__ void __launch_bounds__(256) k_interpolate(const int8_t* __restrict__ msk1,
                                                     const int8_t* __restrict__ msk2,
                                                     const T* __restrict__ x2,
                                                     T* __restrict__ x1, int m2, int n2,
                                                     int n1) {
  const int I2 = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int J2 = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (I2 > n2 - 3 || J2 > m2 - 3) return;
  const T c2[3] = {T(0), T(1), T(0.5)};
  const T c3[5] = {T(0), T(1), T(0.5), T((double)0.3333333333333333333333333333f), T(0.25)};
  const size_t k = at(J2, I2, n2);
  const T a = x2[k], bE = x2[k + 1], bN = x2[k + n2], bNE = x2[k + n2 + 1];
  const int ma = msk2[k], mE = msk2[k + 1], mN = msk2[k + n2], mNE = msk2[k + n2 + 1];
  const size_t f = at(2 * J2 - 2, 2 * I2 - 2, n1);
  const T v00 = msk1[f] > 0 ? a : T(0);
  const T v01 = msk1[f + 1] > 0 ? (a + bE) * c2[ma + mE] : T(0);
  const T v10 = msk1[f + n1] > 0 ? (a + bN) * c2[ma + mN] : T(0);
  const T v11 = msk1[f + n1 + 1] > 0 ? c3[ma + mE + mN + mNE] * (a + bE + bN + bNE) : T(0);
  if (ACC) {
    x1[f] = x1[f] + v00;
    x1[f + 1] = x1[f + 1] + v01;
    x1[f + n1] = x1[f + n1] + v10;
    x1[f + n1 + 1] = x1[f + n1 + 1] + v11;
  } else {
    x1[f] = v00;
    x1[f + 1] = v01;
    x1[f + n1] = v10;
    x1[f + n1 + 1] = v11;
  }
}

// core/gmg/fortran_multigrid.f90:706-811 + fillhalo of the 9 planes.
// Af, Ac: 9 planes.  coef(a,b) = 0.5 * 2^-|a| * 2^-|b| (line 732).
__global__ void k_coarsen(const double* __restrict__ Af, const int8_t* __restrict__ msk1,
                          const int8_t* __restrict__ msk2, double* __restrict__ Ac, int m1,
                          int n1, int m2, int n2) {
  const int I2 = NH + blockIdx.x * blockDim.x + threadIdx.x;
  const int J2 = NH + blockIdx.y * blockDim.y + threadIdx.y;
  if (I2 >= n2 - NH || J2 >= m2 - NH) return;
  const size_t pf = (size_t)m1 * n1, pc = (size_t)m2 * n2;
  const double cw[3] = {0.25, 0.5, 0.25};  // cw[a]*cw[b]*2 = coef
  if (msk2[at(J2, I2, n2)] != 1) {
    for (int l = 0; l < 9; l++) store_img(Ac + l * pc, J2, I2, 0.0, m2, n2);
    return;
  }
  const int J1 = 2 * J2 + 1 - NH, I1 = 2 * I2 + 1 - NH;
  for (int l = 0; l < 9; l++) {
    const int di2 = l % 3 - 1, dj2 = l / 3 - 1;
    double z5[5][5];
    for (int a = 0; a < 5; a++)
      for (int b = 0; b < 5; b++) z5[a][b] = 0.;
    for (int kj = -1; kj <= 1; kj++)
      for (int ki = -1; ki <= 1; ki++) {
        const int ii = 2 * di2 + ki, jj = 2 * dj2 + kj;
        if (ii >= -2 && ii <= 2 && jj >= -2 && jj <= 2 && msk1[at(J1 + jj, I1 + ii, n1)] == 1)
          z5[jj + 2][ii + 2] = 2. * (2. * cw[kj + 1] * cw[ki + 1]);
      }
    double z3[3][3];
    for (int jj = -1; jj <= 1; jj++)
      for (int ii = -1; ii <= 1; ii++) {
        double acc = 0.;
        const size_t f = at(J1 + jj, I1 + ii, n1);
        if (msk1[f] == 1) {
          for (int kj = -1; kj <= 1; kj++)
            for (int ki = -1; ki <= 1; ki++) {
              const int k = (ki + 1) + (kj + 1) * 3;
              acc = acc + Af[k * pf + f] * z5[jj + kj + 2][ii + ki + 2];
            }
        }
        z3[jj + 1][ii + 1] = acc;
      }
    double w = 0.;
    for (int jj = -1; jj <= 1; jj++)
      for (int ii = -1; ii <= 1; ii++)
        w = w + 0.5 * (2. * cw[jj + 1] * cw[ii + 1]) * z3[jj + 1][ii + 1];
    store_img(Ac + l * pc, J2, I2, w, m2, n2);
  }
}

// Level-0 operator (core/gmg/level.py:261-302): st_k * coef * msk_nb * msk,
// then fillhalo.  st: 9 stencil weights (host-computed), coef = 1/(dx dy).
__global__ void k_afinest(const int8_t* __restrict__ msk, double* __restrict__ A, St9 st,
                          double coef, int m, int n) {
  const int I = NH + blockIdx.x * blockDim.x + threadIdx.x;
  const int J = NH + blockIdx.y * blockDim.y + threadIdx.y;
  if (I >= n - NH || J >= m - NH) return;
  const size_t P = (size_t)m * n, c = at(J, I, n);
  for (int k = 0; k < 9; k++) {
    const int di = k % 3 - 1, dj = k / 3 - 1;
    const double v = st.w[k] * coef * (double)msk[at(J + dj, I + di, n)] * (double)msk[c];
    store_img(A + k * P, J, I, v, m, n);
  }
}

// Corner mask (core/operators.py:59-67): 1 where celltocorner(msk) == 1.
__global__ void k_corner_mask(const int8_t* __restrict__ msk, int8_t* __restrict__ mskp, int m,
                              int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= n || j >= m) return;
  const size_t c = at(j, i, n);
  int8_t v = 0;
  if (i < n - 1 && j < m - 1) {
    const double s = 0.25 * ((double)msk[c] + (double)msk[c + 1] + (double)msk[c + n] +
                             (double)msk[c + n + 1]);
    v = s == 1.0 ? 1 : 0;
  }
  mskp[c] = v;
}

// Coarse mask (core/gmg/level.py:233-236): full-weighting restriction of
// the fine mask (coarse mask of ones, so every point), fillhalo, > 0.5.
__global__ void k_coarse_mask(const int8_t* __restrict__ msk1, int8_t* __restrict__ msk2,
                              int m2, int n2, int n1) {
  const int I2 = NH + blockIdx.x * blockDim.x + threadIdx.x;
  const int J2 = NH + blockIdx.y * blockDim.y + threadIdx.y;
  if (I2 >= n2 - NH || J2 >= m2 - NH) return;
  const size_t c = at(2 * J2 + 1 - NH, 2 * I2 + 1 - NH, n1);
  auto X = [&](size_t q) { return (double)msk1[q]; };
  const double val = 0.25 * X(c) + 0.125 * (X(c - 1) + X(c + 1) + X(c - n1) + X(c + n1)) +
                     0.0625 * (X(c - n1 - 1) + X(c - n1 + 1) + X(c + n1 - 1) + X(c + n1 + 1));
  store_img(msk2, J2, I2, (int8_t)(val <= 0.5 ? 0 : 1), m2, n2);
}

// ---------------------------------------------------------------------------
// elementwise
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill_value(T* x, T v, size_t cnt) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < cnt) x[q] = v;
}
template <typename T>
__global__ void k_copy(T* __restrict__ d, const T* __restrict__ s, size_t cnt) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < cnt) d[q] = s[q];
}
// x += y (core/gmg/hierarchy.py:171)
template <typename T>
__global__ void k_add_inplace(T* __restrict__ x, const T* __restrict__ y, size_t cnt) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < cnt) x[q] = x[q] + y[q];
}
// x -= s[0] (core/operators.py:477)
template <typename T>
__global__ void k_sub_dev_scalar(T* __restrict__ x, const T* __restrict__ s, size_t cnt) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < cnt) x[q] = x[q] - s[0];
}

// No-slip wall source (noslip_source_at, core/fortran_operators.f90:221-268)
// on the interior cells; the halo is zeroed (the caller's periodic fill
// rewrites it).
__global__ void k_noslip_source(const int8_t* __restrict__ msk, const double* __restrict__ x,
                                double* __restrict__ y, double cff, int m, int n) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)m * n) return;
  const int j = (int)(q / n), i = (int)(q - (size_t)j * n);
  if (j < NH || j >= m - NH || i < NH || i >= n - NH) {
    y[q] = 0.;
    return;
  }
  y[q] = noslip_source_at(msk, x, j, i, n, cff);
}
// s -= d[0]*mskbc (core/operators.py:275-278)
__global__ void k_sub_mean_mask(double* __restrict__ s, const int8_t* __restrict__ mk,
                                const double* __restrict__ d, size_t cnt) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < cnt) s[q] = s[q] - d[0] * (double)mk[q];
}
// s -= (px*xr0 + py*yr0)*mskbc, px = d[0]/x2bc, py = d[1]/y2bc
// (core/operators.py:280-287)
__global__ void k_noslip_momentum(double* __restrict__ s, const int8_t* __restrict__ mk,
                                  const double* __restrict__ xr0, const double* __restrict__ yr0,
                                  const double* __restrict__ d, double x2bc, double y2bc,
                                  size_t cnt) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= cnt) return;
  const double px = d[0] / x2bc, py = d[1] / y2bc;
  s[q] = s[q] - (px * xr0[q] + py * yr0[q]) * (double)mk[q];
}
// x -= y
__global__ void k_sub_inplace(double* __restrict__ x, const double* __restrict__ y, size_t cnt) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < cnt) x[q] = x[q] - y[q];
}
// x /= c (core/euler.py:112)
__global__ void k_div_scalar(double* __restrict__ x, double c, size_t cnt) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < cnt) x[q] = x[q] / c;
}

// Flux diagnostics (core/fluxes.py:118-170).
// out = 0.5*(u + np.roll(u, 1, axis)): the west (axis 1) or south (axis 0)
// neighbour, periodic over the whole array
template <typename T>
__global__ void k_roll_mean(const T* __restrict__ u, T* __restrict__ out, int axis, int m, int n) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)m * n) return;
  const int j = (int)(q / n), i = (int)(q - (size_t)j * n);
  const size_t w = axis == 1 ? (size_t)j * n + (i == 0 ? n - 1 : i - 1)
                             : (size_t)(j == 0 ? m - 1 : j - 1) * n + i;
  out[q] = T(0.5) * (u[q] + u[w]);
}
// x *= -1
template <typename T>
__global__ void k_negate(T* __restrict__ x, size_t cnt) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < cnt) x[q] = -x[q];
}
// rev = cff*(a + sign*b), irr = cff*(a - sign*b), sign = +-1
template <typename T>
__global__ void k_flux_split(const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ rev,
                             T* __restrict__ irr, T sign, T cff, size_t cnt) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= cnt) return;
  const T sb = sign * b[q];
  rev[q] = cff * (a[q] + sb);
  irr[q] = cff * (a[q] - sb);
}

// RK3_SSP / RK3 combinations (core/timescheme.py:152-180), NumPy order:
// scalar factor, array sums left to right, scaled add.
// mode 0: xs = x + c*k0             (RK3_SSP stage 1: c = dt; RK3: c = dt/3)
// mode 1: xs = x + c*(k0 + k1)      (RK3_SSP stage 2: c = 0.25 dt)
// mode 2: x  = x + c*((k0+k1)+4 k2) (RK3_SSP final:   c = dt/6)
// mode 3: xs = x + c*k1             (RK3 stage 2:     c = 0.5 dt)
// mode 4: x  = x + c*k2             (RK3 final:       c = dt)
template <typename T>
__global__ void k_rk(int mode, T* __restrict__ out, const T* __restrict__ x,
                     const T* __restrict__ k0, const T* __restrict__ k1,
                     const T* __restrict__ k2, T c, size_t cnt) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= cnt) return;
  T r;
  switch (mode) {
    case 0: r = x[q] + c * k0[q]; break;
    case 1: r = x[q] + c * (k0[q] + k1[q]); break;
    case 2: r = x[q] + c * ((k0[q] + k1[q]) + T(4) * k2[q]); break;
    case 3: r = x[q] + c * k1[q]; break;
    default: r = x[q] + c * k2[q]; break;
  }
  out[q] = r;
}

// ---------------------------------------------------------------------------
// reductions (deterministic two-pass: per-workgroup partials in a fixed
// tree order, then one workgroup folds the partials in order)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// per-workgroup sums of NV quantities over interior cells: partial[blk*NV+k]
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], const bool (&is_max)[NV],
                                                   double* partial) {
  __shared__ double sh[NV][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int k = 0; k < NV; k++) {
    double w = is_max[k] ? wave_max(v[k]) : wave_sum(v[k]);
    if (lane == 0) sh[k][wid] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < NV; k++) {
      double a = sh[k][0];
      for (int w = 1; w < (int)(blockDim.x >> 6); w++)
        a = is_max[k] ? fmax(a, sh[k][w]) : a + sh[k][w];
      partial[(size_t)blockIdx.x * NV + k] = a;
    }
  }
}

// sum of x^2 over interior with msk != 0 (computenorm)
__global__ void __launch_bounds__(256) k_norm2_partial(const int8_t* __restrict__ msk,
                                                       const double* __restrict__ x, int m,
                                                       int n, double* partial) {
  const int N = n - 2 * NH;
  const size_t tot = (size_t)(m - 2 * NH) * N;
  double acc = 0.;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t c = at(NH + (int)(q / N), NH + (int)(q % N), n);
    if (msk[c] != 0) acc += x[c] * x[c];
  }
  double v[1] = {acc};
  const bool mx[1] = {false};
  block_reduce_store<1>(v, mx, partial);
}

// diagnostics: {maxu, ke, sum w, sum w^2, d4, d5, d6, d7}
// euler: d4 = dot(w,xr), d5 = dot(w,yr), d6 = sum psi, d7 = dot(w,source)
// boussinesq: d4 = sum b*yr, d5 = sum b, d6 = sum b^2 (msk==1), d7 = 0
__global__ void __launch_bounds__(256) k_diag_partial(const int8_t* __restrict__ msk,
                                                      const double* __restrict__ s, int kind,
                                                      const double* __restrict__ xr,
                                                      const double* __restrict__ yr, int m,
                                                      int n, double* partial) {
  const int N = n - 2 * NH;
  const size_t tot = (size_t)(m - 2 * NH) * N, P = (size_t)m * n;
  const double *w = s, *psi = s + P, *u = s + 2 * P, *v = s + 3 * P, *f4 = s + 4 * P;
  double a[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t c = at(NH + (int)(q / N), NH + (int)(q % N), n);
    const int mk = msk[c];
    if (mk == 1) {
      const double zu = u[c] * u[c] + u[c - 1] * u[c - 1];
      const double zv = v[c] * v[c] + v[c - n] * v[c - n];
      const double um = fabs(u[c] + u[c - 1]), vm = fabs(v[c] + v[c - n]);
      a[0] = fmax(a[0], um + vm);
      a[1] += zu + zv;
      a[2] += w[c];
      a[3] += w[c] * w[c];
    }
    if (kind == 0) {
      if (mk != 0) {
        a[4] += w[c] * xr[c];
        a[5] += w[c] * yr[c];
        a[6] += psi[c];
        a[7] += w[c] * f4[c];
      }
    } else {
      if (mk != 0) a[4] += f4[c] * yr[c];
      if (mk == 1) {
        a[5] += f4[c];
        a[6] += f4[c] * f4[c];
      }
    }
  }
  const bool mx[8] = {true, false, false, false, false, false, false, false};
  block_reduce_store<8>(a, mx, partial);
}

// kemaxu alone: {ke, maxu}
__global__ void __launch_bounds__(256) k_kemaxu_partial(const int8_t* __restrict__ msk,
                                                        const double* __restrict__ u,
                                                        const double* __restrict__ v, int m,
                                                        int n, double* partial) {
  const int N = n - 2 * NH;
  const size_t tot = (size_t)(m - 2 * NH) * N;
  double a[2] = {0., 0.};
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t c = at(NH + (int)(q / N), NH + (int)(q % N), n);
    if (msk[c] == 1) {
      const double zu = u[c] * u[c] + u[c - 1] * u[c - 1];
      const double zv = v[c] * v[c] + v[c - n] * v[c - n];
      a[0] += zu + zv;
      a[1] = fmax(a[1], fabs(u[c] + u[c - 1]) + fabs(v[c] + v[c - n]));
    }
  }
  const bool mx[2] = {false, true};
  block_reduce_store<2>(a, mx, partial);
}

// {sum s*a, sum s*b} over the interior where msk != 0
// (core/fortran_diag.f90:3-30, computedotprod, twice)
__global__ void __launch_bounds__(256) k_dot2_partial(const int8_t* __restrict__ msk,
                                                      const double* __restrict__ s,
                                                      const double* __restrict__ a,
                                                      const double* __restrict__ b, int m, int n,
                                                      double* partial) {
  const int N = n - 2 * NH;
  const size_t tot = (size_t)(m - 2 * NH) * N;
  double acc[2] = {0., 0.};
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t c = at(NH + (int)(q / N), NH + (int)(q % N), n);
    if (msk[c] != 0) {
      acc[0] += s[c] * a[c];
      acc[1] += s[c] * b[c];
    }
  }
  const bool mx[2] = {false, false};
  block_reduce_store<2>(acc, mx, partial);
}

// The three masked sums of core/fortran_diag.f90 over the interior:
//   KIND 0 computedotprod (:3-30)     {sum x*y}        where msk != 0
//   KIND 1 computesum (:63-90)        {sum x}          where msk != 0
//   KIND 2 computesumandnorm (:93-120){sum x, sum x^2} where msk == 1
template <int KIND>
__global__ void __launch_bounds__(256) k_msum_partial(const int8_t* __restrict__ msk,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ y, int m, int n,
                                                      double* partial) {
  const int N = n - 2 * NH;
  const size_t tot = (size_t)(m - 2 * NH) * N;
  constexpr int NV = KIND == 2 ? 2 : 1;
  double acc[NV];
  for (int k = 0; k < NV; k++) acc[k] = 0.;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t c = at(NH + (int)(q / N), NH + (int)(q % N), n);
    const int8_t mk = msk[c];
    if constexpr (KIND == 0) {
      if (mk != 0) acc[0] += x[c] * y[c];
    } else if constexpr (KIND == 1) {
      if (mk != 0) acc[0] += x[c];
    } else {
      if (mk == 1) {
        acc[0] += x[c];
        acc[NV - 1] += x[c] * x[c];
      }
    }
  }
  bool mx[NV];
  for (int k = 0; k < NV; k++) mx[k] = false;
  block_reduce_store<NV>(acc, mx, partial);
}

// fold nblk partial rows of NV values in block order; scale[k] multiplies
// result k (diagnostic constants 0.25 / 0.5).
__global__ void k_fold(const double* __restrict__ partial, int nblk, int nv, MaxMask mm,
                       Scale8 sc, double* __restrict__ out) {
  if (threadIdx.x >= (unsigned)nv) return;
  const int k = threadIdx.x;
  const bool ismax = (mm.bits >> k) & 1;
  double a = partial[k];
  for (int b = 1; b < nblk; b++) a = ismax ? fmax(a, partial[(size_t)b * nv + k]) : a + partial[(size_t)b * nv + k];
  out[k] = a * sc.v[k];
}

// NumPy's pairwise sum of a contiguous run (numpy _core/src/umath
// loops_utils.h.src pairwise_sum), on gathered interior elements.
__device__ double np_pairwise_dev(const double* x, int n, int N, size_t start, size_t cnt) {
  // explicit stack of (start, cnt, partial-state)
  struct Fr {
    size_t s, c;
    double left;
    int stage;
  };
  Fr st[16];
  int sp = 0;
  st[0] = {start, cnt, 0., 0};
  double ret = 0.;
  auto val = [&](size_t q) { return x[at(NH + (int)(q / N), NH + (int)(q % N), n)]; };
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
      // deliver to parent
      while (sp >= 0) {
        Fr& p = st[sp];
        if (p.stage == 1) {  // left done, start right
          p.left = ret;
          p.stage = 2;
          size_t h = p.c / 2;
          h -= h % 8;
          st[++sp] = {p.s + h, p.c - h, 0., 0};
          break;
        } else {  // right done
          ret = p.left + ret;
          sp--;
        }
      }
      continue;
    }
    // split
    size_t h = f.c / 2;
    h -= h % 8;
    f.stage = 1;
    st[sp + 1] = {f.s, h, 0., 0};
    sp++;
  }
  return ret;
}

// one thread per 8192-element chunk of the flattened interior
__global__ void k_npsum_chunks(const double* __restrict__ x, int m, int n, double* partial) {
  const int N = n - 2 * NH;
  const size_t tot = (size_t)(m - 2 * NH) * N;
  const size_t ch = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t s = ch * 8192;
  if (s >= tot) return;
  const size_t cnt = tot - s < 8192 ? tot - s : 8192;
  partial[ch] = np_pairwise_dev(x, n, N, s, cnt);
}
// acc = 0 + chunk0 + chunk1 + ...; out[0] = acc * 1. / div (if div != 0)
__global__ void k_npsum_fold(const double* __restrict__ partial, int nch, double div,
                             double* out) {
  if (threadIdx.x != 0) return;
  double acc = 0.;
  for (int k = 0; k < nch; k++) acc = acc + partial[k];
  out[0] = div != 0. ? acc * 1. / div : acc;
}

// ---------------------------------------------------------------------------
// host-side launch helpers
// ---------------------------------------------------------------------------
static inline dim3 grid2(int cols, int rows) {
  return dim3((unsigned)((cols + 63) / 64), (unsigned)((rows + 3) / 4));
}
static inline dim3 grid1(size_t cnt) { return dim3((unsigned)((cnt + 255) / 256)); }
static const dim3 B2(64, 4), B1(256);

template <typename T>
void launch_adv_upwind(const int8_t* msk, const T* x, T* y, const T* u, const T* v, T* xflx,
                       T* yflx, const double* cst, int method, int order, int m, int n,
                       hipStream_t s) {
  const double dx = cst[0], dy = cst[1], logcosh = cst[2], umax = cst[3], aparab = cst[4];
  UpParams<T> p;
  p.zdx = T(1. / dx);
  p.zdy = T(1. / dy);
  const double a = logcosh / umax;
  p.a = T(a);
  p.KK = T((umax - log(cosh(umax * a)) / a) * 0.5);
  const double u1 = aparab * umax;
  p.u1 = T(u1);
  p.aa = T(1. / (2. * u1));
  p.bb = T(u1 * 0.5);
  p.method = method;
  p.order = order;
  hipLaunchKernelGGL(k_adv_upwind<T>, grid2(n - 2 * NH, m - 2 * NH), B2, 0, s, msk, x, y, u, v,
                     xflx, yflx, p, m, n);
}

template <typename T>
void launch_adv_centered(const int8_t* msk, const T* x, T* y, const T* u, const T* v, T* xflx,
                         T* yflx, const double* cst, int order, int m, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_adv_centered<T>, grid2(n - 2 * NH, m - 2 * NH), B2, 0, s, msk, x, y, u, v,
                     xflx, yflx, T(1. / cst[0]), T(1. / cst[1]), order, m, n);
}

template <typename T>
void launch_fillhalo(T* x, int m, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_fillhalo<T>, grid2(n, m), B2, 0, s, x, m, n);
}
template <typename T>
void launch_celltocorner(const T* xr, T* xp, int m, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_celltocorner<T>, grid2(n - 1, m - 1), B2, 0, s, xr, xp, m, n);
}
template <typename T>
void launch_orthogradient(const int8_t* msk, const T* psi, double dx, double dy, T* u, T* v,
                          int m, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_orthogradient<T>, grid2(n - 2, m - 2), B2, 0, s, msk, psi, T(1. / dx),
                     T(1. / dy), u, v, m, n);
}
template <typename T>
void launch_add_diffusion(const int8_t* msk, const T* t, double dx, double kdiff, T* dt, int m,
                          int n, hipStream_t s) {
  hipLaunchKernelGGL(k_add_diffusion<T>, grid2(n - 2 * NH, m - 2 * NH), B2, 0, s, msk, t,
                     T(kdiff / (dx * dx)), dt, m, n);
}
template <typename T>
void launch_torque(const int8_t* msk, const T* b, double dx, double g, T* dw, int m, int n,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_torque<T>, grid2(n - 2 * NH, m - 2 * NH), B2, 0, s, msk, b,
                     T(0.5 * g / dx), dw, m, n);
}
template <typename T>
void launch_mask_mul(T* x, const int8_t* mk, size_t cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_mask_mul<T>, grid1(cnt), B1, 0, s, x, mk, cnt);
}
template <typename T>
void launch_smooth(const int8_t* msk, const T* A, T* x, const T* b, double omega, T* t, int m,
                   int n, hipStream_t s) {
  const T c1 = T(omega), c2 = T(1. - omega);
  hipLaunchKernelGGL(k_jacobi1<T>, grid2(n - 2 * NH + 2, m - 2 * NH + 2), B2, 0, s, msk, A, x, b,
                     t, c1, c2, m, n);
  hipLaunchKernelGGL(k_jacobi2<T>, grid2(n - 2 * NH, m - 2 * NH), B2, 0, s, msk, A, t, b, x, c1,
                     c2, m, n);
}
template <typename T>
void launch_residual(const int8_t* msk, const T* A, const T* x, const T* b, T* r, int m, int n,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_residual<T>, grid2(n - 2 * NH, m - 2 * NH), B2, 0, s, msk, A, x, b, r, m,
                     n);
}
template <typename T>
void launch_restrict(const int8_t* msk2, const T* x1, T* x2, int m2, int n2, int n1,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_restrict<T>, grid2(n2 - 2 * NH, m2 - 2 * NH), B2, 0, s, msk2, x1, x2, m2,
                     n2, n1);
}
template <typename T>
void launch_interpolate(const int8_t* msk1, const int8_t* msk2, const T* x2, T* x1, int m2,
                        int n2, int n1, bool acc, hipStream_t s) {
  if (acc)
    hipLaunchKernelGGL((k_interpolate<T, true>), grid2(n2 - 3, m2 - 3), B2, 0, s, msk1, msk2, x2,
                       x1, m2, n2, n1);
  else
    hipLaunchKernelGGL((k_interpolate<T, false>), grid2(n2 - 3, m2 - 3), B2, 0, s, msk1, msk2,
                       x2, x1, m2, n2, n1);
}
template <typename T>
void launch_fill_value(T* x, T v, size_t cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_value<T>, grid1(cnt), B1, 0, s, x, v, cnt);
}
template <typename T>
void launch_copy(T* d, const T* src, size_t cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_copy<T>, grid1(cnt), B1, 0, s, d, src, cnt);
}
template <typename T>
void launch_add_inplace(T* x, const T* y, size_t cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_add_inplace<T>, grid1(cnt), B1, 0, s, x, y, cnt);
}
template <typename T>
void launch_sub_dev_scalar(T* x, const T* sc, size_t cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_sub_dev_scalar<T>, grid1(cnt), B1, 0, s, x, sc, cnt);
}
template <typename T>
void launch_rk(int mode, T* out, const T* x, const T* k0, const T* k1, const T* k2, double c,
               size_t cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_rk<T>, grid1(cnt), B1, 0, s, mode, out, x, k0, k1, k2, T(c), cnt);
}
void launch_noslip_source(const int8_t* msk, const double* psi, double* y, double dx, double dy,
                          int m, int n, hipStream_t s) {
  const double cff = 1. / (2 * dx * dy);  // core/fortran_operators.f90:240
  hipLaunchKernelGGL(k_noslip_source, grid1((size_t)m * n), B1, 0, s, msk, psi, y, cff, m, n);
}
void launch_sub_mean_mask(double* x, const int8_t* mk, const double* d, size_t cnt,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_sub_mean_mask, grid1(cnt), B1, 0, s, x, mk, d, cnt);
}
void launch_noslip_momentum(double* x, const int8_t* mk, const double* xr0, const double* yr0,
                            const double* d, double x2bc, double y2bc, size_t cnt,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_noslip_momentum, grid1(cnt), B1, 0, s, x, mk, xr0, yr0, d, x2bc, y2bc,
                     cnt);
}
void launch_sub_inplace(double* x, const double* y, size_t cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_sub_inplace, grid1(cnt), B1, 0, s, x, y, cnt);
}
void launch_div_scalar(double* x, double c, size_t cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_div_scalar, grid1(cnt), B1, 0, s, x, c, cnt);
}
void launch_roll_mean(const double* u, double* out, int axis, int m, int n, hipStream_t s) {
  const size_t cnt = (size_t)m * n;
  hipLaunchKernelGGL(k_roll_mean<double>, grid1(cnt), B1, 0, s, u, out, axis, m, n);
}
void launch_negate(double* x, size_t cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_negate<double>, grid1(cnt), B1, 0, s, x, cnt);
}
void launch_flux_split(const double* a, const double* b, double* rev, double* irr, double sign,
                       double cff, size_t cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_flux_split<double>, grid1(cnt), B1, 0, s, a, b, rev, irr, sign, cff, cnt);
}

void launch_coarsen(const double* Af, const int8_t* msk1, const int8_t* msk2, double* Ac, int m1,
                    int n1, int m2, int n2, hipStream_t s) {
  hipLaunchKernelGGL(k_coarsen, grid2(n2 - 2 * NH, m2 - 2 * NH), B2, 0, s, Af, msk1, msk2, Ac, m1,
                     n1, m2, n2);
}
void launch_afinest(const int8_t* msk, double* A, const St9& st, double coef, int m, int n,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_afinest, grid2(n - 2 * NH, m - 2 * NH), B2, 0, s, msk, A, st, coef, m, n);
}
// core/gmg/hierarchy.py:80-84: A5 -= shift wherever A5 != 0 (whole plane)
__global__ void k_helmholtz_shift(double* __restrict__ A5, size_t cnt, double shift) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < cnt && A5[q] != 0.) A5[q] -= shift;
}
void launch_helmholtz_shift(double* A5, size_t cnt, double shift, hipStream_t s) {
  hipLaunchKernelGGL(k_helmholtz_shift, grid1(cnt), B1, 0, s, A5, cnt, shift);
}
void launch_corner_mask(const int8_t* msk, int8_t* mskp, int m, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_corner_mask, grid2(n, m), B2, 0, s, msk, mskp, m, n);
}
void launch_coarse_mask(const int8_t* msk1, int8_t* msk2, int m2, int n2, int n1, hipStream_t s) {
  hipLaunchKernelGGL(k_coarse_mask, grid2(n2 - 2 * NH, m2 - 2 * NH), B2, 0, s, msk1, msk2, m2, n2,
                     n1);
}

static int red_blocks(int m, int n) {
  const size_t tot = (size_t)(m - 2 * NH) * (n - 2 * NH);
  size_t b = (tot + 255) / 256;
  if (b > (size_t)kRedMaxBlocks) b = kRedMaxBlocks;
  if (b < 1) b = 1;
  return (int)b;
}

void launch_norm2(const int8_t* msk, const double* x, int m, int n, double* partial, double* out,
                  hipStream_t s) {
  const int nb = red_blocks(m, n);
  hipLaunchKernelGGL(k_norm2_partial, dim3(nb), B1, 0, s, msk, x, m, n, partial);
  Scale8 sc;
  for (int k = 0; k < 8; k++) sc.v[k] = 1.;
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(64), 0, s, partial, nb, 1, MaxMask{0}, sc, out);
}
void launch_kemaxu(const int8_t* msk, const double* u, const double* v, int m, int n,
                   double* partial, double* out, hipStream_t s) {
  const int nb = red_blocks(m, n);
  hipLaunchKernelGGL(k_kemaxu_partial, dim3(nb), B1, 0, s, msk, u, v, m, n, partial);
  Scale8 sc;
  for (int k = 0; k < 8; k++) sc.v[k] = 1.;
  sc.v[0] = 0.25;
  sc.v[1] = 0.5;
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(64), 0, s, partial, nb, 2, MaxMask{2}, sc, out);
}
void launch_diag(const int8_t* msk, const double* state, int kind, const double* xr,
                 const double* yr, int m, int n, double* partial, double* out, hipStream_t s) {
  const int nb = red_blocks(m, n);
  hipLaunchKernelGGL(k_diag_partial, dim3(nb), B1, 0, s, msk, state, kind, xr, yr, m, n, partial);
  Scale8 sc;
  for (int k = 0; k < 8; k++) sc.v[k] = 1.;
  sc.v[0] = 0.5;
  sc.v[1] = 0.25;
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(64), 0, s, partial, nb, 8, MaxMask{1}, sc, out);
}
void launch_dot2(const int8_t* msk, const double* x, const double* a, const double* b, int m,
                 int n, double* partial, double* out, hipStream_t s) {
  const int nb = red_blocks(m, n);
  hipLaunchKernelGGL(k_dot2_partial, dim3(nb), B1, 0, s, msk, x, a, b, m, n, partial);
  Scale8 sc;
  for (int k = 0; k < 8; k++) sc.v[k] = 1.;
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(64), 0, s, partial, nb, 2, MaxMask{0}, sc, out);
}
void launch_msum(int kind, const int8_t* msk, const double* x, const double* y, int m, int n,
                 double* partial, double* out, hipStream_t s) {
  const int nb = red_blocks(m, n);
  if (kind == 0)
    hipLaunchKernelGGL(k_msum_partial<0>, dim3(nb), B1, 0, s, msk, x, y, m, n, partial);
  else if (kind == 1)
    hipLaunchKernelGGL(k_msum_partial<1>, dim3(nb), B1, 0, s, msk, x, y, m, n, partial);
  else
    hipLaunchKernelGGL(k_msum_partial<2>, dim3(nb), B1, 0, s, msk, x, y, m, n, partial);
  Scale8 sc;
  for (int k = 0; k < 8; k++) sc.v[k] = 1.;
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(64), 0, s, partial, nb, kind == 2 ? 2 : 1, MaxMask{0},
                     sc, out);
}
void launch_domain_sum(const double* x, int m, int n, double div, double* partial, double* out,
                       hipStream_t s) {
  const size_t tot = (size_t)(m - 2 * NH) * (n - 2 * NH);
  const int nch = (int)((tot + 8191) / 8192);
  hipLaunchKernelGGL(k_npsum_chunks, dim3((nch + 63) / 64), dim3(64), 0, s, x, m, n, partial);
  hipLaunchKernelGGL(k_npsum_fold, dim3(1), dim3(64), 0, s, partial, nch, div, out);
}

template void launch_adv_upwind<double>(const int8_t*, const double*, double*, const double*,
                                        const double*, double*, double*, const double*, int, int,
                                        int, int, hipStream_t);
template void launch_adv_centered<double>(const int8_t*, const double*, double*, const double*,
                                          const double*, double*, double*, const double*, int,
                                          int, int, hipStream_t);
template void launch_fillhalo<double>(double*, int, int, hipStream_t);
template void launch_celltocorner<double>(const double*, double*, int, int, hipStream_t);
template void launch_orthogradient<double>(const int8_t*, const double*, double, double, double*,
                                           double*, int, int, hipStream_t);
template void launch_add_diffusion<double>(const int8_t*, const double*, double, double, double*,
                                           int, int, hipStream_t);
template void launch_torque<double>(const int8_t*, const double*, double, double, double*, int,
                                    int, hipStream_t);
template void launch_mask_mul<double>(double*, const int8_t*, size_t, hipStream_t);
template void launch_smooth<double>(const int8_t*, const double*, double*, const double*, double,
                                    double*, int, int, hipStream_t);
template void launch_residual<double>(const int8_t*, const double*, const double*, const double*,
                                      double*, int, int, hipStream_t);
template void launch_restrict<double>(const int8_t*, const double*, double*, int, int, int,
                                      hipStream_t);
template void launch_interpolate<double>(const int8_t*, const int8_t*, const double*, double*,
                                         int, int, int, bool, hipStream_t);
template void launch_fill_value<double>(double*, double, size_t, hipStream_t);
template void launch_copy<double>(double*, const double*, size_t, hipStream_t);
template void launch_add_inplace<double>(double*, const double*, size_t, hipStream_t);
template void launch_sub_dev_scalar<double>(double*, const double*, size_t, hipStream_t);
template void launch_rk<double>(int, double*, const double*, const double*, const double*,
                                const double*, double, size_t, hipStream_t);

}  // namespace f2d
