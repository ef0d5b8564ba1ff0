// f2d_ops.h -- host-side launchers for the kernels in f2d_ops.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "f2d_kernels.h"

namespace f2d {

constexpr int kRedMaxBlocks = 1024;  // partial-sum slots per reduction

struct MaxMask {
  unsigned bits;  // bit k set: quantity k is a max, else a sum
};
struct Scale8 {
  double v[8];
};
struct St9 {
  double w[9];
};

template <typename T>
void launch_adv_upwind(const int8_t* msk, const T* x, T* y, const T* u, const T* v, T* xflx,
                       T* yflx, const double* cst, int method, int order, int m, int n,
                       hipStream_t s);
template <typename T>
void launch_adv_centered(const int8_t* msk, const T* x, T* y, const T* u, const T* v, T* xflx,
                         T* yflx, const double* cst, int order, int m, int n, hipStream_t s);
template <typename T>
void launch_fillhalo(T* x, int m, int n, hipStream_t s);
template <typename T>
void launch_celltocorner(const T* xr, T* xp, int m, int n, hipStream_t s);
template <typename T>
void launch_orthogradient(const int8_t* msk, const T* psi, double dx, double dy, T* u, T* v, int m,
                          int n, hipStream_t s);
template <typename T>
void launch_add_diffusion(const int8_t* msk, const T* t, double dx, double kdiff, T* dt, int m,
                          int n, hipStream_t s);
template <typename T>
void launch_torque(const int8_t* msk, const T* b, double dx, double g, T* dw, int m, int n,
                   hipStream_t s);
template <typename T>
void launch_mask_mul(T* x, const int8_t* mk, size_t cnt, hipStream_t s);
template <typename T>
void launch_smooth(const int8_t* msk, const T* A, T* x, const T* b, double omega, T* t, int m,
                   int n, hipStream_t s);
template <typename T>
void launch_residual(const int8_t* msk, const T* A, const T* x, const T* b, T* r, int m, int n,
                     hipStream_t s);
template <typename T>
void launch_restrict(const int8_t* msk2, const T* x1, T* x2, int m2, int n2, int n1,
                     hipStream_t s);
template <typename T>
void launch_interpolate(const int8_t* msk1, const int8_t* msk2, const T* x2, T* x1, int m2,
                        int n2, int n1, bool acc, hipStream_t s);
template <typename T>
void launch_fill_value(T* x, T v, size_t cnt, hipStream_t s);
template <typename T>
void launch_copy(T* d, const T* src, size_t cnt, hipStream_t s);
template <typename T>
void launch_add_inplace(T* x, const T* y, size_t cnt, hipStream_t s);
template <typename T>
void launch_sub_dev_scalar(T* x, const T* sc, size_t cnt, hipStream_t s);
template <typename T>
void launch_rk(int mode, T* out, const T* x, const T* k0, const T* k1, const T* k2, double c,
               size_t cnt, hipStream_t s);
void launch_noslip_source(const int8_t* msk, const double* psi, double* y, double dx, double dy,
                          int m, int n, hipStream_t s);
void launch_sub_mean_mask(double* x, const int8_t* mk, const double* d, size_t cnt, hipStream_t s);
void launch_noslip_momentum(double* x, const int8_t* mk, const double* xr0, const double* yr0,
                            const double* d, double x2bc, double y2bc, size_t cnt, hipStream_t s);
void launch_sub_inplace(double* x, const double* y, size_t cnt, hipStream_t s);
void launch_div_scalar(double* x, double c, size_t cnt, hipStream_t s);
void launch_dot2(const int8_t* msk, const double* x, const double* a, const double* b, int m,
                 int n, double* partial, double* out, hipStream_t s);
void launch_roll_mean(const double* u, double* out, int axis, int m, int n, hipStream_t s);
void launch_negate(double* x, size_t cnt, hipStream_t s);
void launch_flux_split(const double* a, const double* b, double* rev, double* irr, double sign,
                       double cff, size_t cnt, hipStream_t s);

void launch_coarsen(const double* Af, const int8_t* msk1, const int8_t* msk2, double* Ac, int m1,
                    int n1, int m2, int n2, hipStream_t s);
void launch_afinest(const int8_t* msk, double* A, const St9& st, double coef, int m, int n,
                    hipStream_t s);
void launch_corner_mask(const int8_t* msk, int8_t* mskp, int m, int n, hipStream_t s);
// A5 -= shift wherever A5 != 0 (the Helmholtz operator of core/gmg/hierarchy.py:80-84)
void launch_helmholtz_shift(double* A5, size_t cnt, double shift, hipStream_t s);
void launch_coarse_mask(const int8_t* msk1, int8_t* msk2, int m2, int n2, int n1, hipStream_t s);

// reductions: partial must hold kRedMaxBlocks*8 doubles; out on the device
void launch_norm2(const int8_t* msk, const double* x, int m, int n, double* partial, double* out,
                  hipStream_t s);
void launch_kemaxu(const int8_t* msk, const double* u, const double* v, int m, int n,
                   double* partial, double* out, hipStream_t s);
void launch_diag(const int8_t* msk, const double* state, int kind, const double* xr,
                 const double* yr, int m, int n, double* partial, double* out, hipStream_t s);
// masked interior sums of core/fortran_diag.f90: kind 0 computedotprod
// (out[0] = sum x*y), 1 computesum (out[0] = sum x), 2 computesumandnorm
// (out[0..1] = sum x, sum x^2); y is read by kind 0 only
void launch_msum(int kind, const int8_t* msk, const double* x, const double* y, int m, int n,
                 double* partial, double* out, hipStream_t s);
// NumPy-order interior sum; out[0] = sum * 1. / div (div == 0: the sum)
void launch_domain_sum(const double* x, int m, int n, double div, double* partial, double* out,
                       hipStream_t s);

}  // namespace f2d
