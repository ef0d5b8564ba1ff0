// f2d_kernels.h -- shared device helpers for the Fluid2d HIP kernels.
//
// Field layout: row-major (m rows = ny + 2*nh, n columns = nx + 2*nh),
// x contiguous, halo width nh = 3.  Masks are int8 0/1.
//
// Periodic images: the reference follows almost every producer with
// fillhalo (core/gmg/fortran_multigrid.f90:365-412), which copies interior
// cell (J,I) into every halo cell whose periodic image it is.  Producers
// here write those copies themselves (store_img); the array is then in the
// state "kernel, then fillhalo" leaves it in.  This needs interiors of at
// least nh cells per axis (checked on the host), so that every halo cell's
// image is an interior cell and is at most one period away.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace f2d {

constexpr int NH = 3;

__device__ __forceinline__ size_t at(int j, int i, int n) {
  return (size_t)j * (size_t)n + (size_t)i;
}

template <typename T>
__device__ __forceinline__ void store_img(T* a, int J, int I, T v, int m, int n) {
  const int M = m - 2 * NH, N = n - 2 * NH;
  int jc[3], ic[3], nj = 1, ni = 1;
  jc[0] = J;
  ic[0] = I;
  if (J + M < m && J + M >= m - NH) jc[nj++] = J + M;
  if (J - M >= 0 && J - M < NH) jc[nj++] = J - M;
  if (I + N < n && I + N >= n - NH) ic[ni++] = I + N;
  if (I - N >= 0 && I - N < NH) ic[ni++] = I - N;
  for (int p = 0; p < nj; p++)
    for (int q = 0; q < ni; q++) a[at(jc[p], ic[q], n)] = v;
}

}  // namespace f2d
