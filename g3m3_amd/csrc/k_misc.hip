// Small kernels: the general banded product of code/sub_banded_tpose.f90
// (for the drop-in entry point; the divB solve itself is matrix-free) and
// batched trilinear sampling of code/cft_trilin_interp.f90:10-66.
#include "g3m3_dev.h"
#include "g3m3_launch.h"

namespace g3m3 {

// y_i = sum_b M(b, i+pos_b) x(i+pos_b); M is the reference's M(7, npts)
__global__ void k_banded(int npts, const int* __restrict__ pos, const float* __restrict__ M,
                         const float* __restrict__ x, float* y) {
    const int i = gtid();
    if (i >= npts) return;
    float t = 0.0f;
#pragma unroll
    for (int b = 0; b < 7; ++b) {
        const int ii = i + pos[b];
        if (ii >= 0 && ii < npts) t = t + M[b + 7 * ii] * x[ii];
    }
    y[i] = t;
}

void launch_banded_matvec(int npts, const int* pos, const float* M, const float* x, float* y,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_banded, dim3((npts + 255) / 256), dim3(256), 0, s, npts, pos, M, x, y);
}

// one point per thread: sxyz[3*p], gridpts[6*p] (Fortran gridpts(3,2)),
// qty[8*p] (Fortran qty(2,2,2))
__global__ void k_trilin(int npts, const float* __restrict__ sxyz, const float* __restrict__ gp,
                         const float* __restrict__ qty, float* out) {
    const int p = gtid();
    if (p >= npts) return;
    const float* s = sxyz + 3 * p;
    const float* g = gp + 6 * p;
    const float* q = qty + 8 * p;
    const float xd = (s[0] - g[0]) / (g[3] - g[0]);
    const float yd = (s[1] - g[1]) / (g[4] - g[1]);
    const float zd = (s[2] - g[2]) / (g[5] - g[2]);
    // q index = ix + 2*iy + 4*iz (0-based)
    const float e00 = q[0] * (1.0f - xd) + q[1] * xd;
    const float e01 = q[4] * (1.0f - xd) + q[5] * xd;
    const float e10 = q[2] * (1.0f - xd) + q[3] * xd;
    const float e11 = q[6] * (1.0f - xd) + q[7] * xd;
    const float f0 = e00 * (1.0f - yd) + e10 * yd;
    const float f1 = e01 * (1.0f - yd) + e11 * yd;
    out[p] = f0 * (1.0f - zd) + f1 * zd;
}

void launch_trilin(int npts, const float* sxyz, const float* gridpts, const float* qty,
                   float* out, hipStream_t s) {
    if (npts <= 0) return;
    hipLaunchKernelGGL(k_trilin, dim3((npts + 255) / 256), dim3(256), 0, s, npts, sxyz, gridpts,
                       qty, out);
}

}  // namespace g3m3
