// divB clean-up: code/sub_divB_correct.f90:25-175 with the conjugate-
// gradient-on-normal-equations solve of code/sub_sparse_solver.f90:19-159.
//
// The 7-band matrix is never stored: its bands are constants (-6 on the
// diagonal, 1 on the six neighbour bands where the neighbour exists), so
// the matrix-vector product of code/sub_banded_tpose.f90:20-30 is evaluated
// in place, adding the bands in the same order (z+1, y+1, x+1, diagonal,
// x-1, y-1, z-1).  Dot products accumulate in double through a fixed-order
// two-stage reduction (deterministic run to run).  Two precision modes:
//   kDivbRef  -- the reference's single-precision scalars (each reduced
//                dot is rounded to float, scalar updates in float);
//   kDivbF64  -- scalars and vector updates in double, stored as float.
#include "g3m3_dev.h"
#include "g3m3_launch.h"

#include <vector>

namespace g3m3 {

constexpr int kRedThreads = 256;

__global__ void k_divb_rhs(Box B, const float* __restrict__ bx, const float* __restrict__ by,
                           const float* __restrict__ bz, float* rhs, float two_rx) {
    int i, j, k;
    const int c = gtid();
    if (!cell_of(B, c, i, j, k)) return;
    if (!B.interior(i, j, k)) {
        rhs[c] = 0.0f;
        return;
    }
    const int Y = B.nx, Z = B.nx * B.ny;
    rhs[c] = ((bx[c + 1] - bx[c - 1]) + (by[c + Y] - by[c - Y]) + (bz[c + Z] - bz[c - Z])) / two_rx;
}

__device__ inline double block_sum(double v) {
    __shared__ double red[kRedThreads];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = kRedThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// y = A x (- b), partial sums of y*y and of u*v (either optional)
__global__ void k_matvec(Box B, const float* __restrict__ x, float* y, const float* __restrict__ b,
                         double* part_yy, const float* __restrict__ u, const float* __restrict__ v,
                         double* part_uv, const float* __restrict__ gplus, double* part_gy) {
    int i, j, k;
    const int c = gtid();
    double yy = 0.0, uv = 0.0, gy = 0.0;
    if (cell_of(B, c, i, j, k)) {
        const int X = 1, Y = B.nx, Z = B.nx * B.ny;
        float t = 0.0f;
        if (k < B.nz - 1) t = t + x[c + Z];
        if (j < B.ny - 1) t = t + x[c + Y];
        if (i < B.nx - 1) t = t + x[c + X];
        t = t + (-6.0f) * x[c];
        if (i > 0) t = t + x[c - X];
        if (j > 0) t = t + x[c - Y];
        if (k > 0) t = t + x[c - Z];
        if (b) t = t - b[c];
        y[c] = t;
        yy = (double)t * (double)t;
        if (u) uv = (double)u[c] * (double)v[c];
        if (gplus) gy = ((double)t + (double)gplus[c]) * (double)t;   // (xi + g) * xi
    }
    const double s1 = block_sum(yy);
    if (threadIdx.x == 0) part_yy[blockIdx.x] = s1;
    if (part_uv) {
        const double s2 = block_sum(uv);
        if (threadIdx.x == 0) part_uv[blockIdx.x] = s2;
    }
    if (part_gy) {
        const double s3 = block_sum(gy);
        if (threadIdx.x == 0) part_gy[blockIdx.x] = s3;
    }
}

// generic dot: partial sums of u*v
__global__ void k_dot(int n, const float* __restrict__ u, const float* __restrict__ v, double* part) {
    const int c = gtid();
    double s = 0.0;
    if (c < n) s = (double)u[c] * (double)v[c];
    const double r = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// out[q] = sum of part[q*stride .. q*stride+nb), fixed order
__global__ void k_finalize(const double* __restrict__ part, int nb, int stride, int nq, double* out) {
    for (int q = 0; q < nq; ++q) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += kRedThreads) s += part[q * stride + b];
        const double r = block_sum(s);
        if (threadIdx.x == 0) out[q] = r;
    }
}

// xi = soln; soln = soln + anum*h
__global__ void k_update_soln(int n, float* soln, float* xi, const float* __restrict__ h,
                              const double* __restrict__ sc, int mode) {
    const int c = gtid();
    if (c >= n) return;
    const float s = soln[c];
    xi[c] = s;
    if (mode == kDivbRef) soln[c] = s + (float)sc[kScAnum] * h[c];
    else soln[c] = (float)((double)s + sc[kScAnum] * (double)h[c]);
}

// g = -xi; h = g + gam*h
__global__ void k_update_dir(int n, float* g, float* h, const float* __restrict__ xi,
                             const double* __restrict__ sc, int mode) {
    const int c = gtid();
    if (c >= n) return;
    const float gn = -xi[c];
    g[c] = gn;
    if (mode == kDivbRef) h[c] = gn + (float)sc[kScGam] * h[c];
    else h[c] = (float)((double)gn + sc[kScGam] * (double)h[c]);
}

__global__ void k_sub_neg(int n, float* xi, const float* __restrict__ b, float* g, int which) {
    const int c = gtid();
    if (c >= n) return;
    if (which == 0) xi[c] = xi[c] - b[c];
    else {
        const float v = -g[c];
        g[c] = v;
        xi[c] = v;   // h = g
    }
}

__global__ void k_zero1(int n, float* dst) {
    const int c = gtid();
    if (c < n) dst[c] = 0.0f;
}

__global__ void k_copy1(int n, float* dst, const float* __restrict__ src) {
    const int c = gtid();
    if (c < n) dst[c] = src[c];
}

// B -= grad(phi), phi = rx^2 * soln, interior (code/sub_divB_correct.f90:147-172)
__global__ void k_divb_apply(Box B, const float* __restrict__ soln, float* bx, float* by, float* bz,
                             float scale, float two_rx, float two_ry, float two_rz) {
    int i, j, k;
    const int c = gtid();
    if (!cell_of(B, c, i, j, k) || !B.interior(i, j, k)) return;
    const int Y = B.nx, Z = B.nx * B.ny;
    const float dbx = (scale * soln[c + 1] - scale * soln[c - 1]) / two_rx;
    const float dby = (scale * soln[c + Y] - scale * soln[c - Y]) / two_ry;
    const float dbz = (scale * soln[c + Z] - scale * soln[c - Z]) / two_rz;
    bx[c] = bx[c] - dbx;
    by[c] = by[c] - dby;
    bz[c] = bz[c] - dbz;
}

// ---------------------------------------------------------------------------
// Host side: the iteration of sparse_solver, recursion included.

namespace {
struct Level {
    float *g, *h, *xi, *xj;
};
}  // namespace

struct DivbSolver {
    Box B;
    int n, nb, mode;
    float* soln;
    const float* rhs;
    double* part;    // 3 * nb partial sums
    double* sc;      // device scalars
    Level lv[4];
    hipStream_t s;
    int redos = 0;
    long iters = 0;
    long max_iters;
    double errlim2, sum_b;
    double residual2 = 0.0;
    double host_sc[8];
    int status = 0;

    dim3 grid() const { return dim3(nb); }

    void read(int nq) {
        hipMemcpyAsync(host_sc, sc, sizeof(double) * nq, hipMemcpyDeviceToHost, s);
        hipStreamSynchronize(s);
    }
    double as_mode(double v) const { return mode == kDivbRef ? (double)(float)v : v; }

    // returns when converged / restarted out / iteration limit
    void solve(int depth, bool recurse, int redo_lim) {
        Level& L = lv[depth];
        redos = recurse ? redo_lim : 0;
        redos += 1;
        const dim3 g(nb), t(kRedThreads);
        // xi = A soln - b, |b|^2, |xi|^2 (:61-71)
        hipLaunchKernelGGL(k_matvec, g, t, 0, s, B, soln, L.xi, rhs, part, rhs, rhs, part + nb,
                           nullptr, nullptr);
        hipLaunchKernelGGL(k_finalize, dim3(1), t, 0, s, part, nb, nb, 2, sc + kScTmp);
        read(kScTmp + 2);
        double resd = as_mode(host_sc[kScTmp]);
        sum_b = as_mode(host_sc[kScTmp + 1]);
        // g = -A xi, h = g (:73-75)
        hipLaunchKernelGGL(k_matvec, g, t, 0, s, B, L.xi, L.g, nullptr, part, nullptr, nullptr,
                           nullptr, nullptr, nullptr);
        hipLaunchKernelGGL(k_sub_neg, g, t, 0, s, n, L.h, nullptr, L.g, 1);
        for (long it = 0; it < 10L * n && iters < max_iters; ++it) {
            ++iters;
            // xi = A h; anum = g.h / |xi|^2 (:82-95)
            hipLaunchKernelGGL(k_matvec, g, t, 0, s, B, L.h, L.xi, nullptr, part, L.g, L.h,
                               part + nb, nullptr, nullptr);
            hipLaunchKernelGGL(k_finalize, dim3(1), t, 0, s, part, nb, nb, 2, sc + kScTmp);
            read(kScTmp + 2);
            const double aden = as_mode(host_sc[kScTmp]);
            double anum = as_mode(host_sc[kScTmp + 1]);
            anum = as_mode(anum / aden);
            host_sc[kScAnum] = anum;
            hipMemcpyAsync(sc + kScAnum, &host_sc[kScAnum], sizeof(double), hipMemcpyHostToDevice, s);
            hipLaunchKernelGGL(k_update_soln, g, t, 0, s, n, soln, L.xi, L.h, sc, mode);
            // xj = A soln - b, residual2 (:101-107)
            hipLaunchKernelGGL(k_matvec, g, t, 0, s, B, soln, L.xj, rhs, part, nullptr, nullptr,
                               nullptr, nullptr, nullptr);
            hipLaunchKernelGGL(k_finalize, dim3(1), t, 0, s, part, nb, nb, 1, sc + kScTmp);
            read(kScTmp + 1);
            const double r2 = as_mode(host_sc[kScTmp]);
            residual2 = r2;
            if (r2 == resd || r2 <= as_mode(sum_b * errlim2)) {
                status = 1;
                return;
            }
            if (r2 > resd) {   // restart from the previous iterate (:118-129)
                hipLaunchKernelGGL(k_copy1, g, t, 0, s, n, soln, L.xi);
                if (redos >= 3 || depth + 1 >= 4) {
                    status = 2;
                    return;
                }
                solve(depth + 1, true, redos);
                if (status != 0) { /* the reference keeps iterating here */ }
            }
            resd = residual2;
            // xi = A xj; gg = |g|^2, dgg = (xi+g).xi (:134-141)
            hipLaunchKernelGGL(k_matvec, g, t, 0, s, B, L.xj, L.xi, nullptr, part, L.g, L.g,
                               part + nb, L.g, part + 2 * nb);
            hipLaunchKernelGGL(k_finalize, dim3(1), t, 0, s, part, nb, nb, 3, sc + kScTmp);
            read(kScTmp + 3);
            const double gg = as_mode(host_sc[kScTmp + 1]);
            const double dgg = as_mode(host_sc[kScTmp + 2]);
            if (gg == 0.0) {
                status = 1;
                return;
            }
            host_sc[kScGam] = as_mode(dgg / gg);
            hipMemcpyAsync(sc + kScGam, &host_sc[kScGam], sizeof(double), hipMemcpyHostToDevice, s);
            hipLaunchKernelGGL(k_update_dir, g, t, 0, s, n, L.g, L.h, L.xi, sc, mode);
        }
        status = 3;   // iteration limit
    }
};

int divb_correct(const Box& B, float* bx, float* by, float* bz, float rx, float errlim, int mode,
                 long max_iters, float* work, double* dwork, long* iters_out, double* res2_out,
                 hipStream_t s) {
    const int n = B.V();
    const int nb = (n + kRedThreads - 1) / kRedThreads;
    DivbSolver S;
    S.B = B;
    S.n = n;
    S.nb = nb;
    S.mode = mode;
    float* rhs = work;
    S.soln = work + (size_t)n;
    S.rhs = rhs;
    for (int l = 0; l < 4; ++l) {
        float* base = work + (size_t)(2 + 4 * l) * n;
        S.lv[l] = Level{base, base + n, base + 2 * (size_t)n, base + 3 * (size_t)n};
    }
    S.sc = dwork;
    S.part = dwork + 16;
    S.s = s;
    S.max_iters = max_iters > 0 ? max_iters : 10L * n;
    S.errlim2 = (double)((float)n * (errlim * errlim));
    if (mode != kDivbRef) S.errlim2 = (double)n * (double)errlim * (double)errlim;
    const dim3 g(nb), t(kRedThreads);
    hipLaunchKernelGGL(k_divb_rhs, g, t, 0, s, B, bx, by, bz, rhs, 2.0f * rx);
    // the reference's soln_vec is an uninitialised automatic array; start from 0
    hipLaunchKernelGGL(k_zero1, g, t, 0, s, n, S.soln);
    S.solve(0, false, 0);
    hipLaunchKernelGGL(k_divb_apply, g, t, 0, s, B, S.soln, bx, by, bz, rx * rx, 2.0f * rx,
                       2.0f * rx, 2.0f * rx);
    if (iters_out) *iters_out = S.iters;
    if (res2_out) *res2_out = S.residual2;
    return S.status;
}

size_t divb_work_floats(int ncells) { return (size_t)18 * ncells; }
size_t divb_work_doubles(int ncells) { return 16 + 3 * (size_t)((ncells + kRedThreads - 1) / kRedThreads); }

}  // namespace g3m3
