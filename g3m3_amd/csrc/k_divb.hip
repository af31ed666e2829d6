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
#include <algorithm>
#include <cmath>
#include <vector>

#include "g3m3_dev.h"
#include "g3m3_launch.h"

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

// Device-side control of the iteration: the scalars of code/
// sub_sparse_solver.f90 live in device memory and each kernel of an
// iteration first reads the status word, so the host can queue a batch of
// iterations and look at the outcome once per batch.
struct DivbCtl {
    int status;          // 0 running, 1 converged / exact, 2 restart wanted
    int pad;
    long long iters;     // completed iterations (all levels)
};

enum { kScResd = 2, kScSumB = 3, kScErr2 = 8, kScR2 = 9 };

// solve status when a HIP call fails (the error is also left in HIP's
// last-error state for the caller's check)
constexpr int kStatusHipFailed = 4;

// device -> host copy of solver control data, waited for; false when HIP
// reports an error
static bool fetch(void* dst, const void* src, size_t bytes, hipStream_t s) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s) == hipSuccess &&
           hipStreamSynchronize(s) == hipSuccess;
}

__device__ inline double as_mode_d(double v, int mode) {
    return mode == kDivbRef ? (double)(float)v : v;
}

// An explicit banded matrix M(nbands, npts) with band offsets pos (the
// general form of code/sub_banded_tpose.f90); M == nullptr selects the
// matrix-free divB Laplacian of the box.
struct Bands {
    const float* M;
    int nbands, npts;
    int pos[8];
};

// y = A x (- b), partial sums of y*y and of u*v (either optional)
__global__ void k_matvec(Box B, Bands Bd, const float* __restrict__ x, float* y,
                         const float* __restrict__ b, double* part_yy, const float* __restrict__ u,
                         const float* __restrict__ v, double* part_uv,
                         const float* __restrict__ gplus, double* part_gy, const DivbCtl* ctl) {
    if (ctl && ctl->status != 0) return;
    int i, j, k;
    const int c = gtid();
    double yy = 0.0, uv = 0.0, gy = 0.0;
    const bool mine = Bd.M ? c < Bd.npts : cell_of(B, c, i, j, k);
    if (mine) {
        float t = 0.0f;
        if (Bd.M) {
            for (int q = 0; q < Bd.nbands; ++q) {
                const int ii = c + Bd.pos[q];
                if (ii >= 0 && ii < Bd.npts) t = t + Bd.M[q + Bd.nbands * (size_t)ii] * x[ii];
            }
        } else {
            const int X = 1, Y = B.nx, Z = B.nx * B.ny;
            if (k < B.nz - 1) t = t + x[c + Z];
            if (j < B.ny - 1) t = t + x[c + Y];
            if (i < B.nx - 1) t = t + x[c + X];
            t = t + (-6.0f) * x[c];
            if (i > 0) t = t + x[c - X];
            if (j > 0) t = t + x[c - Y];
            if (k > 0) t = t + x[c - Z];
        }
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

__device__ inline double sum_parts(const double* __restrict__ part, int nb) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += kRedThreads) s += part[b];
    return block_sum(s);
}

// start of a level (:61-71): resd = |A soln - b|^2, sum_b = |b|^2
__global__ void k_fin_start(const double* __restrict__ part, int nb, double* sc, DivbCtl* ctl,
                            int mode) {
    const double r = sum_parts(part, nb);
    const double b2 = sum_parts(part + nb, nb);
    if (threadIdx.x == 0) {
        sc[kScResd] = as_mode_d(r, mode);
        sc[kScSumB] = as_mode_d(b2, mode);
        ctl->status = 0;
    }
}

// anum = (g.h) / |A h|^2 (:84-95)
__global__ void k_fin_anum(const double* __restrict__ part, int nb, double* sc,
                           const DivbCtl* ctl, int mode) {
    if (ctl->status != 0) return;
    const double aden = as_mode_d(sum_parts(part, nb), mode);
    const double anum = as_mode_d(sum_parts(part + nb, nb), mode);
    if (threadIdx.x == 0) sc[kScAnum] = as_mode_d(anum / aden, mode);
}

// residual2 and the exit / restart tests (:101-129)
__global__ void k_fin_r2(const double* __restrict__ part, int nb, double* sc, DivbCtl* ctl,
                         int mode) {
    if (ctl->status != 0) return;
    const double r2 = as_mode_d(sum_parts(part, nb), mode);
    if (threadIdx.x == 0) {
        sc[kScR2] = r2;
        ctl->iters += 1;
        const double resd = sc[kScResd];
        if (r2 == resd || r2 <= as_mode_d(sc[kScSumB] * sc[kScErr2], mode)) ctl->status = 1;
        else if (r2 > resd) ctl->status = 2;
        else sc[kScResd] = r2;
    }
}

// gam = dgg / gg (:134-146)
__global__ void k_fin_gam(const double* __restrict__ part, int nb, double* sc, DivbCtl* ctl,
                          int mode) {
    if (ctl->status != 0) return;
    const double gg = as_mode_d(sum_parts(part + nb, nb), mode);
    const double dgg = as_mode_d(sum_parts(part + 2 * nb, nb), mode);
    if (threadIdx.x == 0) {
        if (gg == 0.0) ctl->status = 1;
        else sc[kScGam] = as_mode_d(dgg / gg, mode);
    }
}

__global__ void k_divb_init(double* sc, DivbCtl* ctl, double errlim2) {
    if (threadIdx.x == 0) {
        sc[kScErr2] = errlim2;
        ctl->status = 0;
        ctl->iters = 0;
    }
}

// resd = residual2 of the restarted solve, status cleared (after a restart)
__global__ void k_resume(double* sc, DivbCtl* ctl) {
    if (threadIdx.x == 0) {
        sc[kScResd] = sc[kScR2];
        ctl->status = 0;
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

// xi = soln; soln = soln + anum*h
__global__ void k_update_soln(int n, float* soln, float* xi, const float* __restrict__ h,
                              const double* __restrict__ sc, int mode, const DivbCtl* ctl) {
    if (ctl->status != 0) return;
    const int c = gtid();
    if (c >= n) return;
    const float s = soln[c];
    xi[c] = s;
    if (mode == kDivbRef) soln[c] = s + (float)sc[kScAnum] * h[c];
    else soln[c] = (float)((double)s + sc[kScAnum] * (double)h[c]);
}

// g = -xi; h = g + gam*h
__global__ void k_update_dir(int n, float* g, float* h, const float* __restrict__ xi,
                             const double* __restrict__ sc, int mode, const DivbCtl* ctl) {
    if (ctl->status != 0) return;
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
// Host side: the iteration of sparse_solver, recursion included, queued in
// batches of kBatch iterations between looks at the status word.

namespace {
struct Level {
    float *g, *h, *xi, *xj;
};
constexpr int kBatch = 32;
}  // namespace

struct DivbSolver {
    Box B;
    int n, nb, mode;
    float* soln;
    const float* rhs;
    double* part;    // 3 * nb partial sums
    double* sc;      // device scalars
    DivbCtl* ctl;    // device control word
    Level lv[4];
    Bands bands{};
    hipStream_t s;
    int redos = 0;
    long long iters = 0;
    long long max_iters;
    double residual2 = 0.0;
    int status = 0;

    DivbCtl read_ctl() {
        DivbCtl h{};
        if (!fetch(&h, ctl, sizeof(h), s)) h.status = kStatusHipFailed;
        return h;
    }
    double read_sc(int q) {
        double v = 0.0;
        if (!fetch(&v, sc + q, sizeof(double), s)) return std::nan("");
        return v;
    }

    // the tail of an iteration (:131-146): xi = A xj, gg, dgg, new direction
    void tail(const Level& L) {
        const dim3 g(nb), t(kRedThreads);
        hipLaunchKernelGGL(k_matvec, g, t, 0, s, B, bands, L.xj, L.xi, nullptr, part, L.g, L.g,
                           part + nb, L.g, part + 2 * nb, ctl);
        hipLaunchKernelGGL(k_fin_gam, dim3(1), t, 0, s, part, nb, sc, ctl, mode);
        hipLaunchKernelGGL(k_update_dir, g, t, 0, s, n, L.g, L.h, L.xi, sc, mode, ctl);
    }

    void solve(int depth, bool recurse, int redo_lim) {
        Level& L = lv[depth];
        redos = recurse ? redo_lim : 0;
        redos += 1;
        const dim3 g(nb), t(kRedThreads);
        // xi = A soln - b, |xi|^2, |b|^2 (:61-71)
        hipLaunchKernelGGL(k_matvec, g, t, 0, s, B, bands, soln, L.xi, rhs, part, rhs, rhs, part + nb,
                           nullptr, nullptr, nullptr);
        hipLaunchKernelGGL(k_fin_start, dim3(1), t, 0, s, part, nb, sc, ctl, mode);
        // g = -A xi, h = g (:73-75)
        hipLaunchKernelGGL(k_matvec, g, t, 0, s, B, bands, L.xi, L.g, nullptr, part, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr);
        hipLaunchKernelGGL(k_sub_neg, g, t, 0, s, n, L.h, nullptr, L.g, 1);
        const long long level_cap = 10LL * n;
        long long level_done = 0;
        for (;;) {
            const long long start = iters;
            const long long room = std::min(level_cap - level_done, max_iters - iters);
            if (room <= 0) {
                status = 3;   // iteration limit
                return;
            }
            const int nbatch = (int)std::min<long long>(kBatch, room);
            for (int b = 0; b < nbatch; ++b) {
                // xi = A h; anum = g.h / |xi|^2; xi = soln; soln += anum h (:82-99)
                hipLaunchKernelGGL(k_matvec, g, t, 0, s, B, bands, L.h, L.xi, nullptr, part, L.g, L.h,
                                   part + nb, nullptr, nullptr, ctl);
                hipLaunchKernelGGL(k_fin_anum, dim3(1), t, 0, s, part, nb, sc, ctl, mode);
                hipLaunchKernelGGL(k_update_soln, g, t, 0, s, n, soln, L.xi, L.h, sc, mode, ctl);
                // xj = A soln - b, residual2, exit / restart tests (:101-129)
                hipLaunchKernelGGL(k_matvec, g, t, 0, s, B, bands, soln, L.xj, rhs, part, nullptr,
                                   nullptr, nullptr, nullptr, nullptr, ctl);
                hipLaunchKernelGGL(k_fin_r2, dim3(1), t, 0, s, part, nb, sc, ctl, mode);
                tail(L);
            }
            const DivbCtl c = read_ctl();
            if (c.status == kStatusHipFailed) {
                status = kStatusHipFailed;
                return;
            }
            iters = c.iters;
            level_done += iters - start;
            if (c.status == 0) continue;
            residual2 = read_sc(kScR2);
            if (c.status == 1) {
                status = 1;
                return;
            }
            // restart from the previous iterate (:118-129); the gg test of a
            // finished iteration can also end the solve (status 1 above)
            hipLaunchKernelGGL(k_copy1, g, t, 0, s, n, soln, L.xi);
            if (redos >= 3 || depth + 1 >= 4) {
                status = 2;
                return;
            }
            solve(depth + 1, true, redos);
            // the reference continues this level after the nested solve:
            // resd = its residual2, then the rest of the iteration
            hipLaunchKernelGGL(k_resume, dim3(1), t, 0, s, sc, ctl);
            tail(L);
            const DivbCtl c2 = read_ctl();
            if (c2.status == kStatusHipFailed) {
                status = kStatusHipFailed;
                return;
            }
            if (c2.status == 1) {
                status = 1;
                return;
            }
            status = 0;
        }
    }
};

// ---------------------------------------------------------------------------
// Conjugate gradients on the same system (mode kDivbCG).  The 7-band
// matrix A of code/sub_divB_correct.f90:85-123 is symmetric negative
// definite, so plain CG on M = -A, c = -b converges in about the square
// root of CGNR's iteration count.  Vectors FP32, dot products and scalars
// FP64, stopping on the reference's criterion |r|^2 <= |b|^2 npts errlim^2.
// Two fused kernels and two one-block reductions per iteration; the new
// search direction p = r + beta p is formed inside the next matrix product.

enum { kCgRR = 0, kCgAlpha = 1, kCgBeta = 2, kCgB2 = 3, kCgErr2 = 8, kCgTmp = 9 };
constexpr int kCgRestarts = 8;

// p_new = r + beta p_old (r alone on the first iteration) at the 7 points,
// q = M p_new, partial sums of p_new.q
__global__ void k_cg_matvec(Box B, const float* __restrict__ r, const float* __restrict__ pold,
                            float* pnew, float* q, double* part, const double* sc,
                            const DivbCtl* ctl, int first) {
    if (ctl->status != 0) return;
    int i, j, k;
    const int c = gtid();
    double pq = 0.0;
    if (cell_of(B, c, i, j, k)) {
        const double beta = first ? 0.0 : sc[kCgBeta];
        auto P = [&](int n) {
            return first ? r[n] : (float)((double)r[n] + beta * (double)pold[n]);
        };
        const int X = 1, Y = B.nx, Z = B.nx * B.ny;
        const float pc = P(c);
        float t = 6.0f * pc;
        if (k < B.nz - 1) t = t - P(c + Z);
        if (j < B.ny - 1) t = t - P(c + Y);
        if (i < B.nx - 1) t = t - P(c + X);
        if (i > 0) t = t - P(c - X);
        if (j > 0) t = t - P(c - Y);
        if (k > 0) t = t - P(c - Z);
        pnew[c] = pc;
        q[c] = t;
        pq = (double)pc * (double)t;
    }
    const double s1 = block_sum(pq);
    if (threadIdx.x == 0) part[blockIdx.x] = s1;
}

// alpha = rr / p.q
__global__ void k_cg_alpha(const double* __restrict__ part, int nb, double* sc,
                           const DivbCtl* ctl) {
    if (ctl->status != 0) return;
    const double pq = sum_parts(part, nb);
    if (threadIdx.x == 0) sc[kCgAlpha] = sc[kCgRR] / pq;
}

// x += alpha p, r -= alpha q, partial sums of r.r
__global__ void k_cg_update(int n, float* x, float* r, const float* __restrict__ p,
                            const float* __restrict__ q, double* part, const double* sc,
                            const DivbCtl* ctl) {
    if (ctl->status != 0) return;
    const int c = gtid();
    double rr = 0.0;
    if (c < n) {
        const double a = sc[kCgAlpha];
        x[c] = (float)((double)x[c] + a * (double)p[c]);
        const float rn = (float)((double)r[c] - a * (double)q[c]);
        r[c] = rn;
        rr = (double)rn * (double)rn;
    }
    const double s1 = block_sum(rr);
    if (threadIdx.x == 0) part[blockIdx.x] = s1;
}

// rr_new, the stopping test, beta = rr_new / rr
__global__ void k_cg_beta(const double* __restrict__ part, int nb, double* sc, DivbCtl* ctl) {
    if (ctl->status != 0) return;
    const double rn = sum_parts(part, nb);
    if (threadIdx.x == 0) {
        ctl->iters += 1;
        if (rn <= sc[kCgB2] * sc[kCgErr2] || rn == 0.0) ctl->status = 1;
        sc[kCgBeta] = rn / sc[kCgRR];
        sc[kCgRR] = rn;
    }
}

// r = c = -b, x = 0, |c|^2 partials
__global__ void k_cg_init(int n, const float* __restrict__ b, float* x, float* r, double* part) {
    const int c = gtid();
    double bb = 0.0;
    if (c < n) {
        const float v = -b[c];
        r[c] = v;
        x[c] = 0.0f;
        bb = (double)v * (double)v;
    }
    const double s1 = block_sum(bb);
    if (threadIdx.x == 0) part[blockIdx.x] = s1;
}

__global__ void k_sum1(const double* __restrict__ part, int nb, double* out) {
    const double v = sum_parts(part, nb);
    if (threadIdx.x == 0) *out = v;
}

__global__ void k_cg_start(const double* __restrict__ part, int nb, double* sc, DivbCtl* ctl,
                           double err2) {
    const double bb = sum_parts(part, nb);
    if (threadIdx.x == 0) {
        sc[kCgRR] = bb;
        sc[kCgB2] = bb;
        sc[kCgErr2] = err2;
        ctl->status = bb == 0.0 ? 1 : 0;
        ctl->iters = 0;
    }
}

// Resuming from the current solution: r (= A soln - rhs, the residual of
// M soln = c) is already formed and its partial sums are in part; |c|^2,
// the tolerance and the iteration count carry on.
__global__ void k_cg_resume(const double* __restrict__ part, int nb, double* sc, DivbCtl* ctl) {
    const double rr = sum_parts(part, nb);
    if (threadIdx.x == 0) {
        sc[kCgRR] = rr;
        ctl->status = (rr <= sc[kCgB2] * sc[kCgErr2] || rr == 0.0) ? 1 : 0;
    }
}

// CG solve of A soln = rhs, from zero or (from_x) from the solution in soln;
// work: x lives in soln, 4 more vectors (r, p0, p1, q) in `vec`.  Returns 1
// when the recursive residual meets the criterion, 3 at the iteration cap;
// iters_out is the count since the solve from zero began.
static int cg_solve(const Box& B, const float* rhs, float* soln, float* vec, double* sc,
                    DivbCtl* ctl, double* part, double err2, long long max_iters,
                    long long* iters_out, bool from_x, hipStream_t s) {
    const int n = B.V();
    const int nb = (n + kRedThreads - 1) / kRedThreads;
    float* r = vec;
    float* p[2] = {vec + (size_t)n, vec + 2 * (size_t)n};
    float* q = vec + 3 * (size_t)n;
    const dim3 g(nb), t(kRedThreads);
    if (from_x) {
        hipLaunchKernelGGL(k_matvec, g, t, 0, s, B, Bands{}, soln, r, rhs, part, nullptr,
                           nullptr, nullptr, nullptr, nullptr, nullptr);
        hipLaunchKernelGGL(k_cg_resume, dim3(1), t, 0, s, part, nb, sc, ctl);
    } else {
        hipLaunchKernelGGL(k_cg_init, g, t, 0, s, n, rhs, soln, r, part);
        hipLaunchKernelGGL(k_cg_start, dim3(1), t, 0, s, part, nb, sc, ctl, err2);
    }
    long long it = 0;
    int cur = 0;
    for (;;) {
        const int nbatch = (int)std::min<long long>(kBatch, max_iters - it);
        if (nbatch <= 0) break;
        for (int q2 = 0; q2 < nbatch; ++q2) {
            hipLaunchKernelGGL(k_cg_matvec, g, t, 0, s, B, r, p[cur], p[cur ^ 1], q, part, sc, ctl,
                               (it + q2) == 0 ? 1 : 0);
            hipLaunchKernelGGL(k_cg_alpha, dim3(1), t, 0, s, part, nb, sc, ctl);
            hipLaunchKernelGGL(k_cg_update, g, t, 0, s, n, soln, r, p[cur ^ 1], q, part, sc, ctl);
            hipLaunchKernelGGL(k_cg_beta, dim3(1), t, 0, s, part, nb, sc, ctl);
            cur ^= 1;
        }
        it += nbatch;
        DivbCtl h{};
        if (!fetch(&h, ctl, sizeof(h), s)) return kStatusHipFailed;
        if (h.status != 0) {
            if (iters_out) *iters_out = h.iters;
            return 1;
        }
    }
    DivbCtl h{};
    if (!fetch(&h, ctl, sizeof(h), s)) return kStatusHipFailed;
    if (iters_out) *iters_out = h.iters;
    return h.status == 1 ? 1 : 3;
}

// ---------------------------------------------------------------------------
// Residual refinement (mode kDivbRefine): the solution is kept in FP64; each
// outer pass forms r = b - A x in FP64, solves A d = r by the FP32-vector CG
// above to a loose tolerance and adds d in FP64, until the FP64 residual
// meets the reference's criterion.  Reaches residuals FP32 products alone
// cannot (the matrix entries are small integers, exact in FP32).

// r = b - A x in FP64 (x FP64), r rounded to FP32 for the inner solve,
// partial sums of r^2 and (first pass) of b^2
__global__ void k_dres(Box B, const double* __restrict__ x, const float* __restrict__ b,
                       float* rf, double* part, int want_b2) {
    int i, j, k;
    const int c = gtid();
    double rr = 0.0, bb = 0.0;
    if (cell_of(B, c, i, j, k)) {
        const int X = 1, Y = B.nx, Z = B.nx * B.ny;
        double t = 0.0;
        if (k < B.nz - 1) t += x[c + Z];
        if (j < B.ny - 1) t += x[c + Y];
        if (i < B.nx - 1) t += x[c + X];
        t += -6.0 * x[c];
        if (i > 0) t += x[c - X];
        if (j > 0) t += x[c - Y];
        if (k > 0) t += x[c - Z];
        const double r = (double)b[c] - t;
        rf[c] = (float)r;
        rr = r * r;
        bb = (double)b[c] * (double)b[c];
    }
    const double s1 = block_sum(rr);
    if (threadIdx.x == 0) part[blockIdx.x] = s1;
    if (want_b2) {
        const double s2 = block_sum(bb);
        if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = s2;
    }
}

__global__ void k_dzero(int n, double* x) {
    const int c = gtid();
    if (c < n) x[c] = 0.0;
}

__global__ void k_dadd(int n, double* x, const float* __restrict__ d) {
    const int c = gtid();
    if (c < n) x[c] += (double)d[c];
}

__global__ void k_d2f(int n, const double* __restrict__ x, float* y) {
    const int c = gtid();
    if (c < n) y[c] = (float)x[c];
}

__global__ void k_sum2(const double* __restrict__ part, int nb, double* out) {
    const double a = sum_parts(part, nb);
    const double b = sum_parts(part + nb, nb);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
    }
}

// x (FP64, `xd`) solving A x = b to |b - A x|^2 <= |b|^2 * err2; soln gets
// x rounded to FP32.  Returns 1 converged, 3 pass / iteration cap.
static int refine_solve(const Box& B, const float* rhs, float* soln, double* xd, float* rf,
                        float* d, float* vec, double* sc, DivbCtl* ctl, double* part, double err2,
                        long long max_iters, long long* iters_out, double* res2_out,
                        hipStream_t s) {
    const int n = B.V();
    const int nb = (n + kRedThreads - 1) / kRedThreads;
    const dim3 g(nb), t(kRedThreads);
    hipLaunchKernelGGL(k_dzero, g, t, 0, s, n, xd);
    long long total = 0;
    double b2 = 0.0, r2 = 0.0;
    int status = 3;
    for (int pass = 0; pass < 20; ++pass) {
        hipLaunchKernelGGL(k_dres, g, t, 0, s, B, xd, rhs, rf, part, pass == 0 ? 1 : 0);
        hipLaunchKernelGGL(k_sum2, dim3(1), t, 0, s, part, nb, sc + kCgTmp);
        double h[2];
        if (!fetch(h, sc + kCgTmp, sizeof(h), s)) {
            status = kStatusHipFailed;
            break;
        }
        r2 = h[0];
        if (pass == 0) b2 = h[1];
        if (r2 <= b2 * err2) {
            status = 1;
            break;
        }
        if (total >= max_iters) break;
        // inner solve to 1e-4 of the current residual (FP32 vectors)
        long long it = 0;
        if (cg_solve(B, rf, d, vec, sc, ctl, part, 1e-8, max_iters - total, &it, false, s) ==
            kStatusHipFailed) {
            status = kStatusHipFailed;
            break;
        }
        total += it;
        hipLaunchKernelGGL(k_dadd, g, t, 0, s, n, xd, d);
    }
    hipLaunchKernelGGL(k_d2f, g, t, 0, s, n, xd, soln);
    if (iters_out) *iters_out = total;
    if (res2_out) *res2_out = r2;
    return status;
}

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
    S.ctl = reinterpret_cast<DivbCtl*>(dwork + 12);
    S.part = dwork + 16;
    S.s = s;
    S.max_iters = max_iters > 0 ? max_iters : 10LL * n;
    const double errlim2 = mode == kDivbRef ? (double)((float)n * (errlim * errlim))
                                            : (double)n * (double)errlim * (double)errlim;
    const dim3 g(nb), t(kRedThreads);
    hipLaunchKernelGGL(k_divb_init, dim3(1), dim3(64), 0, s, S.sc, S.ctl, errlim2);
    hipLaunchKernelGGL(k_divb_rhs, g, t, 0, s, B, bx, by, bz, rhs, 2.0f * rx);
    if (mode == kDivbRefine) {
        // FP64 solution and residual, FP32-vector inner CG; layout after
        // rhs, soln: CG vectors (4n), inner correction d (n), FP32
        // residual (n), then FP64 x (2n floats)
        float* vec = work + 2 * (size_t)n;
        float* dvec = work + 6 * (size_t)n;
        float* rf = work + 7 * (size_t)n;
        double* xd = reinterpret_cast<double*>(work + 8 * (size_t)n);
        long long it = 0;
        double r2 = 0.0;
        S.status = refine_solve(B, rhs, S.soln, xd, rf, dvec, vec, S.sc, S.ctl, S.part,
                                (double)n * (double)errlim * (double)errlim, S.max_iters, &it,
                                &r2, s);
        S.iters = it;
        S.residual2 = r2;
    } else if (mode == kDivbCG) {
        // plain CG from zero.  The recursive residual drifts from the true
        // one in FP32 vectors, so convergence is decided on the true
        // residual |A soln - b|^2: while it misses the criterion, CG resumes
        // from the solution with the true residual (at most kCgRestarts
        // times within the iteration cap; status 3 if it still misses).
        const double target = (double)n * (double)errlim * (double)errlim;
        long long it = 0;
        float* vec = work + 2 * (size_t)n;
        S.status = cg_solve(B, rhs, S.soln, vec, S.sc, S.ctl, S.part, target, S.max_iters, &it,
                            false, s);
        for (int restart = 0;; ++restart) {
            hipLaunchKernelGGL(k_matvec, g, t, 0, s, B, Bands{}, S.soln, S.lv[3].xi, rhs,
                               S.part, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
            hipLaunchKernelGGL(k_sum1, dim3(1), t, 0, s, S.part, nb, S.sc + kCgTmp);
            S.residual2 = S.read_sc(kCgTmp);
            const double b2 = S.read_sc(kCgB2);
            if (S.status != 1 || S.residual2 <= b2 * target) break;
            if (restart == kCgRestarts || it >= S.max_iters) {
                S.status = 3;
                break;
            }
            S.status = cg_solve(B, rhs, S.soln, vec, S.sc, S.ctl, S.part, target,
                                S.max_iters - it, &it, true, s);
        }
        S.iters = it;
    } else {
        // the reference's soln_vec is an uninitialised automatic array; start from 0
        hipLaunchKernelGGL(k_zero1, g, t, 0, s, n, S.soln);
        S.solve(0, false, 0);
        S.residual2 = S.read_sc(kScR2);
    }
    hipLaunchKernelGGL(k_divb_apply, g, t, 0, s, B, S.soln, bx, by, bz, rx * rx, 2.0f * rx,
                       2.0f * rx, 2.0f * rx);
    if (iters_out) *iters_out = S.iters;
    if (res2_out) *res2_out = S.residual2;
    return S.status;
}

// code/sub_sparse_solver.f90:19-159 on an explicit banded matrix: soln is
// the initial guess on entry and the solution on return
int sparse_solve(int npts, int nbands, const int* pos, const float* M, const float* rhs,
                 float* soln, float errlim, bool recurse, int redo_lim, int mode, long max_iters,
                 float* work, double* dwork, long* iters_out, double* res2_out, hipStream_t s) {
    if (nbands < 1 || nbands > 8) return -1;
    const int n = npts;
    const int nb = (n + kRedThreads - 1) / kRedThreads;
    DivbSolver S;
    S.B = Box{n, 1, 1};
    S.bands.M = M;
    S.bands.nbands = nbands;
    S.bands.npts = npts;
    for (int q = 0; q < nbands; ++q) S.bands.pos[q] = pos[q];
    S.n = n;
    S.nb = nb;
    S.mode = mode;
    S.soln = soln;
    S.rhs = rhs;
    for (int l = 0; l < 4; ++l) {
        float* base = work + (size_t)(4 * l) * n;
        S.lv[l] = Level{base, base + n, base + 2 * (size_t)n, base + 3 * (size_t)n};
    }
    S.sc = dwork;
    S.ctl = reinterpret_cast<DivbCtl*>(dwork + 12);
    S.part = dwork + 16;
    S.s = s;
    S.max_iters = max_iters > 0 ? max_iters : 10LL * n;
    const double errlim2 = mode == kDivbRef ? (double)((float)n * (errlim * errlim))
                                            : (double)n * (double)errlim * (double)errlim;
    hipLaunchKernelGGL(k_divb_init, dim3(1), dim3(64), 0, s, S.sc, S.ctl, errlim2);
    S.solve(0, recurse, redo_lim);
    S.residual2 = S.read_sc(kScR2);
    if (iters_out) *iters_out = S.iters;
    if (res2_out) *res2_out = S.residual2;
    return S.status;
}

size_t sparse_work_floats(int npts) { return (size_t)16 * npts; }

size_t divb_work_floats(int ncells) { return (size_t)18 * ncells; }
size_t divb_work_doubles(int ncells) { return 16 + 3 * (size_t)((ncells + kRedThreads - 1) / kRedThreads); }

}  // namespace g3m3
