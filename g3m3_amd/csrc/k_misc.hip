// Small kernels: the general banded product of code/sub_banded_tpose.f90
// (for the drop-in entry point; the divB solve itself is matrix-free) and
// batched trilinear sampling of code/cft_trilin_interp.f90:10-66.
#include "g3m3_craft.h"
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

// ---------------------------------------------------------------------------
// code/sub_fnd_pres.f90:29-62: parallel / perpendicular / cross pressures
__global__ void k_fnd_pres(int n, const float* __restrict__ px, const float* __restrict__ py,
                           const float* __restrict__ pz, const float* __restrict__ vx,
                           const float* __restrict__ vy, const float* __restrict__ vz,
                           const float* __restrict__ bx, const float* __restrict__ by,
                           const float* __restrict__ bz, float* para, float* perp, float* cross) {
    const int c = gtid();
    if (c >= n) return;
    const float x = px[c], y = py[c], z = pz[c];
    const float presmag = sqrtf(x * x + y * y + z * z) + 1.e-11f;
    const float abx = bx[c], aby = by[c], abz = bz[c];
    const float bmag = sqrtf(abx * abx + aby * aby + abz * abz) + 1.e-12f;
    const float avx = vx[c], avy = vy[c], avz = vz[c];
    const float vcbx = (avy * abz - avz * aby) / bmag;
    const float vcby = -(avx * abz - avz * abx) / bmag;
    const float vcbz = (avx * aby - avy * abx) / bmag;
    const float vcmag = sqrtf(vcbx * vcbx + vcby * vcby + vcbz * vcbz) + 1.e-12f;
    const float ax = abx * x, ay = aby * y, az = abz * z;
    const float pa = sqrtf(ax * ax + ay * ay + az * az) / bmag;
    const float cx = vcbx * x, cy = vcby * y, cz = vcbz * z;
    const float pc = sqrtf(cx * cx + cy * cy + cz * cz) / vcmag;
    para[c] = pa;
    cross[c] = pc;
    perp[c] = sqrtf(fabsf(presmag * presmag - (pa * pa + pc * pc)));
}

void launch_fnd_pres(int n, const float* const p[3], const float* const v[3],
                     const float* const b[3], float* para, float* perp, float* cross,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_fnd_pres, dim3((n + 255) / 256), dim3(256), 0, s, n, p[0], p[1], p[2],
                       v[0], v[1], v[2], b[0], b[1], b[2], para, perp, cross);
}

// ---------------------------------------------------------------------------
// Craft sampling at many points (code/cft_craft_handling.f90:165-195 with
// code/cft_record_aux_data.f90:27-30): findgrid, the aux cube, the 2x2x2
// cube of every field and trilin_interp, one point per thread.  A point
// whose cube would leave its box gets gridpt[0] = 0 and zero samples.
// With trajectories (xcraft, zcraft: (x, y, z, t) per craft) each point is
// first placed at time ut as find_aux_cube places it, and written to sxyz.
__global__ void k_sample(int npts, float* __restrict__ sxyz, const float* __restrict__ xcraft,
                         const float* __restrict__ zcraft, float ut, CraftGrid G, int nx, int ny,
                         int nz, int nq, FieldPtrs F, float* out, int* gridpt) {
    const int p = gtid();
    if (p >= npts) return;
    float s[3];
    if (xcraft) {
        const float xc[4] = {xcraft[4 * p], xcraft[4 * p + 1], xcraft[4 * p + 2],
                             xcraft[4 * p + 3]};
        const float zc[4] = {zcraft[4 * p], zcraft[4 * p + 1], zcraft[4 * p + 2],
                             zcraft[4 * p + 3]};
        craft_position(xc, zc, ut, s);
        for (int a = 0; a < 3; ++a) sxyz[3 * p + a] = s[a];
    } else {
        for (int a = 0; a < 3; ++a) s[a] = sxyz[3 * p + a];
    }
    int gp[4];
    craft_findgrid(G, s, gp);
    const bool ok = gp[0] >= 1 && gp[0] <= nx - 1 && gp[1] >= 1 && gp[1] <= ny - 1 &&
                    gp[2] >= 1 && gp[2] <= nz - 1;
    if (!ok) gp[0] = 0;
    for (int a = 0; a < 4; ++a) gridpt[4 * p + a] = gp[a];
    if (!ok) {
        for (int q = 0; q < nq; ++q) out[q + nq * p] = 0.0f;
        return;
    }
    float v0[3], v1[3];
    craft_cube(G, gp, v0, v1);
    const float xd = (s[0] - v0[0]) / (v1[0] - v0[0]);
    const float yd = (s[1] - v0[1]) / (v1[1] - v0[1]);
    const float zd = (s[2] - v0[2]) / (v1[2] - v0[2]);
    const size_t V = (size_t)nx * ny * nz;
    const size_t c = (size_t)(gp[3] - 1) * V + (size_t)(gp[2] - 1) * nx * ny +
                     (size_t)(gp[1] - 1) * nx + (gp[0] - 1);
    const size_t Y = nx, Z = (size_t)nx * ny;
    for (int q = 0; q < nq; ++q) {
        const float* f = F.p[q];
        const float e00 = f[c] * (1.0f - xd) + f[c + 1] * xd;
        const float e01 = f[c + Z] * (1.0f - xd) + f[c + Z + 1] * xd;
        const float e10 = f[c + Y] * (1.0f - xd) + f[c + Y + 1] * xd;
        const float e11 = f[c + Y + Z] * (1.0f - xd) + f[c + Y + Z + 1] * xd;
        const float f0 = e00 * (1.0f - yd) + e10 * yd;
        const float f1 = e01 * (1.0f - yd) + e11 * yd;
        out[q + nq * p] = f0 * (1.0f - zd) + f1 * zd;
    }
}

void launch_sample(int npts, float* sxyz, const float* xcraft, const float* zcraft, float ut,
                   const CraftGrid& G, int nx, int ny, int nz, int nq, const FieldPtrs& F,
                   float* out, int* gridpt, hipStream_t s) {
    if (npts <= 0) return;
    hipLaunchKernelGGL(k_sample, dim3((npts + 255) / 256), dim3(256), 0, s, npts, sxyz, xcraft,
                       zcraft, ut, G, nx, ny, nz, nq, F, out, gridpt);
}

}  // namespace g3m3

namespace g3m3 {

// ---------------------------------------------------------------------------
// flux_counter's wall sums (code/sub_flux_counter.f90:26-151): thread t sums
// one species' momentum over one wall plane, in the reference's loop order
// (sequential single precision, as the reference accumulates), into
// sums[t]; t = 6*axis + 3*side + species, axis 0: px on i = 2 and nx-1,
// 1: py on j = 2 and ny-1, 2: pz on k = 2 and nz-1 (1-based).
__global__ void k_flux_walls(Box B, FieldPtrs P, float* sums) {
    const int t = threadIdx.x;
    if (t >= 18) return;
    const int axis = t / 6, side = (t / 3) % 2, s = t % 3;
    const float* p = P.p[3 * s + axis];
    float acc = 0.0f;
    if (axis == 0) {
        const int i = side == 0 ? 1 : B.nx - 2;
        for (int k = 0; k < B.nz; ++k)
            for (int j = 0; j < B.ny; ++j) acc = acc + p[B.IX(i, j, k)];
    } else if (axis == 1) {
        const int j = side == 0 ? 1 : B.ny - 2;
        for (int k = 0; k < B.nz; ++k)
            for (int i = 0; i < B.nx; ++i) acc = acc + p[B.IX(i, j, k)];
    } else {
        const int k = side == 0 ? 1 : B.nz - 2;
        for (int j = 0; j < B.ny; ++j)
            for (int i = 0; i < B.nx; ++i) acc = acc + p[B.IX(i, j, k)];
    }
    sums[t] = acc;
}

void launch_flux_walls(const Box& B, const FieldPtrs& P, float* sums, hipStream_t s) {
    hipLaunchKernelGGL(k_flux_walls, dim3(1), dim3(64), 0, s, B, P, sums);
}

}  // namespace g3m3
