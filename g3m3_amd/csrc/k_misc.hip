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
// of one cell from the tensor diagonal (x, y, z), the flow v and the field b
__device__ inline void fnd_pres_point(float x, float y, float z, float avx, float avy, float avz,
                                      float abx, float aby, float abz, float& pa, float& pp,
                                      float& pc) {
    const float presmag = sqrtf(x * x + y * y + z * z) + 1.e-11f;
    const float bmag = sqrtf(abx * abx + aby * aby + abz * abz) + 1.e-12f;
    const float vcbx = (avy * abz - avz * aby) / bmag;
    const float vcby = -(avx * abz - avz * abx) / bmag;
    const float vcbz = (avx * aby - avy * abx) / bmag;
    const float vcmag = sqrtf(vcbx * vcbx + vcby * vcby + vcbz * vcbz) + 1.e-12f;
    const float ax = abx * x, ay = aby * y, az = abz * z;
    pa = sqrtf(ax * ax + ay * ay + az * az) / bmag;
    const float cx = vcbx * x, cy = vcby * y, cz = vcbz * z;
    pc = sqrtf(cx * cx + cy * cy + cz * cz) / vcmag;
    pp = sqrtf(fabsf(presmag * presmag - (pa * pa + pc * pc)));
}

__global__ void k_fnd_pres(int n, const float* __restrict__ px, const float* __restrict__ py,
                           const float* __restrict__ pz, const float* __restrict__ vx,
                           const float* __restrict__ vy, const float* __restrict__ vz,
                           const float* __restrict__ bx, const float* __restrict__ by,
                           const float* __restrict__ bz, float* para, float* perp, float* cross) {
    const int c = gtid();
    if (c >= n) return;
    float pa, pp, pc;
    fnd_pres_point(px[c], py[c], pz[c], vx[c], vy[c], vz[c], bx[c], by[c], bz[c], pa, pp, pc);
    para[c] = pa;
    cross[c] = pc;
    perp[c] = pp;
}

// ---------------------------------------------------------------------------
// The quantities code/sub_write_graphing_data.f90:348-760 tabulates on its
// three cut planes of one box, in its physical units and column order.  One
// thread per plane point: cut 0 the xy plane k = kc (points j, i), cut 1 the
// xz plane j = jc (k, i), cut 2 the yz plane i = ic (k, j), as the
// reference's loops run.  Per point 53 values: plas 8, flow 13, pres 10,
// bande 11, model 11.  Single precision in the reference's expression
// order; log10 is evaluated in double and rounded once.
__device__ inline float log10r(float x) { return (float)log10((double)x); }

__global__ void k_graph_cuts(GraphCuts G, float* __restrict__ out) {
    const int p = gtid();
    const Box& B = G.B;
    const int n0 = B.nx * B.ny, n1 = B.nx * B.nz, n2 = B.ny * B.nz;
    if (p >= n0 + n1 + n2) return;
    int i, j, k;
    if (p < n0) {
        k = G.cut[0];
        j = p / B.nx;
        i = p - j * B.nx;
    } else if (p < n0 + n1) {
        j = G.cut[1];
        k = (p - n0) / B.nx;
        i = (p - n0) - k * B.nx;
    } else {
        i = G.cut[2];
        k = (p - n0 - n1) / B.ny;
        j = (p - n0 - n1) - k * B.ny;
    }
    const int c = B.IX(i, j, k);
    float* o = out + (size_t)53 * p;
    const float ve = G.v_equiv;
    const float unetx = G.tv[0][c] * ve, unety = G.tv[1][c] * ve, unetz = G.tv[2][c] * ve;
    const float net_flow = sqrtf(unetx * unetx + unety * unety + unetz * unetz);
    const float curx_all = G.cur[0][c] * G.cur_equiv;
    const float cury_all = G.cur[1][c] * G.cur_equiv;
    const float curz_all = G.cur[2][c] * G.cur_equiv;
    const float bt[3] = {G.b0[0][c] + G.b[0][c], G.b0[1][c] + G.b[1][c], G.b0[2][c] + G.b[2][c]};
    float rho[3], apres[3], v[3][3], lp[3][3];
    for (int s = 0; s < 3; ++s) {
        rho[s] = G.sp[s][0][c];
        const float px = G.sp[s][1][c], py = G.sp[s][2][c], pz = G.sp[s][3][c];
        const float x = G.sp[s][4][c], y = G.sp[s][5][c], z = G.sp[s][6][c];
        apres[s] = fmaxf(0.0000001f, 0.333f * (x + y + z));
        const float den = fmaxf(rho[s], G.smallbit);
        const float vx = px / den, vy = py / den, vz = pz / den;
        float pa, pp, pc;
        fnd_pres_point(x, y, z, vx, vy, vz, bt[0], bt[1], bt[2], pa, pp, pc);
        lp[s][0] = log10r(pa * G.pres_equiv);
        lp[s][1] = log10r(pp * G.pres_equiv);
        lp[s][2] = log10r(pc * G.pres_equiv);
        v[s][0] = vx * ve;
        v[s][1] = vy * ve;
        v[s][2] = vz * ve;
    }
    const float aepres = fmaxf(0.0000001f, G.epres[c]);
    float qden = fmaxf(rho[0], 0.000001f);
    float hden = fmaxf(0.0001f, rho[1]);
    float oden = fmaxf(0.0001f, rho[2]);
    const float tden = (rho[0] + rho[1] + rho[2]) * G.rho_equiv;
    qden = qden / G.rmass[0] * G.rho_equiv;
    hden = hden / G.rmass[1] * G.rho_equiv;
    oden = oden / G.rmass[2] * G.rho_equiv;
    const float eden = oden + hden + qden + 0.0001f;
    const float te = G.temp_equiv;
    o[0] = log10r(qden);
    o[1] = fmaxf(0.00001f, apres[0] / qden) * te;
    o[2] = log10r(hden);
    o[3] = apres[1] / hden * te;
    o[4] = log10r(oden);
    o[5] = apres[2] / oden * te;
    o[6] = log10r(eden);
    o[7] = aepres / eden * te;
    const float be = G.b_equiv;
    const float abx = be * (G.b[0][c] + G.b0[0][c]);
    const float aby = be * (G.b[1][c] + G.b0[1][c]);
    const float abz = be * (G.b[2][c] + G.b0[2][c]);
    const float bmag = sqrtf(abx * abx + aby * aby + abz * abz);
    const float v_alf = bmag * 1.e-9f / sqrtf(G.mu0 * tden * 1.e6f * G.m_prot);
    for (int s = 0; s < 3; ++s)
        for (int a = 0; a < 3; ++a) o[8 + 3 * s + a] = v[s][a];
    o[17] = unetx;
    o[18] = unety;
    o[19] = unetz;
    o[20] = net_flow * 1.e3f / v_alf;
    for (int s = 0; s < 3; ++s)
        for (int a = 0; a < 3; ++a) o[21 + 3 * s + a] = lp[s][a];
    o[30] = log10r(aepres * G.pres_equiv);
    const float ee = G.e_equiv;
    const float aex = G.e[0][c] * ee, aey = G.e[1][c] * ee, aez = G.e[2][c] * ee;
    o[31] = abx;
    o[32] = aby;
    o[33] = abz;
    o[34] = bmag;
    o[35] = aex;
    o[36] = aey;
    o[37] = aez;
    o[38] = sqrtf(aex * aex + aey * aey + aez * aez);
    o[39] = curx_all;
    o[40] = cury_all;
    o[41] = curz_all;
    const float ri = (G.xmin + G.xs * (float)i) * G.re_equiv;
    const float rj = (G.ymin + G.xs * (float)j) * G.re_equiv;
    const float rk = (G.zmin + G.xs * (float)k) * G.re_equiv;
    o[42] = ri;
    o[43] = rj;
    o[44] = rk;
    o[45] = sqrtf(ri * ri + rj * rj + rk * rk);
    const float b0x = G.b0[0][c], b0y = G.b0[1][c], b0z = G.b0[2][c];
    o[46] = be * G.b[0][c];
    o[47] = be * G.b[1][c];
    o[48] = be * G.b[2][c];
    o[49] = be * b0x;
    o[50] = be * b0y;
    o[51] = be * b0z;
    o[52] = sqrtf(b0x * b0x + b0y * b0y + b0z * b0z) * be;
}

void launch_graph_cuts(const GraphCuts& G, float* out, hipStream_t s) {
    const int n = G.B.nx * G.B.ny + G.B.nx * G.B.nz + G.B.ny * G.B.nz;
    hipLaunchKernelGGL(k_graph_cuts, dim3((n + 255) / 256), dim3(256), 0, s, G, out);
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
