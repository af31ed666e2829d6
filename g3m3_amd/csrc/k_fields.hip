// Field kernels of the two-step Lax-Wendroff push:
//   calcur (J = curl b), total B + fnd_evel, BandE (Ohm's law),
//   push_elec, push_bfld, store_array.
// Each kernel restates one reference loop nest; file:line citations are
// to the reference sources.
#include "g3m3_dev.h"
#include "g3m3_launch.h"

namespace g3m3 {

// ---------------------------------------------------------------------------
// code/sub_calcur.f90:126-149  interior curl, central differences
__global__ void k_calcur_int(Box B, const float* __restrict__ bx, const float* __restrict__ by,
                             const float* __restrict__ bz, float* cx, float* cy, float* cz,
                             float dxt, float dyt, float dzt) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k) || !B.interior(i, j, k)) return;
    const int c = B.IX(i, j, k);
    const int sx = 1, sy = B.nx, sz = B.nx * B.ny;
    cx[c] = (bz[c + sy] - bz[c - sy]) / dyt - (by[c + sz] - by[c - sz]) / dzt;
    cy[c] = (bx[c + sz] - bx[c - sz]) / dzt - (bz[c + sx] - bz[c - sx]) / dxt;
    cz[c] = (by[c + sx] - by[c - sx]) / dxt - (bx[c + sy] - bx[c - sy]) / dyt;
}

// code/sub_calcur.f90:161-209  face interiors copy their inner neighbour.
// The three face sets are disjoint and read only interior cells.
__global__ void k_calcur_faces(Box B, float* cx, float* cy, float* cz) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const bool fi = (i == 0 || i == B.nx - 1), fj = (j == 0 || j == B.ny - 1),
               fk = (k == 0 || k == B.nz - 1);
    if ((int)fi + (int)fj + (int)fk != 1) return;
    int si = i, sj = j, sk = k;
    if (fk) sk = (k == 0) ? 1 : B.nz - 2;
    if (fi) si = (i == 0) ? 1 : B.nx - 2;
    if (fj) sj = (j == 0) ? 1 : B.ny - 2;
    const int c = B.IX(i, j, k), s = B.IX(si, sj, sk);
    cx[c] = cx[s];
    cy[c] = cy[s];
    cz[c] = cz[s];
}

// code/sub_calcur.f90:213-262  edge lines, each from a face cell.
__global__ void k_calcur_edges(Box B, float* cx, float* cy, float* cz) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    const bool fi = (i == 0 || i == nx - 1), fj = (j == 0 || j == ny - 1),
               fk = (k == 0 || k == nz - 1);
    if ((int)fi + (int)fj + (int)fk != 2) return;
    int si = i, sj = j, sk = k;
    if (!fi) {            // lines along x (:214-227); (i,ny,nz) <- (i,ny1,nz)
        if (j == ny - 1 && k == nz - 1) sj = ny - 2;
        else sk = (k == 0) ? 1 : nz - 2;
    } else if (!fj) {     // lines along y (:231-244)
        sk = (k == 0) ? 1 : nz - 2;
    } else {              // lines along z (:248-261)
        sj = (j == 0) ? 1 : ny - 2;
    }
    const int c = B.IX(i, j, k), s = B.IX(si, sj, sk);
    cx[c] = cx[s];
    cy[c] = cy[s];
    cz[c] = cz[s];
}

// code/sub_calcur.f90:264-289  the eight corners, from edge cells.
__device__ inline void calcur_corners1(Box B, float* c) {
    const int nx = B.nx, ny = B.ny, nz = B.nz, n1 = nx - 2, m1 = ny - 2, l1 = nz - 2;
    auto at = [&](int i, int j, int k) -> float& { return c[B.IX(i, j, k)]; };
    at(0, 0, 0) = at(0, 0, 1);
    at(0, ny - 1, 0) = at(0, ny - 1, 1);
    at(0, 0, nz - 1) = (at(0, 0, l1) + at(0, 1, nz - 1)) / 2.0f;
    at(0, ny - 1, nz - 1) = (at(0, ny - 1, l1) + at(0, m1, nz - 1)) / 2.0f;
    at(nx - 1, 0, 0) = (at(nx - 1, 0, 1) + at(nx - 1, 1, 0) + at(n1, 0, 0)) / 3.0f;
    at(nx - 1, ny - 1, 0) = (at(nx - 1, ny - 1, 1) + at(nx - 1, m1, 0) + at(n1, ny - 1, 0)) / 3.0f;
    at(nx - 1, 0, nz - 1) = (at(nx - 1, 0, l1) + at(nx - 1, 1, nz - 1) + at(n1, 0, nz - 1)) / 3.0f;
    at(nx - 1, ny - 1, nz - 1) =
        (at(nx - 1, ny - 1, l1) + at(nx - 1, m1, nz - 1) + at(n1, ny - 1, nz - 1)) / 3.0f;
}
__global__ void k_calcur_corners(Box B, float* cx, float* cy, float* cz) {
    if (threadIdx.x == 0) calcur_corners1(B, cx);
    if (threadIdx.x == 1) calcur_corners1(B, cy);
    if (threadIdx.x == 2) calcur_corners1(B, cz);
}

void launch_calcur(const Box& B, const float* bx, const float* by, const float* bz, float* cx,
                   float* cy, float* cz, float rx, float ry, float rz, hipStream_t s) {
    const int n = B.V();
    const dim3 g((n + 255) / 256), t(256);
    hipLaunchKernelGGL(k_calcur_int, g, t, 0, s, B, bx, by, bz, cx, cy, cz, 2.0f * rx, 2.0f * ry,
                       2.0f * rz);
    hipLaunchKernelGGL(k_calcur_faces, g, t, 0, s, B, cx, cy, cz);
    hipLaunchKernelGGL(k_calcur_edges, g, t, 0, s, B, cx, cy, cz);
    hipLaunchKernelGGL(k_calcur_corners, dim3(1), dim3(64), 0, s, B, cx, cy, cz);
}

// ---------------------------------------------------------------------------
// code/3d_main.f90:2248-2254 (bs = b0 + b, btot) fused with
// code/sub_fnd_evel.f90:318-346 (ion and electron bulk velocity).
__global__ void k_bfield_evel(Box B, const float* __restrict__ bx0, const float* __restrict__ by0,
                              const float* __restrict__ bz0, const float* __restrict__ bx,
                              const float* __restrict__ by, const float* __restrict__ bz,
                              float* bsx, float* bsy, float* bsz, float* btot,
                              const float* __restrict__ qpx, const float* __restrict__ qpy,
                              const float* __restrict__ qpz, const float* __restrict__ qrho,
                              const float* __restrict__ hpx, const float* __restrict__ hpy,
                              const float* __restrict__ hpz, const float* __restrict__ hrho,
                              const float* __restrict__ opx, const float* __restrict__ opy,
                              const float* __restrict__ opz, const float* __restrict__ orho,
                              const float* __restrict__ cx, const float* __restrict__ cy,
                              const float* __restrict__ cz, float* evx, float* evy, float* evz,
                              float* tvx, float* tvy, float* tvz, float rmq, float rmh, float rmo,
                              float reynolds, float smallbit, int do_b, int do_v) {
    const int c = gtid();
    if (c >= B.V()) return;
    if (do_b) {
        const float sx = bx0[c] + bx[c], sy = by0[c] + by[c], sz = bz0[c] + bz[c];
        bsx[c] = sx;
        bsy[c] = sy;
        bsz[c] = sz;
        btot[c] = fmaxf(sqrtf(sx * sx + sy * sy + sz * sz), smallbit);
    }
    if (do_v) {
        const float q = qrho[c], h = hrho[c], o = orho[c];
        const float qden = (q + 0.000001f) / rmq;
        const float hden = (h + 0.000001f) / rmh;
        const float oden = (o + 0.000001f) / rmo;
        const float tot = qden + hden + oden;
        const float tden = q + h + o;
        const float vx = (qpx[c] + hpx[c] + opx[c]) / tden;
        evx[c] = vx - cx[c] / tot / reynolds;
        const float vy = (qpy[c] + hpy[c] + opy[c]) / tden;
        evy[c] = vy - cy[c] / tot / reynolds;
        const float vz = (qpz[c] + hpz[c] + opz[c]) / tden;
        evz[c] = vz - cz[c] / tot / reynolds;
        tvx[c] = vx;
        tvy[c] = vy;
        tvz[c] = vz;
    }
}

void launch_bfield_evel(const Box& B, const float* b0[3], const float* b[3], float* bs[3],
                        float* btot, const float* mom[9], const float* rho[3], const float* cur[3],
                        float* ev[3], float* tv[3], float rmq, float rmh, float rmo,
                        float reynolds, float smallbit, int do_b, int do_v, hipStream_t s) {
    const int n = B.V();
    hipLaunchKernelGGL(k_bfield_evel, dim3((n + 255) / 256), dim3(256), 0, s, B, b0[0], b0[1],
                       b0[2], b[0], b[1], b[2], bs[0], bs[1], bs[2], btot, mom[0], mom[1], mom[2],
                       rho[0], mom[3], mom[4], mom[5], rho[1], mom[6], mom[7], mom[8], rho[2],
                       cur[0], cur[1], cur[2], ev[0], ev[1], ev[2], tv[0], tv[1], tv[2], rmq, rmh,
                       rmo, reynolds, smallbit, do_b, do_v);
}

// ---------------------------------------------------------------------------
// code/sub_BandE.f90:32-84  E = -(ve x B) everywhere, -grad pe / (n Re)
// on the interior.
__global__ void k_bande_main(Box B, float* ex, float* ey, float* ez,
                             const float* __restrict__ bsx, const float* __restrict__ bsy,
                             const float* __restrict__ bsz, const float* __restrict__ evx,
                             const float* __restrict__ evy, const float* __restrict__ evz,
                             const float* __restrict__ ep, const float* __restrict__ q,
                             const float* __restrict__ h, const float* __restrict__ o, float rmq,
                             float rmh, float rmo, float reynolds, float dxt, float dyt,
                             float dzt) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const int c = B.IX(i, j, k);
    const float avx = evx[c], avy = evy[c], avz = evz[c];
    const float abx = bsx[c], aby = bsy[c], abz = bsz[c];
    float Ex = -(avy * abz - avz * aby);
    float Ey = -(avz * abx - avx * abz);
    float Ez = -(avx * aby - avy * abx);
    if (B.interior(i, j, k)) {
        const int sy = B.nx, sz = B.nx * B.ny;
        const float arho = (q[c] / rmq + h[c] / rmh + o[c] / rmo) * reynolds;
        Ex = Ex - ((ep[c + 1] - ep[c - 1]) / dxt) / arho;
        Ey = Ey - ((ep[c + sy] - ep[c - sy]) / dyt) / arho;
        Ez = Ez - ((ep[c + sz] - ep[c - sz]) / dzt) / arho;
    }
    ex[c] = Ex;
    ey[c] = Ey;
    ez[c] = Ez;
}

// code/sub_BandE.f90:88-116  rst*J/n on one inner-boundary list.
// ij is the reference's ij*(mbndry,3,m) array; entry n of box b is
// ij[(b) + mbndry*(axis + 3*n)].
__global__ void k_bande_list(Box B, float* ex, float* ey, float* ez, const float* __restrict__ cx,
                             const float* __restrict__ cy, const float* __restrict__ cz,
                             const float* __restrict__ rst, const float* __restrict__ q,
                             const float* __restrict__ h, const float* __restrict__ o,
                             const int* __restrict__ ij, int mbndry, int b, int num, float rmq,
                             float rmh, float rmo) {
    const int n = gtid();
    if (n >= num) return;
    const int i = ij[b + mbndry * (0 + 3 * n)] - 1;
    const int j = ij[b + mbndry * (1 + 3 * n)] - 1;
    const int k = ij[b + mbndry * (2 + 3 * n)] - 1;
    const int c = B.IX(i, j, k);
    const float eden = q[c] / rmq + h[c] / rmh + o[c] / rmo;
    const float r = rst[c];
    ex[c] = ex[c] + r * cx[c] / eden;
    ey[c] = ey[c] + r * cy[c] / eden;
    ez[c] = ez[c] + r * cz[c] / eden;
}

// Uniform resistivity (no reference counterpart): eta*J/n on the interior,
// after the list term.
__global__ void k_bande_eta(Box B, float* ex, float* ey, float* ez, const float* __restrict__ cx,
                            const float* __restrict__ cy, const float* __restrict__ cz,
                            const float* __restrict__ q, const float* __restrict__ h,
                            const float* __restrict__ o, float rmq, float rmh, float rmo,
                            float eta) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k) || !B.interior(i, j, k)) return;
    const int c = B.IX(i, j, k);
    const float eden = q[c] / rmq + h[c] / rmh + o[c] / rmo;
    ex[c] = ex[c] + eta * cx[c] / eden;
    ey[c] = ey[c] + eta * cy[c] / eden;
    ez[c] = ez[c] + eta * cz[c] / eden;
}

// code/sub_BandE.f90:125-159  flank copies x, then z, then y: every
// boundary cell ends up equal to the interior cell at the clamped index.
__global__ void k_bande_faces(Box B, float* ex, float* ey, float* ez) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k) || B.interior(i, j, k)) return;
    const int si = min(max(i, 1), B.nx - 2), sj = min(max(j, 1), B.ny - 2),
              sk = min(max(k, 1), B.nz - 2);
    const int c = B.IX(i, j, k), s = B.IX(si, sj, sk);
    ex[c] = ex[s];
    ey[c] = ey[s];
    ez[c] = ez[s];
}

void launch_bande(const Box& B, float* e[3], const float* bs[3], const float* cur[3],
                  const float* ev[3], const float* ep, const float* rho[3], const float* rst,
                  float resist, float reynolds, int box, float rmq, float rmh, float rmo,
                  const int* ijmid, int nummid, const int* ijzero, int numzero, int mbndry,
                  float rx, float ry, float rz, float eta, hipStream_t s) {
    const int n = B.V();
    const dim3 g((n + 255) / 256), t(256);
    hipLaunchKernelGGL(k_bande_main, g, t, 0, s, B, e[0], e[1], e[2], bs[0], bs[1], bs[2], ev[0],
                       ev[1], ev[2], ep, rho[0], rho[1], rho[2], rmq, rmh, rmo, reynolds,
                       2.0f * rx, 2.0f * ry, 2.0f * rz);
    if (box <= mbndry && resist < 5000.0f) {
        const float* rb = rst + (size_t)(box - 1) * n;
        if (nummid > 0)
            hipLaunchKernelGGL(k_bande_list, dim3((nummid + 255) / 256), t, 0, s, B, e[0], e[1],
                               e[2], cur[0], cur[1], cur[2], rb, rho[0], rho[1], rho[2], ijmid,
                               mbndry, box - 1, nummid, rmq, rmh, rmo);
        if (numzero > 0)
            hipLaunchKernelGGL(k_bande_list, dim3((numzero + 255) / 256), t, 0, s, B, e[0], e[1],
                               e[2], cur[0], cur[1], cur[2], rb, rho[0], rho[1], rho[2], ijzero,
                               mbndry, box - 1, numzero, rmq, rmh, rmo);
    }
    if (eta != 0.0f)
        hipLaunchKernelGGL(k_bande_eta, g, t, 0, s, B, e[0], e[1], e[2], cur[0], cur[1], cur[2],
                           rho[0], rho[1], rho[2], rmq, rmh, rmo, eta);
    hipLaunchKernelGGL(k_bande_faces, g, t, 0, s, B, e[0], e[1], e[2]);
}

// ---------------------------------------------------------------------------
// code/sub_push_elec.f90:19-50
__global__ void k_push_elec(Box B, const float* __restrict__ w, const float* __restrict__ old,
                            float* ep, const float* __restrict__ evx, const float* __restrict__ evy,
                            const float* __restrict__ evz, float gamma, float gamma1, float delt,
                            float dxt, float dyt, float dzt) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k) || !B.interior(i, j, k)) return;
    const int c = B.IX(i, j, k), sy = B.nx, sz = B.nx * B.ny;
    const float gx = (w[c + 1] - w[c - 1]) / dxt;
    const float gy = (w[c + sy] - w[c - sy]) / dyt;
    const float gz = (w[c + sz] - w[c - sz]) / dzt;
    const float div = ((w[c + 1] * evx[c + 1] - w[c - 1] * evx[c - 1]) / dxt) +
                      ((w[c + sy] * evy[c + sy] - w[c - sy] * evy[c - sy]) / dyt) +
                      ((w[c + sz] * evz[c + sz] - w[c - sz] * evz[c - sz]) / dzt);
    const float e = old[c] - delt * gamma * div;
    ep[c] = e + delt * gamma1 * (evx[c] * gx + evy[c] * gy + evz[c] * gz);
}

void launch_push_elec(const Box& B, const float* w, const float* old, float* ep, const float* ev[3],
                      float gamma, float gamma1, float delt, float rx, float ry, float rz,
                      hipStream_t s) {
    const int n = B.V();
    hipLaunchKernelGGL(k_push_elec, dim3((n + 255) / 256), dim3(256), 0, s, B, w, old, ep, ev[0],
                       ev[1], ev[2], gamma, gamma1, delt, 2.0f * rx, 2.0f * ry, 2.0f * rz);
}

// code/sub_push_bfld.f90:19-47  induction
__global__ void k_push_bfld(Box B, float* bx, float* by, float* bz, const float* __restrict__ ox,
                            const float* __restrict__ oy, const float* __restrict__ oz,
                            const float* __restrict__ ex, const float* __restrict__ ey,
                            const float* __restrict__ ez, float delt, float dxt, float dyt,
                            float dzt) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k) || !B.interior(i, j, k)) return;
    const int c = B.IX(i, j, k), sy = B.nx, sz = B.nx * B.ny;
    bx[c] = ox[c] - delt * (((ez[c + sy] - ez[c - sy]) / dyt) - ((ey[c + sz] - ey[c - sz]) / dzt));
    by[c] = oy[c] + delt * (((ez[c + 1] - ez[c - 1]) / dxt) - ((ex[c + sz] - ex[c - sz]) / dzt));
    bz[c] = oz[c] - delt * (((ey[c + 1] - ey[c - 1]) / dxt) - ((ex[c + sy] - ex[c - sy]) / dyt));
}

void launch_push_bfld(const Box& B, float* b[3], const float* old[3], const float* e[3], float delt,
                      float rx, float ry, float rz, hipStream_t s) {
    const int n = B.V();
    hipLaunchKernelGGL(k_push_bfld, dim3((n + 255) / 256), dim3(256), 0, s, B, b[0], b[1], b[2],
                       old[0], old[1], old[2], e[0], e[1], e[2], delt, 2.0f * rx, 2.0f * ry,
                       2.0f * rz);
}

// ---------------------------------------------------------------------------
// code/sub_store_array.f90:10-17, batched over fields (blockIdx.y)
__global__ void k_copy_fields(int n, FieldPtrs dst, FieldPtrs src) {
    const int c = gtid();
    if (c >= n) return;
    dst.p[blockIdx.y][c] = src.p[blockIdx.y][c];
}

void launch_copy_fields(int n, int nf, const FieldPtrs& dst, const FieldPtrs& src, hipStream_t s) {
    if (nf <= 0) return;
    {
        ProfScope prof_("k_copy_fields", s);
        hipLaunchKernelGGL(k_copy_fields, dim3((n + 255) / 256, nf), dim3(256), 0, s, n, dst, src);
    }
}

// exchange the contents of two field sets (blockIdx.y = field)
__global__ void k_swap_fields(int n, FieldPtrs a, FieldPtrs b) {
    const int c = gtid();
    if (c >= n) return;
    float* pa = a.p[blockIdx.y];
    float* pb = b.p[blockIdx.y];
    const float t = pa[c];
    pa[c] = pb[c];
    pb[c] = t;
}

void launch_swap_fields(int n, int nf, const FieldPtrs& a, const FieldPtrs& b, hipStream_t s) {
    if (nf <= 0) return;
    hipLaunchKernelGGL(k_swap_fields, dim3((n + 255) / 256, nf), dim3(256), 0, s, n, a, b);
}

}  // namespace g3m3
