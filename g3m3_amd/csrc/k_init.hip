// Initialisation on the device:
//   the planet's dipole field (code/sub_dipole.f90:4-52) and its grid form
//   mak_dip_grd (code/sub_mak_dip_grd.f90:4-112);
//   the initialisation block of the main program (code/3d_main.f90:1245-
//   1627): ambient ionospheric plasma, the inner-boundary lists with their
//   saved values, the solar wind, the speed check;
//   the ionospheric resistivity set_resist (code/sub_set_resist.f90:10-63).
//
// Every value is the reference's expression in the reference's order, in
// IEEE single precision with no fused multiply-add.  Integer powers x**n
// are expanded left to right, ((x*x)*x)*..., as flang expands them.  The
// one real power, the ionospheric profile ((r+r_i/2)/(1.5 r_i))**(-alpha_e)
// (code/3d_main.f90:1311-1312), is evaluated in double precision and
// rounded once; a C library's powf may differ from it by an ulp.
#include "g3m3_dev.h"
#include "g3m3_launch.h"

namespace g3m3 {

__host__ __device__ inline float pow5(float x) { return x * x * x * x * x; }

// code/sub_dipole.f90:9-49 given sin/cos of the rotation angle; x1, y1, z1
// are the position minus the dipole centre.  Returns |r| in dipole space.
__host__ __device__ inline float dipole_field(float x1, float y1, float z1, float sin_rot,
                                              float cos_rot, float sin_tilt, float cos_tilt,
                                              float b0, float& abx, float& aby, float& abz) {
    const float xr = x1 * cos_rot + y1 * sin_rot;
    const float yr = -(x1 * sin_rot) + y1 * cos_rot;
    const float xp = xr * cos_tilt - z1 * sin_tilt;
    const float yp = yr;
    const float zp = xr * sin_tilt + z1 * cos_tilt;
    const float x2 = xp * xp, y2 = yp * yp, z2 = zp * zp;
    const float ar = sqrtf(x2 + y2 + z2);
    const float bmag = -(b0 / pow5(ar));
    const float dbx = -(3.0f * bmag * xp * zp);
    const float dby = -(3.0f * bmag * yp * zp);
    const float dbz = bmag * (x2 + y2 - 2.0f * z2);
    const float rbx = dbx * cos_tilt + dbz * sin_tilt;
    const float rby = dby;
    const float rbz = -(dbx * sin_tilt) + dbz * cos_tilt;
    abx = rbx * cos_rot - rby * sin_rot;
    aby = rbx * sin_rot + rby * cos_rot;
    abz = rbz;
    return ar;
}

void dipole_host(float ax, float ay, float az, const RotationGeom& R, float& abx, float& aby,
                 float& abz) {
    dipole_field(ax - R.xdip, ay - R.ydip, az - R.zdip, R.sin_rot, R.cos_rot, R.sin_tilt,
                 R.cos_tilt, R.b0, abx, aby, abz);
}

// ---------------------------------------------------------------------------
// mak_dip_grd: the dipole on every cell of one box, zero inside r_inner-1.5
// (code/sub_mak_dip_grd.f90:29-88)
__global__ void k_mak_dip_grd(Box B, float* bx0, float* by0, float* bz0, InitGeom G,
                              RotationGeom R, float r_inner) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const float az = G.zmin + G.dz * (float)k;
    const float z1 = az - R.zdip;
    const float ay = G.ymin + G.dy * (float)j;
    const float y1 = ay - R.ydip;
    const float ax = G.xmin + G.dx * (float)i;
    const float x1 = ax - R.xdip;
    float abx, aby, abz;
    const float ar = dipole_field(x1, y1, z1, R.sin_rot, R.cos_rot, R.sin_tilt, R.cos_tilt, R.b0,
                                  abx, aby, abz);
    const int c = B.IX(i, j, k);
    const bool keep = ar > r_inner - 1.5f;
    bx0[c] = keep ? abx : 0.0f;
    by0[c] = keep ? aby : 0.0f;
    bz0[c] = keep ? abz : 0.0f;
}

// zero the field at the listed points of box `box` (1-based) of a list
// ij(mbndry, 3, m) (code/sub_mak_dip_grd.f90:93-107)
__global__ void k_zero_at_list(Box B, float* f0, float* f1, float* f2, const int* ij, int mbndry,
                               int box, int num) {
    const int n = gtid();
    if (n >= num) return;
    const int b = box - 1;
    const int i = ij[b + mbndry * (0 + 3 * n)] - 1;
    const int j = ij[b + mbndry * (1 + 3 * n)] - 1;
    const int k = ij[b + mbndry * (2 + 3 * n)] - 1;
    if ((unsigned)i >= (unsigned)B.nx || (unsigned)j >= (unsigned)B.ny ||
        (unsigned)k >= (unsigned)B.nz)
        return;   // a point outside the box is skipped, not written
    const int c = B.IX(i, j, k);
    f0[c] = 0.0f;
    f1[c] = 0.0f;
    f2[c] = 0.0f;
}

void launch_mak_dip_grd(const Box& B, float* bx0, float* by0, float* bz0, const InitGeom& G,
                        const RotationGeom& R, float r_inner, const int* ijzero, int mbndry,
                        int box, int numzero, hipStream_t s) {
    hipLaunchKernelGGL(k_mak_dip_grd, dim3((B.V() + 255) / 256), dim3(256), 0, s, B, bx0, by0,
                       bz0, G, R, r_inner);
    if (box <= mbndry && numzero > 0)
        hipLaunchKernelGGL(k_zero_at_list, dim3((numzero + 255) / 256), dim3(256), 0, s, B, bx0,
                           by0, bz0, ijzero, mbndry, box, numzero);
}

// ---------------------------------------------------------------------------
// set_resist (code/sub_set_resist.f90:24-58): rst = 0 on every box
// 1..mbndry, then b0/resist, 0.5 b0/resist, 0.125 b0/resist at the zero,
// mid and surface points.  The reference takes the point counts of box
// mbndry for every box; so does this.

__global__ void k_fill(float* p, size_t n, float v) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// one list: n over the points, every box per point
// (the point count comes from counts[which] on the device, or from hcount
// when counts is null)
__global__ void k_resist_list(Box B, float* rst, const int* ij, int mbndry, const int* counts,
                              int which, int hcount, float value) {
    const int num = counts ? counts[which] : hcount;
    const int t = gtid();
    if (t >= num * mbndry) return;
    const int n = t / mbndry, b = t % mbndry;
    const int i = ij[b + mbndry * (0 + 3 * n)] - 1;
    const int j = ij[b + mbndry * (1 + 3 * n)] - 1;
    const int k = ij[b + mbndry * (2 + 3 * n)] - 1;
    if ((unsigned)i >= (unsigned)B.nx || (unsigned)j >= (unsigned)B.ny ||
        (unsigned)k >= (unsigned)B.nz)
        return;   // a point outside the box is skipped, not written
    rst[(size_t)b * B.V() + B.IX(i, j, k)] = value;
}

void launch_set_resist(const Box& B, float* rst, int mbndry, float resist, float b0,
                       const int* ijzero, const int* ijmid, const int* ijsrf, const int* dcounts,
                       const int hcounts[3], int max_points, hipStream_t s) {
    const size_t n = (size_t)B.V() * mbndry;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rst, n, 0.0f);
    const unsigned g = (unsigned)(((size_t)max_points * mbndry + 255) / 256);
    if (g == 0) return;
    const float v0 = b0 / resist;
    const float v1 = 0.5f * b0 / resist;
    const float v2 = 0.125f * b0 / resist;
    const int h[3] = {hcounts ? hcounts[0] : 0, hcounts ? hcounts[1] : 0, hcounts ? hcounts[2] : 0};
    hipLaunchKernelGGL(k_resist_list, dim3(g), dim3(256), 0, s, B, rst, ijzero, mbndry, dcounts,
                       2, h[2], v0);
    hipLaunchKernelGGL(k_resist_list, dim3(g), dim3(256), 0, s, B, rst, ijmid, mbndry, dcounts, 1,
                       h[1], v1);
    hipLaunchKernelGGL(k_resist_list, dim3(g), dim3(256), 0, s, B, rst, ijsrf, mbndry, dcounts, 0,
                       h[0], v2);
}

// ---------------------------------------------------------------------------
// The initialisation block (code/3d_main.f90:1238-1633).  Three passes:
// the ambient plasma and dipole on every cell of every box, with each cell
// of a box <= mbndry classified as a zero / mid / surface boundary point;
// the ordered lists of those points with their ambient values; the solar
// wind and the speed check (each a function of the cell alone, so one pass
// gives the reference's three loops' result).

enum { kClsNone = 0, kClsSrf = 1, kClsMid = 2, kClsZero = 3 };

__device__ inline float powd(float x, float e) { return (float)pow((double)x, (double)e); }

__global__ void k_init_ambient(InitPlasma P) {
    int i, j, k;
    const Box& B = P.B;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const int box = blockIdx.y;   // 0-based
    const InitGeom& G = P.geom[box];
    const RotationGeom& R = P.rot;
    const float az = G.zmin + G.dz * (float)k - R.zdip;
    const float ay = G.ymin + G.dy * (float)j - R.ydip;
    const float ax = G.xmin + G.dx * (float)i - R.xdip;
    const float xr = ax * R.cos_rot + ay * R.sin_rot;
    const float yr = -(ax * R.sin_rot) + ay * R.cos_rot;
    const float xp = xr * R.cos_tilt - az * R.sin_tilt;
    const float yp = yr;
    const float zp = xr * R.sin_tilt + az * R.cos_tilt;
    const float ar = sqrtf(xp * xp + yp * yp + zp * zp);
    // the dipole subroutine subtracts the dipole centre again (:1279-1281)
    float abx, aby, abz;
    dipole_field(ax - R.xdip, ay - R.ydip, az - R.zdip, R.sin_rot, R.cos_rot, R.sin_tilt,
                 R.cos_tilt, R.b0, abx, aby, abz);
    if (ar < P.zero_bndry) abx = aby = abz = 0.0f;
    // corotation (:1293-1303)
    const float rx = ax * P.re_equiv;
    const float ry = ay * P.re_equiv;
    const float rd = sqrtf(rx * rx + ry * ry);
    float vfrac = 1.0f;
    if (!(rd < P.r_rot)) {
        const float t = (2.0f * (P.r_rot * P.r_rot)) / (P.r_rot * P.r_rot + rd * rd);
        vfrac = t * t;
    }
    const float rvy = rx * P.v_rot * vfrac;
    const float rvx = -(ry * P.v_rot * vfrac);
    // ionosphere (:1306-1316); the sphere terms of :1318-1332 feed nothing
    float ar_iono = sqrtf(xp * xp + ay * ay + 1.25f * (zp * zp));
    ar_iono = fmaxf(0.0001f, ar_iono);
    float ra = powd((ar_iono + 0.5f * P.r_inner) / (1.5f * P.r_inner), -P.alpha_e);
    const float h15 = 1.5f * P.r_inner, h30 = 3.0f * P.r_inner;
    const float zheight = fmaxf(1.0f, (zp * zp + h15 * h15) / (h30 * h30));
    ra = ra / (zheight * zheight);
    const float rho_iono = fmaxf(P.erho * ra, P.d_min);
    const float arho = fmaxf(rho_iono, P.d_min);
    const float vt = fmaxf(sqrtf(rvy * rvy + rvx * rvx), P.cs_inner);
    // species (:1334-1384)
    const float qrho = arho * P.rmassq / zheight;
    const float qpres = arho * vt;
    const float hrho = 0.05f * qrho * P.rmassh / P.rmassq;
    const float hpres = 0.05f * qpres;
    const float orho = qrho * P.rmasso / P.rmassq;
    const float opres = qpres;
    const float rho[3] = {qrho, hrho, orho};
    const float pres[3] = {qpres, hpres, opres};
    const size_t c = (size_t)box * B.V() + B.IX(i, j, k);
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        float* const* f = P.f + 10 * s;   // rho, presx, presy, presz, px, py, pz, presxy..yz
        f[0][c] = rho[s];
        f[1][c] = pres[s];
        f[2][c] = pres[s];
        f[3][c] = pres[s];
        f[4][c] = rho[s] * rvx;
        f[5][c] = rho[s] * rvy;
        f[6][c] = 0.0f;
        f[7][c] = 0.0f;
        f[8][c] = 0.0f;
        f[9][c] = 0.0f;
    }
    P.f[30][c] = (qpres + hpres + opres) / P.ti_te;
    P.f[31][c] = 0.0f;
    P.f[32][c] = 0.0f;
    P.f[33][c] = 0.0f;
    P.b0f[0][c] = abx;
    P.b0f[1][c] = aby;
    P.b0f[2][c] = abz;
    // boundary points of boxes <= mbndry (:1389-1452)
    if (box < P.mbndry) {
        unsigned char cls = kClsNone;
        if (ar <= P.srf_bndry) cls = (ar < P.zero_bndry) ? kClsZero
                                     : (ar < P.mid_bndry) ? kClsMid : kClsSrf;
        P.cls[c] = cls;
    }
}

// the ordered point lists of one box (block = box): the reference appends
// in loop order, k, j, i, so entry n of a list is the n-th cell of its
// class in memory order.  Each thread scans a contiguous run of cells;
// a block-wide scan of the per-thread counts gives the write offsets.
__global__ void __launch_bounds__(1024) k_init_lists(InitPlasma P) {
    const Box& B = P.B;
    const int box = blockIdx.x, T = blockDim.x, t = threadIdx.x;
    const int V = B.V();
    const int run = (V + T - 1) / T;
    const int c0 = min(V, t * run), c1 = min(V, c0 + run);
    const unsigned char* cls = P.cls + (size_t)box * V;
    int cnt[3] = {0, 0, 0};   // srf, mid, zero
    for (int c = c0; c < c1; ++c)
        if (cls[c] != kClsNone) ++cnt[cls[c] - 1];
    __shared__ int sc[3][1024];
    for (int q = 0; q < 3; ++q) sc[q][t] = cnt[q];
    __syncthreads();
    for (int o = 1; o < T; o <<= 1) {   // inclusive scan, Hillis-Steele
        int v[3];
        for (int q = 0; q < 3; ++q) v[q] = (t >= o) ? sc[q][t - o] : 0;
        __syncthreads();
        for (int q = 0; q < 3; ++q) sc[q][t] += v[q];
        __syncthreads();
    }
    int pos[3];
    for (int q = 0; q < 3; ++q) pos[q] = sc[q][t] - cnt[q];
    const int mb = P.mbndry;
    int* ijs[3] = {P.ijsrf, P.ijmid, P.ijzero};
    float* parms[3] = {P.parm_srf, P.parm_mid, P.parm_zero};
    const int cap[3] = {P.msrf, P.mmid, P.mzero};
    const size_t base = (size_t)box * V;
    for (int c = c0; c < c1; ++c) {
        const int q = cls[c] - 1;
        if (q < 0) continue;
        const int n = pos[q]++;
        if (n >= cap[q]) continue;
        int i, j, k;
        cell_of(B, c, i, j, k);
        int* ij = ijs[q];
        ij[box + mb * (0 + 3 * n)] = i + 1;
        ij[box + mb * (1 + 3 * n)] = j + 1;
        ij[box + mb * (2 + 3 * n)] = k + 1;
        float* pm = parms[q];
        const size_t cc = base + c;
        const float v[7] = {P.f[0][cc],  P.f[10][cc], P.f[20][cc], P.f[1][cc],
                            P.f[11][cc], P.f[21][cc], P.f[30][cc]};
        for (int a = 0; a < 7; ++a) pm[box + mb * (a + 7 * n)] = v[a];
    }
    if (t == T - 1) {   // counts cut at the capacity; a flag when a list overflowed
        for (int q = 0; q < 3; ++q) {
            P.counts[3 * box + q] = min(sc[q][t], cap[q]);
            if (sc[q][t] > cap[q]) P.counts[3 * P.mbndry] = 1;
        }
    }
}

// solar wind (:1464-1587) and the speed check (:1589-1626)
__global__ void k_init_wind(InitPlasma P) {
    int i, j, k;
    const Box& B = P.B;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const int box = blockIdx.y;
    const InitGeom& G = P.geom[box];
    const RotationGeom& R = P.rot;
    const float az = G.zmin + G.dz * (float)k - R.zdip;
    const float ay = G.ymin + G.dy * (float)j - R.ydip;
    const float ax = G.xmin + G.dx * (float)i - R.xdip;
    const float ar = sqrtf(ax * ax + ay * ay + az * az);
    const size_t c = (size_t)box * B.V() + B.IX(i, j, k);
    float* const* f = P.f;
    const float wb = P.wind_bnd, ofrac = P.rho_frac;
    const bool full = (ar >= 1.5f * wb) && (ax < 0.0f);
    const bool ramp = (ar > wb) && (ar < 1.5f * wb) && (ax < 0.0f);
    if (full || ramp) {
        const float d = ramp ? (ar - wb) / (0.5f * wb) : 1.0f;
        // the ramp multiplies each added term by dfrac as its last factor
        auto sc = [&](float v) { return ramp ? v * d : v; };
        f[0][c] = sc(P.srho) + f[0][c];
        f[4][c] = sc(P.spx) + f[4][c];
        f[5][c] = sc(P.spy) + f[5][c];
        f[6][c] = sc(P.spz) + f[6][c];
        const float qr = f[0][c];
        const float avx = f[4][c] / qr, avy = f[5][c] / qr, avz = f[6][c] / qr;
        const float apres = P.cs_wind * P.cs_wind * qr / P.gamma;
        const float qp = sc(0.5f * apres) + f[1][c];
        f[1][c] = qp;
        f[2][c] = qp;
        f[3][c] = qp;
        f[7][c] = f[8][c] = f[9][c] = 0.0f;
        f[30][c] = sc(0.5f * apres / P.ti_te) + f[30][c];
        const float hr = sc(P.srho * P.rho_frac * P.rmassh / P.rmassq) + f[10][c];
        f[10][c] = hr;
        f[14][c] = avx * hr;
        f[15][c] = avy * hr;
        f[16][c] = avz * hr;
        const float hp = sc(0.5f * P.spress * P.rho_frac) + f[11][c];
        f[11][c] = hp;
        f[12][c] = hp;
        f[13][c] = hp;
        f[17][c] = f[18][c] = f[19][c] = 0.0f;
        const float orr = sc(P.srho * ofrac * P.rmasso / P.rmassq) + f[20][c];
        f[20][c] = orr;
        f[24][c] = avx * orr;
        f[25][c] = avy * orr;
        f[26][c] = avz * orr;
        const float op = sc(0.5f * P.spress * ofrac) + f[21][c];
        f[21][c] = op;
        f[22][c] = op;
        f[23][c] = op;
        f[27][c] = f[28][c] = f[29][c] = 0.0f;
    }
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        float* const* g = f + 10 * s;
        const float r = g[0][c];
        const float vx = g[4][c] / r, vy = g[5][c] / r, vz = g[6][c] / r;
        const float spd = sqrtf(vx * vx + vy * vy + vz * vz);
        if (spd > 1.0f) g[4][c] = g[5][c] = g[6][c] = 0.0f;
    }
}

void launch_init_plasma(const InitPlasma& P, int n_grids, hipStream_t s) {
    const dim3 g((P.B.V() + 255) / 256, n_grids);
    hipLaunchKernelGGL(k_init_ambient, g, dim3(256), 0, s, P);
    if (P.mbndry > 0)
        hipLaunchKernelGGL(k_init_lists, dim3(P.mbndry), dim3(1024), 0, s, P);
    hipLaunchKernelGGL(k_init_wind, g, dim3(256), 0, s, P);
}

}  // namespace g3m3

namespace g3m3 {

// ---------------------------------------------------------------------------
// Torus source (code/3d_main.f90:3149-3253): plasma added to box 1 in a
// torus of radius torus_inj_dist around the planet, outside srf_bndry,
// with Gaussian radial and vertical profiles and a field-direction cut.
// The exp profiles are evaluated in double and rounded once (a C library's
// expf may differ by an ulp).  Each cell's injected q, h, o mass times the
// cell volume goes to contrib[3][V] (zero elsewhere); the totals are summed
// afterwards in the reference's loop order.

__device__ inline float expd(float x) { return (float)exp((double)x); }

__global__ void k_torus(TorusParams T) {
    int i, j, k;
    const Box& B = T.B;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const size_t c = B.IX(i, j, k);   // box 1
    const float az = T.zmin + T.dz * (float)k - T.zdip;
    const float ay = T.ymin + T.dy * (float)j - T.ydip;
    const float ax = T.xmin + T.dx * (float)i - T.xdip;
    const float rx = ax * T.re_equiv;
    const float ry = ay * T.re_equiv;
    const float ar = sqrtf((ax + T.xdip) * (ax + T.xdip) + (ay + T.ydip) * (ay + T.ydip) +
                           (az + T.zdip) * (az + T.zdip));
    const float rvy = rx * T.v_rot;
    const float rvx = -(ry * T.v_rot);
    const float corotate = sqrtf(rvx * rvx + rvy * rvy);
    const float ar_tmid = sqrtf(ax * ax + ay * ay) * T.re_equiv;
    const float dr_tmid = fabsf(ar_tmid - T.torus_inj_dist);
    float cq = 0.0f, ch = 0.0f, co = 0.0f;
    if (fabsf(dr_tmid) < 2.0f * T.torus_rad && ar > T.srf_bndry) {
        const float a1 = dr_tmid / (0.5f * T.torus_rad);
        const float rscale = expd(-(a1 * a1));
        const float a2 = (az * T.re_equiv) / (0.5f * T.torus_rad);
        const float zscale = expd(-(a2 * a2));
        const float bsx = T.bx0[c] + T.bx[c], bsy = T.by0[c] + T.by[c], bsz = T.bz0[c] + T.bz[c];
        const float abtot = (bsx * bsx + bsy * bsy) / (bsz * bsz);
        const float dscale = fmaxf(0.0f, 1.0f - 10.0f * abtot);
        const float c2 = corotate * corotate;
        // h, q, o in the reference's order; species 0 q, 1 h, 2 o
        const int order[3] = {1, 0, 2};
        const float den0[3] = {T.denq, T.denh, T.deno};
        const float rm[3] = {T.rmassq, T.rmassh, T.rmasso};
        float del[3] = {0.0f, 0.0f, 0.0f}, dens[3] = {0.0f, 0.0f, 0.0f};
        for (int oi = 0; oi < 3; ++oi) {
            const int s = order[oi];
            float* const* f = T.sp.f[s];   // rho, presx, presy, presz, px, py, pz, presxy, ..
            const float den = den0[s] * rm[s] * rscale * zscale * dscale;
            f[0][c] = f[0][c] + den;
            // h: hden*corotate^2*t_torus; q and o divide by their mass first
            const float dp = (s == 1) ? den * c2 * T.t_torus : (den / rm[s]) * c2 * T.t_torus;
            f[1][c] = f[1][c] + dp;
            f[2][c] = f[2][c] + dp;
            f[3][c] = f[3][c] + dp * T.aniso_factor;
            f[7][c] = f[7][c] + dp;
            f[4][c] = f[4][c] + T.reduct * den * rvx;
            f[5][c] = f[5][c] + T.reduct * den * rvy;
            del[s] = dp;
            dens[s] = den;
        }
        T.epres[c] = T.epres[c] + (del[2] + del[1] + del[0]);
        cq = dens[0] * T.dx * T.dy * T.dz;
        ch = dens[1] * T.dx * T.dy * T.dz;
        co = dens[2] * T.dx * T.dy * T.dz;
    }
    const size_t V = (size_t)B.V();
    T.contrib[c] = cq;
    T.contrib[V + c] = ch;
    T.contrib[2 * V + c] = co;
}

// the injection totals, one thread per species, summed cell by cell in the
// reference's loop order (k, j, i = memory order), in single precision
__global__ void k_torus_totals(const float* contrib, int V, float* tot) {
    const int s = threadIdx.x;
    if (s >= 3) return;
    const float* p = contrib + (size_t)s * V;
    float acc = 0.0f;
    for (int c = 0; c < V; ++c) acc = acc + p[c];
    tot[s] = acc;
}

void launch_torus(const TorusParams& T, float* tot, hipStream_t s) {
    hipLaunchKernelGGL(k_torus, dim3((T.B.V() + 255) / 256), dim3(256), 0, s, T);
    hipLaunchKernelGGL(k_torus_totals, dim3(1), dim3(64), 0, s, T.contrib, T.B.V(), tot);
}

}  // namespace g3m3
