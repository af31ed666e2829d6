// Fused box-step kernels (the native runtime's path, g3m3_run_*).
//
// Same arithmetic as the per-subroutine kernels -- every value is produced
// by the reference's expression in the reference's order -- but each
// kernel covers a whole stage of the box-step, so a field is read from HBM
// about once per stage instead of once per reference loop nest:
//   k_efield    calcur + total B + fnd_evel + BandE (code/sub_calcur.f90,
//               code/3d_main.f90:2248-2254, sub_fnd_evel.f90, sub_BandE.f90)
//   k_push      push_ion x3 species, push_bfld, push_elec (blockIdx.y picks)
//   k_validate  set_rho + set_speed_agrd in one read of the fields
//   k_lap       Lapidus for the 3 species, electrons and B (blockIdx.y)
//   k_fct_*     flux-corrected transport, two passes for all 34 fields
//   k_flanks_shell  coarse -> fine boundary on the 2-layer shell only
#include "g3m3_dev.h"
#include "g3m3_fused.h"

namespace g3m3 {

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---------------------------------------------------------------------------
// E field.  cur at a boundary cell is calcur's face/edge/corner rule; those
// rules resolve to the curl at the clamped index, except the four corners
// at i = nx that average three cells (code/sub_calcur.f90:268-289).  E on
// the boundary is BandE's flank copy: E at the clamped index.

struct CurlCtx {
    Box B;
    const float* __restrict__ bx;
    const float* __restrict__ by;
    const float* __restrict__ bz;
    float dxt, dyt, dzt;
    __device__ void curl(int i, int j, int k, float& cx, float& cy, float& cz) const {
        const int c = B.IX(i, j, k), sx = 1, sy = B.nx, sz = B.nx * B.ny;
        cx = (bz[c + sy] - bz[c - sy]) / dyt - (by[c + sz] - by[c - sz]) / dzt;
        cy = (bx[c + sz] - bx[c - sz]) / dzt - (bz[c + sx] - bz[c - sx]) / dxt;
        cz = (by[c + sx] - by[c - sx]) / dxt - (bx[c + sy] - bx[c - sy]) / dyt;
    }
    __device__ void cur(int i, int j, int k, float& cx, float& cy, float& cz) const {
        const int ci = clampi(i, 1, B.nx - 2), cj = clampi(j, 1, B.ny - 2),
                  ck = clampi(k, 1, B.nz - 2);
        curl(ci, cj, ck, cx, cy, cz);
        if (i == B.nx - 1 && (j == 0 || j == B.ny - 1) && (k == 0 || k == B.nz - 1)) {
            cx = (cx + cx + cx) / 3.0f;
            cy = (cy + cy + cy) / 3.0f;
            cz = (cz + cz + cz) / 3.0f;
        }
    }
};

// E at an interior cell (code/sub_BandE.f90:37-116 + uniform eta)
__device__ inline void efield_interior(const EParams& P, const CurlCtx& C, int i, int j, int k,
                                       float& Ex, float& Ey, float& Ez, float* evo, float* tvo,
                                       float* bso) {
    const Box& B = P.B;
    const int c = B.IX(i, j, k);
    float cx, cy, cz;
    C.curl(i, j, k, cx, cy, cz);
    const float q = P.rho[0][c], h = P.rho[1][c], o = P.rho[2][c];
    const float qden = (q + 0.000001f) / P.rmq;
    const float hden = (h + 0.000001f) / P.rmh;
    const float oden = (o + 0.000001f) / P.rmo;
    const float tot = qden + hden + oden;
    const float tden = q + h + o;
    const float tx = (P.mom[0][c] + P.mom[3][c] + P.mom[6][c]) / tden;
    const float ty = (P.mom[1][c] + P.mom[4][c] + P.mom[7][c]) / tden;
    const float tz = (P.mom[2][c] + P.mom[5][c] + P.mom[8][c]) / tden;
    const float evx = tx - cx / tot / P.reynolds;
    const float evy = ty - cy / tot / P.reynolds;
    const float evz = tz - cz / tot / P.reynolds;
    const float abx = P.b0[0][c] + P.b[0][c], aby = P.b0[1][c] + P.b[1][c],
                abz = P.b0[2][c] + P.b[2][c];
    Ex = -(evy * abz - evz * aby);
    Ey = -(evz * abx - evx * abz);
    Ez = -(evx * aby - evy * abx);
    const int sy = B.nx, sz = B.nx * B.ny;
    const float* ep = P.ep;
    const float arho = (q / P.rmq + h / P.rmh + o / P.rmo) * P.reynolds;
    Ex = Ex - ((ep[c + 1] - ep[c - 1]) / P.dxt) / arho;
    Ey = Ey - ((ep[c + sy] - ep[c - sy]) / P.dyt) / arho;
    Ez = Ez - ((ep[c + sz] - ep[c - sz]) / P.dzt) / arho;
    if (P.rst) {   // list points of box mbndry: rst is zero elsewhere
        const float r = P.rst[c];
        if (r != 0.0f) {
            const float eden = q / P.rmq + h / P.rmh + o / P.rmo;
            Ex = Ex + r * cx / eden;
            Ey = Ey + r * cy / eden;
            Ez = Ez + r * cz / eden;
        }
    }
    if (P.eta != 0.0f) {
        const float eden = q / P.rmq + h / P.rmh + o / P.rmo;
        Ex = Ex + P.eta * cx / eden;
        Ey = Ey + P.eta * cy / eden;
        Ez = Ez + P.eta * cz / eden;
    }
    if (evo) {
        evo[0] = evx;
        evo[1] = evy;
        evo[2] = evz;
        tvo[0] = tx;
        tvo[1] = ty;
        tvo[2] = tz;
        bso[0] = abx;
        bso[1] = aby;
        bso[2] = abz;
    }
}

__global__ void __launch_bounds__(256) k_efield(EParams P) {
    int i, j, k;
    const Box& B = P.B;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const int c = B.IX(i, j, k);
    CurlCtx C{B, P.b[0], P.b[1], P.b[2], P.dxt, P.dyt, P.dzt};
    float Ex, Ey, Ez;
    if (B.interior(i, j, k)) {
        float ev[3], tv[3], bs[3];
        efield_interior(P, C, i, j, k, Ex, Ey, Ez, ev, tv, bs);
        P.ev[0][c] = ev[0];
        P.ev[1][c] = ev[1];
        P.ev[2][c] = ev[2];
        P.tv[0][c] = tv[0];
        P.tv[1][c] = tv[1];
        P.tv[2][c] = tv[2];
        P.bs[0][c] = bs[0];
        P.bs[1][c] = bs[1];
        P.bs[2][c] = bs[2];
    } else {
        // boundary: E copied from the clamped interior cell
        efield_interior(P, C, clampi(i, 1, B.nx - 2), clampi(j, 1, B.ny - 2),
                        clampi(k, 1, B.nz - 2), Ex, Ey, Ez, nullptr, nullptr, nullptr);
        // ev and bs on the boundary (push_elec reads ev there; push_ion the smoothed bs)
        float cx, cy, cz;
        C.cur(i, j, k, cx, cy, cz);
        const float q = P.rho[0][c], h = P.rho[1][c], o = P.rho[2][c];
        const float tot = (q + 0.000001f) / P.rmq + (h + 0.000001f) / P.rmh + (o + 0.000001f) / P.rmo;
        const float tden = q + h + o;
        const float tx = (P.mom[0][c] + P.mom[3][c] + P.mom[6][c]) / tden;
        const float ty = (P.mom[1][c] + P.mom[4][c] + P.mom[7][c]) / tden;
        const float tz = (P.mom[2][c] + P.mom[5][c] + P.mom[8][c]) / tden;
        P.ev[0][c] = tx - cx / tot / P.reynolds;
        P.ev[1][c] = ty - cy / tot / P.reynolds;
        P.ev[2][c] = tz - cz / tot / P.reynolds;
        P.tv[0][c] = tx;
        P.tv[1][c] = ty;
        P.tv[2][c] = tz;
        P.bs[0][c] = P.b0[0][c] + P.b[0][c];
        P.bs[1][c] = P.b0[1][c] + P.b[1][c];
        P.bs[2][c] = P.b0[2][c] + P.b[2][c];
    }
    P.e[0][c] = Ex;
    P.e[1][c] = Ey;
    P.e[2][c] = Ez;
}

// ---------------------------------------------------------------------------
// Push.  blockIdx.y = 0,1,2: ion species q,h,o; 3: push_bfld (+ push_elec
// when P.do_elec).  The species velocity wrk p / max(wrk rho, 1e-6) is
// evaluated where it is needed instead of stored (same bits).

__device__ inline void push_species(const PushParams& P, int s, int i, int j, int k) {
    const Box& B = P.B;
    const int c = B.IX(i, j, k), X = 1, Y = B.nx, Z = B.nx * B.ny;
    const SpeciesSet& W = P.src[s];
    const SpeciesSet& O = P.old[s];
    const SpeciesSet& U = P.out[s];
    const float rmass = P.rmass[s];
    const float delt = P.delt, reynolds = P.reynolds;
    const float dxt = P.dxt, dyt = P.dyt, dzt = P.dzt;
    const float bmax = 0.20f * rmass / delt / reynolds;
    const float* __restrict__ wr = W.f[0];
    const float* __restrict__ wx = W.f[1];
    const float* __restrict__ wy = W.f[2];
    const float* __restrict__ wz = W.f[3];
    const float* __restrict__ wpx = W.f[4];
    const float* __restrict__ wpy = W.f[5];
    const float* __restrict__ wpz = W.f[6];
    const float* __restrict__ wxy = W.f[7];
    const float* __restrict__ wxz = W.f[8];
    const float* __restrict__ wyz = W.f[9];
    // velocities at the centre and the six neighbours (code/sub_push_ion.f90:67-81)
    float vx[7], vy[7], vz[7];
    const int off[7] = {0, X, -X, Y, -Y, Z, -Z};
#pragma unroll
    for (int m = 0; m < 7; ++m) {
        const int p = c + off[m];
        const float a = fmaxf(wr[p], 1.e-6f);
        vx[m] = wpx[p] / a;
        vy[m] = wpy[p] / a;
        vz[m] = wpz[p] / a;
    }
    // index into vx etc: 0 c, 1 +x, 2 -x, 3 +y, 4 -y, 5 +z, 6 -z
    auto flux = [&](const float* __restrict__ w) {
        return ((w[c + X] * vx[1] - w[c - X] * vx[2]) / dxt) +
               ((w[c + Y] * vy[3] - w[c - Y] * vy[4]) / dyt) +
               ((w[c + Z] * vz[5] - w[c - Z] * vz[6]) / dzt);
    };
    U.f[0][c] = O.f[0][c] - delt * (((wpx[c + X] - wpx[c - X]) / dxt) +
                                    ((wpy[c + Y] - wpy[c - Y]) / dyt) +
                                    ((wpz[c + Z] - wpz[c - Z]) / dzt));
    const float aqpx = O.f[4][c] - delt * flux(wpx);
    const float aqpy = O.f[5][c] - delt * flux(wpy);
    const float aqpz = O.f[6][c] - delt * flux(wpz);
    const float* __restrict__ bsx = P.bs[0];
    const float* __restrict__ bsy = P.bs[1];
    const float* __restrict__ bsz = P.bs[2];
    float abx = bsx[c], aby = bsy[c], abz = bsz[c];
    const float bmag0 = fmaxf(sqrtf(abx * abx + aby * aby + abz * abz), P.smallbit);
    const float weight = fmaxf(1.0f, bmag0 / bmax);
    const float skin_depth = reynolds / weight;
    const float tx = P.tv[0][c], ty = P.tv[1][c], tz = P.tv[2][c];
    const float tcx = (ty * abz - tz * aby);
    const float tcy = -(tx * abz - tz * abx);
    const float tcz = (tx * aby - ty * abx);
    const float qcx = (vy[0] * abz - vz[0] * aby);
    const float qcy = -(vx[0] * abz - vz[0] * abx);
    const float qcz = (vx[0] * aby - vy[0] * abx);
    const float gpx = (wx[c + X] - wx[c - X]) / dxt + (wxy[c + Y] - wxy[c - Y]) / dxt +
                      (wxz[c + Z] - wxz[c - Z]) / dxt;
    const float gpy = (wxy[c + X] - wxy[c - X]) / dxt + (wy[c + Y] - wy[c - Y]) / dyt +
                      (wyz[c + Z] - wyz[c - Z]) / dxt;
    const float gpz = (wxz[c + X] - wxz[c - X]) / dxt + (wyz[c + Y] - wyz[c - Y]) / dyt +
                      (wz[c + Z] - wz[c - Z]) / dzt;
    const float dex = (P.e[0][c] + tcx) * reynolds;
    const float dey = (P.e[1][c] + tcy) * reynolds;
    const float dez = (P.e[2][c] + tcz) * reynolds;
    const float dvx = (qcx - tcx) * skin_depth;
    const float dvy = (qcy - tcy) * skin_depth;
    const float dvz = (qcz - tcz) * skin_depth;
    const float wrc = wr[c];
    const float qden = wrc / rmass;
    const float px = aqpx + delt * (qden * (dex + dvx) - gpx);
    const float py = aqpy + delt * (qden * (dey + dvy) - gpy);
    const float pz = aqpz + delt * (qden * (dez + dvz) - gpz);
    const float ax = P.xmin + P.ddx * (float)i;
    const float ay = P.ymin + P.ddy * (float)j;
    const float az = P.zmin + P.ddz * (float)k;
    const float radius = sqrtf(ax * ax + ay * ay + az * az) + 0.0000001f;
    const float tr = P.re_equiv * radius;
    const float g = delt * wrc * P.grav / (tr * tr);
    U.f[4][c] = px - g * ax / radius;
    U.f[5][c] = py - g * ay / radius;
    U.f[6][c] = pz - g * az / radius;

    const float cx = wx[c], cy = wy[c], cz = wz[c];
    const float cxy = wxy[c], cxz = wxz[c], cyz = wyz[c];
    const float px0 = O.f[1][c] - delt * flux(wx) - 2.0f * delt * cx * (vx[1] - vx[2]) / dxt;
    const float py0 = O.f[2][c] - delt * flux(wy) - 2.0f * delt * cy * (vy[3] - vy[4]) / dyt;
    const float pz0 = O.f[3][c] - delt * flux(wz) - 2.0f * delt * cz * (vz[5] - vz[6]) / dzt;
    abx = (6.0f * bsx[c] + bsx[c + X] + bsx[c - X] + bsx[c + Y] + bsx[c - Y] + bsx[c + Z] +
           bsx[c - Z]) / 12.0f;
    aby = (6.0f * bsy[c] + bsy[c + X] + bsy[c - X] + bsy[c + Y] + bsy[c - Y] + bsy[c + Z] +
           bsy[c - Z]) / 12.0f;
    abz = (6.0f * bsz[c] + bsz[c + X] + bsz[c - X] + bsz[c + Y] + bsz[c - Y] + bsz[c + Z] +
           bsz[c - Z]) / 12.0f;
    const float bmag = sqrtf(abx * abx + aby * aby + abz * abz);
    float sf = reynolds;
    if (bmag > bmax) sf = reynolds * (bmax / bmag);
    const float t2 = 2.0f * delt;
    U.f[1][c] = px0 - t2 * cxy * (vx[3] - vx[4]) / dyt - t2 * cxz * (vx[5] - vx[6]) / dzt +
                t2 * sf / rmass * (cxy * abz - cxz * aby);
    U.f[2][c] = py0 - t2 * cxy * (vy[1] - vy[2]) / dxt - t2 * cyz * (vy[5] - vy[6]) / dzt +
                t2 * sf / rmass * (cyz * abx - cxy * abz);
    U.f[3][c] = pz0 - t2 * cxz * (vz[1] - vz[2]) / dxt - t2 * cyz * (vz[3] - vz[4]) / dyt +
                t2 * sf / rmass * (cxz * aby - cyz * abx);
    const float pxy = O.f[7][c] - delt * flux(wxy);
    U.f[7][c] = pxy - delt * cxy * ((vx[1] - vx[2]) / dxt + (vy[3] - vy[4]) / dyt) -
                delt * cy * (vx[3] - vx[4]) / dyt - delt * cx * (vy[1] - vy[2]) / dxt -
                delt * cyz * (vx[5] - vx[6]) / dzt - delt * cxz * (vy[5] - vy[6]) / dzt +
                delt * sf / rmass * (abz * (cy - cx) - aby * cyz + abx * cxz);
    const float pxz = O.f[8][c] - delt * flux(wxz);
    U.f[8][c] = pxz - delt * cxz * ((vx[1] - vx[2]) / dxt + (vz[5] - vz[6]) / dzt) -
                delt * cz * (vx[5] - vx[6]) / dzt - delt * cx * (vz[1] - vz[2]) / dxt -
                delt * cyz * (vx[3] - vx[4]) / dyt - delt * cxy * (vz[3] - vz[4]) / dyt +
                delt * sf / rmass * (aby * (cx - cz) - abx * cxy + abz * cyz);
    const float pyz = O.f[9][c] - delt * flux(wyz);
    U.f[9][c] = pyz - delt * cyz * ((vy[3] - vy[4]) / dyt + (vz[5] - vz[6]) / dzt) -
                delt * cy * (vz[3] - vz[4]) / dyt - delt * cz * (vy[5] - vy[6]) / dzt -
                delt * cxy * (vz[1] - vz[2]) / dxt - delt * cxz * (vy[1] - vy[2]) / dxt +
                delt * sf / rmass * (abx * (cz - cy) - abz * cxz + aby * cxy);
}

__device__ inline void push_field_elec(const PushParams& P, int i, int j, int k) {
    const Box& B = P.B;
    const int c = B.IX(i, j, k), sy = B.nx, sz = B.nx * B.ny;
    const float delt = P.delt, dxt = P.fdxt, dyt = P.fdyt, dzt = P.fdzt;
    const float* __restrict__ ex = P.e[0];
    const float* __restrict__ ey = P.e[1];
    const float* __restrict__ ez = P.e[2];
    // code/sub_push_bfld.f90:87-97
    P.bout[0][c] = P.bold[0][c] - delt * (((ez[c + sy] - ez[c - sy]) / dyt) -
                                          ((ey[c + sz] - ey[c - sz]) / dzt));
    P.bout[1][c] = P.bold[1][c] + delt * (((ez[c + 1] - ez[c - 1]) / dxt) -
                                          ((ex[c + sz] - ex[c - sz]) / dzt));
    P.bout[2][c] = P.bold[2][c] - delt * (((ey[c + 1] - ey[c - 1]) / dxt) -
                                          ((ex[c + sy] - ex[c - sy]) / dyt));
    if (P.do_elec) {   // code/sub_push_elec.f90:31-47
        const float* __restrict__ w = P.wrkepres;
        const float* __restrict__ evx = P.ev[0];
        const float* __restrict__ evy = P.ev[1];
        const float* __restrict__ evz = P.ev[2];
        const float gx = (w[c + 1] - w[c - 1]) / dxt;
        const float gy = (w[c + sy] - w[c - sy]) / dyt;
        const float gz = (w[c + sz] - w[c - sz]) / dzt;
        const float div = ((w[c + 1] * evx[c + 1] - w[c - 1] * evx[c - 1]) / dxt) +
                          ((w[c + sy] * evy[c + sy] - w[c - sy] * evy[c - sy]) / dyt) +
                          ((w[c + sz] * evz[c + sz] - w[c - sz] * evz[c - sz]) / dzt);
        const float e = P.oldepres[c] - delt * P.gamma * div;
        P.epres[c] = e + delt * P.gamma1 * (evx[c] * gx + evy[c] * gy + evz[c] * gz);
    }
}

__global__ void __launch_bounds__(256) k_push(PushParams P) {
    int i, j, k;
    if (!cell_of(P.B, gtid(), i, j, k) || !P.B.interior(i, j, k)) return;
    const int s = blockIdx.y;
    if (s < 3) push_species(P, s, i, j, k);
    else push_field_elec(P, i, j, k);
}

// ---------------------------------------------------------------------------
// Validation: set_rho (all cells) and set_speed_agrd (interior), one read.
// Nothing but the electron cap is written in place: a cell whose species
// values change (floor reset or neighbour average) appends one entry with
// its final values, applied by k_validate_apply after every thread has read
// the unmodified snapshot.

struct RhoFix {
    float r, x, y, z;
    bool hit;
};

__device__ inline RhoFix set_rho_cell(float R, float X, float Y, float Z, float dmin) {
    RhoFix f{R, X, Y, Z, false};
    if (R < dmin || X < 0.0f || Y < 0.0f || Z < 0.0f) {
        f.r = fmaxf(fabsf(R), 1.05f * dmin);
        const float ap = (fabsf(X) + fabsf(Y) + fabsf(Z)) / 3.0f;
        f.x = f.y = f.z = ap;
        f.hit = true;
    }
    return f;
}

__device__ inline float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__global__ void __launch_bounds__(256) k_validate(ValParams Q) {
    const Box& B = Q.B;
    int i, j, k;
    float mx = 0.0f, my = 0.0f, mz = 0.0f, mc = 0.0f, ma = 0.0f;
    const float dmins[3] = {2.e-4f, 1.e-4f, 2.e-5f};
    if (cell_of(B, gtid(), i, j, k)) {
        const int c = B.IX(i, j, k);
        const bool inner = B.interior(i, j, k);
        float rho_final[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            float* const* f = Q.sp[s].f;
            const RhoFix cf = set_rho_cell(f[0][c], f[1][c], f[2][c], f[3][c], dmins[s]);
            rho_final[s] = cf.r;
            bool avg = false;
            if (inner) {
                const float arho = cf.r + Q.smallbit;
                const float avx = fabsf(f[4][c] / arho);
                const float avy = fabsf(f[5][c] / arho);
                const float avz = fabsf(f[6][c] / arho);
                const float v = sqrtf(avx * avx + avy * avy + avz * avz);
                const float cs = sqrtf(0.33f * (cf.x + cf.y + cf.z) / arho);
                mx = fmaxf(mx, avx);
                my = fmaxf(my, avy);
                mz = fmaxf(mz, avz);
                mc = fmaxf(mc, cs);
                avg = (v > Q.vlim * Q.svlim[s]) || (cs > Q.vlim * Q.svlim[s] * Q.sclim[s]);
            }
            if (avg) {
                ValEntry e;
                e.cell = c;
                e.species = s;
                const int nb[6] = {c + 1, c - 1, c + B.nx, c - B.nx, c + B.nx * B.ny,
                                   c - B.nx * B.ny};
                float sr[6], sx[6], sy[6], sz[6], sxy[6], sxz[6], syz[6], spx[6], spy[6], spz[6];
                const float one_m = 1.0f - Q.aniso;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const int p = nb[q];
                    const RhoFix nf = set_rho_cell(f[0][p], f[1][p], f[2][p], f[3][p], dmins[s]);
                    sr[q] = nf.r;
                    spx[q] = f[4][p];
                    spy[q] = f[5][p];
                    spz[q] = f[6][p];
                    if (Q.isotropic) {
                        sx[q] = sy[q] = sz[q] = nf.x;
                        sxy[q] = sxz[q] = syz[q] = 0.0f;
                    } else {
                        const float ap = (nf.x + nf.y + nf.z) / 3.0f;
                        sx[q] = Q.aniso * nf.x + one_m * ap;
                        sy[q] = Q.aniso * nf.y + one_m * ap;
                        sz[q] = Q.aniso * nf.z + one_m * ap;
                        sxy[q] = Q.aniso * (nf.hit ? 0.0f : f[7][p]);
                        sxz[q] = Q.aniso * (nf.hit ? 0.0f : f[8][p]);
                        syz[q] = Q.aniso * (nf.hit ? 0.0f : f[9][p]);
                    }
                }
                auto a6 = [](const float* t) { return (t[0] + t[1] + t[2] + t[3] + t[4] + t[5]) / 6.0f; };
                auto a6i = [&](const float* t) {
                    return Q.iso_flag * (t[0] + t[1] + t[2] + t[3] + t[4] + t[5]) / 6.0f;
                };
                e.v[0] = a6(sr);
                e.v[1] = a6(sx);
                e.v[2] = a6(sy);
                e.v[3] = a6(sz);
                e.v[4] = a6(spx);
                e.v[5] = a6(spy);
                e.v[6] = a6(spz);
                e.v[7] = a6i(sxy);
                e.v[8] = a6i(sxz);
                e.v[9] = a6i(syz);
                rho_final[s] = e.v[0];
                Q.list[atomicAdd(Q.count, 1)] = e;
            } else if (cf.hit) {
                ValEntry e;
                e.cell = c;
                e.species = s;
                e.v[0] = cf.r;
                e.v[1] = cf.x;
                e.v[2] = cf.y;
                e.v[3] = cf.z;
                e.v[4] = f[4][c];
                e.v[5] = f[5][c];
                e.v[6] = f[6][c];
                e.v[7] = e.v[8] = e.v[9] = 0.0f;
                Q.list[atomicAdd(Q.count, 1)] = e;
            }
        }
        // electron pressure: |pe| everywhere, sound-speed cap and Alfven
        // maximum on the interior with the final densities (:404-426)
        const float e0 = Q.ep[c];
        float e1 = fabsf(e0);
        if (inner) {
            const float arho = rho_final[0] + rho_final[1] + rho_final[2] + Q.smallbit;
            const float cs = sqrtf(e1 / arho);
            if (cs > Q.cslim) e1 = e1 * Q.cslim / cs;
            if (Q.partials) {
                const float sx = Q.b0[0][c] + Q.b[0][c], sy = Q.b0[1][c] + Q.b[1][c],
                            sz = Q.b0[2][c] + Q.b[2][c];
                const float bt = fmaxf(sqrtf(sx * sx + sy * sy + sz * sz), Q.smallbit);
                ma = bt / sqrtf(arho);
            }
        }
        if (__float_as_uint(e1) != __float_as_uint(e0)) Q.ep[c] = e1;
    }
    if (Q.partials) {
        // per-block maxima, no same-address atomics
        __shared__ float red[5][4];
        const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
        const float r0 = wave_max(mx), r1 = wave_max(my), r2 = wave_max(mz), r3 = wave_max(mc),
                    r4 = wave_max(ma);
        if (l == 0) {
            red[0][w] = r0;
            red[1][w] = r1;
            red[2][w] = r2;
            red[3][w] = r3;
            red[4][w] = r4;
        }
        __syncthreads();
        if (threadIdx.x < 5) {
            const int q = threadIdx.x;
            const float m = fmaxf(fmaxf(red[q][0], red[q][1]), fmaxf(red[q][2], red[q][3]));
            Q.partials[blockIdx.x * 5 + q] = m;
        }
    }
}

__global__ void k_validate_apply(ValParams Q) {
    const int n = *Q.count;
    for (int e = gtid(); e < n; e += gridDim.x * blockDim.x) {
        const ValEntry& a = Q.list[e];
        float* const* f = Q.sp[a.species].f;
#pragma unroll
        for (int m = 0; m < 10; ++m) f[m][a.cell] = a.v[m];
    }
}

// final maxima of one validation pass -> maxima[0..4]; zero the list counter
__global__ void k_validate_reduce(const float* __restrict__ partials, int nblocks, float* maxima,
                                  int* count) {
    __shared__ float red[5][4];
    float m[5] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < 5; ++q) m[q] = fmaxf(m[q], partials[b * 5 + q]);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const float r = wave_max(m[q]);
        if (l == 0) red[q][w] = r;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        const int q = threadIdx.x;
        maxima[q] = fmaxf(fmaxf(red[q][0], red[q][1]), fmaxf(red[q][2], red[q][3]));
    }
}

__global__ void k_zero_count(int* count) {
    if (threadIdx.x == 0) *count = 0;
}

// ---------------------------------------------------------------------------
// Lapidus.  blockIdx.y = 0,1,2: a species' 10 fields with its own velocity
// p/max(rho, smallbit); 3: epres and b with the total ion velocity.
__global__ void __launch_bounds__(256) k_lap(LapParams L) {
    int i, j, k;
    const Box& B = L.B;
    if (!cell_of(B, gtid(), i, j, k) || !B.interior(i, j, k)) return;
    const int c = B.IX(i, j, k), X = 1, Y = B.nx, Z = B.nx * B.ny;
    const int g = blockIdx.y;
    const int off[7] = {0, X, -X, Y, -Y, Z, -Z};
    float vx[7], vy[7], vz[7];
    if (g < 3) {
        float* const* f = L.src[g].f;
#pragma unroll
        for (int m = 0; m < 7; ++m) {
            const int p = c + off[m];
            const float a = fmaxf(f[0][p], L.smallbit);
            vx[m] = f[4][p] / a;
            vy[m] = f[5][p] / a;
            vz[m] = f[6][p] / a;
        }
    } else {
#pragma unroll
        for (int m = 0; m < 7; ++m) {
            const int p = c + off[m];
            const float tden = L.src[0].f[0][p] + L.src[1].f[0][p] + L.src[2].f[0][p];
            vx[m] = (L.src[0].f[4][p] + L.src[1].f[4][p] + L.src[2].f[4][p]) / tden;
            vy[m] = (L.src[0].f[5][p] + L.src[1].f[5][p] + L.src[2].f[5][p]) / tden;
            vz[m] = (L.src[0].f[6][p] + L.src[1].f[6][p] + L.src[2].f[6][p]) / tden;
        }
    }
    const float uxp = fabsf(vx[1] - vx[0]);
    const float uxm = fabsf(vx[0] - vx[2]);
    const float uyp = fabsf(vy[3] - vy[0]);
    const float uym = fabsf(vy[0] - vy[4]);
    const float uzp = fabsf(vz[5] - vz[0]);
    const float uzm = fabsf(vz[0] - vz[6]);
    const float delt = L.delt;
    auto lap = [&](const float* __restrict__ q, float* w, float chi) {
        const float q0 = q[c];
        w[c] = q0 + chi * (delt * (uxp * (q[c + X] - q0) - uxm * (q0 - q[c - X])) +
                           delt * (uyp * (q[c + Y] - q0) - uym * (q0 - q[c - Y])) +
                           delt * (uzp * (q[c + Z] - q0) - uzm * (q0 - q[c - Z])));
    };
    if (g < 3) {
        lap(L.src[g].f[0], L.dst[g].f[0], L.chirho);
        // the reference's field order within lap_plasma does not matter
        // (each output depends only on its own input)
#pragma unroll
        for (int m = 1; m < 10; ++m) lap(L.src[g].f[m], L.dst[g].f[m], L.chipxyz);
    } else {
        lap(L.ep_src, L.ep_dst, L.chipxyz);
#pragma unroll
        for (int m = 0; m < 3; ++m) lap(L.b_src[m], L.b_dst[m], L.chierg);
    }
}

// ---------------------------------------------------------------------------
// FCT, two passes over all 34 fields (blockIdx.y = field):
//   k_fct_diffuse2: P = w + chi*lap7(old) (interior), w (faces) -> scratch
//   k_fct_limit:    the three limited antidiffusive flux divergences of
//                   code/sub_fcsmooth.f90:279-431 evaluated at each cell
//                   from P and w, written to the output field.

__global__ void __launch_bounds__(256) k_fct_diffuse2(FctParams F) {
    int i, j, k;
    const Box& B = F.B;
    const int f = blockIdx.y;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const int c = B.IX(i, j, k);
    const float* __restrict__ w = F.wrk[f];
    const float* __restrict__ o = F.old[f];
    float* P = F.P + (size_t)f * F.pstride;
    if (B.interior(i, j, k)) {
        const int X = 1, Y = B.nx, Z = B.nx * B.ny;
        P[c] = w[c] + F.chi[f] * (o[c + X] + o[c - X] + o[c + Y] + o[c - Y] + o[c + Z] +
                                  o[c - Z] - 6.0f * o[c]);
    } else {
        P[c] = w[c];
    }
}

__device__ inline float fct_flux2(const float* __restrict__ w, const float* __restrict__ P, int c,
                                  int m, int n, int st, float chi) {
    if (m == 0) {
        c += st;
        m = 1;
    } else if (m == n - 1) {
        c -= st;
        m = n - 2;
    }
    const float dq = w[c + st] - w[c];
    const float d1m = P[c] - P[c - st];
    const float d1p = (m + 1 == n - 1) ? (P[c + st] - P[c]) : (P[c + 2 * st] - P[c + st]);
    const float s = copysignf(1.0f, dq);
    return s * fmaxf(0.0f, fminf(fminf(s * d1m, chi * fabsf(dq)), s * d1p));
}

__global__ void __launch_bounds__(256) k_fct_limit(FctParams F) {
    int i, j, k;
    const Box& B = F.B;
    const int f = blockIdx.y;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const int c = B.IX(i, j, k);
    const float* __restrict__ w = F.wrk[f];
    const float* __restrict__ P = F.P + (size_t)f * F.pstride;
    float* out = F.out[f];
    if (k == 0 || k == B.nz - 1) {     // z faces keep w
        out[c] = w[c];
        return;
    }
    const float chi = F.chi[f];
    const int Y = B.nx, Z = B.nx * B.ny;
    float t;
    if (i > 0 && i < B.nx - 1) {
        const float fx = fct_flux2(w, P, c, i, B.nx, 1, chi);
        const float fxm = fct_flux2(w, P, c - 1, i - 1, B.nx, 1, chi);
        t = P[c] - fx + fxm;
    } else {
        t = w[c];
    }
    if (j > 0 && j < B.ny - 1) {
        const float fy = fct_flux2(w, P, c, j, B.ny, Y, chi);
        const float fym = fct_flux2(w, P, c - Y, j - 1, B.ny, Y, chi);
        t = t - fy + fym;
    }
    const float fz = fct_flux2(w, P, c, k, B.nz, Z, chi);
    const float fzm = fct_flux2(w, P, c - Z, k - 1, B.nz, Z, chi);
    out[c] = t - fz + fzm;
}

// ---------------------------------------------------------------------------
// flanks on the 2-layer shell only.  Shell cells are enumerated slab by
// slab: the four z-planes k in {0,1,nz-2,nz-1}, then for 2<=k<=nz-3 the four
// y-rows j in {0,1,ny-2,ny-1}, then the four x-columns.
__device__ inline bool shell_cell(const Box& B, int t, int& i, int& j, int& k) {
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    const int plane = nx * ny;
    if (t < 4 * plane) {
        const int s = t / plane, r = t % plane;
        k = (s < 2) ? s : nz - 4 + s;
        j = r / nx;
        i = r % nx;
        return true;
    }
    t -= 4 * plane;
    const int nzi = nz - 4;
    const int rows = 4 * nx * nzi;
    if (t < rows) {
        const int kk = t / (4 * nx), r = t % (4 * nx);
        const int s = r / nx;
        i = r % nx;
        j = (s < 2) ? s : ny - 4 + s;
        k = 2 + kk;
        return true;
    }
    t -= rows;
    const int nyi = ny - 4;
    if (t < 4 * nyi * nzi) {
        const int kk = t / (4 * nyi), r = t % (4 * nyi);
        const int s = r / nyi;
        j = 2 + r % nyi;
        i = (s < 2) ? s : nx - 4 + s;
        k = 2 + kk;
        return true;
    }
    return false;
}

__global__ void __launch_bounds__(256) k_flanks_shell(Box B, FlankPtrs Fp, FlankGeom G, float t1,
                                                      float t2) {
    int i, j, k;
    const int f = blockIdx.y;
    if (!shell_cell(B, gtid(), i, j, k)) return;
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    const bool bot = k <= 1, top = k >= nz - 2, fro = j <= 1, bac = j >= ny - 2, lef = i <= 1,
               rig = i >= nx - 2;
    const float ax = G.sxmin + G.sx * (float)i;
    const float ai = 1.0f + (ax - G.bxmin) / G.delx;
    const int ib = (int)ai;
    const float dx = ai - (float)ib, ddx = 1.0f - dx;
    const float ay = G.symin + G.sy * (float)j;
    const float aj = 1.0f + (ay - G.bymin) / G.dely;
    const int jb = (int)aj;
    const float dy = aj - (float)jb, ddy = 1.0f - dy;
    const float az = G.szmin + G.sz * (float)k;
    const float ak = 1.0f + (az - G.bzmin) / G.delz;
    const int kb = (int)ak;
    const float dz = ak - (float)kb, ddz = 1.0f - dz;
    const float* __restrict__ o = Fp.old[f];
    const float* __restrict__ q = Fp.cur[f];
    auto W = [&](int ii, int jj, int kk) {
        const int cc = B.IX(ii - 1, jj - 1, kk - 1);
        return o[cc] * t1 + q[cc] * t2;
    };
    const float I0 = W(ib, jb, kb) * ddx * ddy * ddz + W(ib, jb, kb + 1) * ddx * ddy * dz +
                     W(ib, jb + 1, kb) * ddx * dy * ddz + W(ib, jb + 1, kb + 1) * ddx * dy * dz +
                     W(ib + 1, jb, kb) * dx * ddy * ddz + W(ib + 1, jb, kb + 1) * dx * ddy * dz +
                     W(ib + 1, jb + 1, kb) * dx * dy * ddz +
                     W(ib + 1, jb + 1, kb + 1) * dx * dy * dz;
    float* tgt = Fp.tgt[f];
    const int c = B.IX(i, j, k);
    float w = tgt[c];
    auto blend = [&](bool inner) {
        const float frac = inner ? 0.5f : 0.0f;
        const float frac1 = 1.0f - frac;
        w = frac * w + frac1 * I0;
    };
    if (bot) blend(k == 1);
    if (top) blend(k == nz - 2);
    if (fro) blend(j == 1);
    if (bac) blend(j == ny - 2);
    if (lef) blend(i == 1);
    if (rig) blend(i == nx - 2);
    tgt[c] = w;
}

int shell_cells(const Box& B) {
    return B.V() - (B.nx - 4) * (B.ny - 4) * (B.nz - 4);
}

// ---------------------------------------------------------------------------
// host launchers
void launch_efield(const EParams& P, hipStream_t s) {
    const int n = P.B.V();
    hipLaunchKernelGGL(k_efield, dim3((n + 255) / 256), dim3(256), 0, s, P);
}

void launch_push(const PushParams& P, hipStream_t s) {
    const int n = P.B.V();
    ProfScope prof_("k_push", s);
    hipLaunchKernelGGL(k_push, dim3((n + 255) / 256, 4), dim3(256), 0, s, P);
}

int validate_blocks(const Box& B) { return (B.V() + 255) / 256; }

void launch_validate(const ValParams& Q, float* maxima, hipStream_t s) {
    const int n = Q.B.V();
    const int nb = (n + 255) / 256;
    {
        ProfScope prof_("k_validate", s);
        hipLaunchKernelGGL(k_validate, dim3(nb), dim3(256), 0, s, Q);
    }
    hipLaunchKernelGGL(k_validate_apply, dim3(128), dim3(256), 0, s, Q);
    if (Q.partials && maxima)
        hipLaunchKernelGGL(k_validate_reduce, dim3(1), dim3(256), 0, s, Q.partials, nb, maxima,
                           Q.count);
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, s, Q.count);
}

void launch_lap(const LapParams& L, hipStream_t s) {
    const int n = L.B.V();
    ProfScope prof_("k_lap", s);
    hipLaunchKernelGGL(k_lap, dim3((n + 255) / 256, 4), dim3(256), 0, s, L);
}

void launch_fct2(const FctParams& F, int nf, hipStream_t s) {
    const int n = F.B.V();
    const dim3 g((n + 255) / 256, nf), t(256);
    {
        ProfScope prof_("k_fct_diffuse2", s);
        hipLaunchKernelGGL(k_fct_diffuse2, g, t, 0, s, F);
    }
    ProfScope prof_("k_fct_limit", s);
    hipLaunchKernelGGL(k_fct_limit, g, t, 0, s, F);
}

void launch_flanks_shell(const Box& B, const FlankPtrs& Fp, int nf, const FlankGeom& G, float t1,
                         float t2, hipStream_t s) {
    const int n = shell_cells(B);
    ProfScope prof_("k_flanks_shell", s);
    hipLaunchKernelGGL(k_flanks_shell, dim3((n + 255) / 256, nf), dim3(256), 0, s, B, Fp, G, t1,
                       t2);
}

}  // namespace g3m3
