// Fused box-step kernels (the native runtime's path, g3m3_run_*).
//
// Same arithmetic as the per-subroutine kernels -- every value is produced
// by the reference's expression in the reference's order -- but each
// kernel covers a whole stage of the box-step, so a field is read from HBM
// about once per stage instead of once per reference loop nest:
//   k_efield    calcur + total B + fnd_evel + BandE (code/sub_calcur.f90,
//               code/3d_main.f90:2248-2254, sub_fnd_evel.f90, sub_BandE.f90)
//   k_push      push_ion x3 species, push_bfld, push_elec (launch group picks)
//   k_validate  set_rho + set_speed_agrd in one read of the fields
//   k_lap       Lapidus for the 3 species, electrons and B (launch group)
//   k_fct_fused flux-corrected transport of all 34 fields, one pass
//   k_flanks_table  coarse -> fine boundary on the 2-layer shell only
#include <algorithm>
#include <type_traits>

#include "g3m3_dev.h"
#include "g3m3_fused.h"

namespace g3m3 {

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Launch shape of the fused stage kernels: 256 consecutive cells per block
// (so every wave reads whole, aligned 256-byte runs of a field) and
// blockIdx.y = the group (species / field / B-and-electron group).
//
// XCD-aware order: workgroups are dealt to the 8 XCDs round-robin in
// dispatch order, and each XCD has its own L2.  Block b is given the
// cells of logical block (b % 8) * (nb / 8) + b / 8, so each XCD sweeps a
// contiguous eighth of the box and finds the neighbouring planes a stencil
// reads in its own L2.  The grid is padded to a multiple of 8 blocks.
__device__ inline int xcd_block(int b, int nb) { return (b & 7) * (nb >> 3) + (b >> 3); }

// the cells of logical block lb (blockIdx.y = the group)
__device__ inline bool cell3(const Box& B, int lb, int& i, int& j, int& k, int& g) {
    g = blockIdx.y;
    return cell_of_fast(B, lb * 256 + threadIdx.x, i, j, k);
}
// this launch's logical block: the XCD-aware order of a one-run launch
__device__ inline int lblock() { return xcd_block(blockIdx.x, gridDim.x); }
inline dim3 grid3(const Box& B, int groups) {
    const int nb = (B.V() + 255) / 256;
    return dim3((nb + 7) & ~7, groups, 1);
}
inline dim3 block3() { return dim3(256, 1, 1); }

// Division by a grid step, a species mass or the Reynolds number.  When
// every such divisor is a power of two (P2, checked on the host), x / d is
// computed as x * (1/d): the reciprocal is exact and both operations are
// correctly rounded, so the result is the same float; otherwise a true
// division.
template <bool P2>
__device__ inline float dvp(float x, float d, float r) {
    return P2 ? x * r : x / d;
}

// ---------------------------------------------------------------------------
// E field.  cur at a boundary cell is calcur's face/edge/corner rule; those
// rules resolve to the curl at the clamped index, except the four corners
// at i = nx that average three cells (code/sub_calcur.f90:268-289).  E on
// the boundary is BandE's flank copy: E at the clamped index.

template <bool P2>
struct CurlCtx {
    Box B;
    const float* __restrict__ bx;
    const float* __restrict__ by;
    const float* __restrict__ bz;
    float dxt, dyt, dzt, rdxt, rdyt, rdzt;
    __device__ void curl(int i, int j, int k, float& cx, float& cy, float& cz) const {
        const int c = B.IX(i, j, k), sx = 1, sy = B.nx, sz = B.nx * B.ny;
        auto Dx = [&](float x) { return dvp<P2>(x, dxt, rdxt); };
        auto Dy = [&](float x) { return dvp<P2>(x, dyt, rdyt); };
        auto Dz = [&](float x) { return dvp<P2>(x, dzt, rdzt); };
        cx = Dy(bz[c + sy] - bz[c - sy]) - Dz(by[c + sz] - by[c - sz]);
        cy = Dz(bx[c + sz] - bx[c - sz]) - Dx(bz[c + sx] - bz[c - sx]);
        cz = Dx(by[c + sx] - by[c - sx]) - Dy(bx[c + sy] - bx[c - sy]);
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
template <bool P2>
__device__ inline void efield_interior(const EParams& P, const CurlCtx<P2>& C, int i, int j, int k,
                                       float& Ex, float& Ey, float& Ez, float* evo, float* tvo,
                                       float* bso) {
    const Box& B = P.B;
    const int c = B.IX(i, j, k);
    auto Mq = [&](float x) { return dvp<P2>(x, P.rmq, P.rrmq); };
    auto Mh = [&](float x) { return dvp<P2>(x, P.rmh, P.rrmh); };
    auto Mo = [&](float x) { return dvp<P2>(x, P.rmo, P.rrmo); };
    auto Re = [&](float x) { return dvp<P2>(x, P.reynolds, P.rreynolds); };
    float cx, cy, cz;
    C.curl(i, j, k, cx, cy, cz);
    const float q = P.rho[0][c], h = P.rho[1][c], o = P.rho[2][c];
    const float qden = Mq(q + 0.000001f);
    const float hden = Mh(h + 0.000001f);
    const float oden = Mo(o + 0.000001f);
    const float tot = qden + hden + oden;
    const float tden = q + h + o;
    const float tx = (P.mom[0][c] + P.mom[3][c] + P.mom[6][c]) / tden;
    const float ty = (P.mom[1][c] + P.mom[4][c] + P.mom[7][c]) / tden;
    const float tz = (P.mom[2][c] + P.mom[5][c] + P.mom[8][c]) / tden;
    const float evx = tx - Re(cx / tot);
    const float evy = ty - Re(cy / tot);
    const float evz = tz - Re(cz / tot);
    const float abx = P.b0[0][c] + P.b[0][c], aby = P.b0[1][c] + P.b[1][c],
                abz = P.b0[2][c] + P.b[2][c];
    Ex = -(evy * abz - evz * aby);
    Ey = -(evz * abx - evx * abz);
    Ez = -(evx * aby - evy * abx);
    const int sy = B.nx, sz = B.nx * B.ny;
    const float* ep = P.ep;
    const float arho = (Mq(q) + Mh(h) + Mo(o)) * P.reynolds;
    Ex = Ex - dvp<P2>(ep[c + 1] - ep[c - 1], P.dxt, P.rdxt) / arho;
    Ey = Ey - dvp<P2>(ep[c + sy] - ep[c - sy], P.dyt, P.rdyt) / arho;
    Ez = Ez - dvp<P2>(ep[c + sz] - ep[c - sz], P.dzt, P.rdzt) / arho;
    if (P.rst) {   // list points of box mbndry: rst is zero elsewhere
        const float r = P.rst[c];
        if (r != 0.0f) {
            const float eden = Mq(q) + Mh(h) + Mo(o);
            Ex = Ex + r * cx / eden;
            Ey = Ey + r * cy / eden;
            Ez = Ez + r * cz / eden;
        }
    }
    if (P.eta != 0.0f) {
        const float eden = Mq(q) + Mh(h) + Mo(o);
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

template <bool P2>
__device__ __forceinline__ void efield_block(const EParams& P, int lb) {
    int i, j, k, g;
    const Box& B = P.B;
    if (!cell3(B, lb, i, j, k, g)) return;
    const int c = B.IX(i, j, k);
    const CurlCtx<P2> C{B, P.b[0], P.b[1], P.b[2], P.dxt, P.dyt, P.dzt, P.rdxt, P.rdyt, P.rdzt};
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
        const float tot = dvp<P2>(q + 0.000001f, P.rmq, P.rrmq) +
                          dvp<P2>(h + 0.000001f, P.rmh, P.rrmh) +
                          dvp<P2>(o + 0.000001f, P.rmo, P.rrmo);
        const float tden = q + h + o;
        const float tx = (P.mom[0][c] + P.mom[3][c] + P.mom[6][c]) / tden;
        const float ty = (P.mom[1][c] + P.mom[4][c] + P.mom[7][c]) / tden;
        const float tz = (P.mom[2][c] + P.mom[5][c] + P.mom[8][c]) / tden;
        const float bsx = P.b0[0][c] + P.b[0][c], bsy = P.b0[1][c] + P.b[1][c],
                    bsz = P.b0[2][c] + P.b[2][c];
        P.ev[0][c] = tx - dvp<P2>(cx / tot, P.reynolds, P.rreynolds);
        P.ev[1][c] = ty - dvp<P2>(cy / tot, P.reynolds, P.rreynolds);
        P.ev[2][c] = tz - dvp<P2>(cz / tot, P.reynolds, P.rreynolds);
        P.tv[0][c] = tx;
        P.tv[1][c] = ty;
        P.tv[2][c] = tz;
        P.bs[0][c] = bsx;
        P.bs[1][c] = bsy;
        P.bs[2][c] = bsz;
    }
    P.e[0][c] = Ex;
    P.e[1][c] = Ey;
    P.e[2][c] = Ez;
}

// ---------------------------------------------------------------------------
// Push.  group = 0,1,2: ion species q,h,o; 3: push_bfld (+ push_elec
// when P.do_elec).  The species velocity wrk p / max(wrk rho, 1e-6) is
// evaluated where it is needed instead of stored (same bits).

// the push's time step: from device memory in captured steps (the same
// float the host would pass: 0.5 dt or dt)
__device__ inline float step_delt(const PushParams& P) {
    return P.dt_dev ? P.dt_scale * *P.dt_dev : P.delt;
}

// Stencil directions of the push: 0 the cell, 1 +x, 2 -x, 3 +y, 4 -y,
// 5 +z, 6 -z.  An accessor gives w(m, d) = species array m (0 rho, 1-3
// presx..presz, 4-6 px..pz, 7-9 presxy, presxz, presyz), v(q, d) = the
// velocity component q (px/max(rho, 1e-6) etc.) and b(q, d) = total B
// component q at direction d; push_species_core is the arithmetic, the same
// for every accessor, so every kernel that uses it produces the same bits.
// The operands the push reads at the cell only: the species' old arrays,
// the v x B term and E.  Loaded before anything is stored: on this hardware
// a wait for a load also waits for every store issued before it.
struct PushCentre {
    float o[10], tv[3], e[3], bt;
};

__device__ inline void push_centre_load(const PushParams& P, int s, int c, PushCentre& C) {
    const unsigned ob = (unsigned)c * 4u;
    auto ld = [&](const float* base) {
        return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + ob);
    };
#pragma unroll
    for (int m = 0; m < 10; ++m) C.o[m] = ld(P.old[s].f[m]);
#pragma unroll
    for (int q = 0; q < 3; ++q) C.tv[q] = ld(P.tv[q]);
#pragma unroll
    for (int q = 0; q < 3; ++q) C.e[q] = ld(P.e[q]);
    C.bt = P.btot ? ld(P.btot) : 0.0f;
}

template <bool P2, class A>
__device__ inline void push_species_core(const PushParams& P, int s, int i, int j, int k, int c,
                                         const A& a, const PushCentre& C) {
    const float* o = C.o;
    const float tx = C.tv[0], ty = C.tv[1], tz = C.tv[2];
    const float ex = C.e[0], ey = C.e[1], ez = C.e[2];
    float res[10];
    auto put = [&](int m, float v) { res[m] = v; };
    const float rmass = P.rmass[s], rrmass = P.rrmass[s];
    const float delt = step_delt(P), reynolds = P.reynolds;
    auto Dx = [&](float x) { return dvp<P2>(x, P.dxt, P.rdxt); };
    auto Dy = [&](float x) { return dvp<P2>(x, P.dyt, P.rdyt); };
    auto Dz = [&](float x) { return dvp<P2>(x, P.dzt, P.rdzt); };
    auto Dm = [&](float x) { return dvp<P2>(x, rmass, rrmass); };
    // 0.20*rmass/delt/reynolds (code/sub_push_ion.f90:64)
    const float bmax = P.dt_dev ? 0.20f * rmass / delt / reynolds : P.bmax[s];
    // velocities at the centre and the six neighbours (code/sub_push_ion.f90:67-81)
    auto vx = [&](int d) { return a.v(0, d); };
    auto vy = [&](int d) { return a.v(1, d); };
    auto vz = [&](int d) { return a.v(2, d); };
    auto flux = [&](int m) {
        return Dx(a.w(m, 1) * vx(1) - a.w(m, 2) * vx(2)) +
               Dy(a.w(m, 3) * vy(3) - a.w(m, 4) * vy(4)) +
               Dz(a.w(m, 5) * vz(5) - a.w(m, 6) * vz(6));
    };
    enum { RHO = 0, WX = 1, WY = 2, WZ = 3, PX = 4, PY = 5, PZ = 6, WXY = 7, WXZ = 8, WYZ = 9 };
    put(0, o[0] - delt * (Dx(a.w(PX, 1) - a.w(PX, 2)) + Dy(a.w(PY, 3) - a.w(PY, 4)) +
                               Dz(a.w(PZ, 5) - a.w(PZ, 6))));
    const float aqpx = o[4] - delt * flux(PX);
    const float aqpy = o[5] - delt * flux(PY);
    const float aqpz = o[6] - delt * flux(PZ);
    float abx = a.b(0, 0), aby = a.b(1, 0), abz = a.b(2, 0);
    const float bmag0 =
        P.btot ? C.bt : fmaxf(sqrtf(abx * abx + aby * aby + abz * abz), P.smallbit);
    const float weight = fmaxf(1.0f, bmag0 / bmax);
    const float skin_depth = reynolds / weight;
    const float tcx = (ty * abz - tz * aby);
    const float tcy = -(tx * abz - tz * abx);
    const float tcz = (tx * aby - ty * abx);
    const float qcx = (vy(0) * abz - vz(0) * aby);
    const float qcy = -(vx(0) * abz - vz(0) * abx);
    const float qcz = (vx(0) * aby - vy(0) * abx);
    // the reference divides the y and z cross terms by dxt as well
    // (code/sub_push_ion.f90:155-164)
    const float gpx = Dx(a.w(WX, 1) - a.w(WX, 2)) + Dx(a.w(WXY, 3) - a.w(WXY, 4)) +
                      Dx(a.w(WXZ, 5) - a.w(WXZ, 6));
    const float gpy = Dx(a.w(WXY, 1) - a.w(WXY, 2)) + Dy(a.w(WY, 3) - a.w(WY, 4)) +
                      Dx(a.w(WYZ, 5) - a.w(WYZ, 6));
    const float gpz = Dx(a.w(WXZ, 1) - a.w(WXZ, 2)) + Dy(a.w(WYZ, 3) - a.w(WYZ, 4)) +
                      Dz(a.w(WZ, 5) - a.w(WZ, 6));
    const float dex = (ex + tcx) * reynolds;
    const float dey = (ey + tcy) * reynolds;
    const float dez = (ez + tcz) * reynolds;
    const float dvx = (qcx - tcx) * skin_depth;
    const float dvy = (qcy - tcy) * skin_depth;
    const float dvz = (qcz - tcz) * skin_depth;
    const float wrc = a.w(RHO, 0);
    const float qden = Dm(wrc);
    const float px = aqpx + delt * (qden * (dex + dvx) - gpx);
    const float py = aqpy + delt * (qden * (dey + dvy) - gpy);
    const float pz = aqpz + delt * (qden * (dez + dvz) - gpz);
    const float ax = P.xmin + P.ddx * (float)i;
    const float ay = P.ymin + P.ddy * (float)j;
    const float az = P.zmin + P.ddz * (float)k;
    const float radius = sqrtf(ax * ax + ay * ay + az * az) + 0.0000001f;
    const float tr = P.re_equiv * radius;
    const float g = delt * wrc * P.grav / (tr * tr);
    put(4, px - g * ax / radius);
    put(5, py - g * ay / radius);
    put(6, pz - g * az / radius);

    const float cx = a.w(WX, 0), cy = a.w(WY, 0), cz = a.w(WZ, 0);
    const float cxy = a.w(WXY, 0), cxz = a.w(WXZ, 0), cyz = a.w(WYZ, 0);
    const float px0 = o[1] - delt * flux(WX) - Dx(2.0f * delt * cx * (vx(1) - vx(2)));
    const float py0 = o[2] - delt * flux(WY) - Dy(2.0f * delt * cy * (vy(3) - vy(4)));
    const float pz0 = o[3] - delt * flux(WZ) - Dz(2.0f * delt * cz * (vz(5) - vz(6)));
    abx = (6.0f * a.b(0, 0) + a.b(0, 1) + a.b(0, 2) + a.b(0, 3) + a.b(0, 4) + a.b(0, 5) +
           a.b(0, 6)) / 12.0f;
    aby = (6.0f * a.b(1, 0) + a.b(1, 1) + a.b(1, 2) + a.b(1, 3) + a.b(1, 4) + a.b(1, 5) +
           a.b(1, 6)) / 12.0f;
    abz = (6.0f * a.b(2, 0) + a.b(2, 1) + a.b(2, 2) + a.b(2, 3) + a.b(2, 4) + a.b(2, 5) +
           a.b(2, 6)) / 12.0f;
    const float bmag = sqrtf(abx * abx + aby * aby + abz * abz);
    float sf = reynolds;
    if (bmag > bmax) sf = reynolds * (bmax / bmag);
    const float t2 = 2.0f * delt;
    put(1, px0 - Dy(t2 * cxy * (vx(3) - vx(4))) - Dz(t2 * cxz * (vx(5) - vx(6))) +
               Dm(t2 * sf) * (cxy * abz - cxz * aby));
    put(2, py0 - Dx(t2 * cxy * (vy(1) - vy(2))) - Dz(t2 * cyz * (vy(5) - vy(6))) +
               Dm(t2 * sf) * (cyz * abx - cxy * abz));
    put(3, pz0 - Dx(t2 * cxz * (vz(1) - vz(2))) - Dy(t2 * cyz * (vz(3) - vz(4))) +
               Dm(t2 * sf) * (cxz * aby - cyz * abx));
    const float pxy = o[7] - delt * flux(WXY);
    put(7, pxy - delt * cxy * (Dx(vx(1) - vx(2)) + Dy(vy(3) - vy(4))) -
               Dy(delt * cy * (vx(3) - vx(4))) - Dx(delt * cx * (vy(1) - vy(2))) -
               Dz(delt * cyz * (vx(5) - vx(6))) - Dz(delt * cxz * (vy(5) - vy(6))) +
               Dm(delt * sf) * (abz * (cy - cx) - aby * cyz + abx * cxz));
    const float pxz = o[8] - delt * flux(WXZ);
    put(8, pxz - delt * cxz * (Dx(vx(1) - vx(2)) + Dz(vz(5) - vz(6))) -
               Dz(delt * cz * (vx(5) - vx(6))) - Dx(delt * cx * (vz(1) - vz(2))) -
               Dy(delt * cyz * (vx(3) - vx(4))) - Dy(delt * cxy * (vz(3) - vz(4))) +
               Dm(delt * sf) * (aby * (cx - cz) - abx * cxy + abz * cyz));
    const float pyz = o[9] - delt * flux(WYZ);
    put(9, pyz - delt * cyz * (Dy(vy(3) - vy(4)) + Dz(vz(5) - vz(6))) -
               Dy(delt * cy * (vz(3) - vz(4))) - Dz(delt * cz * (vy(5) - vy(6))) -
               Dx(delt * cxy * (vz(1) - vz(2))) - Dx(delt * cxz * (vy(1) - vy(2))) +
               Dm(delt * sf) * (abx * (cz - cy) - abz * cxz + aby * cxy));
    const unsigned ob = (unsigned)c * 4u;
#pragma unroll
    for (int m = 0; m < 10; ++m)
        *reinterpret_cast<float*>(reinterpret_cast<char*>(P.out[s].f[m]) + ob) = res[m];
}

// The velocity of code/sub_push_ion.f90:67-81 at one point
__device__ inline float push_vel(float p, float rho) { return p / fmaxf(rho, 1.e-6f); }

// Reads every operand from global memory at the cell and its six
// neighbours; the seven velocities are formed by each thread.
struct PushGlobalAcc {
    const float* const* wp;
    const float* const* bs;
    int c, off[7];
    float vel[3][7];
    __device__ PushGlobalAcc(const PushParams& P, int s, int c_)
        : wp(P.src[s].f), bs(P.bs), c(c_) {
        const int X = 1, Y = P.B.nx, Z = P.B.nx * P.B.ny;
        off[0] = 0, off[1] = X, off[2] = -X, off[3] = Y, off[4] = -Y, off[5] = Z, off[6] = -Z;
#pragma unroll
        for (int d = 0; d < 7; ++d) {
            const float r = wp[0][c + off[d]];
            vel[0][d] = push_vel(wp[4][c + off[d]], r);
            vel[1][d] = push_vel(wp[5][c + off[d]], r);
            vel[2][d] = push_vel(wp[6][c + off[d]], r);
        }
    }
    __device__ float w(int m, int d) const { return wp[m][c + off[d]]; }
    __device__ float v(int q, int d) const { return vel[q][d]; }
    __device__ float b(int q, int d) const { return bs[q][c + off[d]]; }
};

template <bool P2>
__device__ inline void push_species(const PushParams& P, int s, int i, int j, int k) {
    const int c = P.B.IX(i, j, k);
    PushCentre C;
    push_centre_load(P, s, c, C);
    const PushGlobalAcc a(P, s, c);
    push_species_core<P2>(P, s, i, j, k, c, a, C);
}

// ---------------------------------------------------------------------------
// The species push marching along k.  A block (4 waves) owns a 62 x 4
// (i, j) tile of one species and a run of MKC k-planes.  Lane l of wave w
// holds cell (i0 - 1 + l, j0 + w): lanes 0 and 63 are the x halo, so a row
// is one aligned 64-wide load and lanes 1..62 produce output.  Waves 0 and
// 1 also load the tile's y-halo rows j0 - 1 and j0 + 4.
//
// Per plane, LDS holds the 15 arrays the stencil reads at x/y neighbours
// (the 3 velocities, the 9 transported species arrays, the 3 B components)
// on the 64 x 6 tile; a 2-slot ring holds planes k and k+1, so the +z
// neighbour is the next slot and the -z neighbour is kept in registers.
// Each point's velocity is formed once, by the thread that loads it.
// Plane k+2 is loaded at the start of plane k's iteration and written into
// plane k's slot after one barrier, so the loads stay in flight during the
// arithmetic.  Loads are clamped into the box; lanes and rows outside the
// interior store nothing.
constexpr int MTX = 64, MOX = 62, MTY = 4, MRY = MTY + 2, MKC = 16, MNA = 15;
constexpr int MNR = 13;     // values loaded per point: rho, arrays 1-9, 3 B components

struct MarchLds {
    float a[2][MNA][MRY][MTX];
};

// LDS index of species array m (1-9) and of B component q
__device__ constexpr int mw(int m) { return m + 2; }
__device__ constexpr int mb(int q) { return 12 + q; }

struct PushMarchAcc {
    const MarchLds& L;
    int cur, nxt, r, l;
    const float* prev;      // the 15 arrays at (i, j, k-1)
    float rho;              // rho at the cell
    __device__ float at(int x, int d) const {
        switch (d) {
            case 0: return L.a[cur][x][r][l];
            case 1: return L.a[cur][x][r][l + 1];
            case 2: return L.a[cur][x][r][l - 1];
            case 3: return L.a[cur][x][r + 1][l];
            case 4: return L.a[cur][x][r - 1][l];
            case 5: return L.a[nxt][x][r][l];
            default: return prev[x];
        }
    }
    __device__ float w(int m, int d) const { return m == 0 ? rho : at(mw(m), d); }
    __device__ float v(int q, int d) const { return at(q, d); }
    __device__ float b(int q, int d) const { return at(mb(q), d); }
};

// base[byte offset]: with a block-uniform base the load takes the base
// from scalar registers and only the 32-bit offset from a vector register
__device__ inline float ldb(const float* base, unsigned off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + off);
}

// the 13 values of one point (byte offset ob), in the order of MNR
__device__ inline void march_load(const PushParams& P, int s, unsigned ob, float* v) {
    const SpeciesSet& W = P.src[s];
#pragma unroll
    for (int m = 0; m < 10; ++m) v[m] = ldb(W.f[m], ob);
#pragma unroll
    for (int q = 0; q < 3; ++q) v[10 + q] = ldb(P.bs[q], ob);
}

__device__ inline void march_store(MarchLds& L, int slot, int row, int l, const float* v) {
    L.a[slot][0][row][l] = push_vel(v[4], v[0]);
    L.a[slot][1][row][l] = push_vel(v[5], v[0]);
    L.a[slot][2][row][l] = push_vel(v[6], v[0]);
#pragma unroll
    for (int m = 1; m < 10; ++m) L.a[slot][mw(m)][row][l] = v[m];
#pragma unroll
    for (int q = 0; q < 3; ++q) L.a[slot][mb(q)][row][l] = v[10 + q];
}

template <bool P2>
__device__ __forceinline__ void push_march_block(const PushParams& P, int ntx, int nty, int nkc, int lb) {
    __shared__ MarchLds L;
    const Box B = P.B;
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    // logical block order: species fastest, so the blocks of one tile's
    // species are neighbours in one XCD's run of blocks and share the
    // tile's b_s, E, tv and btot reads through that XCD's L2
    int t = lb;
    const int s = t % P.nspecies;
    t /= P.nspecies;
    const int tx = t % ntx;
    t /= ntx;
    const int ty = t % nty;
    t /= nty;
    const int kc = t;
    if (kc >= nkc) return;                   // grid padding: whole blocks
    const int l = threadIdx.x, w = threadIdx.y;
    const int i = tx * MOX + l, j0 = 1 + ty * MTY, j = j0 + w;
    const int ic = min(i, nx - 1), jc = min(j, ny - 1);
    const bool out = l >= 1 && l <= MOX && i <= nx - 2 && j <= ny - 2;
    // step 2: the lanes on the box's x faces copy the species' old values
    // there (loaded with the centre operands); k_push copies the rest of
    // the shell.  Each such cell has one owner: lane 0 for i = 0, an output
    // lane -- or lane 63 of the last tile -- for i = nx-1.
    const bool xshell = P.copy_shell && j <= ny - 2 &&
                        ((i == 0 && l == 0) ||
                         (i == nx - 1 && l >= 1 && (l <= MOX || tx == ntx - 1)));
    // the y-halo row this wave loads (waves 0 and 1), and its LDS row
    const bool halo = w < 2;
    const int hj = w == 0 ? j0 - 1 : min(j0 + MTY, ny - 1), hr = w == 0 ? 0 : MRY - 1;
    const int r = w + 1;
    const int kcl = P.march_kc > 0 ? P.march_kc : MKC;
    const int kb = 1 + kc * kcl, ke = min(kb + kcl, nz - 1);
    const int Z = nx * ny;
    const int own0 = ic + nx * jc, halo0 = ic + nx * hj;
    float prev[MNA];
    {
        float v[MNR];
        march_load(P, s, 4u * (own0 + Z * (kb - 1)), v);
        prev[0] = push_vel(v[4], v[0]);
        prev[1] = push_vel(v[5], v[0]);
        prev[2] = push_vel(v[6], v[0]);
#pragma unroll
        for (int m = 1; m < 10; ++m) prev[mw(m)] = v[m];
#pragma unroll
        for (int q = 0; q < 3; ++q) prev[mb(q)] = v[10 + q];
    }
    float rho_cur, rho_nxt;
    {
        float v[MNR];
        march_load(P, s, 4u * (own0 + Z * kb), v);
        march_store(L, 0, r, l, v);
        rho_cur = v[0];
        if (halo) {
            march_load(P, s, 4u * (halo0 + Z * kb), v);
            march_store(L, 0, hr, l, v);
        }
        march_load(P, s, 4u * (own0 + Z * (kb + 1)), v);
        march_store(L, 1, r, l, v);
        rho_nxt = v[0];
        if (halo) {
            march_load(P, s, 4u * (halo0 + Z * (kb + 1)), v);
            march_store(L, 1, hr, l, v);
        }
    }
    __syncthreads();
    int cur = 0;
    for (int k = kb; k < ke; ++k) {
        const int nxt = cur ^ 1;
        // (prefetching these a plane ahead measured no faster: 92.7 against
        // 91.3 us, for 34 more VGPRs)
        PushCentre cc;
        push_centre_load(P, s, own0 + Z * k, cc);
        // plane k+2 (clamped into the box), in flight during plane k
        const int k2 = min(k + 2, nz - 1);
        float vo[MNR], vh[MNR];
        march_load(P, s, 4u * (own0 + Z * k2), vo);
        if (halo) march_load(P, s, 4u * (halo0 + Z * k2), vh);
        if (out) {
            const PushMarchAcc a{L, cur, nxt, r, l, prev, rho_cur};
            push_species_core<P2>(P, s, i, j, k, own0 + Z * k, a, cc);
        } else if (xshell) {
            const unsigned ob = 4u * (unsigned)(own0 + Z * k);
#pragma unroll
            for (int m = 0; m < 10; ++m)
                *reinterpret_cast<float*>(reinterpret_cast<char*>(P.out[s].f[m]) + ob) = cc.o[m];
        }
#pragma unroll
        for (int x = 0; x < MNA; ++x) prev[x] = L.a[cur][x][r][l];
        __syncthreads();                     // plane k's slot is free
        march_store(L, cur, r, l, vo);
        if (halo) march_store(L, cur, hr, l, vh);
        rho_cur = rho_nxt;
        rho_nxt = vo[0];
        cur = nxt;
    }
}

template <bool P2>
__device__ inline void push_field_elec(const PushParams& P, int i, int j, int k) {
    const Box& B = P.B;
    const int c = B.IX(i, j, k), sy = B.nx, sz = B.nx * B.ny;
    const float delt = step_delt(P);
    auto Dx = [&](float x) { return dvp<P2>(x, P.fdxt, P.rfdxt); };
    auto Dy = [&](float x) { return dvp<P2>(x, P.fdyt, P.rfdyt); };
    auto Dz = [&](float x) { return dvp<P2>(x, P.fdzt, P.rfdzt); };
    const float* __restrict__ ex = P.e[0];
    const float* __restrict__ ey = P.e[1];
    const float* __restrict__ ez = P.e[2];
    // code/sub_push_bfld.f90:87-97; stored after every load (see push_shell_copy)
    const float b0 =
        P.bold[0][c] - delt * (Dy(ez[c + sy] - ez[c - sy]) - Dz(ey[c + sz] - ey[c - sz]));
    const float b1 =
        P.bold[1][c] + delt * (Dx(ez[c + 1] - ez[c - 1]) - Dz(ex[c + sz] - ex[c - sz]));
    const float b2 =
        P.bold[2][c] - delt * (Dx(ey[c + 1] - ey[c - 1]) - Dy(ex[c + sy] - ex[c - sy]));
    float ep = 0.0f;
    if (P.do_elec) {   // code/sub_push_elec.f90:31-47
        const float* __restrict__ w = P.wrkepres;
        const float* __restrict__ evx = P.ev[0];
        const float* __restrict__ evy = P.ev[1];
        const float* __restrict__ evz = P.ev[2];
        const float gx = Dx(w[c + 1] - w[c - 1]);
        const float gy = Dy(w[c + sy] - w[c - sy]);
        const float gz = Dz(w[c + sz] - w[c - sz]);
        const float div = Dx(w[c + 1] * evx[c + 1] - w[c - 1] * evx[c - 1]) +
                          Dy(w[c + sy] * evy[c + sy] - w[c - sy] * evy[c - sy]) +
                          Dz(w[c + sz] * evz[c + sz] - w[c - sz] * evz[c - sz]);
        const float e = P.oldepres[c] - delt * P.gamma * div;
        ep = e + delt * P.gamma1 * (evx[c] * gx + evy[c] * gy + evz[c] * gz);
    }
    P.bout[0][c] = b0;
    P.bout[1][c] = b1;
    P.bout[2][c] = b2;
    if (P.do_elec) P.epres[c] = ep;
}

// Cells the push does not write keep their pre-step values.  With the
// runtime's buffer rotation the output buffer holds stale data there, so
// when copy_shell is set the boundary threads copy them from old.
// Every value is loaded before the first store: the compiler cannot prove
// that the output arrays do not alias the inputs, so interleaved copies
// would wait for each load in turn.
__device__ inline void push_shell_copy(const PushParams& P, int s, int c) {
    if (s < 3) {
        float v[10];
#pragma unroll
        for (int m = 0; m < 10; ++m) v[m] = P.old[s].f[m][c];
#pragma unroll
        for (int m = 0; m < 10; ++m) P.out[s].f[m][c] = v[m];
    } else {
        float v[4];
#pragma unroll
        for (int m = 0; m < 3; ++m) v[m] = P.bold[m][c];
        v[3] = P.oldepres[c];
#pragma unroll
        for (int m = 0; m < 3; ++m) P.bout[m][c] = v[m];
        P.epres[c] = v[3];
    }
}

// all four groups' shells of one cell (the march launch): 34 loads, then
// 34 stores
__device__ inline void push_shell_copy_all(const PushParams& P, int c) {
    float v[34];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int m = 0; m < 10; ++m) v[10 * s + m] = P.old[s].f[m][c];
#pragma unroll
    for (int m = 0; m < 3; ++m) v[30 + m] = P.bold[m][c];
    v[33] = P.oldepres[c];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int m = 0; m < 10; ++m) P.out[s].f[m][c] = v[10 * s + m];
#pragma unroll
    for (int m = 0; m < 3; ++m) P.bout[m][c] = v[30 + m];
    P.epres[c] = v[33];
}

// With march set the grid has the B / electron group only, and its
// boundary threads copy the species' shells as well -- except on the x
// faces between the j and k faces, which k_push_march has copied.
template <bool P2>
__device__ __forceinline__ void push_block(const PushParams& P, int lb) {
    int i, j, k, s;
    if (!cell3(P.B, lb, i, j, k, s)) return;
    if (P.march) s = 3;
    if (!P.B.interior(i, j, k)) {
        if (P.copy_shell) {
            const int c = P.B.IX(i, j, k);
            const bool xface = j > 0 && j < P.B.ny - 1 && k > 0 && k < P.B.nz - 1;
            if (P.march && !xface)
                push_shell_copy_all(P, c);
            else
                push_shell_copy(P, s, c);
        }
        return;
    }
    if (s < 3) push_species<P2>(P, s, i, j, k);
    else push_field_elec<P2>(P, i, j, k);
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

// lb: the logical block (its cells); pb: the block's slot in Q.partials
__device__ __forceinline__ void validate_block(const ValParams& Q, int lb, int pb) {
    const Box& B = Q.B;
    int i, j, k, g;
    float mx = 0.0f, my = 0.0f, mz = 0.0f, mc = 0.0f, ma = 0.0f;
    const float dmins[3] = {2.e-4f, 1.e-4f, 2.e-5f};
    if (cell3(B, lb, i, j, k, g)) {
        const int c = B.IX(i, j, k);
        const bool inner = B.interior(i, j, k);
        // every centre operand first: a list append (a store) inside one
        // species' branch would otherwise hold back the next species' loads
        float cv[3][7];
#pragma unroll
        for (int s = 0; s < 3; ++s)
#pragma unroll
            for (int m = 0; m < 7; ++m) cv[s][m] = Q.sp[s].f[m][c];
        const float e0 = Q.ep[c];
        float bt[6] = {0, 0, 0, 0, 0, 0};
        if (Q.partials && inner)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                bt[q] = Q.b0[q][c];
                bt[3 + q] = Q.b[q][c];
            }
        float rho_final[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            float* const* f = Q.sp[s].f;
            const RhoFix cf = set_rho_cell(cv[s][0], cv[s][1], cv[s][2], cv[s][3], dmins[s]);
            rho_final[s] = cf.r;
            bool avg = false;
            if (inner) {
                const float arho = cf.r + Q.smallbit;
                const float avx = fabsf(cv[s][4] / arho);
                const float avy = fabsf(cv[s][5] / arho);
                const float avz = fabsf(cv[s][6] / arho);
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
                Q.list[atomic_add_global(Q.count, 1)] = e;
            } else if (cf.hit) {
                ValEntry e;
                e.cell = c;
                e.species = s;
                e.v[0] = cf.r;
                e.v[1] = cf.x;
                e.v[2] = cf.y;
                e.v[3] = cf.z;
                e.v[4] = cv[s][4];
                e.v[5] = cv[s][5];
                e.v[6] = cv[s][6];
                e.v[7] = e.v[8] = e.v[9] = 0.0f;
                Q.list[atomic_add_global(Q.count, 1)] = e;
            }
        }
        // electron pressure: |pe| everywhere, sound-speed cap and Alfven
        // maximum on the interior with the final densities (:404-426)
        float e1 = fabsf(e0);
        if (inner) {
            const float arho = rho_final[0] + rho_final[1] + rho_final[2] + Q.smallbit;
            const float cs = sqrtf(e1 / arho);
            if (cs > Q.cslim) e1 = e1 * Q.cslim / cs;
            if (Q.partials) {
                const float sx = bt[0] + bt[3], sy = bt[1] + bt[4], sz = bt[2] + bt[5];
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
            Q.partials[pb * 5 + q] = m;
        }
    }
}

// The pass's CFL maxima from the per-block partials (one block)
__device__ inline void validate_reduce_block(const float* __restrict__ partials, int nblocks,
                                             float* maxima) {
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

// Blocks 0 .. kApplyBlocks-1 apply the pass's side-list; with maxima set,
// one more block forms the pass's CFL maxima (independent work, one launch)
constexpr int kApplyBlocks = 128;
__device__ __forceinline__ void validate_apply_block(const ValParams& Q, float* maxima, int nblocks,
                                            int lb) {
    if (lb == kApplyBlocks) {
        validate_reduce_block(Q.partials, nblocks, maxima);
        return;
    }
    const int n = *Q.count;
    // the next pass appends to the other counter; clear it here (no pass
    // is using it now) so no separate zeroing launch is needed
    if (lb == 0 && threadIdx.x == 0) *Q.count_next = 0;
    for (int e = lb * 256 + threadIdx.x; e < n; e += kApplyBlocks * 256) {
        const ValEntry a = Q.list[e];   // whole entry loaded before the stores
        float* const* f = Q.sp[a.species].f;
#pragma unroll
        for (int m = 0; m < 10; ++m) f[m][a.cell] = a.v[m];
    }
}

// ---------------------------------------------------------------------------
// Lapidus.  group = 0,1,2: a species' 10 fields with its own velocity
// p/max(rho, smallbit); 3: epres and b with the total ion velocity.
__device__ inline void lap_body(const LapParams& L, int g, int i, int j, int k) {
    const Box& B = L.B;
    const int c = B.IX(i, j, k), X = 1, Y = B.nx, Z = B.nx * B.ny;
    const int off[7] = {0, X, -X, Y, -Y, Z, -Z};
    const float delt = L.dt_dev ? *L.dt_dev : L.delt;
    auto ldb = [](const float* base, int e) {
        return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) +
                                               4u * (unsigned)e);
    };
    float vx[7], vy[7], vz[7];
    // every operand is loaded before the first store: on this hardware a
    // wait for a load also waits for every store issued before it
    float q[10][7];
    if (g < 3) {
        float* const* f = L.src[g].f;
#pragma unroll
        for (int m = 0; m < 10; ++m)
#pragma unroll
            for (int d = 0; d < 7; ++d) q[m][d] = ldb(f[m], c + off[d]);
#pragma unroll
        for (int d = 0; d < 7; ++d) {
            const float a = fmaxf(q[0][d], L.smallbit);
            vx[d] = q[4][d] / a;
            vy[d] = q[5][d] / a;
            vz[d] = q[6][d] / a;
        }
    } else {
#pragma unroll
        for (int d = 0; d < 7; ++d) {
            q[0][d] = ldb(L.ep_src, c + off[d]);
#pragma unroll
            for (int m = 0; m < 3; ++m) q[1 + m][d] = ldb(L.b_src[m], c + off[d]);
        }
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
    auto lap = [&](const float* t, float chi) {
        const float q0 = t[0];
        return q0 + chi * (delt * (uxp * (t[1] - q0) - uxm * (q0 - t[2])) +
                           delt * (uyp * (t[3] - q0) - uym * (q0 - t[4])) +
                           delt * (uzp * (t[5] - q0) - uzm * (q0 - t[6])));
    };
    const unsigned ob = 4u * (unsigned)c;
    auto st = [&](float* base, float v) {
        *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + ob) = v;
    };
    if (g < 3) {
        float v[10];
        v[0] = lap(q[0], L.chirho);
#pragma unroll
        for (int m = 1; m < 10; ++m) v[m] = lap(q[m], L.chipxyz);
#pragma unroll
        for (int m = 0; m < 10; ++m) st(L.dst[g].f[m], v[m]);
    } else {
        float v[4];
        v[0] = lap(q[0], L.chipxyz);
#pragma unroll
        for (int m = 0; m < 3; ++m) v[1 + m] = lap(q[1 + m], L.chierg);
        st(L.ep_dst, v[0]);
#pragma unroll
        for (int m = 0; m < 3; ++m) st(L.b_dst[m], v[1 + m]);
    }
}

// Lapidus of the three species marching along k, on the push's tiling: a
// block owns a 62 x 4 (i, j) tile of one species and a run of MKC planes;
// LDS holds, per plane, the 10 species arrays and the 3 velocity
// components p / max(rho, smallbit) on the tile with its x and y halo, in
// a 2-slot ring (planes k and k+1); -z comes from registers.  Each velocity
// is formed once, by the thread that loads the point.  Same expression as
// lap_body, so the same bits.
constexpr int LNA = 13;
struct LapMarchLds {
    float a[2][LNA][MRY][MTX];
};

__device__ inline void lap_march_load(const LapParams& L, int s, unsigned ob, float* v) {
#pragma unroll
    for (int m = 0; m < 10; ++m) v[m] = ldb(L.src[s].f[m], ob);
}

__device__ inline void lap_march_store(LapMarchLds& S, int slot, int row, int l, const float* v,
                                       float smallbit) {
#pragma unroll
    for (int m = 0; m < 10; ++m) S.a[slot][m][row][l] = v[m];
    const float a = fmaxf(v[0], smallbit);
    S.a[slot][10][row][l] = v[4] / a;
    S.a[slot][11][row][l] = v[5] / a;
    S.a[slot][12][row][l] = v[6] / a;
}

__device__ __forceinline__ void lap_march_block(const LapParams& L, int ntx, int nty, int nkc, int lb) {
    __shared__ LapMarchLds S;
    const Box B = L.B;
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    int t = lb;
    const int tx = t % ntx;
    t /= ntx;
    const int ty = t % nty;
    t /= nty;
    const int kc = t % nkc;
    const int s = t / nkc;
    if (s >= 3) return;                      // grid padding: whole blocks
    const int l = threadIdx.x, w = threadIdx.y;
    const int i = tx * MOX + l, j0 = 1 + ty * MTY, j = j0 + w;
    const int ic = min(i, nx - 1), jc = min(j, ny - 1);
    const bool out = l >= 1 && l <= MOX && i <= nx - 2 && j <= ny - 2;
    const bool halo = w < 2;
    const int hj = w == 0 ? j0 - 1 : min(j0 + MTY, ny - 1), hr = w == 0 ? 0 : MRY - 1;
    const int r = w + 1;
    const int kb = 1 + kc * MKC, ke = min(kb + MKC, nz - 1);
    const int Z = nx * ny;
    const int own0 = ic + nx * jc, halo0 = ic + nx * hj;
    const float delt = L.dt_dev ? *L.dt_dev : L.delt, sb = L.smallbit;
    float prev[LNA];
    {
        float v[10];
        lap_march_load(L, s, 4u * (own0 + Z * (kb - 1)), v);
#pragma unroll
        for (int m = 0; m < 10; ++m) prev[m] = v[m];
        const float a = fmaxf(v[0], sb);
        prev[10] = v[4] / a;
        prev[11] = v[5] / a;
        prev[12] = v[6] / a;
        lap_march_load(L, s, 4u * (own0 + Z * kb), v);
        lap_march_store(S, 0, r, l, v, sb);
        if (halo) {
            lap_march_load(L, s, 4u * (halo0 + Z * kb), v);
            lap_march_store(S, 0, hr, l, v, sb);
        }
        lap_march_load(L, s, 4u * (own0 + Z * (kb + 1)), v);
        lap_march_store(S, 1, r, l, v, sb);
        if (halo) {
            lap_march_load(L, s, 4u * (halo0 + Z * (kb + 1)), v);
            lap_march_store(S, 1, hr, l, v, sb);
        }
    }
    __syncthreads();
    int cur = 0;
    for (int k = kb; k < ke; ++k) {
        const int nxt = cur ^ 1;
        const int k2 = min(k + 2, nz - 1);
        float vo[10], vh[10];
        lap_march_load(L, s, 4u * (own0 + Z * k2), vo);
        if (halo) lap_march_load(L, s, 4u * (halo0 + Z * k2), vh);
        if (out) {
            // operand x at direction d: 0 the cell, 1 +x, 2 -x, 3 +y, 4 -y, 5 +z, 6 -z
            auto at = [&](int x, int d) {
                switch (d) {
                    case 0: return S.a[cur][x][r][l];
                    case 1: return S.a[cur][x][r][l + 1];
                    case 2: return S.a[cur][x][r][l - 1];
                    case 3: return S.a[cur][x][r + 1][l];
                    case 4: return S.a[cur][x][r - 1][l];
                    case 5: return S.a[nxt][x][r][l];
                    default: return prev[x];
                }
            };
            const float uxp = fabsf(at(10, 1) - at(10, 0));
            const float uxm = fabsf(at(10, 0) - at(10, 2));
            const float uyp = fabsf(at(11, 3) - at(11, 0));
            const float uym = fabsf(at(11, 0) - at(11, 4));
            const float uzp = fabsf(at(12, 5) - at(12, 0));
            const float uzm = fabsf(at(12, 0) - at(12, 6));
            float res[10];
#pragma unroll
            for (int m = 0; m < 10; ++m) {
                const float chi = m == 0 ? L.chirho : L.chipxyz;
                const float q0 = at(m, 0);
                res[m] = q0 + chi * (delt * (uxp * (at(m, 1) - q0) - uxm * (q0 - at(m, 2))) +
                                     delt * (uyp * (at(m, 3) - q0) - uym * (q0 - at(m, 4))) +
                                     delt * (uzp * (at(m, 5) - q0) - uzm * (q0 - at(m, 6))));
            }
            const unsigned ob = 4u * (unsigned)(own0 + Z * k);
#pragma unroll
            for (int m = 0; m < 10; ++m)
                *reinterpret_cast<float*>(reinterpret_cast<char*>(L.dst[s].f[m]) + ob) = res[m];
        }
#pragma unroll
        for (int x = 0; x < LNA; ++x) prev[x] = S.a[cur][x][r][l];
        __syncthreads();                     // plane k's slot is free
        lap_march_store(S, cur, r, l, vo, sb);
        if (halo) lap_march_store(S, cur, hr, l, vh, sb);
        cur = nxt;
    }
}

__device__ __forceinline__ void lap_block(const LapParams& L, int lb) {
    int i, j, k, g;
    const Box& B = L.B;
    if (!cell3(B, lb, i, j, k, g)) return;
    if (L.march) g = 3;
    if (!B.interior(i, j, k)) return;
    lap_body(L, g, i, j, k);
}

// ---------------------------------------------------------------------------
// FCT in one pass (k_fct_fused): a block owns a 63 x FTY (i, j) tile of one
// field and a run of k-planes (16 unless the run's autotuner picked another
// depth), and marches it along k.  Rings of
// three LDS planes hold old (3-cell halo), wrk and the diffused field P
// (2-cell halo), so P never leaves the CU and each input value comes from
// HBM about once.  A wave covers two tile rows: lane tx owns the cell
// i0 - 1 + tx of rows j0 + 2ty and j0 + 2ty + 1.  Lane 0 outputs nothing;
// it supplies the flux at the tile's lower x face, so each x face flux is
// evaluated once and handed to the lane above by a shuffle.  Each thread
// keeps its own column of P (5 planes) and wrk (3 planes) in registers for
// the z fluxes, and carries the z face flux from one plane to the next.
// The arithmetic is the reference's (code/sub_fcsmooth.f90:44-239),
// expression for expression.
//
// Values outside the box: tile loads are clamped into the box, so halo
// cells beyond it hold copies of box cells.  They reach only lanes that
// store nothing, or operands that a select discards (the limiter's far
// difference at the last interior cell), so every stored value equals the
// reference's.

// FTY rows per tile, two per wave: a taller tile spends less of its work on
// the halo (P is made on (FTX + 5)(FTY + 4) cells for FTX * FTY outputs)
#ifndef G3M3_FCT_TILE_ROWS
#define G3M3_FCT_TILE_ROWS 8
#endif
// Word 3 of a raw buffer resource on gfx950: data format 32 bits (bits
// 15-18 = 4), every other field zero (no swizzle, no index stride).
constexpr unsigned kBufferWord3 = 0x00020000u;
constexpr int FTX = 63, FTY = G3M3_FCT_TILE_ROWS, FKC = 16;   // FKC: default k-planes per block
constexpr int FNT = 32 * FTY;                 // threads per block
constexpr int FPX = FTX + 5, FPY = FTY + 4;   // wrk / P: cells i0-3 .. i0+64, j0-2 .. j0+FTY+1
constexpr int FOX = FPX + 2, FOY = FPY + 2;   // old: one more cell on each side
constexpr int FO_N = FOX * FOY, FP_N = FPX * FPY;
constexpr int FO_PT = (FO_N + FNT - 1) / FNT, FP_PT = (FP_N + FNT - 1) / FNT;
static_assert(FTY % 2 == 0 && FP_PT <= 32, "two rows per wave; xyface holds FP_PT bits");

struct FctLds {
    float o[3][FO_N];
    float w[3][FP_N];
    float p[3][FP_N];
};

// the limiter of code/sub_fcsmooth.f90 on one face: dq = wrk difference,
// d1m / d1p = diffused differences below / above
__device__ inline float fct_lim(float dq, float d1m, float d1p, float chi) {
    const float s = copysignf(1.0f, dq);
    return s * fmaxf(0.0f, fminf(fminf(s * d1m, chi * fabsf(dq)), s * d1p));
}

// limited antidiffusive flux of one axis at the face above the cell `base`
// cells along that axis from the output cell (m = the index along the axis
// of that cell, n the extent); W(t) / P(t) read wrk / P t cells along the
// axis.  At the two ends the reference shifts the stencil one cell inwards
// (code/sub_fcsmooth.f90:87-239); the shift and the choice of the far
// difference are selects, so lanes never diverge.
template <class WA, class PA>
__device__ inline float fct_flux(const WA& W, const PA& P, int base, int m, int n, float chi) {
    const int sh = (m == 0) ? 1 : ((m == n - 1) ? -1 : 0);
    const int c = base + sh;
    const float w0 = W(c), w1 = W(c + 1);
    const float pm1 = P(c - 1), p0 = P(c), p1 = P(c + 1), p2 = P(c + 2);
    const float d1p = (m + sh + 1 == n - 1) ? (p1 - p0) : (p2 - p1);
    return fct_lim(w1 - w0, p0 - pm1, d1p, chi);
}

__device__ __forceinline__ void fct_block(const FctParams& F, int ntx, int nty, int nkc, int bl) {
    __shared__ FctLds L;
    const Box& B = F.B;
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    // logical block order: field, k-run, tile row, tile column (fastest)
    const int per_f = ntx * nty * nkc;
    if (bl >= per_f * F.nf) return;
    const int f = bl / per_f;
    const int r = bl % per_f;
    const int kc = r / (ntx * nty);
    const int t = r % (ntx * nty);
    const int i0 = (t % ntx) * FTX, j0 = (t / ntx) * FTY;
    const int kb = kc * F.kc_len, ke = min(nz, kb + F.kc_len);
    // wrk and old through buffer resources: the plane's byte offset goes in
    // a scalar register and each element's in-plane offset in a vector
    // register, so no per-plane address arithmetic runs on the VALU
    const int nbytes = nx * ny * nz * (int)sizeof(float);
    const __amdgpu_buffer_rsrc_t wg =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(F.wrk[f]), (short)0, nbytes,
                                          kBufferWord3);
    const __amdgpu_buffer_rsrc_t og =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(F.old[f]), (short)0, nbytes,
                                          kBufferWord3);
    float* out = F.out[f];
    const float chi = F.chi[f];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = tx + 64 * ty;
    const int plane = nx * ny;
    const int plane_bytes = plane * (int)sizeof(float);

    // this thread's tile elements, fixed for the whole march: clamped
    // in-plane byte offsets of the loads, the P element's centre in the old
    // tile, and whether the P element lies on an x or y face of the box
    unsigned offw[FP_PT], offo[FO_PT];
    int ocw[FP_PT];
    int xyface = 0;
#pragma unroll
    for (int m = 0; m < FP_PT; ++m) {
        const int e = tid + FNT * m;
        const int ii = e % FPX, jj = e / FPX;
        const int gi = i0 - 3 + ii, gj = j0 - 2 + jj;
        offw[m] = (unsigned)(min(max(gj, 0), ny - 1) * nx + min(max(gi, 0), nx - 1)) * 4u;
        ocw[m] = (jj + 1) * FOX + ii + 1;
        if (gi == 0 || gi == nx - 1 || gj == 0 || gj == ny - 1) xyface |= 1 << m;
    }
#pragma unroll
    for (int m = 0; m < FO_PT; ++m) {
        const int e = tid + FNT * m;
        const int ii = e % FOX, jj = e / FOX;
        const int gi = i0 - 4 + ii, gj = j0 - 3 + jj;
        offo[m] = (unsigned)(min(max(gj, 0), ny - 1) * nx + min(max(gi, 0), nx - 1)) * 4u;
    }
    float rw[FP_PT], ro[FO_PT];
    auto fetch_w = [&](int k) {
        const int b = k * plane_bytes;
#pragma unroll
        for (int m = 0; m < FP_PT; ++m)
            rw[m] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(wg, (int)offw[m], b, 0));
    };
    auto fetch_o = [&](int k) {
        const int b = k * plane_bytes;
#pragma unroll
        for (int m = 0; m < FO_PT; ++m)
            ro[m] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(og, (int)offo[m], b, 0));
    };
    auto store_w = [&](float* dst) {
#pragma unroll
        for (int m = 0; m < FP_PT; ++m)
            if (m < FP_PT - 1 || tid + FNT * m < FP_N) dst[tid + FNT * m] = rw[m];
    };
    auto store_o = [&](float* dst) {
#pragma unroll
        for (int m = 0; m < FO_PT; ++m)
            if (m < FO_PT - 1 || tid + FNT * m < FO_N) dst[tid + FNT * m] = ro[m];
    };

    // the cells this thread outputs, and their place in a P / wrk plane
    const int i = i0 - 1 + tx, ja = j0 + 2 * ty, jb = ja + 1;
    const int pa = (2 * ty + 2) * FPX + tx + 2, pb = pa + FPX;
    const bool xin = tx >= 1 && i < nx;
    const bool oka = xin && ja < ny, okb = xin && jb < ny;
    const bool xint = i > 0 && i < nx - 1;

    // own columns: after stage s, Pc*[4] = P(s), ..., Pc*[0] = P(s-4);
    // Wc*[2] = wrk(s-1), Wc*[1] = wrk(s-2), Wc*[0] = wrk(s-3)
    float Pca[5] = {0, 0, 0, 0, 0}, Pcb[5] = {0, 0, 0, 0, 0};
    float Wca[3] = {0, 0, 0}, Wcb[3] = {0, 0, 0};
    float fza_prev = 0.0f, fzb_prev = 0.0f;   // flux above the last plane output
    int fz_plane = -1;

    // output plane k from the slot of its wrk / P plane; columns hold
    // planes k-2 .. k+2 (P) and k-1 .. k+1 (wrk)
    auto output = [&](int k, const float* wk, const float* pk) {
        const size_t cpl = (size_t)k * plane + i;
        if (k == 0 || k == nz - 1) {    // z faces keep w
            if (oka) out[cpl + (size_t)ja * nx] = Wca[1];
            if (okb) out[cpl + (size_t)jb * nx] = Wcb[1];
            return;
        }
        // x: the flux at the face above this lane's cell; the face below
        // comes from the lane below
        auto Wxa = [&](int o) { return wk[pa + o]; };
        auto Pxa = [&](int o) { return pk[pa + o]; };
        auto Wxb = [&](int o) { return wk[pb + o]; };
        auto Pxb = [&](int o) { return pk[pb + o]; };
        const float fxa = fct_flux(Wxa, Pxa, 0, i, nx, chi);
        const float fxb = fct_flux(Wxb, Pxb, 0, i, nx, chi);
        const float fxma = __shfl_up(fxa, 1, 64);
        const float fxmb = __shfl_up(fxb, 1, 64);
        // y: the faces below row a, between the rows and above row b
        auto Wy = [&](int o) { return wk[pa + o * FPX]; };
        auto Py = [&](int o) { return pk[pa + o * FPX]; };
        const float fy0 = fct_flux(Wy, Py, -1, ja - 1, ny, chi);
        const float fy1 = fct_flux(Wy, Py, 0, ja, ny, chi);
        const float fy2 = fct_flux(Wy, Py, 1, jb, ny, chi);
        // z, from the columns: the face above (no end shift for
        // 1 <= k <= nz-2) and the face below
        const bool top = (k + 1 == nz - 1);
        const float fza = fct_lim(Wca[2] - Wca[1], Pca[2] - Pca[1],
                                  top ? (Pca[3] - Pca[2]) : (Pca[4] - Pca[3]), chi);
        const float fzb = fct_lim(Wcb[2] - Wcb[1], Pcb[2] - Pcb[1],
                                  top ? (Pcb[3] - Pcb[2]) : (Pcb[4] - Pcb[3]), chi);
        float fzma, fzmb;
        if (fz_plane == k - 1) {
            fzma = fza_prev;
            fzmb = fzb_prev;
        } else if (k == 1) {   // m = 0: shifted to the face above plane 1
            const bool top1 = (nz == 3);
            fzma = fct_lim(Wca[2] - Wca[1], Pca[2] - Pca[1],
                           top1 ? (Pca[3] - Pca[2]) : (Pca[4] - Pca[3]), chi);
            fzmb = fct_lim(Wcb[2] - Wcb[1], Pcb[2] - Pcb[1],
                           top1 ? (Pcb[3] - Pcb[2]) : (Pcb[4] - Pcb[3]), chi);
        } else {               // the face above plane k - 1
            fzma = fct_lim(Wca[1] - Wca[0], Pca[1] - Pca[0], Pca[3] - Pca[2], chi);
            fzmb = fct_lim(Wcb[1] - Wcb[0], Pcb[1] - Pcb[0], Pcb[3] - Pcb[2], chi);
        }
        fza_prev = fza;
        fzb_prev = fzb;
        fz_plane = k;
        const float txa = xint ? (Pca[2] - fxa + fxma) : Wca[1];
        const float txb = xint ? (Pcb[2] - fxb + fxmb) : Wcb[1];
        const float tya = (ja > 0 && ja < ny - 1) ? (txa - fy1 + fy0) : txa;
        const float tyb = (jb > 0 && jb < ny - 1) ? (txb - fy2 + fy1) : txb;
        if (oka) out[cpl + (size_t)ja * nx] = tya - fza + fzma;
        if (okb) out[cpl + (size_t)jb * nx] = tyb - fzb + fzmb;
    };
    // shift the columns by one plane: P(s) and wrk(s-1) come in
    auto shift = [&](float pa_new, float pb_new, float wa_new, float wb_new) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            Pca[q] = Pca[q + 1];
            Pcb[q] = Pcb[q + 1];
        }
        Pca[4] = pa_new;
        Pcb[4] = pb_new;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            Wca[q] = Wca[q + 1];
            Wcb[q] = Wcb[q + 1];
        }
        Wca[2] = wa_new;
        Wcb[2] = wb_new;
    };

    // P planes pstart..pend are made in order.  Plane p's wrk, P and old
    // live in slot (p - pstart) % 3; during stage s, S0 / S1 / S2 are the
    // slots of planes s / s-1 / s-2 (old(s+1) goes to S2).  Stage s stores
    // wrk(s) and old(s+1) (fetched during stage s-1), makes P(s), starts the
    // loads of the next plane and outputs plane s-2.  The march is unrolled
    // three stages at a time, so every slot is a compile-time constant and
    // each LDS access is a fixed offset from a per-thread base.
    const int pstart = max(0, kb - 2), pend = min(nz - 1, ke + 1);
    if (pstart >= 1) {
        fetch_o(pstart - 1);
        store_o(L.o[2]);
    }
    fetch_o(pstart);
    store_o(L.o[0]);
    fetch_w(pstart);
    if (pstart + 1 <= nz - 1) fetch_o(pstart + 1);
    auto stage = [&](auto rc, int s) {
        constexpr int S0 = decltype(rc)::value, S1 = (S0 + 2) % 3, S2 = (S0 + 1) % 3;
        __syncthreads();   // readers of the slots reused below are done
        store_w(L.w[S0]);
        if (s + 1 <= nz - 1) store_o(L.o[S2]);
        __syncthreads();
        {   // P(s) = w + chi * Lap7(old) inside, w on the box faces
            // (code/sub_fcsmooth.f90:44-83)
            const float* om = L.o[S1];
            const float* oc = L.o[S0];
            const float* op = L.o[S2];
            float* dst = L.p[S0];
            const bool zface = (s == 0 || s == nz - 1);
#pragma unroll
            for (int m = 0; m < FP_PT; ++m) {
                if (m == FP_PT - 1 && tid + FNT * m >= FP_N) break;
                const float wc = rw[m];
                const int o0 = ocw[m];
                const float vi = wc + chi * (oc[o0 + 1] + oc[o0 - 1] + oc[o0 + FOX] +
                                             oc[o0 - FOX] + op[o0] + om[o0] - 6.0f * oc[o0]);
                dst[tid + FNT * m] = (zface || ((xyface >> m) & 1)) ? wc : vi;
            }
        }
        if (s + 1 <= pend) fetch_w(s + 1);   // in flight during the output
        if (s + 2 <= nz - 1) fetch_o(s + 2);
        __syncthreads();
        shift(L.p[S0][pa], L.p[S0][pb], L.w[S1][pa], L.w[S1][pb]);
        const int k = s - 2;
        if (k >= kb && k < ke) output(k, L.w[S2], L.p[S2]);
    };
    // planes past the last P made (only at the top of the box: plane nz-2
    // needs P up to nz-1, plane nz-1 is a z face)
    auto tail = [&](auto rc, int s) {
        constexpr int S0 = decltype(rc)::value, S1 = (S0 + 2) % 3, S2 = (S0 + 1) % 3;
        (void)S0;
        shift(0.0f, 0.0f, L.w[S1][pa], L.w[S1][pb]);
        const int k = s - 2;
        if (k >= kb) output(k, L.w[S2], L.p[S2]);
    };
    using R0 = std::integral_constant<int, 0>;
    using R1 = std::integral_constant<int, 1>;
    using R2 = std::integral_constant<int, 2>;
    for (int s = pstart; s <= pend; s += 3) {
        stage(R0{}, s);
        if (s + 1 <= pend) stage(R1{}, s + 1);
        if (s + 2 <= pend) stage(R2{}, s + 2);
    }
    for (int s = pend + 1; s - 2 < ke; ++s) {
        const int r = (s - pstart) % 3;
        if (r == 0) tail(R0{}, s);
        else if (r == 1) tail(R1{}, s);
        else tail(R2{}, s);
    }
}


// ---------------------------------------------------------------------------
// flanks from a precomputed shell table: one thread per (shell cell, field)
int shell_count(const Box& B) {
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    return 4 * nx * ny + 4 * nx * (nz - 4) + 4 * (ny - 4) * (nz - 4);
}

void build_shell(const Box& B, const FlankGeom& G, ShellPt* out) {
    const int nx = B.nx, ny = B.ny, nz = B.nz, Y = nx, Z = nx * ny;
    int n = 0;
    auto add = [&](int i, int j, int k) {
        ShellPt p;
        p.c = B.IX(i, j, k);
        const float ax = G.sxmin + G.sx * (float)i;
        const float ai = 1.0f + (ax - G.bxmin) / G.delx;
        const int ib = (int)ai;
        const float ay = G.symin + G.sy * (float)j;
        const float aj = 1.0f + (ay - G.bymin) / G.dely;
        const int jb = (int)aj;
        const float az = G.szmin + G.sz * (float)k;
        const float ak = 1.0f + (az - G.bzmin) / G.delz;
        const int kb = (int)ak;
        p.dx = ai - (float)ib;
        p.dy = aj - (float)jb;
        p.dz = ak - (float)kb;
        p.cb = B.IX(ib - 1, jb - 1, kb - 1);
        int f = 0;
        if (k <= 1) f |= 1 | (k == 1 ? 1 << 8 : 0);
        if (k >= nz - 2) f |= 2 | (k == nz - 2 ? 2 << 8 : 0);
        if (j <= 1) f |= 4 | (j == 1 ? 4 << 8 : 0);
        if (j >= ny - 2) f |= 8 | (j == ny - 2 ? 8 << 8 : 0);
        if (i <= 1) f |= 16 | (i == 1 ? 16 << 8 : 0);
        if (i >= nx - 2) f |= 32 | (i == nx - 2 ? 32 << 8 : 0);
        p.flags = f;
        p.sy = Y;
        p.sz = Z;
        out[n++] = p;
    };
    const int ks[4] = {0, 1, nz - 2, nz - 1};
    const int js[4] = {0, 1, ny - 2, ny - 1};
    const int is[4] = {0, 1, nx - 2, nx - 1};
    for (int s = 0; s < 4; ++s)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) add(i, j, ks[s]);
    for (int k = 2; k <= nz - 3; ++k)
        for (int s = 0; s < 4; ++s)
            for (int i = 0; i < nx; ++i) add(i, js[s], k);
    // the x faces: the four cells of a row (i = 0, 1, nx-2, nx-1) are
    // adjacent, so the two cache lines of the row are fetched once
    for (int k = 2; k <= nz - 3; ++k)
        for (int j = 2; j <= ny - 3; ++j)
            for (int s = 0; s < 4; ++s) add(is[s], j, k);
}

// The face a shell cell belongs to, in build_shell's order: the z faces
// take every cell of their layers, the y faces the rest of theirs, the x
// faces what remains.
static int shell_face(int flags) {
    return (flags & 1) ? 0 : (flags & 2) ? 1 : (flags & 4) ? 2 : (flags & 8) ? 3 : (flags & 16) ? 4 : 5;
}

bool build_shell_slab(const Box& B, ShellPt* pts, int npts, std::vector<int>& gather) {
    const int nx = B.nx, ny = B.ny, nz = B.nz, big = 1 << 30;
    int lo[6][3], hi[6][3];
    for (int f = 0; f < 6; ++f)
        for (int a = 0; a < 3; ++a) {
            lo[f][a] = big;
            hi[f][a] = -big;
        }
    auto split = [&](int c, int v[3]) {
        v[0] = c % nx;
        v[1] = (c / nx) % ny;
        v[2] = c / (nx * ny);
    };
    for (int n = 0; n < npts; ++n) {
        const int f = shell_face(pts[n].flags);
        int v[3];
        split(pts[n].cb, v);
        for (int a = 0; a < 3; ++a) {
            lo[f][a] = std::min(lo[f][a], v[a]);
            hi[f][a] = std::max(hi[f][a], v[a] + 1);   // the upper trilinear corner
        }
    }
    int off[6], dim[6][3];
    gather.clear();
    for (int f = 0; f < 6; ++f) {
        off[f] = (int)gather.size();
        if (lo[f][0] > hi[f][0]) continue;              // no cell on this face
        if (lo[f][0] < 0 || lo[f][1] < 0 || lo[f][2] < 0 || hi[f][0] >= nx || hi[f][1] >= ny ||
            hi[f][2] >= nz)
            return false;
        for (int a = 0; a < 3; ++a) dim[f][a] = hi[f][a] - lo[f][a] + 1;
        for (int k = lo[f][2]; k <= hi[f][2]; ++k)
            for (int j = lo[f][1]; j <= hi[f][1]; ++j)
                for (int i = lo[f][0]; i <= hi[f][0]; ++i) gather.push_back(B.IX(i, j, k));
    }
    for (int n = 0; n < npts; ++n) {
        ShellPt& p = pts[n];
        const int f = shell_face(p.flags);
        int v[3];
        split(p.cb, v);
        p.sy = dim[f][0];
        p.sz = dim[f][0] * dim[f][1];
        p.cb = off[f] + (v[0] - lo[f][0]) + p.sy * (v[1] - lo[f][1]) + p.sz * (v[2] - lo[f][2]);
    }
    return true;
}

__global__ void __launch_bounds__(256) k_slab_gather(const int* __restrict__ idx, int n,
                                                     FlankPtrs Fp, float* __restrict__ slab) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const int f = blockIdx.y, c = idx[t];
    float* s = slab + (size_t)2 * f * n;
    const float o = Fp.old[f][c], q = Fp.cur[f][c];
    s[t] = o;
    s[n + t] = q;
}

void launch_slab_gather(const int* idx, int n, const FlankPtrs& Fp, int nf, float* slab,
                        hipStream_t s) {
    if (n <= 0 || nf <= 0) return;
    hipLaunchKernelGGL(k_slab_gather, dim3((n + 255) / 256, nf), dim3(256), 0, s, idx, n, Fp,
                       slab);
}

// Fields per thread: each thread reads its 32-byte table entry once for a
// group of fields (the table is ~3.6 MB at 121x121x61; read once per field
// it was most of this kernel's HBM traffic).  Every load of the group (16
// slab values and the target per field) is issued before the first store:
// on this hardware a wait for a load also waits for every store issued
// before it, so a load-compute-store loop over the fields would serialise
// one memory latency per field.
constexpr int kFlankGroup = 6;

__global__ void __launch_bounds__(256) k_flanks_table(const ShellPt* __restrict__ tab, int npts,
                                                      FlankPtrs Fp, int nf, float t1, float t2) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= npts) return;
    const ShellPt p = tab[t];
    const int X = 1, Y = p.sy, Z = p.sz;
    const float dx = p.dx, dy = p.dy, dz = p.dz;
    const float ddx = 1.0f - dx, ddy = 1.0f - dy, ddz = 1.0f - dz;
    const int c0 = p.cb;
    const int f0 = blockIdx.y * kFlankGroup;
    const int cn[8] = {c0, c0 + Z, c0 + Y, c0 + Y + Z, c0 + X, c0 + X + Z, c0 + X + Y,
                       c0 + X + Y + Z};
    float vo[kFlankGroup][8], vq[kFlankGroup][8], vw[kFlankGroup];
#pragma unroll
    for (int g = 0; g < kFlankGroup; ++g) {
        const int f = f0 + g;
        if (f < nf) {
            const float* __restrict__ o = Fp.old[f];
            const float* __restrict__ q = Fp.cur[f];
#pragma unroll
            for (int n = 0; n < 8; ++n) {
                vo[g][n] = o[cn[n]];
                vq[g][n] = q[cn[n]];
            }
            vw[g] = Fp.tgt[f][p.c];
        }
    }
#pragma unroll
    for (int g = 0; g < kFlankGroup; ++g) {
        const int f = f0 + g;
        if (f >= nf) break;
        float W[8];
#pragma unroll
        for (int n = 0; n < 8; ++n) W[n] = vo[g][n] * t1 + vq[g][n] * t2;
        const float I0 = W[0] * ddx * ddy * ddz + W[1] * ddx * ddy * dz +
                         W[2] * ddx * dy * ddz + W[3] * ddx * dy * dz +
                         W[4] * dx * ddy * ddz + W[5] * dx * ddy * dz +
                         W[6] * dx * dy * ddz + W[7] * dx * dy * dz;
        float w = vw[g];
#pragma unroll
        for (int b = 0; b < 6; ++b) {   // bottom, top, front, back, left, right
            if (p.flags & (1 << b)) {
                const float frac = (p.flags & (256 << b)) ? 0.5f : 0.0f;
                const float frac1 = 1.0f - frac;
                w = frac * w + frac1 * I0;
            }
        }
        Fp.tgt[f][p.c] = w;
    }
}

void launch_flanks_table(const Box& B, const ShellPt* tab, int npts, const FlankPtrs& Fp, int nf,
                         float t1, float t2, hipStream_t s) {
    if (npts <= 0 || nf <= 0) return;
    ProfScope prof_("k_flanks_table", s);
    const int ng = (nf + kFlankGroup - 1) / kFlankGroup;
    // a multiple of 8 blocks per field group: workgroups are dealt to the 8
    // XCDs round-robin, so every field group of a table block runs on the
    // same XCD, whose L2 then holds that block's table entries
    const int nbx = (((npts + 255) / 256) + 7) & ~7;
    hipLaunchKernelGGL(k_flanks_table, dim3(nbx, ng), dim3(256), 0, s, tab, npts,
                       Fp, nf, t1, t2);
}


// ---------------------------------------------------------------------------
// The stage kernels: one-run launches take the logical block from the
// XCD-aware order; the multi-run forms (k_*_runs) find their run and its
// logical block in the launch's concatenation of the runs' blocks.
thread_local Recorder* g_rec = nullptr;

template <bool P2>
__global__ void __launch_bounds__(256) k_efield(EParams P) {
    efield_block<P2>(P, lblock());
}
template <bool P2>
__global__ void __launch_bounds__(256) k_efield_runs(const EParams* __restrict__ A, MultiRun M) {
    int r, lb;
    if (multi_block(M, r, lb)) efield_block<P2>(A[r], lb);
}

template <bool P2>
__global__ void __launch_bounds__(MTX * MTY) k_push_march(PushParams P, int ntx, int nty, int nkc) {
    push_march_block<P2>(P, ntx, nty, nkc, lblock());
}
template <bool P2>
__global__ void __launch_bounds__(MTX * MTY) k_push_march_runs(const MarchArgs* __restrict__ A,
                                                               MultiRun M) {
    int r, lb;
    if (!multi_block(M, r, lb)) return;
    const MarchArgs& a = A[r];
    push_march_block<P2>(a.P, a.ntx, a.nty, a.nkc, lb);
}

template <bool P2>
__global__ void __launch_bounds__(256) k_push(PushParams P) {
    push_block<P2>(P, lblock());
}
template <bool P2>
__global__ void __launch_bounds__(256) k_push_runs(const PushParams* __restrict__ A, MultiRun M) {
    int r, lb;
    if (multi_block(M, r, lb)) push_block<P2>(A[r], lb);
}

__global__ void __launch_bounds__(256) k_validate(ValParams Q) {
    validate_block(Q, lblock(), blockIdx.x);
}
__global__ void __launch_bounds__(256) k_validate_runs(const ValParams* __restrict__ A,
                                                       MultiRun M) {
    int r, lb;
    if (multi_block(M, r, lb)) validate_block(A[r], lb, lb);
}

__global__ void __launch_bounds__(256) k_validate_apply(ValParams Q, float* maxima, int nblocks) {
    validate_apply_block(Q, maxima, nblocks, blockIdx.x);
}
__global__ void __launch_bounds__(256) k_validate_apply_runs(const ApplyArgs* __restrict__ A,
                                                             MultiRun M) {
    int r, lb;
    if (!multi_block(M, r, lb)) return;
    const ApplyArgs& a = A[r];
    validate_apply_block(a.Q, a.maxima, a.nblocks, lb);
}

__global__ void __launch_bounds__(MTX * MTY) k_lap_march(LapParams L, int ntx, int nty, int nkc) {
    lap_march_block(L, ntx, nty, nkc, lblock());
}
__global__ void __launch_bounds__(MTX * MTY) k_lap_march_runs(const LapMarchArgs* __restrict__ A,
                                                              MultiRun M) {
    int r, lb;
    if (!multi_block(M, r, lb)) return;
    const LapMarchArgs& a = A[r];
    lap_march_block(a.L, a.ntx, a.nty, a.nkc, lb);
}

__global__ void __launch_bounds__(256) k_lap(LapParams L) { lap_block(L, lblock()); }
__global__ void __launch_bounds__(256) k_lap_runs(const LapParams* __restrict__ A, MultiRun M) {
    int r, lb;
    if (multi_block(M, r, lb)) lap_block(A[r], lb);
}

// neighbouring tiles on one XCD, unless tuned otherwise
__global__ void __launch_bounds__(FNT, 4) k_fct_fused(FctParams F, int ntx, int nty, int nkc) {
    fct_block(F, ntx, nty, nkc, F.no_remap ? (int)blockIdx.x : lblock());
}
__global__ void __launch_bounds__(FNT, 4) k_fct_fused_runs(const FctArgs* __restrict__ A,
                                                           MultiRun M) {
    int r, lb;
    if (!multi_block(M, r, lb)) return;
    const FctArgs& a = A[r];
    fct_block(a.F, a.ntx, a.nty, a.nkc, lb);
}

size_t kernel_args_size(int id) {
    switch (id) {
    case kKEfield:
    case kKEfieldP2: return sizeof(EParams);
    case kKPushMarch:
    case kKPushMarchP2: return sizeof(MarchArgs);
    case kKPush:
    case kKPushP2: return sizeof(PushParams);
    case kKValidate: return sizeof(ValParams);
    case kKValidateApply: return sizeof(ApplyArgs);
    case kKLapMarch: return sizeof(LapMarchArgs);
    case kKLap: return sizeof(LapParams);
    case kKFct: return sizeof(FctArgs);
    case kKOuter: return sizeof(OuterArgs);
    case kKInner: return sizeof(InnerArgs);
    case kKNextDt: return sizeof(NextDtArgs);
    default: return 0;
    }
}

void launch_multi(int id, const void* args, const MultiRun& M, dim3 block, unsigned gy,
                  hipStream_t s) {
    const dim3 g((unsigned)((M.total + 7) & ~7), gy);
    switch (id) {
    case kKEfield:
        hipLaunchKernelGGL(k_efield_runs<false>, g, block, 0, s, (const EParams*)args, M);
        break;
    case kKEfieldP2:
        hipLaunchKernelGGL(k_efield_runs<true>, g, block, 0, s, (const EParams*)args, M);
        break;
    case kKPushMarch:
        hipLaunchKernelGGL(k_push_march_runs<false>, g, block, 0, s, (const MarchArgs*)args, M);
        break;
    case kKPushMarchP2:
        hipLaunchKernelGGL(k_push_march_runs<true>, g, block, 0, s, (const MarchArgs*)args, M);
        break;
    case kKPush:
        hipLaunchKernelGGL(k_push_runs<false>, g, block, 0, s, (const PushParams*)args, M);
        break;
    case kKPushP2:
        hipLaunchKernelGGL(k_push_runs<true>, g, block, 0, s, (const PushParams*)args, M);
        break;
    case kKValidate:
        hipLaunchKernelGGL(k_validate_runs, g, block, 0, s, (const ValParams*)args, M);
        break;
    case kKValidateApply:
        hipLaunchKernelGGL(k_validate_apply_runs, g, block, 0, s, (const ApplyArgs*)args, M);
        break;
    case kKLapMarch:
        hipLaunchKernelGGL(k_lap_march_runs, g, block, 0, s, (const LapMarchArgs*)args, M);
        break;
    case kKLap:
        hipLaunchKernelGGL(k_lap_runs, g, block, 0, s, (const LapParams*)args, M);
        break;
    case kKFct:
        hipLaunchKernelGGL(k_fct_fused_runs, g, block, 0, s, (const FctArgs*)args, M);
        break;
    default: launch_multi_bounds(id, args, M, block, gy, s); break;
    }
}

// ---------------------------------------------------------------------------
// host launchers
void launch_fct_fused(const FctParams& F, int nf, hipStream_t s) {
    FctParams G = F;
    G.nf = nf;
    if (G.kc_len <= 0) G.kc_len = FKC;
    const int ntx = (F.B.nx + FTX - 1) / FTX, nty = (F.B.ny + FTY - 1) / FTY;
    const int nkc = (F.B.nz + G.kc_len - 1) / G.kc_len;
    const int nb = ntx * nty * nkc * nf;
    if (g_rec) {
        g_rec->add(kKFct, dim3((nb + 7) & ~7), dim3(64, FTY / 2), FctArgs{G, ntx, nty, nkc});
        return;
    }
    ProfScope prof_("k_fct_fused", s);
    hipLaunchKernelGGL(k_fct_fused, dim3((nb + 7) & ~7), dim3(64, FTY / 2), 0, s, G, ntx, nty, nkc);
}

void launch_efield(const EParams& P, hipStream_t s) {
    if (g_rec) {
        g_rec->add(P.pow2 ? kKEfieldP2 : kKEfield, grid3(P.B, 1), block3(), P);
        return;
    }
    if (P.pow2)
        hipLaunchKernelGGL(k_efield<true>, grid3(P.B, 1), block3(), 0, s, P);
    else
        hipLaunchKernelGGL(k_efield<false>, grid3(P.B, 1), block3(), 0, s, P);
}

static void launch_march(const PushParams& P, hipStream_t s) {
    const int kcl = P.march_kc > 0 ? P.march_kc : MKC;
    const int ntx = (P.B.nx - 2 + MOX - 1) / MOX, nty = (P.B.ny - 2 + MTY - 1) / MTY;
    const int nkc = (P.B.nz - 2 + kcl - 1) / kcl, nb = P.nspecies * ntx * nty * nkc;
    const dim3 g((nb + 7) & ~7), b(MTX, MTY);
    if (g_rec) {
        g_rec->add(P.pow2 ? kKPushMarchP2 : kKPushMarch, g, b, MarchArgs{P, ntx, nty, nkc});
        return;
    }
    if (P.pow2)
        hipLaunchKernelGGL(k_push_march<true>, g, b, 0, s, P, ntx, nty, nkc);
    else
        hipLaunchKernelGGL(k_push_march<false>, g, b, 0, s, P, ntx, nty, nkc);
}

void launch_push_species(const PushParams& P, hipStream_t s) { launch_march(P, s); }

void launch_push(const PushParams& P, hipStream_t s) {
    ProfScope prof_("k_push", s);
    if (P.march) launch_march(P, s);
    const int groups = P.march ? 1 : 4;
    if (g_rec) {
        g_rec->add(P.pow2 ? kKPushP2 : kKPush, grid3(P.B, groups), block3(), P);
        return;
    }
    if (P.pow2)
        hipLaunchKernelGGL(k_push<true>, grid3(P.B, groups), block3(), 0, s, P);
    else
        hipLaunchKernelGGL(k_push<false>, grid3(P.B, groups), block3(), 0, s, P);
}

int validate_blocks(const Box& B) {
    const dim3 g = grid3(B, 1);
    return (int)(g.x * g.y * g.z);
}

void launch_validate(const ValParams& Q, float* maxima, hipStream_t s) {
    const int nb = validate_blocks(Q.B);
    const bool red = Q.partials && maxima;
    if (g_rec) {
        g_rec->add(kKValidate, grid3(Q.B, 1), block3(), Q);
        g_rec->add(kKValidateApply, dim3(kApplyBlocks + (red ? 1 : 0)), dim3(256),
                   ApplyArgs{Q, maxima, nb});
        return;
    }
    {
        ProfScope prof_("k_validate", s);
        hipLaunchKernelGGL(k_validate, grid3(Q.B, 1), block3(), 0, s, Q);
    }
    hipLaunchKernelGGL(k_validate_apply, dim3(kApplyBlocks + (red ? 1 : 0)), dim3(256), 0, s, Q,
                       maxima, nb);
}

void launch_lap(const LapParams& L, hipStream_t s) {
    ProfScope prof_("k_lap", s);
    if (L.march) {
        const int ntx = (L.B.nx - 2 + MOX - 1) / MOX, nty = (L.B.ny - 2 + MTY - 1) / MTY;
        const int nkc = (L.B.nz - 2 + MKC - 1) / MKC, nb = 3 * ntx * nty * nkc;
        if (g_rec)
            g_rec->add(kKLapMarch, dim3((nb + 7) & ~7), dim3(MTX, MTY),
                       LapMarchArgs{L, ntx, nty, nkc});
        else
            hipLaunchKernelGGL(k_lap_march, dim3((nb + 7) & ~7), dim3(MTX, MTY), 0, s, L, ntx, nty,
                           nkc);
    }
    if (g_rec) {
        g_rec->add(kKLap, grid3(L.B, L.march ? 1 : 4), block3(), L);
        return;
    }
    hipLaunchKernelGGL(k_lap, grid3(L.B, L.march ? 1 : 4), block3(), 0, s, L);
}



}  // namespace g3m3
