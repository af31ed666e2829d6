// Ion push for one species: code/sub_push_ion.f90:4-362.
// Continuity, momentum (E + tv x B, skin-depth limited Hall term,
// gravity) and the full 6-component pressure tensor with the gyrotropic
// terms, on the interior 2..n-1, from `wrk` (the half-step values) and
// `old` (the step start) into `out`.
#include "g3m3_dev.h"
#include "g3m3_launch.h"

namespace g3m3 {

// code/sub_push_ion.f90:67-81  velocities of the input state, all cells
__global__ void k_ion_vel(int n, const float* __restrict__ r, const float* __restrict__ px,
                          const float* __restrict__ py, const float* __restrict__ pz, float* vx,
                          float* vy, float* vz) {
    const int c = gtid();
    if (c >= n) return;
    const float a = fmaxf(r[c], 1.e-6f);
    vx[c] = px[c] / a;
    vy[c] = py[c] / a;
    vz[c] = pz[c] / a;
}

__global__ void k_push_ion(Box B, IonPtrs P, float rmass, float delt, float grav, float re_equiv,
                           float reynolds, float ddx, float ddy, float ddz, float xmin, float ymin,
                           float zmin) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k) || !B.interior(i, j, k)) return;
    const int c = B.IX(i, j, k), X = 1, Y = B.nx, Z = B.nx * B.ny;
    const float dxt = 2.0f * ddx, dyt = 2.0f * ddy, dzt = 2.0f * ddz;
    const float bmax = 0.20f * rmass / delt / reynolds;
    const float* __restrict__ wr = P.wrk[0];
    const float* __restrict__ wx = P.wrk[1];
    const float* __restrict__ wy = P.wrk[2];
    const float* __restrict__ wz = P.wrk[3];
    const float* __restrict__ wpx = P.wrk[4];
    const float* __restrict__ wpy = P.wrk[5];
    const float* __restrict__ wpz = P.wrk[6];
    const float* __restrict__ wxy = P.wrk[7];
    const float* __restrict__ wxz = P.wrk[8];
    const float* __restrict__ wyz = P.wrk[9];
    const float* __restrict__ qvx = P.qv[0];
    const float* __restrict__ qvy = P.qv[1];
    const float* __restrict__ qvz = P.qv[2];

    // outputs are stored after every load: on this hardware a wait for a
    // load also waits for every store issued before it
    float res[10];
    auto flux = [&](const float* __restrict__ w) {
        return ((w[c + X] * qvx[c + X] - w[c - X] * qvx[c - X]) / dxt) +
               ((w[c + Y] * qvy[c + Y] - w[c - Y] * qvy[c - Y]) / dyt) +
               ((w[c + Z] * qvz[c + Z] - w[c - Z] * qvz[c - Z]) / dzt);
    };

    // continuity (:107-110)
    res[0] = P.old[0][c] - delt * (((wpx[c + X] - wpx[c - X]) / dxt) +
                                        ((wpy[c + Y] - wpy[c - Y]) / dyt) +
                                        ((wpz[c + Z] - wpz[c - Z]) / dzt));
    const float aqpx = P.old[4][c] - delt * flux(wpx);
    const float aqpy = P.old[5][c] - delt * flux(wpy);
    const float aqpz = P.old[6][c] - delt * flux(wpz);

    // momentum sources (:138-188)
    float abx = P.bs[0][c], aby = P.bs[1][c], abz = P.bs[2][c];
    const float bmag0 = P.btot[c];
    const float weight = fmaxf(1.0f, bmag0 / bmax);
    const float skin_depth = reynolds / weight;
    const float tx = P.tv[0][c], ty = P.tv[1][c], tz = P.tv[2][c];
    const float vx = qvx[c], vy = qvy[c], vz = qvz[c];
    const float tcx = (ty * abz - tz * aby);
    const float tcy = -(tx * abz - tz * abx);
    const float tcz = (tx * aby - ty * abx);
    const float qcx = (vy * abz - vz * aby);
    const float qcy = -(vx * abz - vz * abx);
    const float qcz = (vx * aby - vy * abx);
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
    float px = aqpx + delt * (qden * (dex + dvx) - gpx);
    float py = aqpy + delt * (qden * (dey + dvy) - gpy);
    float pz = aqpz + delt * (qden * (dez + dvz) - gpz);

    // gravity (:202-206)
    const float ax = xmin + ddx * (float)i;
    const float ay = ymin + ddy * (float)j;
    const float az = zmin + ddz * (float)k;
    const float radius = sqrtf(ax * ax + ay * ay + az * az) + 0.0000001f;
    const float tr = re_equiv * radius;
    const float g = delt * wrc * grav / (tr * tr);
    res[4] = px - g * ax / radius;
    res[5] = py - g * ay / radius;
    res[6] = pz - g * az / radius;

    // diagonal pressures (:210-238)
    const float cx = wx[c], cy = wy[c], cz = wz[c];
    const float cxy = wxy[c], cxz = wxz[c], cyz = wyz[c];
    const float px0 = P.old[1][c] - delt * flux(wx) - 2.0f * delt * cx * (qvx[c + X] - qvx[c - X]) / dxt;
    const float py0 = P.old[2][c] - delt * flux(wy) - 2.0f * delt * cy * (qvy[c + Y] - qvy[c - Y]) / dyt;
    const float pz0 = P.old[3][c] - delt * flux(wz) - 2.0f * delt * cz * (qvz[c + Z] - qvz[c - Z]) / dzt;

    // smoothed B for the gyrotropic terms (:242-254)
    const float* __restrict__ bsx = P.bs[0];
    const float* __restrict__ bsy = P.bs[1];
    const float* __restrict__ bsz = P.bs[2];
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

    res[1] = px0 - t2 * cxy * (qvx[c + Y] - qvx[c - Y]) / dyt -
                  t2 * cxz * (qvx[c + Z] - qvx[c - Z]) / dzt +
                  t2 * sf / rmass * (cxy * abz - cxz * aby);
    res[2] = py0 - t2 * cxy * (qvy[c + X] - qvy[c - X]) / dxt -
                  t2 * cyz * (qvy[c + Z] - qvy[c - Z]) / dzt +
                  t2 * sf / rmass * (cyz * abx - cxy * abz);
    res[3] = pz0 - t2 * cxz * (qvz[c + X] - qvz[c - X]) / dxt -
                  t2 * cyz * (qvz[c + Y] - qvz[c - Y]) / dyt +
                  t2 * sf / rmass * (cxz * aby - cyz * abx);

    // off-diagonal pressures (:285-356)
    const float pxy = P.old[7][c] - delt * flux(wxy);
    res[7] = pxy - delt * cxy * ((qvx[c + X] - qvx[c - X]) / dxt + (qvy[c + Y] - qvy[c - Y]) / dyt) -
                  delt * cy * (qvx[c + Y] - qvx[c - Y]) / dyt -
                  delt * cx * (qvy[c + X] - qvy[c - X]) / dxt -
                  delt * cyz * (qvx[c + Z] - qvx[c - Z]) / dzt -
                  delt * cxz * (qvy[c + Z] - qvy[c - Z]) / dzt +
                  delt * sf / rmass * (abz * (cy - cx) - aby * cyz + abx * cxz);
    const float pxz = P.old[8][c] - delt * flux(wxz);
    res[8] = pxz - delt * cxz * ((qvx[c + X] - qvx[c - X]) / dxt + (qvz[c + Z] - qvz[c - Z]) / dzt) -
                  delt * cz * (qvx[c + Z] - qvx[c - Z]) / dzt -
                  delt * cx * (qvz[c + X] - qvz[c - X]) / dxt -
                  delt * cyz * (qvx[c + Y] - qvx[c - Y]) / dyt -
                  delt * cxy * (qvz[c + Y] - qvz[c - Y]) / dyt +
                  delt * sf / rmass * (aby * (cx - cz) - abx * cxy + abz * cyz);
    const float pyz = P.old[9][c] - delt * flux(wyz);
    res[9] = pyz - delt * cyz * ((qvy[c + Y] - qvy[c - Y]) / dyt + (qvz[c + Z] - qvz[c - Z]) / dzt) -
                  delt * cy * (qvz[c + Y] - qvz[c - Y]) / dyt -
                  delt * cz * (qvy[c + Z] - qvy[c - Z]) / dzt -
                  delt * cxy * (qvz[c + X] - qvz[c - X]) / dxt -
                  delt * cxz * (qvy[c + X] - qvy[c - X]) / dxt +
                  delt * sf / rmass * (abx * (cz - cy) - abz * cxz + aby * cxy);
#pragma unroll
    for (int m = 0; m < 10; ++m) P.out[m][c] = res[m];
}

void launch_push_ion(const Box& B, const IonPtrs& P, float rmass, float delt, float grav,
                     float re_equiv, float reynolds, const float gmin[3], const float gmax[3],
                     hipStream_t s) {
    const int n = B.V();
    const dim3 g((n + 255) / 256), t(256);
    hipLaunchKernelGGL(k_ion_vel, g, t, 0, s, n, P.wrk[0], P.wrk[4], P.wrk[5], P.wrk[6], P.qv[0],
                       P.qv[1], P.qv[2]);
    const float ddx = (gmax[0] - gmin[0]) / (float)(B.nx - 1);
    const float ddy = (gmax[1] - gmin[1]) / (float)(B.ny - 1);
    const float ddz = (gmax[2] - gmin[2]) / (float)(B.nz - 1);
    {
        ProfScope prof_("k_push_ion", s);
        hipLaunchKernelGGL(k_push_ion, g, t, 0, s, B, P, rmass, delt, grav, re_equiv, reynolds, ddx,
                       ddy, ddz, gmin[0], gmin[1], gmin[2]);
    }
}

}  // namespace g3m3
