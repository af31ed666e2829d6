// Smoothing passes of the box-step:
//   velocities for Lapidus (code/3d_main.f90:2698-2748) and fnd_vtot
//   (code/sub_fnd_vtot.f90:288-303);
//   Lapidus artificial viscosity (code/sub_lap_plasma.f90:37-144,
//   sub_lap_elec.f90:164-195, sub_lap_bfld.f90:217-262).
#include "g3m3_dev.h"
#include "g3m3_launch.h"

namespace g3m3 {

// v = p / max(rho, dmin) on every cell
__global__ void k_velocity(int n, const float* __restrict__ r, const float* __restrict__ px,
                           const float* __restrict__ py, const float* __restrict__ pz, float* vx,
                           float* vy, float* vz, float dmin) {
    const int c = gtid();
    if (c >= n) return;
    const float a = fmaxf(r[c], dmin);
    vx[c] = px[c] / a;
    vy[c] = py[c] / a;
    vz[c] = pz[c] / a;
}

void launch_velocity(int n, const float* r, const float* p[3], float* v[3], float dmin,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_velocity, dim3((n + 255) / 256), dim3(256), 0, s, n, r, p[0], p[1], p[2],
                       v[0], v[1], v[2], dmin);
}

// total ion velocity (code/sub_fnd_vtot.f90:288-303)
__global__ void k_vtot(int n, const float* __restrict__ qpx, const float* __restrict__ qpy,
                       const float* __restrict__ qpz, const float* __restrict__ qr,
                       const float* __restrict__ hpx, const float* __restrict__ hpy,
                       const float* __restrict__ hpz, const float* __restrict__ hr,
                       const float* __restrict__ opx, const float* __restrict__ opy,
                       const float* __restrict__ opz, const float* __restrict__ orr, float* vx,
                       float* vy, float* vz) {
    const int c = gtid();
    if (c >= n) return;
    const float tden = qr[c] + hr[c] + orr[c];
    vx[c] = (qpx[c] + hpx[c] + opx[c]) / tden;
    vy[c] = (qpy[c] + hpy[c] + opy[c]) / tden;
    vz[c] = (qpz[c] + hpz[c] + opz[c]) / tden;
}

void launch_vtot(int n, const float* mom[9], const float* rho[3], float* v[3], hipStream_t s) {
    hipLaunchKernelGGL(k_vtot, dim3((n + 255) / 256), dim3(256), 0, s, n, mom[0], mom[1], mom[2],
                       rho[0], mom[3], mom[4], mom[5], rho[1], mom[6], mom[7], mom[8], rho[2],
                       v[0], v[1], v[2]);
}

// Lapidus: w = q + chi*( dt*(|dv+| dq+ - |dv-| dq-) summed over x, y, z )
__global__ void k_lapidus(Box B, LapPtrs L, int nf, const float* __restrict__ vx,
                          const float* __restrict__ vy, const float* __restrict__ vz, float delt) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k) || !B.interior(i, j, k)) return;
    const int c = B.IX(i, j, k), X = 1, Y = B.nx, Z = B.nx * B.ny;
    const float uxp = fabsf(vx[c + X] - vx[c]);
    const float uxm = fabsf(vx[c] - vx[c - X]);
    const float uyp = fabsf(vy[c + Y] - vy[c]);
    const float uym = fabsf(vy[c] - vy[c - Y]);
    const float uzp = fabsf(vz[c + Z] - vz[c]);
    const float uzm = fabsf(vz[c] - vz[c - Z]);
    // five fields at a time, stored after their loads: on this hardware a
    // wait for a load also waits for every store issued before it
    constexpr int G = 5;
    for (int f0 = 0; f0 < nf; f0 += G) {
        float w[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (f0 + u >= nf) break;
            const float* __restrict__ q = L.src[f0 + u];
            const float q0 = q[c];
            w[u] = q0 + L.chi[f0 + u] * (delt * (uxp * (q[c + X] - q0) - uxm * (q0 - q[c - X])) +
                                         delt * (uyp * (q[c + Y] - q0) - uym * (q0 - q[c - Y])) +
                                         delt * (uzp * (q[c + Z] - q0) - uzm * (q0 - q[c - Z])));
        }
#pragma unroll
        for (int u = 0; u < G; ++u)
            if (f0 + u < nf) L.dst[f0 + u][c] = w[u];
    }
}

void launch_lapidus(const Box& B, const LapPtrs& L, int nf, const float* v[3], float delt,
                    hipStream_t s) {
    const int n = B.V();
    {
        ProfScope prof_("k_lapidus", s);
        hipLaunchKernelGGL(k_lapidus, dim3((n + 255) / 256), dim3(256), 0, s, B, L, nf, v[0], v[1],
                       v[2], delt);
    }
}


}  // namespace g3m3
