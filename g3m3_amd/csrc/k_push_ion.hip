// Ion push for one species: code/sub_push_ion.f90:4-362.
// Continuity, momentum (E + tv x B, skin-depth limited Hall term,
// gravity) and the full 6-component pressure tensor with the gyrotropic
// terms, on the interior 2..n-1, from `wrk` (the half-step values) and
// `old` (the step start) into `out`.
#include "g3m3_dev.h"
#include "g3m3_fused.h"
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

// The push itself is the native runtime's k-marching species kernel
// (k_push_march) for this one species: the same arithmetic, expression for
// expression, with |B| read from the caller's btot as the reference reads it.
void launch_push_ion(const Box& B, const IonPtrs& I, float rmass, float delt, float grav,
                     float re_equiv, float reynolds, const float gmin[3], const float gmax[3],
                     hipStream_t s) {
    const int n = B.V();
    // qv is one of the subroutine's outputs
    hipLaunchKernelGGL(k_ion_vel, dim3((n + 255) / 256), dim3(256), 0, s, n, I.wrk[0], I.wrk[4],
                       I.wrk[5], I.wrk[6], I.qv[0], I.qv[1], I.qv[2]);
    PushParams P{};
    P.B = B;
    for (int m = 0; m < 10; ++m) {
        P.src[0].f[m] = const_cast<float*>(I.wrk[m]);
        P.old[0].f[m] = const_cast<float*>(I.old[m]);
        P.out[0].f[m] = I.out[m];
    }
    for (int q = 0; q < 3; ++q) {
        P.bs[q] = I.bs[q];
        P.tv[q] = I.tv[q];
        P.e[q] = I.e[q];
    }
    P.btot = I.btot;
    P.nspecies = 1;
    // one species alone fills the chip only with shorter k runs
    P.march_kc = 8;
    P.march = 1;
    P.rmass[0] = rmass;
    P.rrmass[0] = 1.0f / rmass;
    P.bmax[0] = 0.20f * rmass / delt / reynolds;   // code/sub_push_ion.f90:64
    P.delt = delt;
    P.reynolds = reynolds;
    P.ddx = (gmax[0] - gmin[0]) / (float)(B.nx - 1);
    P.ddy = (gmax[1] - gmin[1]) / (float)(B.ny - 1);
    P.ddz = (gmax[2] - gmin[2]) / (float)(B.nz - 1);
    P.dxt = 2.0f * P.ddx;
    P.dyt = 2.0f * P.ddy;
    P.dzt = 2.0f * P.ddz;
    P.rdxt = 1.0f / P.dxt;
    P.rdyt = 1.0f / P.dyt;
    P.rdzt = 1.0f / P.dzt;
    P.pow2 = (is_pow2(rmass) && is_pow2(P.dxt) && is_pow2(P.dyt) && is_pow2(P.dzt)) ? 1 : 0;
    P.xmin = gmin[0];
    P.ymin = gmin[1];
    P.zmin = gmin[2];
    P.re_equiv = re_equiv;
    P.grav = grav;
    P.dt_scale = 1.0f;
    ProfScope prof_("k_push_ion", s);
    launch_push_species(P, s);
}

}  // namespace g3m3
