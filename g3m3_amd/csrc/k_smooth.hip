// Smoothing passes of the box-step:
//   velocities for Lapidus (code/3d_main.f90:2698-2748) and fnd_vtot
//   (code/sub_fnd_vtot.f90:288-303);
//   Lapidus artificial viscosity (code/sub_lap_plasma.f90:37-144,
//   sub_lap_elec.f90:164-195, sub_lap_bfld.f90:217-262);
//   flux-corrected transport (code/sub_fcsmooth.f90:7-242) on up to 34
//   fields per launch.
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
    for (int f = 0; f < nf; ++f) {
        const float* __restrict__ q = L.src[f];
        const float q0 = q[c];
        L.dst[f][c] = q0 + L.chi[f] * (delt * (uxp * (q[c + X] - q0) - uxm * (q0 - q[c - X])) +
                                       delt * (uyp * (q[c + Y] - q0) - uym * (q0 - q[c - Y])) +
                                       delt * (uzp * (q[c + Z] - q0) - uzm * (q0 - q[c - Z])));
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

// ---------------------------------------------------------------------------
// FCT.  For one field with w = unsmoothed (wrk), o = step-start (old):
//   P   = w + chi*lap7(o) on the interior, w on every face (:226-275)
//   dq  = forward differences of w, d1 = of P, last copied (:279-298)
//   fn  = s*max(0, min(s*d1[-1], chi*|dq|, s*d1[+1])) on 2..n-1,
//         ends copied (:300-317)
//   out = P - fnx + fnx[-1] (x interior), then -fny + fny[-1] (y interior),
//         then -fnz + fnz[-1] (z interior); z faces keep w (:319-431).
// Three passes: P into the output array, the three limited fluxes into
// scratch, then the flux divergence in place.

__global__ void k_fct_diffuse(Box B, FctPtrs F, int nf) {
    int i, j, k;
    const int f = blockIdx.y;
    if (f >= nf || !cell_of(B, gtid(), i, j, k)) return;
    const int c = B.IX(i, j, k);
    const float* __restrict__ w = F.wrk[f];
    const float* __restrict__ o = F.old[f];
    if (B.interior(i, j, k)) {
        const int X = 1, Y = B.nx, Z = B.nx * B.ny;
        F.out[f][c] = w[c] + F.chi[f] * (o[c + X] + o[c - X] + o[c + Y] + o[c - Y] + o[c + Z] +
                                         o[c - Z] - 6.0f * o[c]);
    } else {
        F.out[f][c] = w[c];
    }
}

// limited antidiffusive flux at position m along an axis of length n and
// stride st (index m of that axis, 1..n-2 computed, ends copied)
__device__ inline float fct_flux(const float* __restrict__ w, const float* __restrict__ P, int c,
                                 int m, int n, int st, float chi) {
    if (m == 0) {
        c += st;
        m = 1;
    } else if (m == n - 1) {
        c -= st;
        m = n - 2;
    }
    // forward differences at m-1, m, m+1 (the last one copied from n-2)
    const float dq = w[c + st] - w[c];
    const float d1m = P[c] - P[c - st];
    const float d1p = (m + 1 == n - 1) ? (P[c + st] - P[c]) : (P[c + 2 * st] - P[c + st]);
    const float s = copysignf(1.0f, dq);
    return s * fmaxf(0.0f, fminf(fminf(s * d1m, chi * fabsf(dq)), s * d1p));
}

__global__ void k_fct_flux(Box B, FctPtrs F, int nf, float* fnx, float* fny, float* fnz,
                           size_t fstride) {
    int i, j, k;
    const int f = blockIdx.y;
    if (f >= nf || !cell_of(B, gtid(), i, j, k)) return;
    const int c = B.IX(i, j, k);
    const float* __restrict__ w = F.wrk[f];
    const float* __restrict__ P = F.out[f];
    const float chi = F.chi[f];
    const size_t o = (size_t)f * fstride;
    fnx[o + c] = fct_flux(w, P, c, i, B.nx, 1, chi);
    fny[o + c] = fct_flux(w, P, c, j, B.ny, B.nx, chi);
    fnz[o + c] = fct_flux(w, P, c, k, B.nz, B.nx * B.ny, chi);
}

__global__ void k_fct_apply(Box B, FctPtrs F, int nf, const float* __restrict__ fnx,
                            const float* __restrict__ fny, const float* __restrict__ fnz,
                            size_t fstride) {
    int i, j, k;
    const int f = blockIdx.y;
    if (f >= nf || !cell_of(B, gtid(), i, j, k)) return;
    if (k == 0 || k == B.nz - 1) return;          // z faces keep w (= P there)
    const int c = B.IX(i, j, k);
    const size_t o = (size_t)f * fstride;
    const float* gx = fnx + o;
    const float* gy = fny + o;
    const float* gz = fnz + o;
    float t;
    if (i > 0 && i < B.nx - 1) t = F.out[f][c] - gx[c] + gx[c - 1];
    else t = F.wrk[f][c];
    if (j > 0 && j < B.ny - 1) t = t - gy[c] + gy[c - B.nx];
    F.out[f][c] = t - gz[c] + gz[c - B.nx * B.ny];
}

void launch_fct(const Box& B, const FctPtrs& F, int nf, float* scratch, hipStream_t s) {
    const int n = B.V();
    const size_t fs = (size_t)n;
    float* fnx = scratch;
    float* fny = scratch + (size_t)nf * fs;
    float* fnz = scratch + (size_t)2 * nf * fs;
    const dim3 g((n + 255) / 256, nf), t(256);
    {
        ProfScope prof_("k_fct_diffuse", s);
        hipLaunchKernelGGL(k_fct_diffuse, g, t, 0, s, B, F, nf);
    }
    {
        ProfScope prof_("k_fct_flux", s);
        hipLaunchKernelGGL(k_fct_flux, g, t, 0, s, B, F, nf, fnx, fny, fnz, fs);
    }
    {
        ProfScope prof_("k_fct_apply", s);
        hipLaunchKernelGGL(k_fct_apply, g, t, 0, s, B, F, nf, fnx, fny, fnz, fs);
    }
}

size_t fct_scratch_floats(int ncells, int nf) { return (size_t)3 * nf * ncells; }

}  // namespace g3m3
