// Adjoints of the box-step pieces that couple the plasma to the field and
// to the electron pressure: fnd_evel's ion and electron velocities
// (code/sub_fnd_evel.f90:4-57), BandE's pressure-gradient and resistive
// terms (code/sub_BandE.f90:55-84 and the uniform resistivity), the total-B
// magnitude (code/3d_main.f90:2247-2255), push_elec (code/sub_push_elec.f90:
// 4-53), and bndry_outer on every field kind (code/sub_bndry_outer.f90:
// 50-728) with the upstream values as controls.  Same conventions as
// k_adjoint.hip: transposes of the forward kernels' linearisations,
// gathers where a stencil is read, adjoints accumulated.
#include "g3m3_dev.h"
#include "g3m3_launch.h"

namespace g3m3 {

namespace {

__device__ inline double full_block_sum(double v) {
    __shared__ double red[256];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

}  // namespace

// BandE's body, all of it: E = -(ev x bs) everywhere; on the interior
// - grad pe / (n Re) + eta J / n, n = q/mq + h/mh + o/mo.  From E_bar
// (face copies already transposed): ev_bar = E_bar x bs (assigned), bs_bar,
// J_bar (eta term), eta_bar (per-block double sums), rho bars through n,
// and g[d] = -E_bar_d / (dd_d n Re), the coefficient of pe's centred
// difference, for the gather of pe's adjoint.
__global__ void k_adj_bande_full(EfAdj A) {
    int i, j, k;
    double et = 0.0;
    if (cell_of(A.B, gtid(), i, j, k)) {
        const Box& B = A.B;
        const int c = B.IX(i, j, k), n = B.V();
        const float Ex = A.eb[0][c], Ey = A.eb[1][c], Ez = A.eb[2][c];
        const float bx = A.bs[0][c], by = A.bs[1][c], bz = A.bs[2][c];
        const float vx = A.ev[0][c], vy = A.ev[1][c], vz = A.ev[2][c];
        A.evb[0][c] = Ey * bz - Ez * by;
        A.evb[1][c] = Ez * bx - Ex * bz;
        A.evb[2][c] = Ex * by - Ey * bx;
        A.bsb[0][c] += vy * Ez - vz * Ey;
        A.bsb[1][c] += vz * Ex - vx * Ez;
        A.bsb[2][c] += vx * Ey - vy * Ex;
        float g[3] = {0.0f, 0.0f, 0.0f};
        if (B.interior(i, j, k)) {
            const int st[3] = {1, B.nx, B.nx * B.ny};
            const float dd[3] = {A.dxt, A.dyt, A.dzt};
            const float eden = A.rho[0][c] / A.rmq + A.rho[1][c] / A.rmh + A.rho[2][c] / A.rmo;
            const float arho = eden * A.reynolds;
            const float Eb[3] = {Ex, Ey, Ez};
            float edenB = 0.0f, arhoB = 0.0f;
            double e2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                const float J = A.cur[d][c];
                const float grad = (A.ep[c + st[d]] - A.ep[c - st[d]]) / dd[d];
                arhoB += Eb[d] * grad / (arho * arho);
                g[d] = -Eb[d] / (dd[d] * arho);
                A.curb[d][c] += A.eta * Eb[d] / eden;
                edenB -= Eb[d] * A.eta * J / (eden * eden);
                e2 += (double)Eb[d] * J;
            }
            et = e2 / eden;
            edenB += arhoB * A.reynolds;
            A.rhob[0][c] += edenB / A.rmq;
            A.rhob[1][c] += edenB / A.rmh;
            A.rhob[2][c] += edenB / A.rmo;
        }
        for (int d = 0; d < 3; ++d) A.scr[(size_t)d * n + c] = g[d];
    }
    const double s = full_block_sum(et);
    if (threadIdx.x == 0) atomicAdd(A.etab, s);
}

// pe's adjoint from the gradient terms: ep_bar(m) += sum_d g_d(m-d) - g_d(m+d)
__global__ void k_adj_bande_ep(EfAdj A) {
    int i, j, k;
    if (!cell_of(A.B, gtid(), i, j, k)) return;
    const Box& B = A.B;
    const int m = B.IX(i, j, k), n = B.V();
    const int st[3] = {1, B.nx, B.nx * B.ny}, pos[3] = {i, j, k}, len[3] = {B.nx, B.ny, B.nz};
    float acc = 0.0f;
    for (int d = 0; d < 3; ++d) {
        const float* g = A.scr + (size_t)d * n;
        if (pos[d] - 1 >= 0) acc += g[m - st[d]];
        if (pos[d] + 1 < len[d]) acc -= g[m + st[d]];
    }
    A.epb[m] += acc;
}

// fnd_evel: ev = tv - J / tot / Re, tot = sum (rho_s + 1e-6) / m_s, with
// tv = sum p / sum rho.  From ev_bar and the caller's tv_bar (tv's other
// uses): J_bar, and the plasma's adjoints through tot and tv.
__global__ void k_adj_evel_full(EfAdj A) {
    const int c = gtid();
    if (c >= A.B.V()) return;
    const float q = A.rho[0][c], h = A.rho[1][c], o = A.rho[2][c];
    const float tot = (q + 0.000001f) / A.rmq + (h + 0.000001f) / A.rmh +
                      (o + 0.000001f) / A.rmo;
    const float tden = q + h + o;
    float totB = 0.0f, tdenB = 0.0f;
    for (int d = 0; d < 3; ++d) {
        const float evb = A.evb[d][c];
        const float J = A.cur[d][c];
        A.curb[d][c] -= evb / tot / A.reynolds;
        totB += evb * J / (tot * tot * A.reynolds);
        const float tvb = A.tvb[d][c] + evb;
        const float tv = (A.mom[d][c] + A.mom[3 + d][c] + A.mom[6 + d][c]) / tden;
        for (int s = 0; s < 3; ++s) A.momb[3 * s + d][c] += tvb / tden;
        tdenB -= tvb * tv / tden;
    }
    A.rhob[0][c] += totB / A.rmq + tdenB;
    A.rhob[1][c] += totB / A.rmh + tdenB;
    A.rhob[2][c] += totB / A.rmo + tdenB;
}

// btot = max(|bs|, smallbit): bs_bar += btot_bar bs / |bs| where |bs| > smallbit
__global__ void k_adj_btot(int n, const float* __restrict__ bx, const float* __restrict__ by,
                           const float* __restrict__ bz, const float* __restrict__ btb,
                           float smallbit, float* bxb, float* byb, float* bzb) {
    const int c = gtid();
    if (c >= n) return;
    const float x = bx[c], y = by[c], z = bz[c];
    const float m = sqrtf(x * x + y * y + z * z);
    if (m > smallbit) {
        const float g = btb[c] / m;
        bxb[c] += g * x;
        byb[c] += g * y;
        bzb[c] += g * z;
    }
}

// v = sum_s p_s / sum_s rho_s (fnd_vtot, and fnd_evel's tv), transposed
__global__ void k_adj_vtot(int n, VtotAdj A) {
    const int c = gtid();
    if (c >= n) return;
    const float tden = A.rho[0][c] + A.rho[1][c] + A.rho[2][c];
    float tdenB = 0.0f;
    for (int d = 0; d < 3; ++d) {
        const float vb = A.vb[d][c];
        const float v = (A.mom[d][c] + A.mom[3 + d][c] + A.mom[6 + d][c]) / tden;
        for (int s = 0; s < 3; ++s) A.momb[3 * s + d][c] += vb / tden;
        tdenB -= vb * v / tden;
    }
    for (int s = 0; s < 3; ++s) A.rhob[s][c] += tdenB;
}

// push_elec on interior cells c:
//   pe = old - dt g (sum_d (w ev_d)[c+d] - (w ev_d)[c-d]) / dd_d)
//        + dt g1 (ev . grad w)[c],  grad_d w = (w[c+d] - w[c-d]) / dd_d
// gathered at every cell m: w_bar, ev_bar and old_bar.
__global__ void k_adj_push_elec(PeAdj A) {
    int i, j, k;
    if (!cell_of(A.B, gtid(), i, j, k)) return;
    const Box& B = A.B;
    const int m = B.IX(i, j, k);
    const int st[3] = {1, B.nx, B.nx * B.ny};
    const float dd[3] = {A.dxt, A.dyt, A.dzt};
    const float f1 = A.delt * A.gamma, f2 = A.delt * A.gamma1;
    auto in = [&](int d, int s) {
        int q[3] = {i, j, k};
        q[d] += s;
        return B.interior(q[0], q[1], q[2]);
    };
    const float wm = A.w[m];
    float wB = 0.0f;
    if (B.interior(i, j, k)) {
        const float ob = A.ob[m];
        A.oldb[m] += ob;
        for (int d = 0; d < 3; ++d) {      // ev_d at m in the compression term
            const float grad = (A.w[m + st[d]] - A.w[m - st[d]]) / dd[d];
            A.evb[d][m] += f2 * ob * grad;
        }
    }
    for (int d = 0; d < 3; ++d) {
        const float evm = A.ev[d][m];
        float prodB = 0.0f;                // adjoint of (w ev_d)[m]
        if (in(d, -1)) {                   // c = m - d reads (w ev_d)[c + d] and w[c + d]
            const int c = m - st[d];
            prodB -= f1 * A.ob[c] / dd[d];
            wB += f2 * A.ob[c] * A.ev[d][c] / dd[d];
        }
        if (in(d, 1)) {                    // c = m + d reads (w ev_d)[c - d] and w[c - d]
            const int c = m + st[d];
            prodB += f1 * A.ob[c] / dd[d];
            wB -= f2 * A.ob[c] * A.ev[d][c] / dd[d];
        }
        wB += prodB * evm;
        A.evb[d][m] += prodB * wm;
    }
    A.wb[m] += wB;
}

// bndry_outer on any field kind, transposed (bndry_outer's eight passes in
// reverse; within a pass written and read cells are disjoint).  Pass 0 (the
// inflow face) sends its adjoint, times the field's coefficient, to its
// control (sums in double).  blockIdx.y = field.
// element t of pass `pass` (1..7) of field f
__device__ inline void adj_outer_pass(const Box& B, const OuterAdj& O, int f, int pass, int t) {
    float* q = O.f[f];
    const int kind = O.kind[f];
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    auto take = [&](int i, int j, int k) {
        const int c = B.IX(i, j, k);
        const float v = q[c];
        q[c] = 0.0f;
        return v;
    };
    // dst = s * |src| (s = +-1) or dst = src: the source's adjoint
    auto give = [&](int i, int j, int k, float v, int sgn_abs) {
        const int c = B.IX(i, j, k);
        if (sgn_abs != 0) {
            const float x = O.val[f][c];
            v *= (float)sgn_abs * (float)((x > 0.0f) - (x < 0.0f));
        }
        q[c] += v;    // no two threads of a phase give to one cell
    };
    switch (pass) {
    case 1: {
        if (t >= ny * nz) return;
        const int j = t % ny, k = t / ny;
        give(nx - 2, j, k, take(nx - 1, j, k), kind == kPX ? 1 : 0);
        break;
    }
    case 2: {
        if (t >= (nx - 1) * nz) return;
        const int i = 1 + t % (nx - 1), k = t / (nx - 1);
        give(i, 1, k, take(i, 0, k), kind == kPY ? -1 : 0);
        give(i, ny - 2, k, take(i, ny - 1, k), kind == kPY ? 1 : 0);
        break;
    }
    case 3: {
        if (t >= (nx - 1) * ny) return;
        const int i = 1 + t % (nx - 1), j = t / (nx - 1);
        give(i, j, nz - 2, take(i, j, nz - 1), kind == kPZ ? 1 : 0);
        break;
    }
    case 4: {
        if (t >= (nx - 1) * ny) return;
        const int i = 1 + t % (nx - 1), j = t / (nx - 1);
        give(i, j, 1, take(i, j, 0), kind == kPZ ? -1 : 0);
        break;
    }
    case 5: {
        if (t >= nz - 2) return;
        const int k = 1 + t;
        const float lo = take(nx - 1, 0, k), hi = take(nx - 1, ny - 1, k);
        if (kind == kBX) {
            give(nx - 2, 0, k, lo, 0);
            give(nx - 2, ny - 1, k, hi, 0);
        } else {
            give(nx - 2, 0, k, 0.5f * lo, 0);
            give(nx - 1, 1, k, 0.5f * lo, 0);
            give(nx - 2, ny - 1, k, 0.5f * hi, 0);
            give(nx - 1, ny - 2, k, 0.5f * hi, 0);
        }
        break;
    }
    case 6: {
        if (t >= ny - 2) return;
        const int j = 1 + t;
        give(nx - 1, j, 1, take(nx - 1, j, 0), 0);
        if (kind != kEP) give(nx - 2, j, nz - 1, take(nx - 1, j, nz - 1), 0);
        break;
    }
    case 7: {
        if (t >= nx - 1) return;
        const int i = 1 + t;
        give(i, 0, 1, take(i, 0, 0), 0);
        give(i, ny - 1, 1, take(i, ny - 1, 0), 0);
        give(i, 0, nz - 2, take(i, 0, nz - 1), 0);
        give(i, ny - 1, nz - 2, take(i, ny - 1, nz - 1), 0);
        break;
    }
    }
}

// The passes in reverse, in the forward's five phases: 7 and 6, then 5,
// then 4 and 3, then 2, then 1 and 0 (the inflow face, whose adjoint goes
// to its control, summed in double).  The two passes of a phase take and
// give disjoint cells, so they run together and every sum keeps its order.
// blockIdx.y = field.
__global__ void k_adj_outer_phase_k(Box B, OuterAdj O, int phase, double* ctlb) {
    const int f = blockIdx.y;
    const int t = gtid();
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    switch (phase) {
    case 0:
        if (t < nx - 1) adj_outer_pass(B, O, f, 7, t);
        if (t < ny - 2) adj_outer_pass(B, O, f, 6, t);
        break;
    case 1:
        if (t < nz - 2) adj_outer_pass(B, O, f, 5, t);
        break;
    case 2:
        if (t < (nx - 1) * ny) {
            adj_outer_pass(B, O, f, 4, t);
            adj_outer_pass(B, O, f, 3, t);
        }
        break;
    case 3:
        if (t < (nx - 1) * nz) adj_outer_pass(B, O, f, 2, t);
        break;
    default: {
        double v = 0.0;
        if (t < ny * nz) {
            adj_outer_pass(B, O, f, 1, t);
            float* q = O.f[f];
            const int c = B.IX(0, t % ny, t / ny);
            v = q[c];
            q[c] = 0.0f;
        }
        const double s = full_block_sum(v);
        if (threadIdx.x == 0 && O.ctl[f] >= 0) atomicAdd(ctlb + O.ctl[f], s * (double)O.coef[f]);
        break;
    }
    }
}

// --- launches ---------------------------------------------------------------

void launch_adj_bande_plasma(const EfAdj& A, hipStream_t s) {
    const int n = A.B.V();
    const dim3 g((n + 255) / 256), t(256);
    hipLaunchKernelGGL(k_adj_bande_full, g, t, 0, s, A);
    hipLaunchKernelGGL(k_adj_bande_ep, g, t, 0, s, A);
}

void launch_adj_evel_plasma(const EfAdj& A, hipStream_t s) {
    const int n = A.B.V();
    hipLaunchKernelGGL(k_adj_evel_full, dim3((n + 255) / 256), dim3(256), 0, s, A);
}

void launch_adj_btot(int n, const float* b[3], const float* btb, float smallbit, float* bb[3],
                     hipStream_t s) {
    hipLaunchKernelGGL(k_adj_btot, dim3((n + 255) / 256), dim3(256), 0, s, n, b[0], b[1], b[2],
                       btb, smallbit, bb[0], bb[1], bb[2]);
}

void launch_adj_vtot(int n, const VtotAdj& A, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_vtot, dim3((n + 255) / 256), dim3(256), 0, s, n, A);
}

void launch_adj_push_elec(const PeAdj& A, hipStream_t s) {
    const int n = A.B.V();
    hipLaunchKernelGGL(k_adj_push_elec, dim3((n + 255) / 256), dim3(256), 0, s, A);
}

void launch_adj_outer(const Box& B, const OuterAdj& O, int nf, double* ctlb, hipStream_t s) {
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    const int ext[5] = {(ny - 2 > nx - 1) ? ny - 2 : nx - 1, nz - 2, (nx - 1) * ny,
                        (nx - 1) * nz, ny * nz};
    for (int p = 0; p < 5; ++p)
        hipLaunchKernelGGL(k_adj_outer_phase_k, dim3((ext[p] + 255) / 256, nf), dim3(256), 0, s,
                           B, O, p, ctlb);
}

}  // namespace g3m3
