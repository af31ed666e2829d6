// Validation passes run after every stage of the box-step:
//   set_rho        (code/sub_set_rho.f90:5-104)   density floors, pressure reset
//   set_speed_agrd (code/sub_set_speed_agrd.f90:6-435)  speed limits with
//                  neighbour averaging from a snapshot, electron sound-speed
//                  cap, and the CFL maxima.
//
// set_speed_agrd's averaging reads the *snapshot* (pre-modification) values
// of the six neighbours.  Here the snapshot is the input arrays themselves:
// the flag kernel never writes a species field, it appends each flagged
// (cell, species) with its ten averaged values to a side list, and a second
// kernel scatters the list.  Flagged cells are rare, so the pass reads the
// fields once and writes almost nothing.
#include "g3m3_dev.h"
#include "g3m3_launch.h"

namespace g3m3 {

// ---------------------------------------------------------------------------
__global__ void k_set_rho(int n, SpeciesPtrs S, float* ep) {
    const int c = gtid();
    if (c >= n) return;
    const float mins[3] = {2.e-4f, 1.e-4f, 2.e-5f};
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        float* r = S.f[s][0];
        float* px = S.f[s][1];
        float* py = S.f[s][2];
        float* pz = S.f[s][3];
        const float R = r[c], X = px[c], Y = py[c], Z = pz[c];
        if (R < mins[s] || X < 0.0f || Y < 0.0f || Z < 0.0f) {
            r[c] = fmaxf(fabsf(R), 1.05f * mins[s]);
            const float ap = (fabsf(X) + fabsf(Y) + fabsf(Z)) / 3.0f;
            px[c] = ap;
            py[c] = ap;
            pz[c] = ap;
            S.f[s][7][c] = 0.0f;
            S.f[s][8][c] = 0.0f;
            S.f[s][9][c] = 0.0f;
        }
    }
    const float e = ep[c];
    if (signbit(e)) ep[c] = fabsf(e);
}

void launch_set_rho(int n, const SpeciesPtrs& S, float* ep, hipStream_t s) {
    hipLaunchKernelGGL(k_set_rho, dim3((n + 255) / 256), dim3(256), 0, s, n, S, ep);
}

// ---------------------------------------------------------------------------
// Maxima of non-negative values.  Each block writes its own (no
// same-address atomics: thousands of blocks on one L2 line serialise), and
// one block reduces them.  Blocks are 256 threads (4 waves).
__device__ inline float wave_maxf(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// dst[q] = the block's maximum of v[q]
template <int N>
__device__ inline void block_max_store(const float (&v)[N], float* dst) {
    __shared__ float red[N][4];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const float r = wave_maxf(v[q]);
        if (l == 0) red[q][w] = r;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        const int q = threadIdx.x;
        dst[q] = fmaxf(fmaxf(red[q][0], red[q][1]), fmaxf(red[q][2], red[q][3]));
    }
}

// maxima[q] = max over blocks of partials[5 b + q], q < 5 (one block)
__global__ void __launch_bounds__(256) k_speed_reduce(const float* __restrict__ partials, int nb,
                                                      float* maxima) {
    float m[5] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += 256)
#pragma unroll
        for (int q = 0; q < 5; ++q) m[q] = fmaxf(m[q], partials[5 * b + q]);
    block_max_store<5>(m, maxima);
}

__global__ void k_zero_ints(int* p, int n) {
    const int c = gtid();
    if (c < n) p[c] = 0;
}

// One entry of the averaging side list.
struct AvgEntry {
    int cell;
    int species;
    float v[10];
};

// code/sub_set_speed_agrd.f90:88-184 (and the h, o copies :188-381)
// maxima layout: [0]=pxmax [1]=pymax [2]=pzmax [3]=csmax [4]=alfmax
__global__ void __launch_bounds__(256) k_speed_flag(Box B, SpeciesPtrs S, SpeedParams Q,
                                                    AvgEntry* list, int* count, float* partials) {
    int i, j, k;
    float mx = 0.0f, my = 0.0f, mz = 0.0f, mc = 0.0f;
    if (cell_of(B, gtid(), i, j, k) && B.interior(i, j, k)) {
        const int c = B.IX(i, j, k);
        const int nb[6] = {c + 1, c - 1, c + B.nx, c - B.nx, c + B.nx * B.ny, c - B.nx * B.ny};
        const float one_m = 1.0f - Q.aniso;
        // every centre operand first: a list append (a store) inside one
        // species' branch would otherwise hold back the next species' loads
        float cv[3][7];
#pragma unroll
        for (int s = 0; s < 3; ++s)
#pragma unroll
            for (int m = 0; m < 7; ++m) cv[s][m] = S.f[s][m][c];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            float* const* f = S.f[s];
            const float arho = cv[s][0] + Q.smallbit;
            const float avx = fabsf(cv[s][4] / arho);
            const float avy = fabsf(cv[s][5] / arho);
            const float avz = fabsf(cv[s][6] / arho);
            const float v = sqrtf(avx * avx + avy * avy + avz * avz);
            const float cs = sqrtf(0.33f * (cv[s][1] + cv[s][2] + cv[s][3]) / arho);
            const float vl = Q.vlim * Q.svlim[s];
            if (v > vl || cs > Q.vlim * Q.svlim[s] * Q.sclim[s]) {
                AvgEntry e;
                e.cell = c;
                e.species = s;
                // rho, px, py, pz: plain snapshot averages
                const int plain[4] = {0, 4, 5, 6};
                const int slot[4] = {0, 4, 5, 6};
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const float* g = f[plain[m]];
                    e.v[slot[m]] = (g[nb[0]] + g[nb[1]] + g[nb[2]] + g[nb[3]] + g[nb[4]] + g[nb[5]]) / 6.0f;
                }
                // diagonal pressures, anisotropy-limited snapshot (:99-109)
                float sd[3][6];
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const int p = nb[q];
                    const float X = f[1][p], Y = f[2][p], Z = f[3][p];
                    if (Q.isotropic) {
                        sd[0][q] = X;
                        sd[1][q] = X;
                        sd[2][q] = X;
                    } else {
                        const float ap = (X + Y + Z) / 3.0f;
                        sd[0][q] = Q.aniso * X + one_m * ap;
                        sd[1][q] = Q.aniso * Y + one_m * ap;
                        sd[2][q] = Q.aniso * Z + one_m * ap;
                    }
                }
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    e.v[1 + d] = (sd[d][0] + sd[d][1] + sd[d][2] + sd[d][3] + sd[d][4] + sd[d][5]) / 6.0f;
                // off-diagonal pressures: iso_flag * sum(aniso*P)/6
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const float* g = f[7 + d];
                    float sn[6];
#pragma unroll
                    for (int q = 0; q < 6; ++q) sn[q] = Q.isotropic ? 0.0f : Q.aniso * g[nb[q]];
                    e.v[7 + d] = Q.iso_flag * (sn[0] + sn[1] + sn[2] + sn[3] + sn[4] + sn[5]) / 6.0f;
                }
                const int slotidx = atomicAdd(count, 1);
                list[slotidx] = e;
            }
            mx = fmaxf(mx, avx);
            my = fmaxf(my, avy);
            mz = fmaxf(mz, avz);
            mc = fmaxf(mc, cs);
        }
    }
    const float m4[4] = {mx, my, mz, mc};
    block_max_store<4>(m4, partials + 5 * blockIdx.x);
}

__global__ void k_speed_apply(SpeciesPtrs S, const AvgEntry* list, const int* count) {
    const int n = *count;
    for (int e = gtid(); e < n; e += gridDim.x * blockDim.x) {
        const AvgEntry& a = list[e];
        float* const* f = S.f[a.species];
#pragma unroll
        for (int m = 0; m < 10; ++m) f[m][a.cell] = a.v[m];
    }
}

// code/sub_set_speed_agrd.f90:385-432  total field magnitude everywhere,
// electron cs cap and Alfven maximum on the interior (after averaging).
__global__ void k_speed_elec(Box B, const float* __restrict__ bx0, const float* __restrict__ by0,
                             const float* __restrict__ bz0, const float* __restrict__ bx,
                             const float* __restrict__ by, const float* __restrict__ bz,
                             float* btot, const float* __restrict__ q, const float* __restrict__ h,
                             const float* __restrict__ o, float* ep, float smallbit, float cslim,
                             float* partials) {
    int i, j, k;
    float ma = 0.0f;
    if (cell_of(B, gtid(), i, j, k)) {
        const int c = B.IX(i, j, k);
        const float sx = bx0[c] + bx[c], sy = by0[c] + by[c], sz = bz0[c] + bz[c];
        const float bt = fmaxf(sqrtf(sx * sx + sy * sy + sz * sz), smallbit);
        if (btot) btot[c] = bt;
        if (B.interior(i, j, k)) {
            const float arho = q[c] + h[c] + o[c] + smallbit;
            const float e = ep[c];
            const float cs = sqrtf(e / arho);
            if (cs > cslim) ep[c] = e * cslim / cs;
            ma = bt / sqrtf(arho);
        }
    }
    const float m1[1] = {ma};
    block_max_store<1>(m1, partials + 5 * blockIdx.x + 4);
}

void launch_set_speed(const Box& B, const SpeciesPtrs& S, const SpeedParams& Q, const float* b0[3],
                      const float* b[3], float* btot, float* ep, float* maxima, void* list,
                      int* count, float* partials, hipStream_t s) {
    const int n = B.V(), nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, count, 1);
    {
        ProfScope prof_("k_speed_flag", s);
        hipLaunchKernelGGL(k_speed_flag, dim3(nb), dim3(256), 0, s, B, S, Q, (AvgEntry*)list, count,
                           partials);
    }
    hipLaunchKernelGGL(k_speed_apply, dim3(256), dim3(256), 0, s, S, (const AvgEntry*)list, count);
    hipLaunchKernelGGL(k_speed_elec, dim3(nb), dim3(256), 0, s, B, b0[0], b0[1], b0[2], b[0], b[1],
                       b[2], btot, S.f[0][0], S.f[1][0], S.f[2][0], ep, Q.smallbit, Q.cslim,
                       partials);
    hipLaunchKernelGGL(k_speed_reduce, dim3(1), dim3(256), 0, s, partials, nb, maxima);
}

size_t speed_list_bytes(int ncells) { return (size_t)3 * ncells * sizeof(AvgEntry); }
size_t speed_partials_bytes(int ncells) { return (size_t)5 * ((ncells + 255) / 256) * sizeof(float); }

}  // namespace g3m3
