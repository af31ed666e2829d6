// Boundary conditions and nested-grid coupling:
//   bndry_outer   (code/sub_bndry_outer.f90:5-731)  outermost-box faces/edges
//   bndry_inner   (code/sub_bndry_inner.f90:5-268)  planet index lists
//   flanks        (code/sub_flanks.f90:5-303)       coarse -> fine 2-layer shell
//   corer         (code/sub_corer.f90:5-55)         fine -> coarse core
//   flanks_synced (code/sub_flanks_synced.f90:5-141) coarse -> fine faces
#include "g3m3_dev.h"
#include "g3m3_fused.h"
#include "g3m3_launch.h"

namespace g3m3 {

// ---------------------------------------------------------------------------
// bndry_outer (code/sub_bndry_outer.f90:5-731): eight ordered passes over
// the faces, edges and corner lines of each field.
// The eight passes in five phases: 0 and 1 (the i = 1 and i = nx faces)
// touch disjoint cells and read nothing the other writes, and so do 3 and 4
// (z top / bottom) and 6 and 7 (back-wall lines j = 2..ny-1 / x-edge lines
// j = 1, ny); every other pass reads what the one before it wrote.  Within
// a phase no cell is both read and written, so its cells go in any order.
constexpr int kOuterPhases = 5;
__host__ __device__ inline int outer_phase_cells(const Box& B, int p) {
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    switch (p) {
    case 0: return ny * nz;
    case 1: return (nx - 1) * nz;
    case 2: return (nx - 1) * ny;
    case 3: return nz - 2;
    default: return (ny - 2 > nx - 1) ? ny - 2 : nx - 1;
    }
}

// Cells t0, t0 + stride, ... of phase p of field f.  The cells' loads are
// issued kOuterUnroll at a time before their stores, so a thread with
// several cells waits about one memory latency for them.
constexpr int kOuterUnroll = 8;
__device__ __forceinline__ void outer_phase(const Box& B, const OuterPtrs& O, int f, int p,
                                            int t0, int stride) {
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    float* q = O.f[f];
    const int kind = O.kind[f];
    auto at = [&](int i, int j, int k) -> float& { return q[B.IX(i, j, k)]; };
    constexpr int U = kOuterUnroll;
    auto sweep = [&](int n, auto load, auto store) {
        for (int b = t0; b < n; b += stride * U) {
            decltype(load(0)) v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = b + u * stride;
                if (t < n) v[u] = load(t);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = b + u * stride;
                if (t < n) store(t, v[u]);
            }
        }
    };
    switch (p) {
    case 0:   // passes 0 (:50-103) and 1 (:107-154), all j, k
        sweep(ny * nz,
              [&](int t) {
                  const int j = t % ny, k = t / ny;
                  const float v0 =
                      kind >= kBX ? O.val[f] - O.b0[kind - kBX][B.IX(0, j, k)] : O.val[f];
                  return float2{v0, at(nx - 2, j, k)};
              },
              [&](int t, float2 v) {
                  const int j = t % ny, k = t / ny;
                  at(0, j, k) = v.x;
                  at(nx - 1, j, k) = (kind == kPX) ? fabsf(v.y) : v.y;
              });
        break;
    case 1:   // pass 2: y faces, i=2..nx, all k (:158-252)
        sweep((nx - 1) * nz,
              [&](int t) {
                  const int i = 1 + t % (nx - 1), k = t / (nx - 1);
                  return float2{at(i, 1, k), at(i, ny - 2, k)};
              },
              [&](int t, float2 v) {
                  const int i = 1 + t % (nx - 1), k = t / (nx - 1);
                  at(i, 0, k) = (kind == kPY) ? -fabsf(v.x) : v.x;
                  at(i, ny - 1, k) = (kind == kPY) ? fabsf(v.y) : v.y;
              });
        break;
    case 2:   // passes 3 (z top, :256-307) and 4 (z bottom, :311-359), i=2..nx, all j
        sweep((nx - 1) * ny,
              [&](int t) {
                  const int i = 1 + t % (nx - 1), j = t / (nx - 1);
                  return float2{at(i, j, nz - 2), at(i, j, 1)};
              },
              [&](int t, float2 v) {
                  const int i = 1 + t % (nx - 1), j = t / (nx - 1);
                  at(i, j, nz - 1) = (kind == kPZ) ? fabsf(v.x) : v.x;
                  at(i, j, 0) = (kind == kPZ) ? -fabsf(v.y) : v.y;
              });
        break;
    case 3:   // pass 5: corner lines of the back wall, k=2..nz-1 (:363-466)
        for (int t = t0; t < nz - 2; t += stride) {
            const int k = 1 + t;
            if (kind == kBX) {
                at(nx - 1, 0, k) = at(nx - 2, 0, k);
                at(nx - 1, ny - 1, k) = at(nx - 2, ny - 1, k);
            } else {
                at(nx - 1, 0, k) = (at(nx - 2, 0, k) + at(nx - 1, 1, k)) / 2.0f;
                at(nx - 1, ny - 1, k) = (at(nx - 2, ny - 1, k) + at(nx - 1, ny - 2, k)) / 2.0f;
            }
        }
        break;
    default:   // passes 6 (back-wall lines j=2..ny-1, :470-566) and 7 (x-edge lines, :570-728)
        for (int t = t0; t < ny - 2; t += stride) {
            const int j = 1 + t;
            at(nx - 1, j, 0) = at(nx - 1, j, 1);
            if (kind != kEP) at(nx - 1, j, nz - 1) = at(nx - 2, j, nz - 1);
        }
        for (int t = t0; t < nx - 1; t += stride) {
            const int i = 1 + t;
            at(i, 0, 0) = at(i, 0, 1);
            at(i, ny - 1, 0) = at(i, ny - 1, 1);
            at(i, 0, nz - 1) = at(i, 0, nz - 2);
            at(i, ny - 1, nz - 1) = at(i, ny - 1, nz - 2);
        }
        break;
    }
}

// One run: a launch per phase, blockIdx.y = the field, one cell per thread.
__global__ void __launch_bounds__(256) k_outer_phase(Box B, OuterPtrs O, int p) {
    outer_phase(B, O, blockIdx.y, p, blockIdx.x * 256 + threadIdx.x, gridDim.x * 256);
}

// Many runs (multi-run launch): one block per (run, field) runs the five
// phases with barriers between them -- one launch instead of five.
constexpr int kOuterThreads = 1024;
__device__ __forceinline__ void outer_all_block(const Box& B, const OuterPtrs& O, int f) {
    for (int p = 0; p < kOuterPhases; ++p) {
        outer_phase(B, O, f, p, threadIdx.x, blockDim.x);
        __syncthreads();
    }
}

void launch_bndry_outer(const Box& B, const OuterPtrs& O, int nf, hipStream_t s) {
    if (g_rec) {
        g_rec->add(kKOuter, dim3(nf), dim3(kOuterThreads), OuterArgs{B, O});
        return;
    }
    for (int p = 0; p < kOuterPhases; ++p) {
        const int n = outer_phase_cells(B, p);
        hipLaunchKernelGGL(k_outer_phase, dim3((n + 255) / 256, nf), dim3(256), 0, s, B, O, p);
    }
}

// ---------------------------------------------------------------------------
// bndry_inner.  which: 0 = surface (:78-140), 1 = atmosphere (:143-201),
// 2 = interior (:204-265).  Field order in I.f: q,h,o x
// (rho,presx,presy,presz,px,py,pz,presxy,presxz,presyz), epres, bx, by, bz.
__device__ inline void inner_point(const Box& B, const InnerPtrs& I, int which,
                                   const int* __restrict__ ij, const float* __restrict__ parm,
                                   int n, int mbndry, int b, float xmin, float ymin, float dx,
                                   float dy, float re_equiv, float v_rot) {
    const int i = ij[b + mbndry * (0 + 3 * n)] - 1;
    const int j = ij[b + mbndry * (1 + 3 * n)] - 1;
    const int k = ij[b + mbndry * (2 + 3 * n)] - 1;
    const int c = B.IX(i, j, k);
    auto P = [&](int comp) { return parm[b + mbndry * (comp + 7 * n)]; };
    const float ay = ymin + dy * (float)j;
    const float ax = xmin + dx * (float)i;
    const float rx = ax * re_equiv;
    const float ry = ay * re_equiv;
    const float vfrac = 1.0f;
    const float rvy = rx * v_rot * vfrac;
    const float rvx = -ry * v_rot * vfrac;
    const float rvz = 0.0f;
    float den[3];
    for (int s = 0; s < 3; ++s) {
        float* const* f = I.f + 10 * s;
        den[s] = (which == 0) ? fmaxf(P(s), f[0][c]) : P(s);
        f[0][c] = den[s];
        f[1][c] = P(3 + s);
        f[2][c] = P(3 + s);
        f[3][c] = P(3 + s);
        f[4][c] = den[s] * rvx;
        f[5][c] = den[s] * rvy;
        f[6][c] = den[s] * rvz;
        f[7][c] = 0.0f;
        f[8][c] = 0.0f;
        f[9][c] = 0.0f;
    }
    I.f[30][c] = P(6);
    if (which == 2) {
        I.f[31][c] = 0.0f;
        I.f[32][c] = 0.0f;
        I.f[33][c] = 0.0f;
    }
}

// which >= 0: list `which` of A.Ls; which = -1: the three lists in one
// launch (blockIdx.y = list), only for lists with no cell in common, where
// their order does not matter
__device__ __forceinline__ void inner_block(const InnerArgs& A, int lb) {
    const int w = A.which >= 0 ? A.which : (int)blockIdx.y;
    const int n = lb * 256 + threadIdx.x;
    if (n >= A.Ls.num[w]) return;
    inner_point(A.B, A.I, w, A.Ls.ij[w], A.Ls.parm[w], n, A.mbndry, A.b, A.xmin, A.ymin, A.dx,
                A.dy, A.re_equiv, A.v_rot);
}

__global__ void __launch_bounds__(256) k_inner_lists(InnerArgs A) { inner_block(A, blockIdx.x); }

void launch_bndry_inner(const Box& B, const InnerPtrs& I, const int* ij[3], const float* parm[3],
                        const int num[3], int mbndry, int box, float xmin, float ymin, float dx,
                        float dy, float re_equiv, float v_rot, hipStream_t s, bool disjoint) {
    InnerArgs A{};
    A.B = B;
    A.I = I;
    int nmax = 0;
    for (int w = 0; w < 3; ++w) {
        A.Ls.ij[w] = ij[w];
        A.Ls.parm[w] = parm[w];
        A.Ls.num[w] = num[w] > 0 ? num[w] : 0;
        nmax = num[w] > nmax ? num[w] : nmax;
    }
    A.mbndry = mbndry;
    A.b = box - 1;
    A.xmin = xmin;
    A.ymin = ymin;
    A.dx = dx;
    A.dy = dy;
    A.re_equiv = re_equiv;
    A.v_rot = v_rot;
    auto go = [&](dim3 g) {
        if (g_rec) g_rec->add(kKInner, g, dim3(256), A);
        else hipLaunchKernelGGL(k_inner_lists, g, dim3(256), 0, s, A);
    };
    if (disjoint) {
        A.which = -1;
        if (nmax > 0) go(dim3((nmax + 255) / 256, 3));
        return;
    }
    for (int w = 0; w < 3; ++w)
        if (num[w] > 0) {
            A.which = w;
            go(dim3((num[w] + 255) / 256));
        }
}

__global__ void __launch_bounds__(kOuterThreads) k_outer_all_runs(const OuterArgs* __restrict__ A,
                                                                  MultiRun M) {
    int r, lb;
    if (!multi_block(M, r, lb)) return;
    const OuterArgs& a = A[r];
    outer_all_block(a.B, a.O, lb);
}

__global__ void __launch_bounds__(256) k_inner_runs(const InnerArgs* __restrict__ A, MultiRun M) {
    int r, lb;
    if (multi_block(M, r, lb)) inner_block(A[r], lb);
}

void launch_multi_bounds(int id, const void* args, const MultiRun& M, dim3 block, unsigned gy,
                         hipStream_t s) {
    const dim3 g((unsigned)((M.total + 7) & ~7), gy);
    if (id == kKOuter)
        hipLaunchKernelGGL(k_outer_all_runs, g, block, 0, s, (const OuterArgs*)args, M);
    else if (id == kKInner)
        hipLaunchKernelGGL(k_inner_runs, g, block, 0, s, (const InnerArgs*)args, M);
}

// ---------------------------------------------------------------------------
// flanks: coarse box mb -> fine box ms, two layers on every face.
// The time interpolation work = old*t1 + cur*t2 is evaluated at the eight
// coarse corners of each shell cell instead of over the whole coarse box;
// the panels' blends are applied in the reference's order (bottom, top,
// front, back, left, right), so edge cells see the same sequence.
// Shell cell t of a box with every extent >= 5 -> (i, j, k): the four z
// layers (k = 0, 1, nz-2, nz-1; all i, j), then per plane k = 2..nz-3 the
// four y rows (j = 0, 1, ny-2, ny-1; all i), then per (k, j) with
// j = 2..ny-3 the four x cells (i = 0, 1, nx-2, nx-1).  False past the end.
__device__ inline bool shell2_cell(const Box& B, int t, int& i, int& j, int& k) {
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    const int zl = nx * ny, yl = 4 * nx, xl = 4 * (ny - 4);
    auto layer = [](int s, int n) { return s < 2 ? s : n - 4 + s; };
    if (t < 4 * zl) {
        const int s = t / zl, r = t - s * zl;
        k = layer(s, nz);
        j = r / nx;
        i = r - j * nx;
        return true;
    }
    t -= 4 * zl;
    if (t < (nz - 4) * yl) {
        k = 2 + t / yl;
        const int r = t - (k - 2) * yl, s = r / nx;
        i = r - s * nx;
        j = layer(s, ny);
        return true;
    }
    t -= (nz - 4) * yl;
    if (t >= (nz - 4) * xl) return false;
    k = 2 + t / xl;
    const int r = t - (k - 2) * xl;
    j = 2 + r / 4;
    i = layer(r & 3, nx);
    return true;
}

int shell2_count(const Box& B) {
    return 4 * B.nx * B.ny + (B.nz - 4) * 4 * B.nx + (B.nz - 4) * 4 * (B.ny - 4);
}

// shell: threads enumerate the 2-layer shell (every extent >= 5) instead of
// the whole box
template <bool shell>
__global__ void k_flanks(Box B, FlankPtrs Fp, FlankGeom G, float t1, float t2) {
    int i, j, k;
    const int f = blockIdx.y;
    if (shell ? !shell2_cell(B, gtid(), i, j, k) : !cell_of(B, gtid(), i, j, k)) return;
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    const bool bot = k <= 1, top = k >= nz - 2, fro = j <= 1, bac = j >= ny - 2, lef = i <= 1,
               rig = i >= nx - 2;
    if (!(bot || top || fro || bac || lef || rig)) return;
    // coarse position of this fine cell (:50-55, 60-65, 68-73)
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

void launch_flanks(const Box& B, const FlankPtrs& Fp, int nf, const FlankGeom& G, float t1,
                   float t2, hipStream_t s) {
    const int n = B.V();
    {
        ProfScope prof_("k_flanks", s);
        if (B.nx >= 5 && B.ny >= 5 && B.nz >= 5)
            hipLaunchKernelGGL(k_flanks<true>, dim3((shell2_count(B) + 255) / 256, nf), dim3(256),
                               0, s, B, Fp, G, t1, t2);
        else
            hipLaunchKernelGGL(k_flanks<false>, dim3((n + 255) / 256, nf), dim3(256), 0, s, B, Fp,
                               G, t1, t2);
    }
}

// ---------------------------------------------------------------------------
// corer: for every nskip-th fine cell, coarse(ib,jb,kb) = (6c + 6 nbrs)/12
__global__ void k_corer(Box B, FieldPtrs fine, FieldPtrs coarse, int nskip, int nbx1, int nby1,
                        int nbz1, int ni, int nj, int nk) {
    const int f = blockIdx.y;
    const int t = gtid();
    if (t >= ni * nj * nk) return;
    const int a = t % ni, bq = (t / ni) % nj, cq = t / (ni * nj);
    const int is = 1 + nskip + a * nskip, js = 1 + nskip + bq * nskip, ks = 1 + nskip + cq * nskip;
    const int ib = nbx1 + (is - 1) / nskip, jb = nby1 + (js - 1) / nskip,
              kb = nbz1 + (ks - 1) / nskip;
    const float* r = fine.p[f];
    const int c = B.IX(is - 1, js - 1, ks - 1), X = 1, Y = B.nx, Z = B.nx * B.ny;
    coarse.p[f][B.IX(ib - 1, jb - 1, kb - 1)] =
        (6.0f * r[c] + r[c + X] + r[c - X] + r[c + Y] + r[c - Y] + r[c + Z] + r[c - Z]) / 12.0f;
}

void launch_corer(const Box& B, const FieldPtrs& fine, const FieldPtrs& coarse, int nf, int nskip,
                  int nbx1, int nby1, int nbz1, hipStream_t s) {
    // number of fine indices 1+nskip, 1+2*nskip, ... <= n-nskip (1-based)
    auto cnt = [&](int n) { return (n - nskip - (1 + nskip)) / nskip + 1; };
    const int ci = cnt(B.nx), cj = cnt(B.ny), ck = cnt(B.nz);
    if (ci <= 0 || cj <= 0 || ck <= 0) return;
    const int tot = ci * cj * ck;
    hipLaunchKernelGGL(k_corer, dim3((tot + 255) / 256, nf), dim3(256), 0, s, B, fine, coarse,
                       nskip, nbx1, nby1, nbz1, ci, cj, ck);
}

// ---------------------------------------------------------------------------
// flanks_synced: a fine face cell takes the bilinear value of the last
// panel that contains it (top/bottom, then front/back, then left/right).
__global__ void k_flanks_synced(Box B, FieldPtrs fine, FieldPtrs coarse, FlankGeom G, int kb_top,
                                int kb_bot, int jb_top, int jb_bot, int ib_top, int ib_bot) {
    int i, j, k;
    const int f = blockIdx.y;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    const bool xf = (i == 0 || i == nx - 1), yf = (j == 0 || j == ny - 1),
               zf = (k == 0 || k == nz - 1);
    if (!(xf || yf || zf)) return;
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
    const float* Bc = coarse.p[f];
    auto at = [&](int ii, int jj, int kk) { return Bc[B.IX(ii - 1, jj - 1, kk - 1)]; };
    float v;
    if (xf) {
        const int ibb = (i == nx - 1) ? ib_top : ib_bot;
        v = at(ibb, jb, kb) * ddy * ddz + at(ibb, jb + 1, kb) * dy * ddz +
            at(ibb, jb, kb + 1) * ddy * dz + at(ibb, jb + 1, kb + 1) * dy * dz;
    } else if (yf) {
        const int jbb = (j == ny - 1) ? jb_top : jb_bot;
        v = at(ib, jbb, kb) * ddx * ddz + at(ib, jbb, kb + 1) * ddx * dz +
            at(ib + 1, jbb, kb) * dx * ddz + at(ib + 1, jbb, kb + 1) * dx * dz;
    } else {
        const int kbb = (k == nz - 1) ? kb_top : kb_bot;
        v = at(ib, jb, kbb) * ddx * ddy + at(ib, jb + 1, kbb) * ddx * dy +
            at(ib + 1, jb, kbb) * dx * ddy + at(ib + 1, jb + 1, kbb) * dx * dy;
    }
    fine.p[f][B.IX(i, j, k)] = v;
}

void launch_flanks_synced(const Box& B, const FieldPtrs& fine, const FieldPtrs& coarse, int nf,
                          const FlankGeom& G, const int kbi[6], hipStream_t s) {
    const int n = B.V();
    hipLaunchKernelGGL(k_flanks_synced, dim3((n + 255) / 256, nf), dim3(256), 0, s, B, fine, coarse,
                       G, kbi[0], kbi[1], kbi[2], kbi[3], kbi[4], kbi[5]);
}

}  // namespace g3m3
