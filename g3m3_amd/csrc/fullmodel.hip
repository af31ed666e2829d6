// One box-step of the outermost box (box = n_grids = 1, no planet lists)
// with every field evolving, and its discrete adjoint.
//
// Forward (code/3d_main.f90:2163-3030 for that box; the validation calls
// set_rho/set_speed_agrd are left out -- on the model's states they change
// nothing, which the parity test against the oracle's whole box-step
// checks -- and the time step is fixed):
//   stage 1: E-field from S (calcur, total B, fnd_evel, BandE + eta);
//            W1 = push_ion x3 and push_bfld from S with delt/2 (W1's
//            pressure keeps the stale WE: the reference's step-1 push_elec
//            writes the state, which step 2 overwrites); bndry_outer(W1)
//   stage 2: E-field from W1; S2 = push_ion x3, push_elec (source W1's
//            pressure) and push_bfld with old = S and delt; bndry_outer(S2)
//   Lapidus: W3 over W1 = lap_plasma per species with its velocity, and
//            lap_elec / lap_bfld with the total ion velocity; bndry_outer(W3)
//   FCT:     S' = fcsmooth(old = S, wrk = W3) on the 34 fields;
//            bndry_outer(S'); WE' = W3's electron pressure.
// The state is S (34 fields, canonical order) and WE (the stale wrk
// electron pressure the reference carries from one step's Lapidus to the
// next step's push_elec and step-2 BandE): 35 arrays of nx*ny*nz.
// Controls ctl = {eta, sbx, sby, sbz, srho, spx, spy, spz, spress}: the
// resistivity and the upstream values bndry_outer imposes.
//
// An inner box of a nested pair (g3m3_full_model_set_parent) bounds each
// stage by the flanks from its parent instead (code/sub_flanks.f90), as the
// reference's box-step does for box < n_grids: step 1's set at step 1's
// t_grid; the Lapidus set and the FCT output at the full step's t_grid.
// Step 2's flanks go to wrk, where the Lapidus pass and its own flanks
// overwrite them before any read, so the stage-2 set keeps the state's
// shell: no boundary pass there.  The corer (code/sub_corer.f90) restricts
// the box into its parent's core.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <algorithm>
#include <cstring>
#include <utility>
#include <vector>

#include "../../include/g3m3.h"
#include "g3m3_fused.h"

using namespace g3m3;

namespace {

constexpr int NFLD = 34, NST = 35;       // fields, state arrays (fields + WE)
constexpr int EP = 30, BX = 31, WE = 34;
// tape: S (35) | W1 (34) | S2 (34) | W3 (34) | S' (34) | E-field of stage 1 (16) | of stage 2
constexpr int T_S = 0, T_W1 = 35, T_S2 = 69, T_W3 = 103, T_S4 = 137, T_EF1 = 171, T_EF2 = 187,
              T_N = 203;

// boundary cell t of a box (0 <= t < bnd_cells): the k faces, then the j
// faces of interior k, then the i faces of interior j and k
__host__ __device__ inline int bnd_cells(const Box& B) {
    return 2 * B.nx * B.ny + 2 * B.nx * (B.nz - 2) + 2 * (B.ny - 2) * (B.nz - 2);
}
__device__ inline int bnd_cell(const Box& B, int t) {
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    const int a = nx * ny, b = nx * (nz - 2), d = (ny - 2) * (nz - 2);
    if (t < 2 * a) {
        const int k = t < a ? 0 : nz - 1, r = t < a ? t : t - a;
        return B.IX(r % nx, r / nx, k);
    }
    t -= 2 * a;
    if (t < 2 * b) {
        const int j = t < b ? 0 : ny - 1, r = t < b ? t : t - b;
        return B.IX(r % nx, j, 1 + r / nx);
    }
    t -= 2 * b;
    const int i = t < d ? 0 : nx - 1, r = t < d ? t : t - d;
    return B.IX(i, 1 + r % (ny - 2), 1 + r / (ny - 2));
}

// dst += src on the boundary cells (blockIdx.y = field)
__global__ void k_add_bnd(Box B, FieldPtrs dst, FieldPtrs src) {
    const int t = gtid();
    if (t >= bnd_cells(B)) return;
    const int c = bnd_cell(B, t), f = blockIdx.y;
    dst.p[f][c] += src.p[f][c];
}
// dst = src on the boundary cells (blockIdx.y = field)
__global__ void k_copy_bnd(Box B, FieldPtrs dst, FieldPtrs src) {
    const int t = gtid();
    if (t >= bnd_cells(B)) return;
    const int c = bnd_cell(B, t), f = blockIdx.y;
    dst.p[f][c] = src.p[f][c];
}
// zero n cells of each listed array (blockIdx.y = array)
__global__ void k_zero_list(int n, FieldPtrs p) {
    const int c = gtid();
    if (c < n) p.p[blockIdx.y][c] = 0.0f;
}
__global__ void k_add_all(int n, FieldPtrs dst, FieldPtrs src) {
    const int c = gtid();
    if (c >= n) return;
    dst.p[blockIdx.y][c] += src.p[blockIdx.y][c];
}

struct EF {   // one stage's E-field intermediates (16 arrays)
    float *J[3], *bs[3], *btot, *ev[3], *tv[3], *E[3];
};
EF ef_at(float* base, int n) {
    EF e;
    int q = 0;
    auto take = [&]() { return base + (size_t)n * q++; };
    for (int a = 0; a < 3; ++a) e.J[a] = take();
    for (int a = 0; a < 3; ++a) e.bs[a] = take();
    e.btot = take();
    for (int a = 0; a < 3; ++a) e.ev[a] = take();
    for (int a = 0; a < 3; ++a) e.tv[a] = take();
    for (int a = 0; a < 3; ++a) e.E[a] = take();
    return e;
}

struct FullModel {
    g3m3_full_model_desc d;
    Box B;
    int n = 0;
    hipStream_t s = nullptr;
    float* mem = nullptr;
    EF ef;
    float* qv[3];
    float* vel[3];
    float* own_tape = nullptr;
    // adjoint scratch
    float *A = nullptr, *W3B = nullptr, *S2B = nullptr, *W1B = nullptr;
    float *EB[3], *EVB[3], *CURB[3], *BSB[3], *BTB, *TVB[3], *VB[3];
    float *fct_scr, *lap_scr, *ion_scr, *ef_scr;
    // as an inner box: its 2-layer shell's cells with their parent-box
    // corners and weights, and the corer's geometry
    ShellPt* shell = nullptr;
    int nshell = 0;
    int c_nskip = 0, c_nbx1 = 0, c_nby1 = 0, c_nbz1 = 0;
    // the flanks' adjoint as a gather: the parent cells the shell's
    // trilinear corners touch (fc_cell), and for each its (shell point,
    // weight) pairs fc_p/fc_w[fc_start[i] .. fc_start[i+1]); fc_I0b holds
    // each shell point's interpolation adjoint per field
    int* fc_cell = nullptr;
    int* fc_start = nullptr;
    int* fc_p = nullptr;
    float* fc_w = nullptr;
    float* fc_I0b = nullptr;
    int nfc = 0;
};

// the flanks of one inner-box step: the parent's states (34 fields each,
// canonical order) and the time weights of step 1's t_grid (t1a, t2a) and
// of the later stages' (t1b, t2b); for the adjoint, the parent adjoints the
// flanks add into
struct Flank {
    const float* p_old;
    const float* p_cur;
    float t1a, t2a, t1b, t2b;
    float* p_old_bar;
    float* p_cur_bar;
};

float* fld(float* base, int n, int f) { return base + (size_t)n * f; }

bool launched(const char* where) {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return true;
    fprintf(stderr, "g3m3_full_model: %s: %s\n", where, hipGetErrorString(e));
    return false;
}

void zero(FullModel* M, float* p, size_t floats) {
    hipMemsetAsync(p, 0, sizeof(float) * floats, M->s);
}
void zero_vb(FullModel* M) {
    FieldPtrs z{};
    for (int a = 0; a < 3; ++a) z.p[a] = M->VB[a];
    hipLaunchKernelGGL(k_zero_list, dim3((M->n + 255) / 256, 3), dim3(256), 0, M->s, M->n, z);
}
void copyf(FullModel* M, float* dst, const float* src, int nf) {
    hipMemcpyAsync(dst, src, sizeof(float) * (size_t)M->n * nf, hipMemcpyDeviceToDevice, M->s);
}
// the boundary cells of nf fields: a set whose interior the next pass
// writes whole (the pushes, the Lapidus passes) takes only its boundary
// from the set it starts from
void copy_bnd(FullModel* M, float* dst, const float* src, int nf) {
    FieldPtrs d{}, s{};
    for (int f = 0; f < nf; ++f) {
        d.p[f] = dst + (size_t)M->n * f;
        s.p[f] = const_cast<float*>(src) + (size_t)M->n * f;
    }
    hipLaunchKernelGGL(k_copy_bnd, dim3((bnd_cells(M->B) + 255) / 256, nf), dim3(256), 0, M->s,
                       M->B, d, s);
}

// inflow values of the 34 fields from the controls, as the runtime forms
// them (code/sub_bndry_outer.f90:50-103), with each value's control index
// and derivative
struct Inflow {
    float val[NFLD];
    int kind[NFLD];
    int ctl[NFLD];
    float coef[NFLD];
};
Inflow inflow(const g3m3_full_model_desc& d, const float* c) {
    Inflow I;
    const float srho = c[4], spx = c[5], spy = c[6], spz = c[7], spress = c[8];
    const float ofrac = d.rho_frac * d.o_conc;
    const float frac_h = d.rmassh / d.rmassq, frac_o = d.rmasso / d.rmassq;
    const float pq = 0.5f * spress, rf = d.rho_frac;
    const float sval[3][10] = {
        {srho, pq, pq, pq, spx, spy, spz, 0.f, 0.f, 0.f},
        {srho * rf * frac_h, 0.5f * spress * rf, 0.5f * spress * rf, 0.5f * spress * rf,
         spx * rf * frac_h, spy * rf * frac_h, spz * rf * frac_h, 0.f, 0.f, 0.f},
        {srho * ofrac * frac_o, 0.5f * spress * ofrac, 0.5f * spress * ofrac,
         0.5f * spress * ofrac, spx * ofrac * frac_o, spy * ofrac * frac_o,
         spz * ofrac * frac_o, 0.f, 0.f, 0.f}};
    const float scale[3] = {1.0f, rf * frac_h, ofrac * frac_o};
    const float pscale[3] = {0.5f, 0.5f * rf, 0.5f * ofrac};
    const int kinds[10] = {kRHO, kPRES, kPRES, kPRES, kPX, kPY, kPZ, kOFF, kOFF, kOFF};
    const int ctls[10] = {4, 8, 8, 8, 5, 6, 7, -1, -1, -1};
    for (int sp = 0; sp < 3; ++sp)
        for (int m = 0; m < 10; ++m) {
            const int f = 10 * sp + m;
            I.val[f] = sval[sp][m];
            I.kind[f] = kinds[m];
            I.ctl[f] = ctls[m];
            I.coef[f] = (m >= 1 && m <= 3) ? pscale[sp] : (m == 0 || m >= 4 && m <= 6)
                                                              ? scale[sp] : 0.0f;
        }
    I.val[EP] = pq / d.ti_te;
    I.kind[EP] = kEP;
    I.ctl[EP] = 8;
    I.coef[EP] = 0.5f / d.ti_te;
    for (int a = 0; a < 3; ++a) {
        I.val[BX + a] = c[1 + a];
        I.kind[BX + a] = kBX + a;
        I.ctl[BX + a] = 1 + a;
        I.coef[BX + a] = 1.0f;
    }
    return I;
}

// The flanks' adjoint (code/sub_flanks.f90, transposed), first pass: per
// shell cell, for each field w = frac w + (1 - frac) I0 over the panels
// containing the cell (bottom, top, front, back, left, right) makes the
// cell's adjoint the product of the fracs times its own, and I0's adjoint
// the cell's adjoint times sum_j (1 - frac_j) prod_{l > j} frac_l, stored
// point-major: I0b[t * NFLD + f].
__global__ void k_adj_flanks_cells(const ShellPt* __restrict__ tab, int npts, FieldPtrs xb,
                                   float* __restrict__ I0b) {
    const int t = gtid();
    if (t >= npts) return;
    const ShellPt p = tab[t];
    float pw = 1.0f, pI = 0.0f;
#pragma unroll
    for (int b = 0; b < 6; ++b)
        if (p.flags & (1 << b)) {
            const float frac = (p.flags & (256 << b)) ? 0.5f : 0.0f;
            pI = frac * pI + (1.0f - frac);
            pw *= frac;
        }
    float wb[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) wb[f] = xb.p[f][p.c];
    float* __restrict__ dst = I0b + (size_t)t * NFLD;
#pragma unroll
    for (int f = 0; f < NFLD; f += 2)
        *reinterpret_cast<float2*>(dst + f) = make_float2(pI * wb[f], pI * wb[f + 1]);
#pragma unroll
    for (int f = 0; f < NFLD; ++f) xb.p[f][p.c] = pw * wb[f];
}

// Second pass, per parent cell the shell's corners touch (ascending), all
// fields: the adjoint of work = old t1 + cur t2 at that cell is the sum over
// the shell points that interpolate from it of their trilinear weight times
// I0's adjoint (each point's row of NFLD values read once); old's adjoint
// gains t1 times it, cur's t2 times it.
__global__ void __launch_bounds__(64) k_adj_flanks_gather(
    const int* __restrict__ cell, const int* __restrict__ start, const int* __restrict__ pt,
    const float* __restrict__ wt, int ncell, const float* __restrict__ I0b, FieldPtrs ob,
    FieldPtrs cb, float t1, float t2) {
    const int i = gtid();
    if (i >= ncell) return;
    float acc[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) acc[f] = 0.0f;
    const int e1 = start[i + 1];
    for (int e = start[i]; e < e1; ++e) {
        const float w = wt[e];
        const float2* __restrict__ src =
            reinterpret_cast<const float2*>(I0b + (size_t)pt[e] * NFLD);
#pragma unroll
        for (int h = 0; h < NFLD / 2; ++h) {
            const float2 v = src[h];
            acc[2 * h] += w * v.x;
            acc[2 * h + 1] += w * v.y;
        }
    }
    const int c = cell[i];
    float o[NFLD], q[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) {
        o[f] = ob.p[f][c];
        q[f] = cb.p[f][c];
    }
#pragma unroll
    for (int f = 0; f < NFLD; ++f) {
        ob.p[f][c] = o[f] + t1 * acc[f];
        cb.p[f][c] = q[f] + t2 * acc[f];
    }
}

// The corer's adjoint (code/sub_corer.f90:34-53, transposed), as a gather
// over this box's cells: the parent core cell fed by a fine centre (0-based
// fine index a nskip, 1 <= a <= count on each axis) hands 6/12 of its
// adjoint to the centre and 1/12 to each of the centre's six neighbours, so
// a fine cell collects 6/12 of its own centre's adjoint (if it is one) and
// 1/12 of each centre one step away along one axis.  The parent core's
// adjoint is zeroed afterwards (k_zero_core: the corer overwrote it).
// One thread per cell, all fields.
struct CorerGeom {
    int nskip, nb[3], cnt[3];
};
__device__ inline bool corer_centre(const CorerGeom& G, int ax, int q, int& a) {
    if (q < G.nskip) return false;
    a = udiv_small(q, G.nskip);
    return a * G.nskip == q && a <= G.cnt[ax];
}
__global__ void k_adj_corer(Box B, FieldPtrs fb, FieldPtrs cb, CorerGeom G) {
    int x[3];
    if (!cell_of_fast(B, gtid(), x[0], x[1], x[2])) return;
    int a[3];
    bool al[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) al[d] = corer_centre(G, d, x[d], a[d]);
    // the parent cell of the centre with fine indices (a0, a1, a2) nskip:
    // nb1 + a (1-based) on each axis
    auto parent = [&](int i, int j, int k) {
        return B.IX(G.nb[0] + i - 1, G.nb[1] + j - 1, G.nb[2] + k - 1);
    };
    // the parent cells this cell gathers from and their weights: its own
    // centre's (6/12), each centre one step away along one axis (1/12)
    int pc[7];
    float pw[7];
    int np = 0;
    if (al[0] && al[1] && al[2]) {
        pc[np] = parent(a[0], a[1], a[2]);
        pw[np++] = 6.0f;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int e = (d + 1) % 3, g = (d + 2) % 3;
        if (!(al[e] && al[g])) continue;
#pragma unroll
        for (int s = -1; s <= 1; s += 2) {
            int c;
            if (!corer_centre(G, d, x[d] + s, c)) continue;
            int idx[3];
            idx[d] = c;
            idx[e] = a[e];
            idx[g] = a[g];
            pc[np] = parent(idx[0], idx[1], idx[2]);
            pw[np++] = 1.0f;
        }
    }
    if (np == 0) return;
    const int m = B.IX(x[0], x[1], x[2]);
    for (int f = 0; f < NFLD; ++f) {
        const float* __restrict__ src = cb.p[f];
        float acc = 0.0f;
        for (int q = 0; q < np; ++q)
            acc += pw[q] == 6.0f ? 6.0f * src[pc[q]] / 12.0f : src[pc[q]] / 12.0f;
        fb.p[f][m] += acc;
    }
}
__global__ void k_zero_core(Box B, FieldPtrs cb, CorerGeom G) {
    const int t = gtid(), f = blockIdx.y;
    const int ni = G.cnt[0], nj = G.cnt[1], nk = G.cnt[2];
    if (t >= ni * nj * nk) return;
    const int i = 1 + t % ni, j = 1 + (t / ni) % nj, k = 1 + t / (ni * nj);
    cb.p[f][B.IX(G.nb[0] + i - 1, G.nb[1] + j - 1, G.nb[2] + k - 1)] = 0.0f;
}

// corer counts: fine indices 1+nskip, 1+2 nskip, ... <= n - nskip (1-based)
int corer_count(int n, int nskip) { return (n - nskip - (1 + nskip)) / nskip + 1; }

void flanks(FullModel* M, float* set, const Flank& Fk, float t1, float t2) {
    FlankPtrs P{};
    for (int f = 0; f < NFLD; ++f) {
        P.tgt[f] = fld(set, M->n, f);
        P.old[f] = fld((float*)Fk.p_old, M->n, f);
        P.cur[f] = fld((float*)Fk.p_cur, M->n, f);
    }
    launch_flanks_table(M->B, M->shell, M->nshell, P, NFLD, t1, t2, M->s);
}

void flanks_adj(FullModel* M, float* bar, const Flank& Fk, float t1, float t2) {
    FieldPtrs xb{}, ob{}, cb{};
    for (int f = 0; f < NFLD; ++f) {
        xb.p[f] = fld(bar, M->n, f);
        ob.p[f] = fld(Fk.p_old_bar, M->n, f);
        cb.p[f] = fld(Fk.p_cur_bar, M->n, f);
    }
    hipLaunchKernelGGL(k_adj_flanks_cells, dim3((M->nshell + 255) / 256), dim3(256), 0, M->s,
                       M->shell, M->nshell, xb, M->fc_I0b);
    hipLaunchKernelGGL(k_adj_flanks_gather, dim3((M->nfc + 63) / 64), dim3(64), 0, M->s,
                       M->fc_cell, M->fc_start, M->fc_p, M->fc_w, M->nfc, M->fc_I0b, ob, cb, t1,
                       t2);
}

void free_parent(FullModel* M) {
    void* p[] = {M->shell, M->fc_cell, M->fc_start, M->fc_p, M->fc_w, M->fc_I0b};
    for (void* q : p)
        if (q) (void)hipFree(q);
    M->shell = nullptr;
    M->fc_cell = M->fc_start = M->fc_p = nullptr;
    M->fc_w = M->fc_I0b = nullptr;
    M->nshell = M->nfc = 0;
}

void outer(FullModel* M, float* set, const Inflow& I) {
    OuterPtrs O{};
    for (int f = 0; f < NFLD; ++f) {
        O.f[f] = fld(set, M->n, f);
        O.kind[f] = I.kind[f];
        O.val[f] = I.val[f];
    }
    for (int a = 0; a < 3; ++a) O.b0[a] = M->d.b0[a];
    launch_bndry_outer(M->B, O, NFLD, M->s);
}

void outer_adj(FullModel* M, float* bar, const float* post, const Inflow& I, double* ctlb) {
    OuterAdj O{};
    for (int f = 0; f < NFLD; ++f) {
        O.f[f] = fld(bar, M->n, f);
        O.val[f] = post + (size_t)M->n * f;
        O.kind[f] = I.kind[f];
        O.ctl[f] = I.ctl[f];
        O.coef[f] = I.coef[f];
    }
    launch_adj_outer(M->B, O, NFLD, ctlb, M->s);
}

// E-field intermediates of a set (fields in canonical order)
void efield(FullModel* M, const float* set, float eta) {
    const g3m3_full_model_desc& d = M->d;
    const int n = M->n;
    const float r = d.rx;
    EF& e = M->ef;
    const float* b[3] = {fld((float*)set, n, BX), fld((float*)set, n, BX + 1),
                         fld((float*)set, n, BX + 2)};
    const float* rho[3];
    const float* mom[9];
    for (int sp = 0; sp < 3; ++sp) {
        rho[sp] = fld((float*)set, n, 10 * sp);
        for (int a = 0; a < 3; ++a) mom[3 * sp + a] = fld((float*)set, n, 10 * sp + 4 + a);
    }
    launch_calcur(M->B, b[0], b[1], b[2], e.J[0], e.J[1], e.J[2], r, r, r, M->s);
    const float* J[3] = {e.J[0], e.J[1], e.J[2]};
    const float* b0[3] = {d.b0[0], d.b0[1], d.b0[2]};
    launch_bfield_evel(M->B, b0, b, e.bs, e.btot, mom, rho, J, e.ev, e.tv, d.rmassq, d.rmassh,
                       d.rmasso, d.reynolds, d.smallbit, 1, 1, M->s);
    const float* bs[3] = {e.bs[0], e.bs[1], e.bs[2]};
    const float* ev[3] = {e.ev[0], e.ev[1], e.ev[2]};
    launch_bande(M->B, e.E, bs, J, ev, fld((float*)set, n, EP), rho, nullptr, 0.0f, d.reynolds,
                 1, d.rmassq, d.rmassh, d.rmasso, nullptr, 0, nullptr, 0, 0, r, r, r, eta, M->s);
}

void push_species(FullModel* M, float* out, const float* old, const float* src, float delt) {
    const g3m3_full_model_desc& d = M->d;
    const int n = M->n;
    const float rm[3] = {d.rmassq, d.rmassh, d.rmasso};
    const float gmin[3] = {d.xmin, d.ymin, d.zmin}, gmax[3] = {d.xmax, d.ymax, d.zmax};
    for (int sp = 0; sp < 3; ++sp) {
        IonPtrs P{};
        for (int f = 0; f < 10; ++f) {
            P.out[f] = fld(out, n, 10 * sp + f);
            P.old[f] = fld((float*)old, n, 10 * sp + f);
            P.wrk[f] = fld((float*)src, n, 10 * sp + f);
        }
        for (int a = 0; a < 3; ++a) {
            P.bs[a] = M->ef.bs[a];
            P.e[a] = M->ef.E[a];
            P.qv[a] = M->qv[a];
            P.tv[a] = M->ef.tv[a];
        }
        P.btot = M->ef.btot;
        launch_push_ion(M->B, P, rm[sp], delt, d.grav, d.re_equiv, d.reynolds, gmin, gmax, M->s);
    }
}

void push_b(FullModel* M, float* out, const float* old, float delt) {
    const int n = M->n;
    float* b[3] = {fld(out, n, BX), fld(out, n, BX + 1), fld(out, n, BX + 2)};
    const float* o[3] = {fld((float*)old, n, BX), fld((float*)old, n, BX + 1),
                         fld((float*)old, n, BX + 2)};
    const float* E[3] = {M->ef.E[0], M->ef.E[1], M->ef.E[2]};
    const float r = M->d.rx;
    launch_push_bfld(M->B, b, o, E, delt, r, r, r, M->s);
}

}  // namespace

extern "C" {

void* g3m3_full_model_create(const g3m3_full_model_desc* desc) {
    if (!desc || desc->nx < 5 || desc->ny < 5 || desc->nz < 5 ||
        (long long)desc->nx * desc->ny * desc->nz >= kFastDivMax) {
        fprintf(stderr, "g3m3_full_model_create: box extents out of range\n");
        return nullptr;
    }
    FullModel* M = new FullModel();
    M->d = *desc;
    M->B = Box{desc->nx, desc->ny, desc->nz};
    M->n = M->B.V();
    const size_t n = M->n;
    // qv 3 | vel 3 | own tape T_N | A 34 | W3B, S2B, W1B 34 each |
    // EB, EVB, CURB, BSB 3 each, BTB 1, TVB 3, VB 3 | fct | lap 6 | ion 15 | ef 3
    const size_t total = 3 + 3 + T_N + 34 * 4 + 12 + 1 + 6 +
                         adj_fct_scratch_floats(M->B) / n + 6 + 15 + 3;
    if (hipMalloc(&M->mem, sizeof(float) * total * n) != hipSuccess) {
        fprintf(stderr, "g3m3_full_model_create: out of device memory\n");
        delete M;
        return nullptr;
    }
    float* p = M->mem;
    auto take = [&](size_t k) {
        float* r = p;
        p += k * n;
        return r;
    };
    for (int a = 0; a < 3; ++a) M->qv[a] = take(1);
    for (int a = 0; a < 3; ++a) M->vel[a] = take(1);
    M->own_tape = take(T_N);
    M->A = take(34);
    M->W3B = take(34);
    M->S2B = take(34);
    M->W1B = take(34);
    for (int a = 0; a < 3; ++a) M->EB[a] = take(1);
    for (int a = 0; a < 3; ++a) M->EVB[a] = take(1);
    for (int a = 0; a < 3; ++a) M->CURB[a] = take(1);
    for (int a = 0; a < 3; ++a) M->BSB[a] = take(1);
    M->BTB = take(1);
    for (int a = 0; a < 3; ++a) M->TVB[a] = take(1);
    for (int a = 0; a < 3; ++a) M->VB[a] = take(1);
    M->fct_scr = take(adj_fct_scratch_floats(M->B) / n);
    M->lap_scr = take(6);
    M->ion_scr = take(15);
    M->ef_scr = take(3);
    zero(M, M->mem, total * n);
    if (!launched("create") || hipStreamSynchronize(M->s) != hipSuccess) {
        hipFree(M->mem);
        delete M;
        return nullptr;
    }
    return M;
}

void g3m3_full_model_destroy(void* model) {
    FullModel* M = (FullModel*)model;
    if (!M) return;
    hipStreamSynchronize(M->s);
    hipFree(M->mem);
    free_parent(M);
    delete M;
}

void g3m3_full_model_set_stream(void* model, void* hip_stream) {
    if (model) ((FullModel*)model)->s = (hipStream_t)hip_stream;
}

long long g3m3_full_model_tape_floats(void* model) {
    return model ? (long long)T_N * ((FullModel*)model)->n : 0;
}

// one step, outermost (Fk null: bndry_outer) or inner (the flanks of Fk)
static int step_impl(void* model, const float* s_in, float* s_out, float* tape,
                     const float* ctl, const Flank* Fk) {
    FullModel* M = (FullModel*)model;
    if (!M || !s_in || !s_out || !ctl || s_in == s_out) return 1;
    if (Fk && (!M->shell || !Fk->p_old || !Fk->p_cur)) return 1;
    const g3m3_full_model_desc& d = M->d;
    const int n = M->n;
    const float eta = ctl[0], delt = d.delt;
    const Inflow I = inflow(d, ctl);
    float* T = tape ? tape : M->own_tape;
    float* S = fld(T, n, T_S);
    float* W1 = fld(T, n, T_W1);
    float* S2 = fld(T, n, T_S2);
    float* W3 = fld(T, n, T_W3);
    float* S4 = fld(T, n, T_S4);
    copyf(M, S, s_in, NST);
    // stage 1
    M->ef = ef_at(fld(T, n, T_EF1), n);
    efield(M, S, eta);
    copy_bnd(M, W1, S, NFLD);
    copyf(M, fld(W1, n, EP), fld(S, n, WE), 1);
    push_species(M, W1, S, S, 0.5f * delt);
    push_b(M, W1, S, 0.5f * delt);
    if (Fk) flanks(M, W1, *Fk, Fk->t1a, Fk->t2a);
    else outer(M, W1, I);
    // stage 2
    M->ef = ef_at(fld(T, n, T_EF2), n);
    efield(M, W1, eta);
    copy_bnd(M, S2, S, NFLD);
    push_species(M, S2, S, W1, delt);
    {
        const float* ev[3] = {M->ef.ev[0], M->ef.ev[1], M->ef.ev[2]};
        const float r = d.rx;
        launch_push_elec(M->B, fld(W1, n, EP), fld(S, n, EP), fld(S2, n, EP), ev, d.gamma,
                         d.gamma1, delt, r, r, r, M->s);
    }
    push_b(M, S2, S, delt);
    if (!Fk) outer(M, S2, I);
    // Lapidus into a copy of W1
    copy_bnd(M, W3, W1, NFLD);
    for (int sp = 0; sp < 3; ++sp) {
        const float* pp[3] = {fld(S2, n, 10 * sp + 4), fld(S2, n, 10 * sp + 5),
                              fld(S2, n, 10 * sp + 6)};
        launch_velocity(n, fld(S2, n, 10 * sp), pp, M->vel, d.smallbit, M->s);
        LapPtrs L{};
        for (int f = 0; f < 10; ++f) {
            L.src[f] = fld(S2, n, 10 * sp + f);
            L.dst[f] = fld(W3, n, 10 * sp + f);
            L.chi[f] = f == 0 ? d.chirho : d.chipxyz;
        }
        const float* v[3] = {M->vel[0], M->vel[1], M->vel[2]};
        launch_lapidus(M->B, L, 10, v, delt, M->s);
    }
    {
        const float* mom[9];
        const float* rho[3];
        for (int sp = 0; sp < 3; ++sp) {
            rho[sp] = fld(S2, n, 10 * sp);
            for (int a = 0; a < 3; ++a) mom[3 * sp + a] = fld(S2, n, 10 * sp + 4 + a);
        }
        launch_vtot(n, mom, rho, M->vel, M->s);
        LapPtrs L{};
        for (int f = 0; f < 4; ++f) {
            L.src[f] = fld(S2, n, EP + f);
            L.dst[f] = fld(W3, n, EP + f);
            L.chi[f] = f == 0 ? d.chipxyz : d.chierg;
        }
        const float* v[3] = {M->vel[0], M->vel[1], M->vel[2]};
        launch_lapidus(M->B, L, 4, v, delt, M->s);
    }
    if (Fk) flanks(M, W3, *Fk, Fk->t1b, Fk->t2b);
    else outer(M, W3, I);
    // flux correction
    FctParams F{};
    F.B = M->B;
    for (int f = 0; f < NFLD; ++f) {
        F.out[f] = fld(S4, n, f);
        F.old[f] = fld(S, n, f);
        F.wrk[f] = fld(W3, n, f);
        F.chi[f] = f >= BX ? d.diferg : d.difrho;
    }
    launch_fct_fused(F, NFLD, M->s);
    if (Fk) flanks(M, S4, *Fk, Fk->t1b, Fk->t2b);
    else outer(M, S4, I);
    copyf(M, s_out, S4, NFLD);
    copyf(M, fld(s_out, n, WE), fld(W3, n, EP), 1);
    return launched("step") ? 0 : 2;
}

namespace {

// the adjoint of one stage's push and E-field: from the stage output's
// adjoint `ob` (34 fields), accumulate into the source set's adjoint `sb`
// and the old set's adjoint `oldb`; src is the stage's source set, old S
void stage_adj(FullModel* M, const EF& ef, const float* src, const float* old, float* ob,
               float* sb, float* oldb, float delt, bool elec, double* ctl_bar, float eta) {
    const g3m3_full_model_desc& d = M->d;
    const int n = M->n;
    const float r = d.rx;
    hipStream_t s = M->s;
    {
        FieldPtrs z{};
        for (int a = 0; a < 3; ++a) {
            z.p[a] = M->BSB[a];
            z.p[3 + a] = M->TVB[a];
            z.p[6 + a] = M->CURB[a];
        }
        z.p[9] = M->BTB;
        hipLaunchKernelGGL(k_zero_list, dim3((n + 255) / 256, 10), dim3(256), 0, s, n, z);
    }
    // push_bfld: old's adjoint, E's adjoint (assigned)
    {
        const float* bb[3] = {fld(ob, n, BX), fld(ob, n, BX + 1), fld(ob, n, BX + 2)};
        float* obb[3] = {fld(oldb, n, BX), fld(oldb, n, BX + 1), fld(oldb, n, BX + 2)};
        launch_adj_push_bfld(M->B, bb, obb, M->EB, delt, r, r, r, s);
    }
    // push_ion, each species
    const float rm[3] = {d.rmassq, d.rmassh, d.rmasso};
    for (int sp = 0; sp < 3; ++sp) {
        IonAdj A{};
        A.B = M->B;
        for (int f = 0; f < 10; ++f) {
            A.src[f] = fld((float*)src, n, 10 * sp + f);
            A.ob[f] = fld(ob, n, 10 * sp + f);
            A.sb[f] = fld(sb, n, 10 * sp + f);
            A.oldb[f] = fld(oldb, n, 10 * sp + f);
        }
        for (int a = 0; a < 3; ++a) {
            A.bs[a] = ef.bs[a];
            A.e[a] = ef.E[a];
            A.tv[a] = ef.tv[a];
            A.bsb[a] = M->BSB[a];
            A.eb[a] = M->EB[a];
            A.tvb[a] = M->TVB[a];
        }
        A.btot = ef.btot;
        A.btotb = M->BTB;
        A.scr = M->ion_scr;
        A.rmass = rm[sp];
        A.delt = delt;
        A.grav = d.grav;
        A.re_equiv = d.re_equiv;
        A.reynolds = d.reynolds;
        A.xmin = d.xmin;
        A.ymin = d.ymin;
        A.zmin = d.zmin;
        A.ddx = (d.xmax - d.xmin) / (float)(d.nx - 1);
        A.ddy = (d.ymax - d.ymin) / (float)(d.ny - 1);
        A.ddz = (d.zmax - d.zmin) / (float)(d.nz - 1);
        A.dxt = 2.0f * A.ddx;
        A.dyt = 2.0f * A.ddy;
        A.dzt = 2.0f * A.ddz;
        A.bmax = 0.20f * A.rmass / A.delt / A.reynolds;
        launch_adj_push_ion(A, s);
    }
    // BandE's face copies, then its body with the plasma terms
    launch_adj_bande_faces(M->B, M->EB, s);
    EfAdj E{};
    E.B = M->B;
    for (int a = 0; a < 3; ++a) {
        E.eb[a] = M->EB[a];
        E.bs[a] = ef.bs[a];
        E.ev[a] = ef.ev[a];
        E.cur[a] = ef.J[a];
        E.tvb[a] = M->TVB[a];
        E.evb[a] = M->EVB[a];
        E.bsb[a] = M->BSB[a];
        E.curb[a] = M->CURB[a];
    }
    for (int sp = 0; sp < 3; ++sp) {
        E.rho[sp] = fld((float*)src, n, 10 * sp);
        E.rhob[sp] = fld(sb, n, 10 * sp);
        for (int a = 0; a < 3; ++a) {
            E.mom[3 * sp + a] = fld((float*)src, n, 10 * sp + 4 + a);
            E.momb[3 * sp + a] = fld(sb, n, 10 * sp + 4 + a);
        }
    }
    E.ep = fld((float*)src, n, EP);
    E.epb = fld(sb, n, EP);
    E.etab = ctl_bar;
    E.scr = M->ef_scr;
    E.eta = eta;
    E.rmq = d.rmassq;
    E.rmh = d.rmassh;
    E.rmo = d.rmasso;
    E.reynolds = d.reynolds;
    E.dxt = E.dyt = E.dzt = 2.0f * r;
    // BandE's body assigns ev's adjoint; push_elec's part is added to it
    // before fnd_evel's transpose reads it
    launch_adj_bande_plasma(E, s);
    if (elec) {
        PeAdj P{};
        P.B = M->B;
        P.w = fld((float*)src, n, EP);
        for (int a = 0; a < 3; ++a) {
            P.ev[a] = ef.ev[a];
            P.evb[a] = M->EVB[a];
        }
        P.ob = fld(ob, n, EP);
        P.oldb = fld(oldb, n, EP);
        P.wb = fld(sb, n, EP);
        P.delt = delt;
        P.gamma = d.gamma;
        P.gamma1 = d.gamma1;
        P.dxt = P.dyt = P.dzt = 2.0f * r;
        launch_adj_push_elec(P, s);
    }
    launch_adj_evel_plasma(E, s);
    // total B: btot, then bs = b0 + b
    const float* bsc[3] = {ef.bs[0], ef.bs[1], ef.bs[2]};
    launch_adj_btot(n, bsc, M->BTB, d.smallbit, M->BSB, s);
    float* bb[3] = {fld(sb, n, BX), fld(sb, n, BX + 1), fld(sb, n, BX + 2)};
    const float* bsb[3] = {M->BSB[0], M->BSB[1], M->BSB[2]};
    launch_add3(n, bsb, bb, s);
    launch_adj_calcur(M->B, M->CURB, bb, r, r, r, s);
}

}  // namespace

static int adjoint_impl(void* model, const float* tape, const float* sout_bar, float* sin_bar,
                        double* ctl_bar, const float* ctl, const Flank* Fk) {
    FullModel* M = (FullModel*)model;
    if (!M || !tape || !sout_bar || !sin_bar || !ctl_bar || !ctl) return 1;
    if (Fk && (!M->shell || !Fk->p_old_bar || !Fk->p_cur_bar)) return 1;
    const g3m3_full_model_desc& d = M->d;
    const int n = M->n;
    const float eta = ctl[0], delt = d.delt;
    const Inflow I = inflow(d, ctl);
    hipStream_t s = M->s;
    float* T = const_cast<float*>(tape);
    const float* S = fld(T, n, T_S);
    const float* W1 = fld(T, n, T_W1);
    const float* S2 = fld(T, n, T_S2);
    const float* W3 = fld(T, n, T_W3);
    const float* S4 = fld(T, n, T_S4);
    // SB's 34 fields are first written (assigned) by the FCT's adjoint, its
    // WE by the last pass; S2B's by the Lapidus passes' adjoints
    float* SB = sin_bar;
    zero(M, M->W1B, (size_t)NFLD * n);
    copyf(M, M->A, sout_bar, NFLD);
    // S' = bndry_outer(fcsmooth(S, W3)) (or the flanks)
    if (Fk) flanks_adj(M, M->A, *Fk, Fk->t1b, Fk->t2b);
    else outer_adj(M, M->A, S4, I, ctl_bar);
    for (int f0 = 0; f0 < NFLD;) {   // (assigns W3B)
        // groups of up to three fields of one coefficient
        const int end = f0 < BX ? BX : NFLD;
        const int nf = end - f0 < 3 ? end - f0 : 3;
        const float* old[3];
        const float* wrk[3];
        float *ob[3], *wb[3], *oldb[3];
        for (int q = 0; q < 3; ++q) {
            const int f = f0 + (q < nf ? q : 0);
            old[q] = fld((float*)S, n, f);
            wrk[q] = fld((float*)W3, n, f);
            ob[q] = fld(M->A, n, f);
            wb[q] = fld(M->W3B, n, f);
            oldb[q] = fld(SB, n, f);
        }
        launch_adj_fct(M->B, old, wrk, ob, wb, oldb, f0 >= BX ? d.diferg : d.difrho,
                       M->fct_scr, s, nf, true);
        f0 += nf;
    }
    // WE' = W3's pressure after its boundary pass
    {
        FieldPtrs dst{}, src{};
        dst.p[0] = fld(M->W3B, n, EP);
        src.p[0] = const_cast<float*>(fld((float*)sout_bar, n, WE));
        hipLaunchKernelGGL(k_add_all, dim3((n + 255) / 256, 1), dim3(256), 0, s, n, dst, src);
    }
    if (Fk) flanks_adj(M, M->W3B, *Fk, Fk->t1b, Fk->t2b);
    else outer_adj(M, M->W3B, W3, I, ctl_bar);
    // W3 = Lapidus of S2 on the interior, W1 on the boundary
    {
        FieldPtrs dst{}, src{};
        for (int f = 0; f < NFLD; ++f) {
            dst.p[f] = fld(M->W1B, n, f);
            src.p[f] = fld(M->W3B, n, f);
        }
        hipLaunchKernelGGL(k_add_bnd, dim3((bnd_cells(M->B) + 255) / 256, NFLD), dim3(256), 0,
                           s, M->B, dst, src);
    }
    for (int sp = 0; sp < 3; ++sp) {
        const float* pp[3] = {fld((float*)S2, n, 10 * sp + 4), fld((float*)S2, n, 10 * sp + 5),
                              fld((float*)S2, n, 10 * sp + 6)};
        launch_velocity(n, fld((float*)S2, n, 10 * sp), pp, M->vel, d.smallbit, s);
        zero_vb(M);
        LapAdj L{};
        L.B = M->B;
        for (int f = 0; f < 10; ++f) {
            L.q[f] = fld((float*)S2, n, 10 * sp + f);
            L.wb[f] = fld(M->W3B, n, 10 * sp + f);
            L.qb[f] = fld(M->S2B, n, 10 * sp + f);
            L.chi[f] = f == 0 ? d.chirho : d.chipxyz;
        }
        for (int a = 0; a < 3; ++a) {
            L.v[a] = M->vel[a];
            L.vb[a] = M->VB[a];
        }
        L.delt = delt;
        L.scr = M->lap_scr;
        L.nf = 10;
        L.assign = true;
        launch_adj_lap_species(L, s);
        const float* vb[3] = {M->VB[0], M->VB[1], M->VB[2]};
        float* pb[3] = {fld(M->S2B, n, 10 * sp + 4), fld(M->S2B, n, 10 * sp + 5),
                        fld(M->S2B, n, 10 * sp + 6)};
        launch_adj_velocity(n, fld((float*)S2, n, 10 * sp), pp, vb, d.smallbit,
                            fld(M->S2B, n, 10 * sp), pb, s);
    }
    {
        VtotAdj V{};
        const float* mom[9];
        const float* rho[3];
        for (int sp = 0; sp < 3; ++sp) {
            rho[sp] = fld((float*)S2, n, 10 * sp);
            V.rho[sp] = rho[sp];
            V.rhob[sp] = fld(M->S2B, n, 10 * sp);
            for (int a = 0; a < 3; ++a) {
                mom[3 * sp + a] = fld((float*)S2, n, 10 * sp + 4 + a);
                V.mom[3 * sp + a] = mom[3 * sp + a];
                V.momb[3 * sp + a] = fld(M->S2B, n, 10 * sp + 4 + a);
            }
        }
        launch_vtot(n, mom, rho, M->vel, s);
        zero_vb(M);
        LapAdj L{};
        L.B = M->B;
        for (int f = 0; f < 4; ++f) {
            L.q[f] = fld((float*)S2, n, EP + f);
            L.wb[f] = fld(M->W3B, n, EP + f);
            L.qb[f] = fld(M->S2B, n, EP + f);
            L.chi[f] = f == 0 ? d.chipxyz : d.chierg;
        }
        for (int a = 0; a < 3; ++a) {
            L.v[a] = M->vel[a];
            L.vb[a] = M->VB[a];
            V.vb[a] = M->VB[a];
        }
        L.delt = delt;
        L.scr = M->lap_scr;
        L.nf = 4;
        L.assign = true;
        launch_adj_lap_species(L, s);
        launch_adj_vtot(n, V, s);
    }
    // S2 = bndry_outer(push from W1 with old = S) (an inner box: no boundary pass)
    if (!Fk) outer_adj(M, M->S2B, S2, I, ctl_bar);
    {
        FieldPtrs dst{}, src{};
        for (int f = 0; f < NFLD; ++f) {
            dst.p[f] = fld(SB, n, f);
            src.p[f] = fld(M->S2B, n, f);
        }
        hipLaunchKernelGGL(k_add_bnd, dim3((bnd_cells(M->B) + 255) / 256, EP + 1), dim3(256), 0,
                           s, M->B, dst, src);
    }
    stage_adj(M, ef_at(fld(T, n, T_EF2), n), W1, S, M->S2B, M->W1B, SB, delt, true, ctl_bar,
              eta);
    // W1 = bndry_outer(push from S with delt/2; its pressure is WE) (or the flanks)
    if (Fk) flanks_adj(M, M->W1B, *Fk, Fk->t1a, Fk->t2a);
    else outer_adj(M, M->W1B, W1, I, ctl_bar);
    {
        FieldPtrs dst{}, src{};
        copyf(M, fld(SB, n, WE), fld(M->W1B, n, EP), 1);
        zero(M, fld(M->W1B, n, EP), n);
        for (int f = 0; f < NFLD; ++f) {
            dst.p[f] = fld(SB, n, f);
            src.p[f] = fld(M->W1B, n, f);
        }
        hipLaunchKernelGGL(k_add_bnd, dim3((bnd_cells(M->B) + 255) / 256, EP + 1), dim3(256), 0,
                           s, M->B, dst, src);
    }
    stage_adj(M, ef_at(fld(T, n, T_EF1), n), S, S, M->W1B, SB, SB, 0.5f * delt, false, ctl_bar,
              eta);
    return launched("adjoint") ? 0 : 2;
}

int g3m3_full_model_step(void* model, const float* s_in, float* s_out, float* tape,
                         const float* ctl) {
    return step_impl(model, s_in, s_out, tape, ctl, nullptr);
}

int g3m3_full_model_adjoint(void* model, const float* tape, const float* sout_bar,
                            float* sin_bar, double* ctl_bar, const float* ctl) {
    return adjoint_impl(model, tape, sout_bar, sin_bar, ctl_bar, ctl, nullptr);
}

int g3m3_full_model_set_parent(void* model, const float* geom) {
    FullModel* M = (FullModel*)model;
    if (!M || !geom) return 1;
    const g3m3_full_model_desc& d = M->d;
    FlankGeom G;
    // this box (fine, s*) and its parent (coarse, b*): origin and spacing
    G.sxmin = geom[0];
    G.symin = geom[2];
    G.szmin = geom[4];
    G.sx = (geom[1] - geom[0]) / (d.nx - 1.0f);
    G.sy = (geom[3] - geom[2]) / (d.ny - 1.0f);
    G.sz = (geom[5] - geom[4]) / (d.nz - 1.0f);
    G.bxmin = geom[6];
    G.bymin = geom[8];
    G.bzmin = geom[10];
    G.delx = (geom[7] - geom[6]) / (d.nx - 1.0f);
    G.dely = (geom[9] - geom[8]) / (d.ny - 1.0f);
    G.delz = (geom[11] - geom[10]) / (d.nz - 1.0f);
    const int n = shell_count(M->B);
    std::vector<ShellPt> h(n);
    build_shell(M->B, G, h.data());
    for (const ShellPt& p : h) {   // every trilinear corner inside the parent box
        const int x = p.cb % d.nx, y = (p.cb / d.nx) % d.ny, z = p.cb / (d.nx * d.ny);
        if (p.cb < 0 || x + 1 >= d.nx || y + 1 >= d.ny || z + 1 >= d.nz) {
            fprintf(stderr, "g3m3_full_model_set_parent: the shell reads outside the parent\n");
            return 1;
        }
    }
    // the corer's geometry (code/sub_corer.f90:17-32)
    M->c_nskip = (int)(G.delx / G.sx);
    M->c_nbx1 = (int)(1.0f + (G.sxmin - G.bxmin) / G.delx + 0.5f);
    M->c_nby1 = (int)(1.0f + (G.symin - G.bymin) / G.dely + 0.5f);
    M->c_nbz1 = (int)(1.0f + (G.szmin - G.bzmin) / G.delz + 0.5f);
    if (M->c_nskip < 1) return 1;
    {   // the corer's parent cells nb1 + 1 .. nb1 + count (1-based) inside the parent
        const int nb1[3] = {M->c_nbx1, M->c_nby1, M->c_nbz1}, ext[3] = {d.nx, d.ny, d.nz};
        for (int a = 0; a < 3; ++a) {
            const int cnt = corer_count(ext[a], M->c_nskip);
            if (cnt > 0 && (nb1[a] < 0 || nb1[a] + cnt > ext[a])) {
                fprintf(stderr, "g3m3_full_model_set_parent: the core lies outside the parent\n");
                return 1;
            }
        }
    }
    // the gather table of the flanks' adjoint: each touched parent cell with
    // its (shell point, trilinear weight) pairs, the weights as the forward
    // kernel forms them
    std::vector<std::vector<std::pair<int, float>>> by_cell;
    std::vector<int> slot((size_t)M->n, -1), cells;
    for (int t = 0; t < n; ++t) {
        const ShellPt& p = h[t];
        const float dx = p.dx, dy = p.dy, dz = p.dz;
        const float ddx = 1.0f - dx, ddy = 1.0f - dy, ddz = 1.0f - dz;
        const int X = 1, Y = p.sy, Z = p.sz, c0 = p.cb;
        const int cn[8] = {c0, c0 + Z, c0 + Y, c0 + Y + Z, c0 + X, c0 + X + Z, c0 + X + Y,
                           c0 + X + Y + Z};
        const float w8[8] = {ddx * ddy * ddz, ddx * ddy * dz, ddx * dy * ddz, ddx * dy * dz,
                             dx * ddy * ddz,  dx * ddy * dz,  dx * dy * ddz,  dx * dy * dz};
        for (int q = 0; q < 8; ++q) {
            if (slot[cn[q]] < 0) {
                slot[cn[q]] = (int)cells.size();
                cells.push_back(cn[q]);
                by_cell.emplace_back();
            }
            by_cell[slot[cn[q]]].push_back({t, w8[q]});
        }
    }
    // parent cells in ascending order: neighbouring threads of the gather
    // update neighbouring cells
    std::vector<int> order(cells.size());
    for (size_t i = 0; i < cells.size(); ++i) order[i] = (int)i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return cells[a] < cells[b]; });
    std::vector<int> start(cells.size() + 1, 0), pts, sorted_cells(cells.size());
    std::vector<float> wts;
    for (size_t i = 0; i < cells.size(); ++i) {
        sorted_cells[i] = cells[order[i]];
        for (const auto& e : by_cell[order[i]]) {
            pts.push_back(e.first);
            wts.push_back(e.second);
        }
        start[i + 1] = (int)pts.size();
    }
    cells.swap(sorted_cells);
    free_parent(M);
    auto up = [](void** dst, const void* src, size_t bytes) {
        return hipMalloc(dst, bytes) == hipSuccess &&
               hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!up((void**)&M->shell, h.data(), sizeof(ShellPt) * n) ||
        !up((void**)&M->fc_cell, cells.data(), sizeof(int) * cells.size()) ||
        !up((void**)&M->fc_start, start.data(), sizeof(int) * start.size()) ||
        !up((void**)&M->fc_p, pts.data(), sizeof(int) * pts.size()) ||
        !up((void**)&M->fc_w, wts.data(), sizeof(float) * wts.size()) ||
        hipMalloc(&M->fc_I0b, sizeof(float) * (size_t)n * NFLD) != hipSuccess) {
        free_parent(M);
        return 2;
    }
    M->nshell = n;
    M->nfc = (int)cells.size();
    return 0;
}

int g3m3_full_model_step_inner(void* model, const float* s_in, float* s_out, float* tape,
                               const float* ctl, const float* p_old, const float* p_cur,
                               const float* t12) {
    if (!t12) return 1;
    const Flank Fk{p_old, p_cur, t12[0], t12[1], t12[2], t12[3], nullptr, nullptr};
    return step_impl(model, s_in, s_out, tape, ctl, &Fk);
}

int g3m3_full_model_adjoint_inner(void* model, const float* tape, const float* sout_bar,
                                  float* sin_bar, double* ctl_bar, const float* ctl,
                                  const float* t12, float* p_old_bar, float* p_cur_bar) {
    if (!t12) return 1;
    const Flank Fk{nullptr, nullptr, t12[0], t12[1], t12[2], t12[3], p_old_bar, p_cur_bar};
    return adjoint_impl(model, tape, sout_bar, sin_bar, ctl_bar, ctl, &Fk);
}

int g3m3_full_model_corer(void* model, const float* s, float* p_state) {
    FullModel* M = (FullModel*)model;
    if (!M || !s || !p_state || !M->shell) return 1;
    FieldPtrs fine{}, coarse{};
    for (int f = 0; f < NFLD; ++f) {
        fine.p[f] = fld((float*)s, M->n, f);
        coarse.p[f] = fld(p_state, M->n, f);
    }
    launch_corer(M->B, fine, coarse, NFLD, M->c_nskip, M->c_nbx1, M->c_nby1, M->c_nbz1, M->s);
    return launched("corer") ? 0 : 2;
}

int g3m3_full_model_corer_adjoint(void* model, float* s_bar, float* p_bar) {
    FullModel* M = (FullModel*)model;
    if (!M || !s_bar || !p_bar || !M->shell) return 1;
    FieldPtrs fb{}, cb{};
    for (int f = 0; f < NFLD; ++f) {
        fb.p[f] = fld(s_bar, M->n, f);
        cb.p[f] = fld(p_bar, M->n, f);
    }
    CorerGeom G;
    G.nskip = M->c_nskip;
    G.nb[0] = M->c_nbx1;
    G.nb[1] = M->c_nby1;
    G.nb[2] = M->c_nbz1;
    G.cnt[0] = corer_count(M->B.nx, G.nskip);
    G.cnt[1] = corer_count(M->B.ny, G.nskip);
    G.cnt[2] = corer_count(M->B.nz, G.nskip);
    if (G.cnt[0] > 0 && G.cnt[1] > 0 && G.cnt[2] > 0) {
        hipLaunchKernelGGL(k_adj_corer, dim3((M->n + 255) / 256), dim3(256), 0, M->s, M->B, fb,
                           cb, G);
        hipLaunchKernelGGL(k_zero_core, dim3((G.cnt[0] * G.cnt[1] * G.cnt[2] + 255) / 256, NFLD),
                           dim3(256), 0, M->s, M->B, cb, G);
    }
    return launched("corer adjoint") ? 0 : 2;
}

}  // extern "C"
