// Native runtime of the hot path: one run (all boxes, all fields) resident
// in HBM, advanced a full step per call with the fused kernels.
//
// The host logic restates code/3d_main.f90:1925-3095 -- dt
// synchronisation across boxes, the subcycling schedule, the box-step call
// sequence with the reference's boundary quirks, nested-grid sync -- and
// issues only kernel launches; the one device->host read per full step is
// the CFL maxima of each box, for t_stepnew (code/3d_main.f90:3038).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/g3m3.h"
#include "g3m3_fused.h"

using namespace g3m3;

namespace {

enum Fld {   // canonical field order
    QRHO = 0, HRHO = 10, ORHO = 20, EPRES = 30, BX = 31, BY = 32, BZ = 33
};
constexpr int NF = 34;
// the main loop's smallest time step (code/3d_main.f90:116) and the
// run error a smaller (or non-finite) one latches
constexpr float kTStepMin = 1.e-6f;
constexpr int kErrStepCollapsed = -2;

// one box's slice of every field
struct Set {
    float* f[NF];
};

struct Run {
    g3m3_run_desc d;
    Box B;
    int n;        // cells per box
    hipStream_t s = nullptr;
    // scratch (one box)
    float* scr = nullptr;         // e(3) ev(3) tv(3) bs(3)
    ValEntry* list = nullptr;
    int* count = nullptr;
    float* partials = nullptr;
    float* maxima = nullptr;      // 5 per box
    float* rst_eff = nullptr;     // rst on BandE's list points of box mbndry
    double t = 0.0;
    std::vector<double> t_old, t_new, t_step, t_stepnew;
    // the main loop's tilt and universal time, now and at the step's start
    // (code/3d_main.f90:1966-1977)
    float tilt = 0.0f, old_tilt = 0.0f;
    double ut = 0.0, utold = 0.0;
    long long cell_updates = 0, box_steps = 0;
    float iso_flag = 1.0f;
    int err = 0;
    // Buffer rotation: instead of copying the state into old at the start of
    // a box-step (code/3d_main.f90:2193-2229) the two slices swap roles.
    // rot[b] set: box b's state lives in the old array and its old state in
    // the state array.
    std::vector<char> rot;
    int vpass = 0;                // which validation counter the next pass uses
    std::vector<int> stepped;     // boxes stepped by the last issued full step
    // the box-steps of the last issued full step in the order taken: box,
    // t_old and t_new after the step's clock update (the speeds.dat record
    // of code/3d_main.f90:2167-2168)
    struct StepRec {
        int box;
        double t_old, t_new;
    };
    std::vector<StepRec> step_log;
    bool lists_disjoint = false;  // the planet lists share no cell: one launch for all three
    // flank shell tables of the inner boxes (geometry is fixed for a run)
    ShellPt* shell[G3M3_MAX_GRIDS] = {};
    int nshell = 0;
    // The coarse cells inner box b+1's shell reads from box b+2, gathered
    // (old and state, every field) into a dense slab once per coarse
    // change: the shell's x faces would otherwise fetch a cache line per
    // 2-3 coarse values on each of the six flank passes a coarse step
    // serves.  slab_ok[b] clears whenever box b+2's state or old may change.
    float* slab[G3M3_MAX_GRIDS] = {};
    int* slab_idx[G3M3_MAX_GRIDS] = {};
    int nslab[G3M3_MAX_GRIDS] = {};
    bool slab_ok[G3M3_MAX_GRIDS] = {};
    // captured full steps of a single-grid run: gexec[p] is one full step
    // starting from rotation parity p; the time step and the run time live
    // on the device (dt_dev[0] = the next step's dt, [1] = t_stepnew,
    // [2] = the dt of the step last taken; t_dev = t)
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    float* dt_dev = nullptr;
    double* t_dev = nullptr;
    float* host_buf = nullptr;    // pinned: t_stepnew and t read back
    bool capturing = false;       // kernels read the time step from dt_dev
    bool own_stream = false;
    // the FCT's launch shape, picked by tune_fct: k-planes per block and
    // block order, with the measured time of each candidate
    int fct_kc = 16, fct_no_remap = 0;
    // species push, picked by tune_push: 1 = k_push_march, 0 = one thread
    // per cell (k_push), with the measured time of each
    int push_march = 1;
    float push_us[2] = {};
    float fct_us[6] = {};

    Set slice(float* const* arrays, int box) const {
        Set S;
        for (int f = 0; f < NF; ++f) S.f[f] = arrays[f] + (size_t)(box - 1) * n;
        return S;
    }
    Set state(int box) const { return slice(rot[box - 1] ? d.old : d.state, box); }
    Set old(int box) const { return slice(rot[box - 1] ? d.state : d.old, box); }
    Set wrk(int box) const { return slice(d.wrk, box); }
};

// true when e is hipSuccess; otherwise the run's first error is recorded
bool ok_hip(Run* R, hipError_t e, const char* where) {
    if (e == hipSuccess) return true;
    if (R->err == 0) {
        R->err = (int)e;
        fprintf(stderr, "g3m3 run: %s: %s\n", where, hipGetErrorString(e));
    }
    return false;
}

void check(Run* R, const char* where) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess && R->err == 0) {
        R->err = (int)e;
        fprintf(stderr, "g3m3 run: %s: %s\n", where, hipGetErrorString(e));
    }
}

bool in_box(const Box& B, int i, int j, int k) {
    return i >= 0 && i < B.nx && j >= 0 && j < B.ny && k >= 0 && k < B.nz;
}

// x is +-2^n (so 1/x is exact and x / y == x * (1/y) bit for bit)

SpeciesSet species(const Set& S, int s) {
    SpeciesSet Q;
    for (int m = 0; m < 10; ++m) Q.f[m] = S.f[10 * s + m];
    return Q;
}

// -- stage pieces ---------------------------------------------------------------
void efield(Run* R, const Set& src, int box) {
    const g3m3_run_desc& d = R->d;
    EParams P{};
    P.B = R->B;
    for (int c = 0; c < 3; ++c) {
        P.b[c] = src.f[BX + c];
        P.b0[c] = (c == 0 ? d.bx0 : c == 1 ? d.by0 : d.bz0) + (size_t)(box - 1) * R->n;
        P.rho[c] = src.f[10 * c];
        for (int m = 0; m < 3; ++m) P.mom[3 * c + m] = src.f[10 * c + 4 + m];
        P.e[c] = R->scr + (size_t)(0 + c) * R->n;
        P.ev[c] = R->scr + (size_t)(3 + c) * R->n;
        P.tv[c] = R->scr + (size_t)(6 + c) * R->n;
        P.bs[c] = R->scr + (size_t)(9 + c) * R->n;
    }
    P.ep = src.f[EPRES];
    P.rst = (box <= d.mbndry && d.resist < 5000.0f) ? R->rst_eff : nullptr;
    P.rmq = d.rmassq;
    P.rmh = d.rmassh;
    P.rmo = d.rmasso;
    P.reynolds = d.reynolds;
    const float rx = d.xspac[box - 1];
    P.dxt = 2.0f * rx;
    P.dyt = 2.0f * rx;
    P.dzt = 2.0f * rx;
    P.eta = d.eta;
    const float ds[7] = {P.rmq, P.rmh, P.rmo, P.reynolds, P.dxt, P.dyt, P.dzt};
    float* rs[7] = {&P.rrmq, &P.rrmh, &P.rrmo, &P.rreynolds, &P.rdxt, &P.rdyt, &P.rdzt};
    bool p2 = true;
    for (int q = 0; q < 7; ++q) {
        *rs[q] = 1.0f / ds[q];
        p2 = p2 && is_pow2(ds[q]);
    }
    P.pow2 = p2 ? 1 : 0;
    launch_efield(P, R->s);
}

void push(Run* R, const Set& out, const Set& old, const Set& src, int box, float delt,
          float dt_scale, int do_elec, int copy_shell) {
    const g3m3_run_desc& d = R->d;
    PushParams P{};
    P.B = R->B;
    for (int s = 0; s < 3; ++s) {
        P.src[s] = species(src, s);
        P.old[s] = species(old, s);
        P.out[s] = species(out, s);
    }
    for (int c = 0; c < 3; ++c) {
        P.e[c] = R->scr + (size_t)(0 + c) * R->n;
        P.ev[c] = R->scr + (size_t)(3 + c) * R->n;
        P.tv[c] = R->scr + (size_t)(6 + c) * R->n;
        P.bs[c] = R->scr + (size_t)(9 + c) * R->n;
        P.bout[c] = out.f[BX + c];
        P.bold[c] = old.f[BX + c];
    }
    P.wrkepres = R->wrk(box).f[EPRES];
    P.oldepres = old.f[EPRES];
    P.epres = out.f[EPRES];
    P.rmass[0] = d.rmassq;
    P.rmass[1] = d.rmassh;
    P.rmass[2] = d.rmasso;
    P.delt = delt;
    P.reynolds = d.reynolds;
    P.smallbit = d.smallbit;
    const int b = box - 1;
    P.ddx = (d.grd_xmax[b] - d.grd_xmin[b]) / (float)(d.nx - 1);
    P.ddy = (d.grd_ymax[b] - d.grd_ymin[b]) / (float)(d.ny - 1);
    P.ddz = (d.grd_zmax[b] - d.grd_zmin[b]) / (float)(d.nz - 1);
    P.dxt = 2.0f * P.ddx;
    P.dyt = 2.0f * P.ddy;
    P.dzt = 2.0f * P.ddz;
    const float rx = d.xspac[b];
    P.fdxt = 2.0f * rx;
    P.fdyt = 2.0f * rx;
    P.fdzt = 2.0f * rx;
    P.xmin = d.grd_xmin[b];
    P.ymin = d.grd_ymin[b];
    P.zmin = d.grd_zmin[b];
    P.re_equiv = d.re_equiv;
    P.grav = d.grav;
    P.gamma = d.gamma;
    P.gamma1 = d.gamma1;
    bool p2 = true;
    for (int q = 0; q < 3; ++q) {
        P.bmax[q] = 0.20f * P.rmass[q] / delt / d.reynolds;
        P.rrmass[q] = 1.0f / P.rmass[q];
        p2 = p2 && is_pow2(P.rmass[q]);
    }
    const float ds[6] = {P.dxt, P.dyt, P.dzt, P.fdxt, P.fdyt, P.fdzt};
    float* rs[6] = {&P.rdxt, &P.rdyt, &P.rdzt, &P.rfdxt, &P.rfdyt, &P.rfdzt};
    for (int q = 0; q < 6; ++q) {
        *rs[q] = 1.0f / ds[q];
        p2 = p2 && is_pow2(ds[q]);
    }
    P.pow2 = p2 ? 1 : 0;
    P.do_elec = do_elec;
    P.copy_shell = copy_shell;
    P.march = R->push_march;
    P.nspecies = 3;
    P.btot = nullptr;
    P.dt_dev = R->capturing ? R->dt_dev : nullptr;
    P.dt_scale = dt_scale;
    launch_push(P, R->s);
}

void outer_bc(Run* R, const Set& set) {
    const g3m3_run_desc& d = R->d;
    const int box = d.n_grids;
    const float ofrac = d.rho_frac * d.o_conc;
    const float frac_h = d.rmassh / d.rmassq;
    const float frac_o = d.rmasso / d.rmassq;
    const float pq = 0.5f * d.spress;
    OuterPtrs O{};
    const float rf = d.rho_frac;
    const float sval[3][10] = {
        {d.srho, pq, pq, pq, d.spx, d.spy, d.spz, 0.f, 0.f, 0.f},
        {d.srho * rf * frac_h, 0.5f * d.spress * rf, 0.5f * d.spress * rf, 0.5f * d.spress * rf,
         d.spx * rf * frac_h, d.spy * rf * frac_h, d.spz * rf * frac_h, 0.f, 0.f, 0.f},
        {d.srho * ofrac * frac_o, 0.5f * d.spress * ofrac, 0.5f * d.spress * ofrac,
         0.5f * d.spress * ofrac, d.spx * ofrac * frac_o, d.spy * ofrac * frac_o,
         d.spz * ofrac * frac_o, 0.f, 0.f, 0.f}};
    const int kinds[10] = {kRHO, kPRES, kPRES, kPRES, kPX, kPY, kPZ, kOFF, kOFF, kOFF};
    for (int s = 0; s < 3; ++s)
        for (int m = 0; m < 10; ++m) {
            O.f[10 * s + m] = set.f[10 * s + m];
            O.kind[10 * s + m] = kinds[m];
            O.val[10 * s + m] = sval[s][m];
        }
    O.f[EPRES] = set.f[EPRES];
    O.kind[EPRES] = kEP;
    O.val[EPRES] = pq / d.ti_te;
    const float sb[3] = {d.sbx_wind, d.sby_wind, d.sbz_wind};
    for (int c = 0; c < 3; ++c) {
        O.f[BX + c] = set.f[BX + c];
        O.kind[BX + c] = kBX + c;
        O.val[BX + c] = sb[c];
    }
    O.b0[0] = d.bx0 + (size_t)(box - 1) * R->n;
    O.b0[1] = d.by0 + (size_t)(box - 1) * R->n;
    O.b0[2] = d.bz0 + (size_t)(box - 1) * R->n;
    launch_bndry_outer(R->B, O, NF, R->s);
}

FlankGeom geom(const g3m3_run_desc& d, int box) {
    const int mb = box, ms = box - 1;
    FlankGeom G;
    G.delx = (d.grd_xmax[mb] - d.grd_xmin[mb]) / (d.nx - 1.0f);
    G.dely = (d.grd_ymax[mb] - d.grd_ymin[mb]) / (d.ny - 1.0f);
    G.delz = (d.grd_zmax[mb] - d.grd_zmin[mb]) / (d.nz - 1.0f);
    G.sx = (d.grd_xmax[ms] - d.grd_xmin[ms]) / (d.nx - 1.0f);
    G.sy = (d.grd_ymax[ms] - d.grd_ymin[ms]) / (d.ny - 1.0f);
    G.sz = (d.grd_zmax[ms] - d.grd_zmin[ms]) / (d.nz - 1.0f);
    G.sxmin = d.grd_xmin[ms];
    G.symin = d.grd_ymin[ms];
    G.szmin = d.grd_zmin[ms];
    G.bxmin = d.grd_xmin[mb];
    G.bymin = d.grd_ymin[mb];
    G.bzmin = d.grd_zmin[mb];
    return G;
}

// box's state or old may have changed: the slab of the box inside it is stale
void coarse_changed(Run* R, int box) {
    if (box >= 2) R->slab_ok[box - 2] = false;
}
void all_coarse_changed(Run* R) {
    for (bool& v : R->slab_ok) v = false;
}

void flanks(Run* R, const Set& tgt, int box, double t_grid) {
    const int mb = box;   // 0-based index of box+1
    const float t1 = (float)((R->t_new[mb] - t_grid) / (R->t_new[mb] - R->t_old[mb]));
    const float t2 = 1.0f - t1;
    const int b = box - 1, n = R->nslab[b];
    FlankPtrs P{};
    if (!R->slab_ok[b]) {
        const Set cur = R->state(box + 1), old = R->old(box + 1);
        for (int f = 0; f < NF; ++f) {
            P.cur[f] = cur.f[f];
            P.old[f] = old.f[f];
        }
        launch_slab_gather(R->slab_idx[b], n, P, NF, R->slab[b], R->s);
        R->slab_ok[b] = true;
    }
    for (int f = 0; f < NF; ++f) {
        P.tgt[f] = tgt.f[f];
        P.old[f] = R->slab[b] + (size_t)2 * f * n;
        P.cur[f] = R->slab[b] + (size_t)(2 * f + 1) * n;
    }
    launch_flanks_table(R->B, R->shell[b], R->nshell, P, NF, t1, t2, R->s);
}

void inner_bc(Run* R, const Set& set) {
    const g3m3_run_desc& d = R->d;
    const int box = d.mbndry;
    InnerPtrs I{};
    // InnerPtrs order: q,h,o x (rho,presx,presy,presz,px,py,pz,presxy,presxz,presyz),
    // epres, bx, by, bz -- the canonical order
    for (int f = 0; f < NF; ++f) I.f[f] = set.f[f];
    const int* ij[3] = {d.ijsrf, d.ijmid, d.ijzero};
    const float* parm[3] = {d.parm_srf, d.parm_mid, d.parm_zero};
    const int num[3] = {d.numsrf, d.nummid, d.numzero};
    const int b = box - 1;
    const float dx = (d.grd_xmax[b] - d.grd_xmin[b]) / (float)(d.nx - 1.0f);
    const float dy = (d.grd_ymax[b] - d.grd_ymin[b]) / (float)(d.ny - 1.0f);
    launch_bndry_inner(R->B, I, ij, parm, num, d.mbndry, box, d.grd_xmin[b], d.grd_ymin[b], dx, dy,
                       d.re_equiv, d.v_rot, R->s, R->lists_disjoint);
}

void validate(Run* R, const Set& set, int box, bool want_max) {
    const g3m3_run_desc& d = R->d;
    ValParams Q{};
    Q.B = R->B;
    for (int s = 0; s < 3; ++s) Q.sp[s] = species(set, s);
    Q.ep = set.f[EPRES];
    for (int c = 0; c < 3; ++c) {
        Q.b[c] = set.f[BX + c];
        Q.b0[c] = (c == 0 ? d.bx0 : c == 1 ? d.by0 : d.bz0) + (size_t)(box - 1) * R->n;
    }
    Q.list = R->list;
    Q.count = R->count + 16 * R->vpass;           // separate 64-byte lines
    Q.count_next = R->count + 16 * (R->vpass ^ 1);
    R->vpass ^= 1;
    Q.partials = want_max ? R->partials : nullptr;
    if (d.isotropic) R->iso_flag = 0.0f;
    Q.aniso = d.aniso_limit;
    Q.smallbit = d.smallbit;
    Q.vlim = d.vlim_factor * (float)box;
    Q.cslim = d.cslim;
    Q.iso_flag = R->iso_flag;
    Q.isotropic = d.isotropic ? 1 : 0;
    Q.svlim[0] = d.qvlim;
    Q.svlim[1] = d.hvlim;
    Q.svlim[2] = d.ovlim;
    Q.sclim[0] = d.qclim;
    Q.sclim[1] = d.hclim;
    Q.sclim[2] = d.oclim;
    launch_validate(Q, want_max ? R->maxima + 5 * (box - 1) : nullptr, R->s);
}

void set_rho_only(Run* R, const Set& set) {
    SpeciesPtrs S{};
    for (int s = 0; s < 3; ++s)
        for (int m = 0; m < 10; ++m) S.f[s][m] = set.f[10 * s + m];
    launch_set_rho(R->n, S, set.f[EPRES], R->s);
}

void bounds(Run* R, const Set& set, const Set& flank_tgt, int box, double t_grid,
            bool do_flanks = true) {
    if (box == R->d.n_grids) outer_bc(R, set);
    else if (do_flanks) flanks(R, flank_tgt, box, t_grid);
    if (box <= R->d.mbndry) inner_bc(R, set);
}

// into wrk, or into *dst (the autotuner's scratch)
void lapidus(Run* R, int box, float delt, const Set* dst = nullptr) {
    const g3m3_run_desc& d = R->d;
    LapParams L{};
    L.B = R->B;
    const Set st = R->state(box), wk = dst ? *dst : R->wrk(box);
    for (int s = 0; s < 3; ++s) {
        L.src[s] = species(st, s);
        L.dst[s] = species(wk, s);
    }
    L.ep_src = st.f[EPRES];
    L.ep_dst = wk.f[EPRES];
    for (int c = 0; c < 3; ++c) {
        L.b_src[c] = st.f[BX + c];
        L.b_dst[c] = wk.f[BX + c];
    }
    L.chirho = d.chirho;
    L.chipxyz = d.chipxyz;
    L.chierg = d.chierg;
    L.delt = delt;
    L.smallbit = d.smallbit;
    L.dt_dev = R->capturing ? R->dt_dev : nullptr;
    L.march = R->push_march;
    launch_lap(L, R->s);
}

void fct(Run* R, int box) {
    const g3m3_run_desc& d = R->d;
    FctParams F{};
    F.B = R->B;
    const Set st = R->state(box), ol = R->old(box), wk = R->wrk(box);
    for (int f = 0; f < NF; ++f) {
        F.out[f] = st.f[f];
        F.old[f] = ol.f[f];
        F.wrk[f] = wk.f[f];
        F.chi[f] = (f >= BX) ? d.diferg : d.difrho;
    }
    F.kc_len = R->fct_kc;
    F.no_remap = R->fct_no_remap;
    launch_fct_fused(F, NF, R->s);
}

// Autotuning of the FCT's launch shape: the k-planes per block (more
// planes re-load less halo, fewer planes put more blocks in flight) and the
// block order (XCD-contiguous or plain) are timed on the run's own box --
// box 1's arrays as inputs, a scratch output -- and the fastest is kept.
// Every candidate gives the same bits.
constexpr int kFctDepths[3] = {8, 16, 32};

// Times one step-2 push and one Lapidus pass of box 1 (every output into
// scratch) with each family of species kernels and keeps the faster.  Both
// give the same bits.
void tune_push(Run* R) {
    float* out = nullptr;
    if (hipMalloc(&out, sizeof(float) * (size_t)NF * R->n) != hipSuccess) return;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    Set o;
    for (int f = 0; f < NF; ++f) o.f[f] = out + (size_t)f * R->n;
    const Set ol = R->old(1), wk = R->wrk(1);
    float best = 1e30f;
    const int keep = R->push_march;
    int pick = keep;
    for (int m = 0; ok && m < 2; ++m) {
        R->push_march = m;
        push(R, o, ol, wk, 1, 1.0e-3f, 1.0f, 1, 1);   // warm-up
        lapidus(R, 1, 1.0e-3f, &o);
        ok = hipEventRecord(e0, R->s) == hipSuccess;
        for (int q = 0; q < 3; ++q) {
            push(R, o, ol, wk, 1, 1.0e-3f, 1.0f, 1, 1);
            lapidus(R, 1, 1.0e-3f, &o);
        }
        float ms = 0.0f;
        ok = ok && hipEventRecord(e1, R->s) == hipSuccess &&
             hipEventSynchronize(e1) == hipSuccess &&
             hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        if (!ok) break;
        R->push_us[m] = ms * 1000.0f / 3.0f;
        if (R->push_us[m] < best) {
            best = R->push_us[m];
            pick = m;
        }
    }
    R->push_march = ok ? pick : keep;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
}

void tune_fct(Run* R) {
    float* out = nullptr;
    if (hipMalloc(&out, sizeof(float) * (size_t)NF * R->n) != hipSuccess) return;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    const g3m3_run_desc& d = R->d;
    FctParams F{};
    F.B = R->B;
    const Set ol = R->old(1), wk = R->wrk(1);
    for (int f = 0; f < NF; ++f) {
        F.out[f] = out + (size_t)f * R->n;
        F.old[f] = ol.f[f];
        F.wrk[f] = wk.f[f];
        F.chi[f] = (f >= BX) ? d.diferg : d.difrho;
    }
    float best = 1e30f;
    for (int c = 0; ok && c < 3; ++c)
        for (int rm = 0; ok && rm < 2; ++rm) {
            F.kc_len = kFctDepths[c];
            F.no_remap = rm;
            launch_fct_fused(F, NF, R->s);   // warm-up
            ok = hipEventRecord(e0, R->s) == hipSuccess;
            for (int q = 0; q < 3; ++q) launch_fct_fused(F, NF, R->s);
            float ms = 0.0f;
            ok = ok && hipEventRecord(e1, R->s) == hipSuccess &&
                 hipEventSynchronize(e1) == hipSuccess &&
                 hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
            if (!ok) break;
            const float us = ms * 1000.0f / 3.0f;
            R->fct_us[2 * c + rm] = us;
            if (us < best) {
                best = us;
                R->fct_kc = kFctDepths[c];
                R->fct_no_remap = rm;
            }
        }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
}

void corer(Run* R, int box) {
    const g3m3_run_desc& d = R->d;
    coarse_changed(R, box + 1);
    const int mb = box, ms = box - 1;
    const float sx = (d.grd_xmax[ms] - d.grd_xmin[ms]) / (d.nx - 1.0f);
    const float dx = (d.grd_xmax[mb] - d.grd_xmin[mb]) / (d.nx - 1.0f);
    const float dy = (d.grd_ymax[mb] - d.grd_ymin[mb]) / (d.ny - 1.0f);
    const float dz = (d.grd_zmax[mb] - d.grd_zmin[mb]) / (d.nz - 1.0f);
    const int nskip = (int)(dx / sx);
    const int nbx1 = (int)(1.0f + (d.grd_xmin[ms] - d.grd_xmin[mb]) / dx + 0.5f);
    const int nby1 = (int)(1.0f + (d.grd_ymin[ms] - d.grd_ymin[mb]) / dy + 0.5f);
    const int nbz1 = (int)(1.0f + (d.grd_zmin[ms] - d.grd_zmin[mb]) / dz + 0.5f);
    FieldPtrs fine{}, coarse{};
    const Set fs = R->state(box), cs = R->state(box + 1);
    for (int f = 0; f < NF; ++f) {
        fine.p[f] = fs.f[f];
        coarse.p[f] = cs.f[f];
    }
    launch_corer(R->B, fine, coarse, NF, nskip, nbx1, nby1, nbz1, R->s);
}

// b0 of one box for the tilt and rotation at t_old + dt_off, in the
// reference's mixed precision: before step 1 at the half step
// (code/3d_main.f90:2173-2187), before step 2 at the full step (:2465-2478)
void tilt_box(Run* R, int box, float dt_off) {
    const g3m3_run_desc& d = R->d;
    const int b = box - 1;
    const float pi = 3.14159274f;   // 4*atan(1.0) in single precision
    const double tb = R->t_old[b] + (double)dt_off;
    const float atilt = (float)((double)R->old_tilt + tb * (double)d.dtilt);
    const float sin_tilt = std::sin(atilt * pi / 180.0f);   // the C library's sinf / cosf
    const float cos_tilt = std::cos(atilt * pi / 180.0f);
    const float ut2 = (float)(R->utold + tb * (double)d.t_equiv / 3600.0);
    // a planet without rotation (period 0) has no rotation angle; the
    // reference's division by 0 there has no defined integer result
    float rot_angle = 0.0f;
    if (d.planet_per != 0.0f) {
        const int nrot2 = (int)(ut2 / d.planet_per);
        const float rot_hr2 = ut2 - (float)nrot2 * d.planet_per;
        rot_angle = 2.0f * pi * rot_hr2 / d.planet_per;
    }
    RotationGeom rot;
    rot.xdip = d.xdip;
    rot.ydip = d.ydip;
    rot.zdip = d.zdip;
    rot.sin_rot = std::sin(rot_angle);
    rot.cos_rot = std::cos(rot_angle);
    rot.sin_tilt = sin_tilt;
    rot.cos_tilt = cos_tilt;
    rot.b0 = d.dip_b0;
    InitGeom G;
    G.xmin = d.grd_xmin[b];
    G.ymin = d.grd_ymin[b];
    G.zmin = d.grd_zmin[b];
    G.dx = (d.grd_xmax[b] - d.grd_xmin[b]) / (d.nx - 1.0f);
    G.dy = (d.grd_ymax[b] - d.grd_ymin[b]) / (d.ny - 1.0f);
    G.dz = (d.grd_zmax[b] - d.grd_zmin[b]) / (d.nz - 1.0f);
    const size_t off = (size_t)b * R->n;
    launch_mak_dip_grd(R->B, const_cast<float*>(d.bx0) + off, const_cast<float*>(d.by0) + off,
                       const_cast<float*>(d.bz0) + off, G, rot, d.r_inner, d.ijzero, d.mbndry,
                       box, box <= d.mbndry ? d.numzero : 0, R->s);
}

// -- code/3d_main.f90:2163-3065 ----------------------------------------------------
void box_step(Run* R, int box) {
    g3m3_run_desc& d = R->d;
    const int b = box - 1;
    R->t_old[b] = R->t_new[b];
    R->t_new[b] = R->t_old[b] + R->t_step[b];
    R->step_log.push_back(Run::StepRec{box, R->t_old[b], R->t_new[b]});
    const float delt = (float)R->t_step[b];
    const float delt2 = delt / 2.0f;
    if (d.tilting) tilt_box(R, box, delt2);
    // store_array x34 (:2193-2229) by rotation: the state becomes old and
    // step 2 writes the new state into the other slice (its push copies the
    // cells it does not compute from old)
    R->rot[b] ^= 1;
    coarse_changed(R, box);
    const Set st = R->state(box), ol = R->old(box), wk = R->wrk(box);
    auto clamp_t = [&](double tg) {
        if (tg > R->t_new[b + 1]) tg = R->t_new[b + 1];
        if (tg < R->t_old[b + 1]) tg = R->t_old[b + 1];
        return tg;
    };
    const bool inner_box = box < d.n_grids;
    // step 1 (:2243-2458), reading the state from old (same values); its
    // push_elec only writes the state epres interior, which step 2
    // overwrites before any read, so it is not issued
    efield(R, ol, box);
    push(R, wk, ol, ol, box, 0.5f * delt, 0.5f, 0, 0);
    double tg = inner_box ? clamp_t(R->t_old[b] + (double)delt2) : 0.0;
    bounds(R, wk, wk, box, tg);
    validate(R, wk, box, false);
    // step 2 (:2476-2684).  Inner boxes put these flanks into wrk, where
    // nothing reads them: the Lapidus stage overwrites the wrk interior
    // (layer 2 of the shell) and its own flanks, same t_grid, set layer 1
    // to the same interpolant with weight 1 (w = 0*w + I0) before any read.
    // So the call is not issued (results equal up to the sign of an exact
    // zero).
    if (d.tilting) tilt_box(R, box, delt);
    efield(R, wk, box);
    push(R, st, ol, wk, box, delt, 1.0f, 1, 1);
    tg = inner_box ? clamp_t(R->t_old[b] + (double)delt) : 0.0;
    bounds(R, st, wk, box, tg, false);
    validate(R, st, box, false);
    // Lapidus (:2692-2890), same t_grid
    lapidus(R, box, delt);
    bounds(R, wk, wk, box, tg);
    validate(R, wk, box, false);
    // FCT (:2896-3030)
    fct(R, box);
    bounds(R, st, st, box, tg);
    validate(R, st, box, true);
    // nested sync (:3045-3063)
    if (inner_box && std::fabs(R->t_new[b] - R->t_new[b + 1]) <= 0.005 * R->t_step[b]) {
        corer(R, box);
        R->t_old[b + 1] = R->t_new[b + 1];
    }
    R->cell_updates += R->n;
    R->box_steps += 1;
    check(R, "box_step");
}

void read_tstepnew(Run* R, const std::vector<int>& boxes) {
    std::vector<float> h(5 * R->d.n_grids);
    if (!ok_hip(R,
                hipMemcpyAsync(h.data(), R->maxima, h.size() * sizeof(float),
                               hipMemcpyDeviceToHost, R->s),
                "read CFL maxima") ||
        !ok_hip(R, hipStreamSynchronize(R->s), "read CFL maxima"))
        return;
    for (int box : boxes) {
        const float* m = &h[5 * (box - 1)];
        float f = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2] + m[3] * m[3] + m[4] * m[4]);
        f = std::max(0.0f, f);
        const float st = R->d.stepsz * R->d.xspac[box - 1] / f;
        R->t_stepnew[box - 1] = (double)std::min(st, R->d.t_step_max);
    }
}

// The next time step of a single-grid run from the CFL maxima of its last
// validation pass, as read_tstepnew forms it on the host, and the run time
// advanced by the step just taken (code/3d_main.f90:1929-1968, 3038).
// dt[3] becomes 1 (and stays so) when a step is formed below t_step_min or
// not finite, where the main loop would stop (:1969-1973).
__global__ void k_next_dt(const float* m, float stepsz, float xspac, float t_step_max,
                          float* dt, double* t) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *t = *t + (double)dt[0];
    dt[2] = dt[0];
    float f = sqrtf(m[0] * m[0] + m[1] * m[1] + m[2] * m[2] + m[3] * m[3] + m[4] * m[4]);
    f = (0.0f < f) ? f : 0.0f;              // std::max(0.0f, f)
    const float st = stepsz * xspac / f;
    const float nd = (t_step_max < st) ? t_step_max : st;   // std::min(st, t_step_max)
    dt[1] = nd;
    dt[0] = nd;
    if (!(nd >= kTStepMin && nd <= 3.0e38f)) dt[3] = 1.0f;
}

__global__ void k_set_dt(float* dt, double* t, float dt0, double t0) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    dt[0] = dt0;
    dt[1] = dt0;
    dt[2] = dt0;
    dt[3] = 0.0f;
    *t = t0;
}

// the device time step and run time of captured or batched steps
bool graph_scalars(Run* R) {
    if (R->dt_dev) return true;
    return ok_hip(R, hipMalloc(&R->dt_dev, 4 * sizeof(float)), "graph scalars") &&
           ok_hip(R, hipMalloc(&R->t_dev, sizeof(double)), "graph scalars") &&
           ok_hip(R, hipHostMalloc(&R->host_buf, 4 * sizeof(double)), "graph scalars");
}

// Captures the two full steps (one per rotation parity) of a single-grid
// run; host bookkeeping touched by the capture is restored
bool build_graphs(Run* R) {
    // captured steps replay fixed kernel arguments: a tilting dipole's
    // rotation is formed on the host per box-step
    if (R->d.n_grids != 1 || R->d.tilting) return false;
    if (!R->s) {
        if (!ok_hip(R, hipStreamCreate(&R->s), "graph stream")) return false;
        R->own_stream = true;
    }
    if (!graph_scalars(R)) return false;
    const std::vector<double> t_old = R->t_old, t_new = R->t_new;
    const long long cu = R->cell_updates, bs = R->box_steps;
    const int vpass = R->vpass;
    bool ok = true;
    for (int k = 0; ok && k < 2; ++k) {
        const int p = R->rot[0];
        R->capturing = true;
        ok = ok_hip(R, hipStreamBeginCapture(R->s, hipStreamCaptureModeThreadLocal),
                    "begin capture");
        if (ok) {
            box_step(R, 1);
            hipLaunchKernelGGL(k_next_dt, dim3(1), dim3(64), 0, R->s, R->maxima, R->d.stepsz,
                               R->d.xspac[0], R->d.t_step_max, R->dt_dev, R->t_dev);
            ok = ok_hip(R, hipStreamEndCapture(R->s, &R->graph[p]), "end capture") &&
                 ok_hip(R, hipGraphInstantiate(&R->gexec[p], R->graph[p], nullptr, nullptr, 0),
                        "instantiate");
        }
        R->capturing = false;
    }
    R->t_old = t_old;
    R->t_new = t_new;
    R->cell_updates = cu;
    R->box_steps = bs;
    if (R->vpass != vpass) ok = false;   // a full step keeps the counter parity
    return ok && R->err == 0;
}


// ---------------------------------------------------------------------------
// Many single-grid runs stepped in lockstep, one launch per stage for all
// of them.  At creation every run's box-step is recorded, for each rotation
// parity, as the launches (kernel, grid, block, arguments) its own
// captured step would issue; launch i of every run then becomes one
// multi-run launch whose grid concatenates the runs' grids (runs whose
// kernel differs at a position -- power-of-two spacings or not -- form one
// launch per kernel).  Each run's blocks read that run's own arguments and
// touch only its own arrays and scratch, so a run's results are those of
// its own captured step, bit for bit.

__global__ void k_next_dt_runs(const NextDtArgs* __restrict__ A, MultiRun M) {
    int r, lb;
    if (!multi_block(M, r, lb) || threadIdx.x != 0) return;
    const NextDtArgs& a = A[r];
    float* dt = a.dt;
    double* t = a.t;
    const float* m = a.m;
    *t = *t + (double)dt[0];
    dt[2] = dt[0];
    float f = sqrtf(m[0] * m[0] + m[1] * m[1] + m[2] * m[2] + m[3] * m[3] + m[4] * m[4]);
    f = (0.0f < f) ? f : 0.0f;
    const float st = a.stepsz * a.xspac / f;
    const float nd = (a.t_step_max < st) ? a.t_step_max : st;
    dt[1] = nd;
    dt[0] = nd;
    if (!(nd >= kTStepMin && nd <= 3.0e38f)) dt[3] = 1.0f;
}

struct MultiNode {
    int id = 0;
    dim3 block;
    unsigned gy = 1;
    int nruns = 0, total = 0;
    int* d_start = nullptr;        // nruns + 1
    unsigned char* d_args[2] = {nullptr, nullptr};   // per batch phase
};

struct FusedBatch {
    std::vector<Run*> runs;
    std::vector<int> rot0;         // each run's rotation parity at creation
    std::vector<MultiNode> nodes;
    hipStream_t s = nullptr;
    bool own_stream = false;
    int phase = 0;                 // steps taken since creation, mod 2
    bool issued = false;           // steps issued since the last finish
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int err = 0;
};

bool batch_hip(FusedBatch* F, hipError_t e, const char* where) {
    if (e == hipSuccess) return true;
    if (!F->err) {
        F->err = (int)e;
        fprintf(stderr, "g3m3 batch: %s: %s\n", where, hipGetErrorString(e));
    }
    return false;
}

// one box-step of run R from rotation parity p, recorded (nothing runs)
bool record_step(Run* R, int p, std::vector<LaunchRec>& out) {
    const std::vector<double> t_old = R->t_old, t_new = R->t_new;
    const long long cu = R->cell_updates, bs = R->box_steps;
    const int vpass = R->vpass;
    const size_t nlog = R->step_log.size();
    const char rot = R->rot[0];
    const int march = R->push_march;
    Recorder rec;
    R->rot[0] = (char)p;
    R->push_march = 1;             // every run on the same kernel family (same bits)
    R->capturing = true;
    g_rec = &rec;
    box_step(R, 1);
    rec.add(kKNextDt, dim3(1), dim3(64),
            NextDtArgs{R->maxima, R->d.stepsz, R->d.xspac[0], R->d.t_step_max, R->dt_dev,
                       R->t_dev});
    g_rec = nullptr;
    R->capturing = false;
    R->push_march = march;
    R->rot[0] = rot;
    R->t_old = t_old;
    R->t_new = t_new;
    R->cell_updates = cu;
    R->box_steps = bs;
    R->step_log.resize(nlog);
    const bool same_vpass = R->vpass == vpass;
    R->vpass = vpass;
    out = std::move(rec.recs);
    return same_vpass && R->err == 0;
}

void launch_node(const MultiNode& N, int phase, hipStream_t s) {
    const MultiRun M{N.d_start, N.nruns, N.total};
    if (N.id == kKNextDt)
        hipLaunchKernelGGL(k_next_dt_runs, dim3((unsigned)((N.total + 7) & ~7)), N.block, 0, s,
                           (const NextDtArgs*)N.d_args[phase], M);
    else
        launch_multi(N.id, N.d_args[phase], M, N.block, N.gy, s);
}

void free_batch(FusedBatch* F) {
    for (MultiNode& N : F->nodes) {
        if (N.d_start) (void)hipFree(N.d_start);
        for (unsigned char* a : N.d_args)
            if (a) (void)hipFree(a);
    }
    for (int q = 0; q < 2; ++q) {
        if (F->gexec[q]) (void)hipGraphExecDestroy(F->gexec[q]);
        if (F->graph[q]) (void)hipGraphDestroy(F->graph[q]);
    }
    if (F->own_stream && F->s) (void)hipStreamDestroy(F->s);
    delete F;
}

FusedBatch* build_batch(Run** runs, int n, hipStream_t stream) {
    FusedBatch* F = new FusedBatch();
    F->s = stream;
    if (!F->s) {
        if (!batch_hip(F, hipStreamCreate(&F->s), "stream")) {
            free_batch(F);
            return nullptr;
        }
        F->own_stream = true;
    }
    // recs[p][r]: run r's launches from its parity at batch phase p
    std::vector<std::vector<std::vector<LaunchRec>>> recs(2);
    for (int r = 0; r < n; ++r) {
        Run* R = runs[r];
        if (!R || R->err || R->d.n_grids != 1 || R->d.tilting || !graph_scalars(R)) {
            fprintf(stderr, "g3m3_batch_create: run %d is not a single-grid run without "
                            "tilting, or has failed\n", r);
            free_batch(F);
            return nullptr;
        }
        F->runs.push_back(R);
        F->rot0.push_back(R->rot[0]);
        for (int ph = 0; ph < 2; ++ph) {
            std::vector<LaunchRec> v;
            if (!record_step(R, (R->rot[0] + ph) & 1, v)) {
                fprintf(stderr, "g3m3_batch_create: recording run %d failed\n", r);
                free_batch(F);
                return nullptr;
            }
            recs[ph].push_back(std::move(v));
        }
    }
    // the same sequence of stage launches in every run and phase
    const size_t nl = recs[0][0].size();
    for (int ph = 0; ph < 2; ++ph)
        for (int r = 0; r < n; ++r) {
            const std::vector<LaunchRec>& v = recs[ph][r];
            bool same = v.size() == nl;
            for (size_t i = 0; same && i < nl; ++i) {
                const LaunchRec &a = v[i], &b = recs[0][0][i];
                same = a.block.x == b.block.x && a.block.y == b.block.y &&
                       a.grid.y == b.grid.y && kernel_args_size(a.id) == kernel_args_size(b.id);
            }
            if (!same) {
                fprintf(stderr, "g3m3_batch_create: run %d issues another launch sequence\n", r);
                free_batch(F);
                return nullptr;
            }
        }
    for (size_t i = 0; i < nl; ++i) {
        // one node per kernel met at this position, runs in batch order
        std::vector<int> ids;
        for (int r = 0; r < n; ++r)
            if (std::find(ids.begin(), ids.end(), recs[0][r][i].id) == ids.end())
                ids.push_back(recs[0][r][i].id);
        for (int id : ids) {
            MultiNode N;
            N.id = id;
            N.block = recs[0][0][i].block;
            N.gy = recs[0][0][i].grid.y;
            std::vector<int> start{0};
            std::vector<unsigned char> blob[2];
            const size_t sz = kernel_args_size(id);
            for (int r = 0; r < n; ++r) {
                if (recs[0][r][i].id != id) continue;
                for (int ph = 0; ph < 2; ++ph) {
                    const LaunchRec& L = recs[ph][r][i];
                    if (L.id != id || L.args.size() != sz) {
                        fprintf(stderr, "g3m3_batch_create: run %d changes kernel between "
                                        "parities\n", r);
                        free_batch(F);
                        return nullptr;
                    }
                    blob[ph].insert(blob[ph].end(), L.args.begin(), L.args.end());
                }
                start.push_back(start.back() + (int)recs[0][r][i].grid.x);
                ++N.nruns;
            }
            N.total = start.back();
            bool ok = batch_hip(F, hipMalloc(&N.d_start, sizeof(int) * start.size()), "alloc") &&
                      batch_hip(F, hipMemcpy(N.d_start, start.data(), sizeof(int) * start.size(),
                                             hipMemcpyHostToDevice), "copy");
            for (int ph = 0; ok && ph < 2; ++ph)
                ok = batch_hip(F, hipMalloc(&N.d_args[ph], blob[ph].size()), "alloc") &&
                     batch_hip(F, hipMemcpy(N.d_args[ph], blob[ph].data(), blob[ph].size(),
                                            hipMemcpyHostToDevice), "copy");
            F->nodes.push_back(N);
            if (!ok) {
                free_batch(F);
                return nullptr;
            }
        }
    }
    // one captured graph per phase: every node of one lockstep box-step
    for (int ph = 0; ph < 2; ++ph) {
        bool ok = batch_hip(F, hipStreamBeginCapture(F->s, hipStreamCaptureModeThreadLocal),
                            "begin capture");
        if (ok) {
            for (const MultiNode& N : F->nodes) launch_node(N, ph, F->s);
            ok = batch_hip(F, hipStreamEndCapture(F->s, &F->graph[ph]), "end capture") &&
                 batch_hip(F, hipGraphInstantiate(&F->gexec[ph], F->graph[ph], nullptr, nullptr,
                                                  0),
                           "instantiate");
        }
        if (!ok) {
            free_batch(F);
            return nullptr;
        }
    }
    return F;
}

}  // namespace

extern "C" {

void* g3m3_run_create(const g3m3_run_desc* desc) {
    if (!desc || desc->n_grids < 1 || desc->n_grids > G3M3_MAX_GRIDS) return nullptr;
    Run* R = new Run();
    R->d = *desc;
    R->B = Box{desc->nx, desc->ny, desc->nz};
    R->n = R->B.V();
    const int n = R->n, ng = desc->n_grids;
    if (n >= kFastDivMax || desc->nx < 5 || desc->ny < 5 || desc->nz < 5) {
        fprintf(stderr, "g3m3_run_create: box extents out of range\n");
        delete R;
        return nullptr;
    }
    bool ok = hipMalloc(&R->scr, sizeof(float) * 12 * (size_t)n) == hipSuccess &&
              hipMalloc(&R->list, sizeof(ValEntry) * 3 * (size_t)n) == hipSuccess &&
              hipMalloc(&R->count, 128) == hipSuccess &&
              hipMalloc(&R->partials, sizeof(float) * 5 * (size_t)validate_blocks(R->B)) ==
                  hipSuccess &&
              hipMalloc(&R->maxima, sizeof(float) * 5 * ng) == hipSuccess;
    if (ok && desc->mbndry >= 1 && desc->rst) {
        // rst restricted to BandE's lists (mid and zero) of box mbndry
        ok = hipMalloc(&R->rst_eff, sizeof(float) * (size_t)n) == hipSuccess;
        if (ok) {
            std::vector<float> h(n, 0.0f), rst(n);
            ok = hipMemcpy(rst.data(), desc->rst + (size_t)(desc->mbndry - 1) * n,
                           sizeof(float) * n, hipMemcpyDeviceToHost) == hipSuccess;
            const int mb = desc->mbndry, b = desc->mbndry - 1;
            const int* lists[2] = {desc->ijmid, desc->ijzero};
            const int nums[2] = {desc->nummid, desc->numzero};
            for (int l = 0; ok && l < 2; ++l) {
                if (nums[l] <= 0) continue;
                std::vector<int> ij((size_t)mb * 3 * nums[l]);
                ok = hipMemcpy(ij.data(), lists[l], sizeof(int) * ij.size(),
                               hipMemcpyDeviceToHost) == hipSuccess;
                for (int p = 0; ok && p < nums[l]; ++p) {
                    const int i = ij[b + mb * (0 + 3 * p)] - 1, j = ij[b + mb * (1 + 3 * p)] - 1,
                              k = ij[b + mb * (2 + 3 * p)] - 1;
                    if (!in_box(R->B, i, j, k)) {
                        ok = false;
                        break;
                    }
                    const int c = R->B.IX(i, j, k);
                    h[c] = rst[c];
                }
            }
            ok = ok && hipMemcpy(R->rst_eff, h.data(), sizeof(float) * n,
                                 hipMemcpyHostToDevice) == hipSuccess;
        }
    }
    if (ok && desc->mbndry >= 1) {
        // do the planet lists of box mbndry share a cell?  (their order
        // matters only if they do)
        const int mb = desc->mbndry, b = desc->mbndry - 1;
        const int* lists[3] = {desc->ijsrf, desc->ijmid, desc->ijzero};
        const int nums[3] = {desc->numsrf, desc->nummid, desc->numzero};
        std::vector<int> cells;
        for (int l = 0; ok && l < 3; ++l) {
            if (nums[l] <= 0) continue;
            std::vector<int> ij((size_t)mb * 3 * nums[l]);
            ok = hipMemcpy(ij.data(), lists[l], sizeof(int) * ij.size(),
                           hipMemcpyDeviceToHost) == hipSuccess;
            for (int p = 0; ok && p < nums[l]; ++p) {
                const int i = ij[b + mb * (0 + 3 * p)] - 1, j = ij[b + mb * (1 + 3 * p)] - 1,
                          k = ij[b + mb * (2 + 3 * p)] - 1;
                ok = in_box(R->B, i, j, k);
                if (ok) cells.push_back(R->B.IX(i, j, k));
            }
        }
        std::sort(cells.begin(), cells.end());
        R->lists_disjoint = std::adjacent_find(cells.begin(), cells.end()) == cells.end();
    }
    R->nshell = shell_count(R->B);
    for (int box = 1; ok && box < ng; ++box) {
        const int b = box - 1;
        std::vector<ShellPt> h(R->nshell);
        std::vector<int> gather;
        build_shell(R->B, geom(R->d, box), h.data());
        ok = build_shell_slab(R->B, h.data(), R->nshell, gather);
        if (!ok) fprintf(stderr, "g3m3_run_create: box %d's shell reads outside box %d\n", box,
                         box + 1);
        R->nslab[b] = (int)gather.size();
        ok = ok && hipMalloc(&R->shell[b], sizeof(ShellPt) * h.size()) == hipSuccess &&
             hipMemcpy(R->shell[b], h.data(), sizeof(ShellPt) * h.size(),
                       hipMemcpyHostToDevice) == hipSuccess &&
             hipMalloc(&R->slab_idx[b], sizeof(int) * gather.size()) == hipSuccess &&
             hipMemcpy(R->slab_idx[b], gather.data(), sizeof(int) * gather.size(),
                       hipMemcpyHostToDevice) == hipSuccess &&
             hipMalloc(&R->slab[b], sizeof(float) * 2 * NF * gather.size()) == hipSuccess;
    }
    ok = ok && hipMemset(R->count, 0, 128) == hipSuccess &&
         hipMemset(R->maxima, 0, sizeof(float) * 5 * ng) == hipSuccess;
    if (!ok) {
        fprintf(stderr,
                "g3m3_run_create: device allocation or copy failed, or a planet list point "
                "lies outside the box\n");
        g3m3_run_destroy(R);
        return nullptr;
    }
    R->t_old.assign(ng, 0.0);
    R->t_new.assign(ng, 0.0);
    R->t_step.assign(ng, 0.0);
    R->t_stepnew.assign(ng, 0.0);
    R->tilt = R->d.tilt;
    R->old_tilt = R->d.tilt;
    R->ut = R->d.ut;
    R->utold = R->d.ut;
    R->rot.assign(ng, 0);
    tune_fct(R);
    tune_push(R);
    return R;
}

void g3m3_run_destroy(void* run) {
    Run* R = (Run*)run;
    if (!R) return;
    // teardown: errors here have no caller left to report them to
    (void)hipStreamSynchronize(R->s);
    for (void* p : {(void*)R->scr, (void*)R->list, (void*)R->count,
                    (void*)R->partials, (void*)R->maxima, (void*)R->rst_eff})
        if (p) (void)hipFree(p);
    for (ShellPt* p : R->shell)
        if (p) (void)hipFree(p);
    for (float* p : R->slab)
        if (p) (void)hipFree(p);
    for (int* p : R->slab_idx)
        if (p) (void)hipFree(p);
    for (int q = 0; q < 2; ++q) {
        if (R->gexec[q]) (void)hipGraphExecDestroy(R->gexec[q]);
        if (R->graph[q]) (void)hipGraphDestroy(R->graph[q]);
    }
    if (R->dt_dev) (void)hipFree(R->dt_dev);
    if (R->t_dev) (void)hipFree(R->t_dev);
    if (R->host_buf) (void)hipHostFree(R->host_buf);
    if (R->own_stream) (void)hipStreamDestroy(R->s);
    delete R;
}

void g3m3_run_set_stream(void* run, void* stream) { ((Run*)run)->s = (hipStream_t)stream; }

int g3m3_run_init_dt(void* run) {
    Run* R = (Run*)run;
    all_coarse_changed(R);   // the caller may have written the arrays
    std::vector<int> boxes;
    for (int box = R->d.n_grids; box >= 1; --box) {
        R->t_old[box - 1] = 0.0;
        R->t_new[box - 1] = 0.0;
        validate(R, R->state(box), box, true);
        boxes.push_back(box);
    }
    read_tstepnew(R, boxes);
    check(R, "init_dt");
    return R->err;
}

int g3m3_run_full_step(void* run, const double* dt_override) {
    const int e = g3m3_run_full_step_async(run, dt_override);
    if (e) return e;
    return g3m3_run_finish(run);
}

int g3m3_run_full_step_async(void* run, const double* dt_override) {
    Run* R = (Run*)run;
    all_coarse_changed(R);   // the caller may have written the arrays
    const int n = R->d.n_grids;
    std::vector<double>& ts = R->t_step;
    // dt synchronisation (code/3d_main.f90:1929-1968)
    float smallest = 0.0f;
    int m_small = n;
    for (int box = n; box >= 1; --box) {
        const int b = box - 1;
        ts[b] = R->t_stepnew[b];
        R->t_old[b] = 0.0;
        R->t_new[b] = 0.0;
        if (box == n) {
            smallest = (float)ts[b];
            m_small = box;
        } else {
            if (ts[b] > ts[b + 1]) ts[b] = ts[b + 1];
            if ((double)smallest > ts[b]) {
                smallest = (float)ts[b];
                m_small = box;
            }
        }
    }
    for (int box = 1; box <= n; ++box) {
        const int b = box - 1;
        if (box <= m_small) {
            ts[b] = (double)smallest;
        } else {
            const float astep = (float)(ts[b] / ts[b - 1] + 0.50);
            int nsteps = (int)astep;
            nsteps = std::min(std::max(nsteps, 1), 2);
            ts[b] = ts[b - 1] * nsteps;
        }
    }
    if (dt_override)
        for (int b = 0; b < n; ++b) ts[b] = dt_override[b];
    // the main loop's clock (code/3d_main.f90:1966-1977): delt is single
    // precision there, t and ut double, tilt single
    const float delt_main = (float)ts[n - 1];
    // the main loop stops on a step below t_step_min (:1969-1973); a step
    // that is not a finite number stops the run too
    if (!(std::isfinite(delt_main) && delt_main >= kTStepMin) ||
        !(std::isfinite(ts[m_small - 1]) && ts[m_small - 1] > 0.0)) {
        fprintf(stderr, "g3m3 run: time step %g below t_step_min %g or not finite\n",
                (double)delt_main, (double)kTStepMin);
        if (!R->err) R->err = kErrStepCollapsed;
        return R->err;
    }
    const int m_step = (int)(ts[n - 1] / ts[m_small - 1]);
    R->step_log.clear();
    R->utold = R->ut;
    R->old_tilt = R->tilt;
    R->t += (double)delt_main;
    R->ut = R->t * (double)R->d.t_equiv / 3600.0;
    R->tilt = R->tilt + R->d.dtilt * delt_main;
    std::vector<int> stepped;
    for (int ms = 0; ms < m_step; ++ms) {
        for (int box = n; box >= 1; --box) {
            const int b = box - 1;
            const bool yes = box == 1 || (R->t_old[b] == R->t_new[b] && R->t_new[b] < ts[n - 1] &&
                                          std::fabs(R->t_new[b] - R->t_new[0]) <= 0.005 * ts[0]);
            if (yes) {
                box_step(R, box);
                if (std::find(stepped.begin(), stepped.end(), box) == stepped.end())
                    stepped.push_back(box);
            }
        }
    }
    // final sync (:3069-3095)
    R->t_old[0] = R->t_new[0];
    for (int box = 1; box < n; ++box) {
        corer(R, box);
        set_rho_only(R, R->state(box));
        coarse_changed(R, box);
    }
    R->stepped = stepped;
    check(R, "full_step");
    return R->err;
}

int g3m3_run_finish(void* run) {
    Run* R = (Run*)run;
    read_tstepnew(R, R->stepped);
    R->stepped.clear();
    check(R, "finish");
    return R->err;
}

int g3m3_run_graph_steps(void* run, int nsteps) {
    Run* R = (Run*)run;
    if (R->err) return R->err;
    if (R->d.n_grids != 1 || R->d.tilting) {
        fprintf(stderr, "g3m3_run_graph_steps: single-grid runs without tilting only\n");
        return R->err = -1;
    }
    if (!R->gexec[0] || !R->gexec[1]) {
        if (!build_graphs(R)) {
            if (!R->err) R->err = -1;
            return R->err;
        }
    }
    // this step's dt and the run time, as the host path would use them
    const float dt0 = (float)R->t_stepnew[0];
    if (!(std::isfinite(dt0) && dt0 >= kTStepMin)) {
        fprintf(stderr, "g3m3 run: time step %g below t_step_min or not finite\n", (double)dt0);
        return R->err = kErrStepCollapsed;
    }
    hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(64), 0, R->s, R->dt_dev, R->t_dev, dt0, R->t);
    for (int q = 0; q < nsteps; ++q) {
        if (!ok_hip(R, hipGraphLaunch(R->gexec[(int)R->rot[0]], R->s), "graph launch")) break;
        R->rot[0] ^= 1;
        R->cell_updates += R->n;
        R->box_steps += 1;
    }
    check(R, "graph_steps");
    return R->err;
}

int g3m3_run_graph_finish(void* run) {
    Run* R = (Run*)run;
    if (!R->host_buf) return R->err;
    double* h = (double*)R->host_buf;
    if (ok_hip(R, hipMemcpyAsync(h, R->t_dev, sizeof(double), hipMemcpyDeviceToHost, R->s),
               "graph finish") &&
        ok_hip(R, hipMemcpyAsync(h + 1, R->dt_dev, 4 * sizeof(float), hipMemcpyDeviceToHost,
                                 R->s),
               "graph finish") &&
        ok_hip(R, hipStreamSynchronize(R->s), "graph finish")) {
        R->t = h[0];
        const float* f = (const float*)(h + 1);
        R->t_stepnew[0] = (double)f[1];
        R->t_step[0] = (double)f[2];
        if (f[3] != 0.0f && !R->err) {
            fprintf(stderr, "g3m3 run: a captured step formed a time step below t_step_min "
                            "or not finite\n");
            R->err = kErrStepCollapsed;
        }
    }
    check(R, "graph_finish");
    return R->err;
}

int g3m3_run_fct_shape(void* run, int* kc_len, int* no_remap, float* cand_us) {
    Run* R = (Run*)run;
    if (kc_len) *kc_len = R->fct_kc;
    if (no_remap) *no_remap = R->fct_no_remap;
    if (cand_us)
        for (int q = 0; q < 6; ++q) cand_us[q] = R->fct_us[q];
    return 0;
}

int g3m3_run_set_fct_shape(void* run, int kc_len, int no_remap) {
    Run* R = (Run*)run;
    if (kc_len < 1) return 1;
    R->fct_kc = kc_len;
    R->fct_no_remap = no_remap ? 1 : 0;
    for (int q = 0; q < 2; ++q) {   // captured steps hold the old shape
        if (R->gexec[q]) (void)hipGraphExecDestroy(R->gexec[q]);
        if (R->graph[q]) (void)hipGraphDestroy(R->graph[q]);
        R->gexec[q] = nullptr;
        R->graph[q] = nullptr;
    }
    return 0;
}

int g3m3_run_species_kernels(void* run, int* march, float* cand_us) {
    Run* R = (Run*)run;
    if (march) *march = R->push_march;
    if (cand_us)
        for (int q = 0; q < 2; ++q) cand_us[q] = R->push_us[q];
    return 0;
}

int g3m3_run_set_species_kernels(void* run, int march) {
    Run* R = (Run*)run;
    if (march != 0 && march != 1) return 1;
    R->push_march = march;
    for (int q = 0; q < 2; ++q) {   // captured steps hold the old kernel
        if (R->gexec[q]) (void)hipGraphExecDestroy(R->gexec[q]);
        if (R->graph[q]) (void)hipGraphDestroy(R->graph[q]);
        R->gexec[q] = nullptr;
        R->graph[q] = nullptr;
    }
    return 0;
}

int g3m3_run_sync_arrays(void* run) {
    Run* R = (Run*)run;
    all_coarse_changed(R);   // the caller may have written the arrays
    for (int box = 1; box <= R->d.n_grids; ++box) {
        if (!R->rot[box - 1]) continue;
        const Set a = R->slice(R->d.state, box), o = R->slice(R->d.old, box);
        FieldPtrs pa{}, po{};
        for (int f = 0; f < NF; ++f) {
            pa.p[f] = a.f[f];
            po.p[f] = o.f[f];
        }
        launch_swap_fields(R->n, NF, pa, po, R->s);
        R->rot[box - 1] = 0;
    }
    // the caller reads or writes the arrays next, possibly on another stream
    ok_hip(R, hipStreamSynchronize(R->s), "sync_arrays");
    check(R, "sync_arrays");
    return R->err;
}

void g3m3_run_set_time(void* run, double t) { ((Run*)run)->t = t; }

int g3m3_run_box_step(void* run, int box) {
    Run* R = (Run*)run;
    all_coarse_changed(R);   // the caller may have written the arrays
    box_step(R, box);
    return R->err;
}

int g3m3_run_step_log(void* run, int cap, int* box, double* t_old, double* t_new) {
    const Run* R = (const Run*)run;
    const int n = (int)R->step_log.size();
    for (int q = 0; q < n && q < cap; ++q) {
        if (box) box[q] = R->step_log[q].box;
        if (t_old) t_old[q] = R->step_log[q].t_old;
        if (t_new) t_new[q] = R->step_log[q].t_new;
    }
    return n;
}

void g3m3_run_info(void* run, double* t, double* t_step, double* t_stepnew, double* t_old,
                   double* t_new, long long* cell_updates, long long* box_steps) {
    Run* R = (Run*)run;
    const int n = R->d.n_grids;
    if (t) *t = R->t;
    for (int b = 0; b < n; ++b) {
        if (t_step) t_step[b] = R->t_step[b];
        if (t_stepnew) t_stepnew[b] = R->t_stepnew[b];
        if (t_old) t_old[b] = R->t_old[b];
        if (t_new) t_new[b] = R->t_new[b];
    }
    if (cell_updates) *cell_updates = R->cell_updates;
    if (box_steps) *box_steps = R->box_steps;
}


void* g3m3_batch_create(void* const* runs, int nruns, void* stream) {
    if (!runs || nruns < 1) return nullptr;
    std::vector<Run*> rs(nruns);
    for (int r = 0; r < nruns; ++r) {
        rs[r] = (Run*)runs[r];
        // each run's own pending work is done before the batch reads it
        if (rs[r] && rs[r]->s) (void)hipStreamSynchronize(rs[r]->s);
    }
    return build_batch(rs.data(), nruns, (hipStream_t)stream);
}

int g3m3_batch_steps(void* batch, int nsteps) {
    FusedBatch* F = (FusedBatch*)batch;
    if (!F) return -1;
    if (F->err) return F->err;
    for (size_t r = 0; r < F->runs.size(); ++r) {
        Run* R = F->runs[r];
        if (R->err) return F->err = R->err;
        // the recorded arguments follow each run's buffer rotation from the
        // batch's creation: a run stepped on its own since then no longer
        // matches them
        if (R->rot[0] != (char)(F->rot0[r] ^ F->phase)) {
            fprintf(stderr, "g3m3 batch: run %d was stepped outside the batch\n", (int)r);
            return F->err = -1;
        }
        // the device already holds the step and time of steps issued since
        // the last finish; otherwise they come from the host
        if (F->issued) continue;
        const float dt0 = (float)R->t_stepnew[0];
        if (!(std::isfinite(dt0) && dt0 >= kTStepMin)) {
            fprintf(stderr, "g3m3 batch: time step %g below t_step_min or not finite\n",
                    (double)dt0);
            return F->err = kErrStepCollapsed;
        }
        if (R->s && R->s != F->s) (void)hipStreamSynchronize(R->s);
        hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(64), 0, F->s, R->dt_dev, R->t_dev, dt0, R->t);
    }
    F->issued = true;
    for (int q = 0; q < nsteps; ++q) {
        if (!batch_hip(F, hipGraphLaunch(F->gexec[F->phase], F->s), "graph launch")) break;
        F->phase ^= 1;
        for (Run* R : F->runs) {
            R->rot[0] ^= 1;
            R->cell_updates += R->n;
            R->box_steps += 1;
        }
    }
    batch_hip(F, hipGetLastError(), "batch steps");
    return F->err;
}

int g3m3_batch_finish(void* batch) {
    FusedBatch* F = (FusedBatch*)batch;
    if (!F) return -1;
    if (F->err) return F->err;
    for (Run* R : F->runs) {
        double* h = (double*)R->host_buf;
        if (!batch_hip(F, hipMemcpyAsync(h, R->t_dev, sizeof(double), hipMemcpyDeviceToHost, F->s),
                       "finish") ||
            !batch_hip(F, hipMemcpyAsync(h + 1, R->dt_dev, 4 * sizeof(float),
                                         hipMemcpyDeviceToHost, F->s),
                       "finish"))
            return F->err;
    }
    if (!batch_hip(F, hipStreamSynchronize(F->s), "finish")) return F->err;
    F->issued = false;
    for (Run* R : F->runs) {
        const double* h = (const double*)R->host_buf;
        const float* f = (const float*)(h + 1);
        R->t = h[0];
        R->t_stepnew[0] = (double)f[1];
        R->t_step[0] = (double)f[2];
        if (f[3] != 0.0f && !R->err) {
            fprintf(stderr, "g3m3 batch: a step formed a time step below t_step_min or not "
                            "finite\n");
            R->err = kErrStepCollapsed;
        }
        if (R->err && !F->err) F->err = R->err;
    }
    return F->err;
}

int g3m3_batch_launches(void* batch) {
    FusedBatch* F = (FusedBatch*)batch;
    return F ? (int)F->nodes.size() : 0;
}

void g3m3_batch_destroy(void* batch) {
    FusedBatch* F = (FusedBatch*)batch;
    if (!F) return;
    (void)hipStreamSynchronize(F->s);
    free_batch(F);
}

}  // extern "C"
