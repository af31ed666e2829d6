// The magnetic-field half of a box-step on one outermost box with the
// plasma held fixed, and its discrete adjoint.
//
// Forward (the field calls of code/3d_main.f90:2243-2324, 2486-2565,
// 2574-2590, 2765-2790, 2896-2945 for box = n_grids = 1, mbndry = 0):
//   stage 1: J1 = calcur(b), bs1 = b0 + b, ev1 = fnd_evel(J1),
//            E1 = BandE(bs1, J1, ev1) + eta J1 / n,
//            w1 = push_bfld(old = b, E1, delt/2), bndry_outer(w1)
//   stage 2: the same from w1 -> J2, bs2, ev2, E2;
//            b2 = push_bfld(old = b, E2, delt), bndry_outer(b2)
//   smoothing: W = lap_bfld(b2) over w1, bndry_outer(W);
//            b' = fcsmooth(old = b, wrk = W, diferg), bndry_outer(b')
// Each push writes the interior of an array that holds b (w1, b2) or w1
// (W); bndry_outer then overwrites every boundary cell, so the step is a
// function of b, eta and the upstream field sb alone.  The plasma (ion
// densities and momenta, electron pressure) is fixed, so tv and the
// Lapidus velocity (fnd_vtot, code/sub_fnd_vtot.f90) are computed once.
//
// The adjoint takes the tape the forward step wrote -- b, J, bs and ev of
// both stages and W -- and runs the stages' adjoint kernels (k_adjoint.hip)
// in reverse, returning b's adjoint and adding the adjoints of eta and sb
// (in double) to ctl_bar.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../include/g3m3.h"
#include "g3m3_fused.h"

using namespace g3m3;

namespace {

// tape slots, 3 arrays of V cells each
enum Tape { T_B = 0, T_J1, T_BS1, T_EV1, T_J2, T_BS2, T_EV2, T_W, T_SLOTS };

struct FieldModel {
    g3m3_field_model_desc d;
    Box B;
    int n = 0;
    hipStream_t s = nullptr;
    float* v[3] = {};      // Lapidus velocity (fnd_vtot of the fixed plasma)
    float* tv[3] = {};     // fnd_evel's ion velocity (written, unused)
    float* btot = nullptr;
    float* e[3] = {};
    float* w1[3] = {};     // stage-1 field
    float* b2[3] = {};     // stage-2 field
    float* fwd = nullptr;  // tape slots for a step without a tape
    float* adj = nullptr;  // adjoint scratch
    float* mem = nullptr;
};

float* slot(float* base, int n, int t, int c) { return base + (size_t)n * (3 * t + c); }

bool launched(const char* where) {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return true;
    fprintf(stderr, "g3m3_field_model: %s: %s\n", where, hipGetErrorString(e));
    return false;
}

void outer_b(FieldModel* M, float* b[3], const float ctl[4]) {
    OuterPtrs O{};
    for (int c = 0; c < 3; ++c) {
        O.f[c] = b[c];
        O.kind[c] = kBX + c;
        O.val[c] = ctl[1 + c];
        O.b0[c] = M->d.b0[c];
    }
    launch_bndry_outer(M->B, O, 3, M->s);
}

void copy3(FieldModel* M, float* dst[3], const float* src[3]) {
    for (int c = 0; c < 3; ++c)
        hipMemcpyAsync(dst[c], src[c], sizeof(float) * M->n, hipMemcpyDeviceToDevice, M->s);
}

// J, bs, ev and E of one stage from field b
void efield(FieldModel* M, const float* b[3], float* J[3], float* bs[3], float* ev[3], float eta) {
    g3m3_field_model_desc& d = M->d;
    const float r = d.rx;
    launch_calcur(M->B, b[0], b[1], b[2], J[0], J[1], J[2], r, r, r, M->s);
    const float* Jc[3] = {J[0], J[1], J[2]};
    launch_bfield_evel(M->B, d.b0, b, bs, M->btot, d.mom, d.rho, Jc, ev, M->tv, d.rmassq,
                       d.rmassh, d.rmasso, d.reynolds, d.smallbit, 1, 1, M->s);
    const float* bsc[3] = {bs[0], bs[1], bs[2]};
    const float* evc[3] = {ev[0], ev[1], ev[2]};
    launch_bande(M->B, M->e, bsc, Jc, evc, d.epres, d.rho, nullptr, 0.0f, d.reynolds, 1,
                 d.rmassq, d.rmassh, d.rmasso, nullptr, 0, nullptr, 0, 0, r, r, r, eta, M->s);
}

}  // namespace

extern "C" {

void* g3m3_field_model_create(const g3m3_field_model_desc* desc) {
    if (!desc || desc->nx < 5 || desc->ny < 5 || desc->nz < 5 ||
        (long long)desc->nx * desc->ny * desc->nz >= kFastDivMax) {
        fprintf(stderr, "g3m3_field_model_create: box extents out of range\n");
        return nullptr;
    }
    FieldModel* M = new FieldModel();
    M->d = *desc;
    M->B = Box{desc->nx, desc->ny, desc->nz};
    M->n = M->B.V();
    const size_t n = M->n;
    // v 3, tv 3, btot 1, e 3, w1 3, b2 3 | fwd tape 3*T_SLOTS | adjoint scratch
    const size_t fixed = 16 * n, tape = (size_t)3 * T_SLOTS * n;
    const size_t adj = 8 * 3 * n + adj_fct_scratch_floats(M->B);
    if (hipMalloc(&M->mem, sizeof(float) * (fixed + tape + adj)) != hipSuccess) {
        fprintf(stderr, "g3m3_field_model_create: out of device memory\n");
        delete M;
        return nullptr;
    }
    float* p = M->mem;
    for (int c = 0; c < 3; ++c) M->v[c] = p + n * c;
    for (int c = 0; c < 3; ++c) M->tv[c] = p + n * (3 + c);
    M->btot = p + n * 6;
    for (int c = 0; c < 3; ++c) M->e[c] = p + n * (7 + c);
    for (int c = 0; c < 3; ++c) M->w1[c] = p + n * (10 + c);
    for (int c = 0; c < 3; ++c) M->b2[c] = p + n * (13 + c);
    M->fwd = p + fixed;
    M->adj = M->fwd + tape;
    hipMemsetAsync(M->mem, 0, sizeof(float) * (fixed + tape + adj), M->s);
    launch_vtot(M->n, M->d.mom, M->d.rho, M->v, M->s);
    if (!launched("create") || hipStreamSynchronize(M->s) != hipSuccess) {
        hipFree(M->mem);
        delete M;
        return nullptr;
    }
    return M;
}

void g3m3_field_model_destroy(void* model) {
    FieldModel* M = (FieldModel*)model;
    if (!M) return;
    hipStreamSynchronize(M->s);
    hipFree(M->mem);
    delete M;
}

void g3m3_field_model_set_stream(void* model, void* hip_stream) {
    if (model) ((FieldModel*)model)->s = (hipStream_t)hip_stream;
}

long long g3m3_field_model_tape_floats(void* model) {
    return model ? (long long)3 * T_SLOTS * ((FieldModel*)model)->n : 0;
}

int g3m3_field_model_step(void* model, const float* b_in, float* b_out, float* tape,
                          const float* ctl) {
    FieldModel* M = (FieldModel*)model;
    if (!M || !b_in || !b_out || !ctl || b_in == b_out) return 1;
    const int n = M->n;
    g3m3_field_model_desc& d = M->d;
    const float r = d.rx, eta = ctl[0];
    float* T = tape ? tape : M->fwd;
    float* J[3], *bs[3], *ev[3], *W[3], *bo[3];
    const float* b[3];
    for (int c = 0; c < 3; ++c) {
        b[c] = b_in + (size_t)n * c;
        bo[c] = b_out + (size_t)n * c;
        W[c] = slot(T, n, T_W, c);
    }
    if (tape) {
        float* tb[3] = {slot(T, n, T_B, 0), slot(T, n, T_B, 1), slot(T, n, T_B, 2)};
        copy3(M, tb, b);
    }
    // stage 1
    for (int c = 0; c < 3; ++c) {
        J[c] = slot(T, n, T_J1, c);
        bs[c] = slot(T, n, T_BS1, c);
        ev[c] = slot(T, n, T_EV1, c);
    }
    efield(M, b, J, bs, ev, eta);
    const float* ec[3] = {M->e[0], M->e[1], M->e[2]};
    copy3(M, M->w1, b);
    launch_push_bfld(M->B, M->w1, b, ec, 0.5f * d.delt, r, r, r, M->s);
    outer_b(M, M->w1, ctl);
    // stage 2
    const float* w1c[3] = {M->w1[0], M->w1[1], M->w1[2]};
    for (int c = 0; c < 3; ++c) {
        J[c] = slot(T, n, T_J2, c);
        bs[c] = slot(T, n, T_BS2, c);
        ev[c] = slot(T, n, T_EV2, c);
    }
    efield(M, w1c, J, bs, ev, eta);
    copy3(M, M->b2, b);
    launch_push_bfld(M->B, M->b2, b, ec, d.delt, r, r, r, M->s);
    outer_b(M, M->b2, ctl);
    // Lapidus into a copy of the stage-1 field, then flux correction
    copy3(M, W, w1c);
    LapPtrs L{};
    for (int c = 0; c < 3; ++c) {
        L.src[c] = M->b2[c];
        L.dst[c] = W[c];
        L.chi[c] = d.chierg;
    }
    const float* vc[3] = {M->v[0], M->v[1], M->v[2]};
    launch_lapidus(M->B, L, 3, vc, d.delt, M->s);
    outer_b(M, W, ctl);
    FctParams F{};
    F.B = M->B;
    for (int c = 0; c < 3; ++c) {
        F.out[c] = bo[c];
        F.old[c] = b[c];
        F.wrk[c] = W[c];
        F.chi[c] = d.diferg;
    }
    launch_fct_fused(F, 3, M->s);
    outer_b(M, bo, ctl);
    return launched("step") ? 0 : 2;
}

int g3m3_field_model_adjoint(void* model, const float* tape, const float* bout_bar, float* bin_bar,
                             double* ctl_bar, const float* ctl) {
    FieldModel* M = (FieldModel*)model;
    if (!M || !tape || !bout_bar || !bin_bar || !ctl_bar || !ctl) return 1;
    const int n = M->n;
    g3m3_field_model_desc& d = M->d;
    const float r = d.rx, eta = ctl[0];
    float* T = const_cast<float*>(tape);
    float* S = M->adj;
    auto a3 = [&](int k, float* out[3]) {
        for (int c = 0; c < 3; ++c) out[c] = S + (size_t)n * (3 * k + c);
    };
    auto t3 = [&](int k, const float* out[3]) {
        for (int c = 0; c < 3; ++c) out[c] = slot(T, n, k, c);
    };
    float *A[3], *WB[3], *B2B[3], *W1B[3], *EB[3], *EVB[3], *CB[3], *BSB[3], *BN[3];
    a3(0, A);
    a3(1, WB);
    a3(2, B2B);
    a3(3, W1B);
    a3(4, EB);
    a3(5, EVB);
    a3(6, CB);
    a3(7, BSB);
    float* fct_scr = S + (size_t)24 * n;
    const float *bn[3], *W[3], *J[3], *bs[3], *ev[3], *bo_bar[3];
    t3(T_B, bn);
    t3(T_W, W);
    for (int c = 0; c < 3; ++c) {
        BN[c] = bin_bar + (size_t)n * c;
        bo_bar[c] = bout_bar + (size_t)n * c;
    }
    double* sbb = ctl_bar + 1;
    hipStream_t s = M->s;
    copy3(M, A, bo_bar);
    hipMemsetAsync(bin_bar, 0, sizeof(float) * 3 * n, s);
    hipMemsetAsync(W1B[0], 0, sizeof(float) * 3 * n, s);
    // b' = bndry_outer(fcsmooth(b, W))
    launch_adj_outer_b(M->B, A, sbb, s);
    launch_adj_fct(M->B, bn, W, A, WB, BN, d.diferg, fct_scr, s);
    // W = bndry_outer(lap_bfld(b2) over w1)
    launch_adj_outer_b(M->B, WB, sbb, s);
    const float* WBc[3] = {WB[0], WB[1], WB[2]};
    const float* vc[3] = {M->v[0], M->v[1], M->v[2]};
    launch_adj_lap(M->B, WBc, vc, d.chierg, d.delt, B2B, W1B, s);
    // stages 2 and 1
    for (int stage = 2; stage >= 1; --stage) {
        float** fb = stage == 2 ? B2B : W1B;       // the stage's output adjoint
        launch_adj_outer_b(M->B, fb, sbb, s);
        const float* fbc[3] = {fb[0], fb[1], fb[2]};
        launch_adj_push_bfld(M->B, fbc, BN, EB, stage == 2 ? d.delt : 0.5f * d.delt, r, r, r, s);
        t3(stage == 2 ? T_J2 : T_J1, J);
        t3(stage == 2 ? T_BS2 : T_BS1, bs);
        t3(stage == 2 ? T_EV2 : T_EV1, ev);
        hipMemsetAsync(CB[0], 0, sizeof(float) * 3 * n, s);
        hipMemsetAsync(BSB[0], 0, sizeof(float) * 3 * n, s);
        launch_adj_bande(M->B, EB, bs, ev, J, d.rho, d.rmassq, d.rmassh, d.rmasso, eta, BSB, EVB,
                         CB, ctl_bar, s);
        const float* EVBc[3] = {EVB[0], EVB[1], EVB[2]};
        launch_adj_evel(M->B, EVBc, d.rho, d.rmassq, d.rmassh, d.rmasso, d.reynolds, CB, s);
        // the stage read w1 (stage 2) or b (stage 1)
        float** into = stage == 2 ? W1B : BN;
        const float* BSBc[3] = {BSB[0], BSB[1], BSB[2]};
        launch_add3(n, BSBc, into, s);
        launch_adj_calcur(M->B, CB, into, r, r, r, s);
    }
    return launched("adjoint") ? 0 : 2;
}

}  // extern "C"
