// Discrete adjoints of the ion push of one species (push_ion,
// code/sub_push_ion.f90:4-362), of the species Lapidus smoothing
// (lap_plasma, code/sub_lap_plasma.f90:4-147) and of the species velocity
// the main program forms for it (v = p / max(rho, smallbit),
// code/3d_main.f90:2698-2703).  Each is the transpose of the forward
// kernel's linearisation about the forward inputs, in two passes without
// atomics: pass 1, per interior cell c, differentiates c's own outputs with
// respect to the values c reads at its own position and writes, for every
// centred difference c takes, the coefficient that difference's adjoint
// carries; pass 2, per cell m, gathers those coefficients from the cells
// whose stencils read m.  All adjoints accumulate (+=) into the caller's
// arrays.  Box-local arrays of nx*ny*nz cells; fields in the order rho,
// presx, presy, presz, px, py, pz, presxy, presxz, presyz.
#include "g3m3_dev.h"
#include "g3m3_launch.h"

namespace g3m3 {

namespace {

constexpr float kDmin = 1.e-6f;       // push_ion's density floor for velocities
enum { R_ = 0, X_ = 1, Y_ = 2, Z_ = 3, PX_ = 4, PY_ = 5, PZ_ = 6, XY_ = 7, XZ_ = 8, YZ_ = 9 };

// scratch of push_ion_adjoint: D[a][d] (9) the coefficient of
// (qv_a[c+d] - qv_a[c-d]), S[a] (3) a twelfth of the smoothed field's
// adjoint, QL[a] (3) the adjoint of qv_a at c from c's own terms
constexpr int kIonScr = 15;

__device__ inline float qv_at(const IonAdj& A, int a, int m) {
    return A.src[PX_ + a][m] / fmaxf(A.src[R_][m], kDmin);
}

}  // namespace

__global__ void k_adj_ion_local(IonAdj A) {
    int i, j, k;
    if (!cell_of(A.B, gtid(), i, j, k)) return;
    const Box& B = A.B;
    const int c = B.IX(i, j, k), n = B.V();
    float* D = A.scr;                    // D[(3a+d)*n + c]
    float* S = A.scr + 9 * (size_t)n;
    float* QL = A.scr + 12 * (size_t)n;
    if (!B.interior(i, j, k)) {
        for (int q = 0; q < kIonScr; ++q) A.scr[(size_t)q * n + c] = 0.0f;
        return;
    }
    const int st[3] = {1, B.nx, B.nx * B.ny};
    const float dt = A.delt, dd[3] = {A.dxt, A.dyt, A.dzt};
    // Every operand, and every adjoint this cell accumulates into, is loaded
    // before the first store: on this hardware a wait for a load also waits
    // for every store issued before it.  The old set's and the source set's
    // adjoints may be one array (a stage whose old is its source): then the
    // two increments land on it in the order old, source.
    float ob[10], olb[10], sbv[10];
    for (int f = 0; f < 10; ++f) {
        ob[f] = A.ob[f][c];
        olb[f] = A.oldb[f][c];
        sbv[f] = A.sb[f] == A.oldb[f] ? 0.0f : A.sb[f][c];
    }
    float ebv[3], bsbv[3], tvbv[3];
    for (int a = 0; a < 3; ++a) {
        ebv[a] = A.eb[a][c];
        bsbv[a] = A.bsb[a][c];
        tvbv[a] = A.tvb[a][c];
    }
    const float btbv = A.btotb[c];
    const float r = A.src[R_][c];
    const float x = A.src[X_][c], y = A.src[Y_][c], z = A.src[Z_][c];
    const float xy = A.src[XY_][c], xz = A.src[XZ_][c], yz = A.src[YZ_][c];
    const float aqc = fmaxf(r, kDmin);
    const float v[3] = {A.src[PX_][c] / aqc, A.src[PY_][c] / aqc, A.src[PZ_][c] / aqc};
    float dQ[3][3];                      // dQ[a][d] = qv_a(c+d) - qv_a(c-d)
    for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d) dQ[a][d] = qv_at(A, a, c + st[d]) - qv_at(A, a, c - st[d]);
    float b[3], t[3], e[3], ab[3];
    for (int a = 0; a < 3; ++a) {
        b[a] = A.bs[a][c];
        t[a] = A.tv[a][c];
        e[a] = A.e[a][c];
        const float* s = A.bs[a];
        ab[a] = (6.0f * s[c] + s[c + 1] + s[c - 1] + s[c + st[1]] + s[c - st[1]] + s[c + st[2]] +
                 s[c - st[2]]) / 12.0f;
    }
    const float re = A.reynolds, rm = A.rmass, bmax = A.bmax;
    const float bmag = A.btot[c];
    const float weight = fmaxf(1.0f, bmag / bmax);
    const float skin = re / weight;
    const float bm2 = sqrtf(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
    const float sf = bm2 > bmax ? re * (bmax / bm2) : re;
    const float tc[3] = {t[1] * b[2] - t[2] * b[1], t[2] * b[0] - t[0] * b[2],
                         t[0] * b[1] - t[1] * b[0]};
    const float qc[3] = {v[1] * b[2] - v[2] * b[1], v[2] * b[0] - v[0] * b[2],
                         v[0] * b[1] - v[1] * b[0]};
    const float qden = r / rm;
    // gravity's position (code/sub_push_ion.f90:202-206)
    const float pos[3] = {A.xmin + A.ddx * (float)i, A.ymin + A.ddy * (float)j,
                          A.zmin + A.ddz * (float)k};
    const float radius = sqrtf(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]) + 0.0000001f;
    const float tr = A.re_equiv * radius;
    const float gfac = dt * A.grav / (tr * tr);

    float rB = 0.0f, xB = 0.0f, yB = 0.0f, zB = 0.0f, xyB = 0.0f, xzB = 0.0f, yzB = 0.0f;
    float vB[3] = {0, 0, 0}, bB[3] = {0, 0, 0}, tB[3] = {0, 0, 0}, abB[3] = {0, 0, 0};
    float Dc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    float skinB = 0.0f, qdenB = 0.0f, sfB = 0.0f;
    // momenta: p_a = old + ... + dt (qden (de_a + dv_a) - gp_a) - g pos_a / radius
    float tcB[3], qcB[3];
    for (int a = 0; a < 3; ++a) {
        const float PA = ob[PX_ + a];
        const float de = (e[a] + tc[a]) * re, dv = (qc[a] - tc[a]) * skin;
        const float deB = dt * qden * PA;
        ebv[a] += deB * re;
        tcB[a] = deB * re - deB * skin;
        qcB[a] = deB * skin;
        skinB += deB * (qc[a] - tc[a]);
        qdenB += dt * PA * (de + dv);
        rB -= PA * gfac * pos[a] / radius;
    }
    rB += qdenB / rm;
    const bool btot_on = bmag / bmax > 1.0f;
    const float btb_new = btot_on ? btbv + -skinB * re / (weight * weight) / bmax : btbv;
    // tc = t x b, qc = v x b
    tB[0] = b[1] * tcB[2] - b[2] * tcB[1];
    tB[1] = b[2] * tcB[0] - b[0] * tcB[2];
    tB[2] = b[0] * tcB[1] - b[1] * tcB[0];
    bB[0] = tcB[1] * t[2] - tcB[2] * t[1] + qcB[1] * v[2] - qcB[2] * v[1];
    bB[1] = tcB[2] * t[0] - tcB[0] * t[2] + qcB[2] * v[0] - qcB[0] * v[2];
    bB[2] = tcB[0] * t[1] - tcB[1] * t[0] + qcB[0] * v[1] - qcB[1] * v[0];
    vB[0] = b[1] * qcB[2] - b[2] * qcB[1];
    vB[1] = b[2] * qcB[0] - b[0] * qcB[2];
    vB[2] = b[0] * qcB[1] - b[1] * qcB[0];
    // diagonal pressures (:210-280)
    const float h2 = 2.0f * dt, k2 = 2.0f * dt * sf / rm, k1 = dt * sf / rm;
    const float Xb = ob[X_], Yb = ob[Y_], Zb = ob[Z_];
    xB += -h2 * Xb * dQ[0][0] / dd[0];
    Dc[0][0] += -h2 * x * Xb / dd[0];
    xyB += -h2 * Xb * dQ[0][1] / dd[1];
    Dc[0][1] += -h2 * xy * Xb / dd[1];
    xzB += -h2 * Xb * dQ[0][2] / dd[2];
    Dc[0][2] += -h2 * xz * Xb / dd[2];
    xyB += k2 * Xb * ab[2];
    abB[2] += k2 * Xb * xy;
    xzB -= k2 * Xb * ab[1];
    abB[1] -= k2 * Xb * xz;
    sfB += 2.0f * dt / rm * Xb * (xy * ab[2] - xz * ab[1]);

    yB += -h2 * Yb * dQ[1][1] / dd[1];
    Dc[1][1] += -h2 * y * Yb / dd[1];
    xyB += -h2 * Yb * dQ[1][0] / dd[0];
    Dc[1][0] += -h2 * xy * Yb / dd[0];
    yzB += -h2 * Yb * dQ[1][2] / dd[2];
    Dc[1][2] += -h2 * yz * Yb / dd[2];
    yzB += k2 * Yb * ab[0];
    abB[0] += k2 * Yb * yz;
    xyB -= k2 * Yb * ab[2];
    abB[2] -= k2 * Yb * xy;
    sfB += 2.0f * dt / rm * Yb * (yz * ab[0] - xy * ab[2]);

    zB += -h2 * Zb * dQ[2][2] / dd[2];
    Dc[2][2] += -h2 * z * Zb / dd[2];
    xzB += -h2 * Zb * dQ[2][0] / dd[0];
    Dc[2][0] += -h2 * xz * Zb / dd[0];
    yzB += -h2 * Zb * dQ[2][1] / dd[1];
    Dc[2][1] += -h2 * yz * Zb / dd[1];
    xzB += k2 * Zb * ab[1];
    abB[1] += k2 * Zb * xz;
    yzB -= k2 * Zb * ab[0];
    abB[0] -= k2 * Zb * yz;
    sfB += 2.0f * dt / rm * Zb * (xz * ab[1] - yz * ab[0]);
    // off-diagonal pressures (:285-356)
    const float XYb = ob[XY_], XZb = ob[XZ_], YZb = ob[YZ_];
    xyB += -dt * XYb * (dQ[0][0] / dd[0] + dQ[1][1] / dd[1]);
    Dc[0][0] += -dt * xy * XYb / dd[0];
    Dc[1][1] += -dt * xy * XYb / dd[1];
    yB += -dt * XYb * dQ[0][1] / dd[1];
    Dc[0][1] += -dt * y * XYb / dd[1];
    xB += -dt * XYb * dQ[1][0] / dd[0];
    Dc[1][0] += -dt * x * XYb / dd[0];
    yzB += -dt * XYb * dQ[0][2] / dd[2];
    Dc[0][2] += -dt * yz * XYb / dd[2];
    xzB += -dt * XYb * dQ[1][2] / dd[2];
    Dc[1][2] += -dt * xz * XYb / dd[2];
    abB[2] += k1 * XYb * (y - x);
    yB += k1 * XYb * ab[2];
    xB -= k1 * XYb * ab[2];
    abB[1] -= k1 * XYb * yz;
    yzB -= k1 * XYb * ab[1];
    abB[0] += k1 * XYb * xz;
    xzB += k1 * XYb * ab[0];
    sfB += dt / rm * XYb * (ab[2] * (y - x) - ab[1] * yz + ab[0] * xz);

    xzB += -dt * XZb * (dQ[0][0] / dd[0] + dQ[2][2] / dd[2]);
    Dc[0][0] += -dt * xz * XZb / dd[0];
    Dc[2][2] += -dt * xz * XZb / dd[2];
    zB += -dt * XZb * dQ[0][2] / dd[2];
    Dc[0][2] += -dt * z * XZb / dd[2];
    xB += -dt * XZb * dQ[2][0] / dd[0];
    Dc[2][0] += -dt * x * XZb / dd[0];
    yzB += -dt * XZb * dQ[0][1] / dd[1];
    Dc[0][1] += -dt * yz * XZb / dd[1];
    xyB += -dt * XZb * dQ[2][1] / dd[1];
    Dc[2][1] += -dt * xy * XZb / dd[1];
    abB[1] += k1 * XZb * (x - z);
    xB += k1 * XZb * ab[1];
    zB -= k1 * XZb * ab[1];
    abB[0] -= k1 * XZb * xy;
    xyB -= k1 * XZb * ab[0];
    abB[2] += k1 * XZb * yz;
    yzB += k1 * XZb * ab[2];
    sfB += dt / rm * XZb * (ab[1] * (x - z) - ab[0] * xy + ab[2] * yz);

    yzB += -dt * YZb * (dQ[1][1] / dd[1] + dQ[2][2] / dd[2]);
    Dc[1][1] += -dt * yz * YZb / dd[1];
    Dc[2][2] += -dt * yz * YZb / dd[2];
    yB += -dt * YZb * dQ[2][1] / dd[1];
    Dc[2][1] += -dt * y * YZb / dd[1];
    zB += -dt * YZb * dQ[1][2] / dd[2];
    Dc[1][2] += -dt * z * YZb / dd[2];
    xyB += -dt * YZb * dQ[2][0] / dd[0];
    Dc[2][0] += -dt * xy * YZb / dd[0];
    xzB += -dt * YZb * dQ[1][0] / dd[0];
    Dc[1][0] += -dt * xz * YZb / dd[0];
    abB[0] += k1 * YZb * (z - y);
    zB += k1 * YZb * ab[0];
    yB -= k1 * YZb * ab[0];
    abB[2] -= k1 * YZb * xz;
    xzB -= k1 * YZb * ab[2];
    abB[1] += k1 * YZb * xy;
    xyB += k1 * YZb * ab[1];
    sfB += dt / rm * YZb * (ab[0] * (z - y) - ab[2] * xz + ab[1] * xy);
    // sf = re bmax / |ab| where |ab| > bmax
    if (bm2 > bmax) {
        const float bm2B = -sfB * re * bmax / (bm2 * bm2);
        for (int a = 0; a < 3; ++a) abB[a] += bm2B * ab[a] / bm2;
    }
    // out = old + ... on the interior: old's adjoint first
    for (int f = 0; f < 10; ++f) olb[f] += ob[f];
    const float srcB[10] = {rB, xB, yB, zB, 0.0f, 0.0f, 0.0f, xyB, xzB, yzB};
    for (int f = 0; f < 10; ++f) {
        if (f >= PX_ && f <= PZ_) {
            A.oldb[f][c] = olb[f];       // the momenta's source adjoint comes in the gather
        } else if (A.sb[f] == A.oldb[f]) {
            A.oldb[f][c] = olb[f] + srcB[f];
        } else {
            A.oldb[f][c] = olb[f];
            A.sb[f][c] = sbv[f] + srcB[f];
        }
    }
    for (int a = 0; a < 3; ++a) {
        A.eb[a][c] = ebv[a];
        A.bsb[a][c] = bsbv[a] + bB[a];
        A.tvb[a][c] = tvbv[a] + tB[a];
        for (int d = 0; d < 3; ++d) D[(size_t)(3 * a + d) * n + c] = Dc[a][d];
        S[(size_t)a * n + c] = abB[a] / 12.0f;
        QL[(size_t)a * n + c] = vB[a];
    }
    if (btot_on) A.btotb[c] = btb_new;
}

__global__ void k_adj_ion_gather(IonAdj A) {
    int i, j, k;
    if (!cell_of(A.B, gtid(), i, j, k)) return;
    const Box& B = A.B;
    const int m = B.IX(i, j, k), n = B.V();
    const int st[3] = {1, B.nx, B.nx * B.ny};
    const int pos[3] = {i, j, k}, len[3] = {B.nx, B.ny, B.nz};
    const float dt = A.delt, dd[3] = {A.dxt, A.dyt, A.dzt};
    const float* D = A.scr;
    const float* S = A.scr + 9 * (size_t)n;
    const float* QL = A.scr + 12 * (size_t)n;
    auto inb = [&](int d, int s) { return pos[d] + s >= 0 && pos[d] + s < len[d]; };
    auto interior_nb = [&](int d, int s) {
        int q[3] = {i, j, k};
        q[d] += s;
        return B.interior(q[0], q[1], q[2]);
    };
    // the output adjoint times -dt at the neighbour m + s e_d, zero unless
    // that neighbour is interior (only interior cells are outputs)
    auto Fn = [&](int f, int d, int s) {
        return interior_nb(d, s) ? -dt * A.ob[f][m + s * st[d]] : 0.0f;
    };
    auto Dn = [&](int a, int d, int s) {
        return inb(d, s) ? D[(size_t)(3 * a + d) * n + m + s * st[d]] : 0.0f;
    };
    const float aq = fmaxf(A.src[R_][m], kDmin);
    float qv[3], w[10];
    for (int f = 0; f < 10; ++f) w[f] = A.src[f][m];
    for (int a = 0; a < 3; ++a) qv[a] = w[PX_ + a] / aq;
    float add[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    float qvB[3];
    for (int a = 0; a < 3; ++a) qvB[a] = QL[(size_t)a * n + m];
    // fluxes (w qv_d)[m] for the nine fields with a flux term
    for (int f = 1; f < 10; ++f)
        for (int d = 0; d < 3; ++d) {
            const float g = (Fn(f, d, -1) - Fn(f, d, 1)) / dd[d];
            add[f] += g * qv[d];
            qvB[d] += g * w[f];
        }
    // continuity: rho reads the momenta's centred differences
    for (int d = 0; d < 3; ++d) add[PX_ + d] += (Fn(R_, d, -1) - Fn(R_, d, 1)) / dd[d];
    // pressure gradients in the momenta (the reference divides all of gpx's
    // and the z term of gpy's differences by dxt)
    add[X_] += (Fn(PX_, 0, -1) - Fn(PX_, 0, 1)) / dd[0];
    add[XY_] += (Fn(PX_, 1, -1) - Fn(PX_, 1, 1)) / dd[0];
    add[XZ_] += (Fn(PX_, 2, -1) - Fn(PX_, 2, 1)) / dd[0];
    add[XY_] += (Fn(PY_, 0, -1) - Fn(PY_, 0, 1)) / dd[0];
    add[Y_] += (Fn(PY_, 1, -1) - Fn(PY_, 1, 1)) / dd[1];
    add[YZ_] += (Fn(PY_, 2, -1) - Fn(PY_, 2, 1)) / dd[0];
    add[XZ_] += (Fn(PZ_, 0, -1) - Fn(PZ_, 0, 1)) / dd[0];
    add[YZ_] += (Fn(PZ_, 1, -1) - Fn(PZ_, 1, 1)) / dd[1];
    add[Z_] += (Fn(PZ_, 2, -1) - Fn(PZ_, 2, 1)) / dd[2];
    // velocity differences
    for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d) qvB[a] += Dn(a, d, -1) - Dn(a, d, 1);
    // qv = p / max(rho, dmin)
    float rB = 0.0f;
    for (int a = 0; a < 3; ++a) {
        add[PX_ + a] += qvB[a] / aq;
        rB -= qvB[a] * qv[a] / aq;
    }
    if (A.src[R_][m] > kDmin) add[R_] += rB;
    // the smoothed field of the gyrotropic terms: (6 b[c] + 6 neighbours)/12,
    // read (with the accumulators) before the first store
    float acc[3], sbv[10], bsbv[3];
    for (int a = 0; a < 3; ++a) {
        const float* s = S + (size_t)a * n;
        acc[a] = 6.0f * s[m];
        for (int d = 0; d < 3; ++d) {
            if (inb(d, -1)) acc[a] += s[m - st[d]];
            if (inb(d, 1)) acc[a] += s[m + st[d]];
        }
        bsbv[a] = A.bsb[a][m];
    }
    for (int f = 0; f < 10; ++f) sbv[f] = A.sb[f][m];
    for (int f = 0; f < 10; ++f) A.sb[f][m] = sbv[f] + add[f];
    for (int a = 0; a < 3; ++a) A.bsb[a][m] = bsbv[a] + acc[a];
}

// lap_plasma, pass 1: per interior cell c, the adjoints the six |dv| carry
// to the velocity (sign of the difference times the sum over fields of the
// output adjoint times chi dt times the field's difference)
__global__ void k_adj_lapsp_coef(LapAdj L) {
    int i, j, k;
    if (!cell_of(L.B, gtid(), i, j, k)) return;
    const Box& B = L.B;
    const int c = B.IX(i, j, k), n = B.V();
    if (!B.interior(i, j, k)) {
        for (int q = 0; q < 6; ++q) L.scr[(size_t)q * n + c] = 0.0f;
        return;
    }
    const int st[3] = {1, B.nx, B.nx * B.ny};
    auto sgn = [](float x) { return (float)((x > 0.0f) - (x < 0.0f)); };
    for (int d = 0; d < 3; ++d) {
        const float* v = L.v[d];
        const float sp = sgn(v[c + st[d]] - v[c]), sm = sgn(v[c] - v[c - st[d]]);
        float ap = 0.0f, am = 0.0f;
        for (int f = 0; f < L.nf; ++f) {
            const float g = L.chi[f] * L.delt * L.wb[f][c];
            const float* q = L.q[f];
            ap += g * (q[c + st[d]] - q[c]);
            am -= g * (q[c] - q[c - st[d]]);
        }
        L.scr[(size_t)(2 * d) * n + c] = ap * sp;
        L.scr[(size_t)(2 * d + 1) * n + c] = am * sm;
    }
}

// pass 2: per cell m, the fields' adjoints (gathered as in lap_bfld's
// adjoint, each field with its own chi) and the velocity's
__global__ void k_adj_lapsp_gather(LapAdj L) {
    int i, j, k;
    if (!cell_of(L.B, gtid(), i, j, k)) return;
    const Box& B = L.B;
    const int m = B.IX(i, j, k), n = B.V();
    const int st[3] = {1, B.nx, B.nx * B.ny};
    const int pos[3] = {i, j, k}, len[3] = {B.nx, B.ny, B.nz};
    auto in = [&](int d, int s) {
        int q[3] = {i, j, k};
        q[d] += s;
        return B.interior(q[0], q[1], q[2]);
    };
    const bool im = B.interior(i, j, k);
    // |dv| on the faces around m, as the forward forms them
    float up[3] = {0, 0, 0}, um[3] = {0, 0, 0};   // |v(m+d) - v(m)|, |v(m) - v(m-d)|
    for (int d = 0; d < 3; ++d) {
        const float* v = L.v[d];
        if (pos[d] + 1 < len[d]) up[d] = fabsf(v[m + st[d]] - v[m]);
        if (pos[d] - 1 >= 0) um[d] = fabsf(v[m] - v[m - st[d]]);
    }
    for (int f = 0; f < L.nf; ++f) {
        const float g = L.chi[f] * L.delt;
        const float* wb = L.wb[f];
        float acc = 0.0f;
        if (im) acc += wb[m] * (1.0f - g * (up[0] + um[0] + up[1] + um[1] + up[2] + um[2]));
        for (int d = 0; d < 3; ++d) {
            // c = m - d reads q[c + d] = q[m] with |v(c+d) - v(c)| = um[d]
            if (in(d, -1)) acc += g * um[d] * wb[m - st[d]];
            if (in(d, 1)) acc += g * up[d] * wb[m + st[d]];
        }
        L.qb[f][m] = (L.assign ? 0.0f : L.qb[f][m]) + acc;
    }
    for (int d = 0; d < 3; ++d) {
        const float* P = L.scr + (size_t)(2 * d) * n;       // |v(c+d) - v(c)|
        const float* M = L.scr + (size_t)(2 * d + 1) * n;   // |v(c) - v(c-d)|
        float acc = -P[m] + M[m];
        if (pos[d] - 1 >= 0) acc += P[m - st[d]];
        if (pos[d] + 1 < len[d]) acc -= M[m + st[d]];
        L.vb[d][m] += acc;
    }
}

// v = p / max(rho, smallbit): p_bar += v_bar / den, rho_bar -= sum v_bar v / den
__global__ void k_adj_velocity(int n, const float* __restrict__ rho, const float* __restrict__ px,
                               const float* __restrict__ py, const float* __restrict__ pz,
                               const float* __restrict__ vxb, const float* __restrict__ vyb,
                               const float* __restrict__ vzb, float smallbit, float* rhob,
                               float* pxb, float* pyb, float* pzb) {
    const int c = gtid();
    if (c >= n) return;
    const float r = rho[c], den = fmaxf(r, smallbit);
    const float a = vxb[c], b = vyb[c], d = vzb[c];
    pxb[c] += a / den;
    pyb[c] += b / den;
    pzb[c] += d / den;
    if (r > smallbit) rhob[c] -= (a * (px[c] / den) + b * (py[c] / den) + d * (pz[c] / den)) / den;
}

size_t adj_ion_scratch_floats(const Box& B) { return (size_t)kIonScr * B.V(); }
size_t adj_lapsp_scratch_floats(const Box& B) { return (size_t)6 * B.V(); }

void launch_adj_push_ion(const IonAdj& A, hipStream_t s) {
    const int n = A.B.V();
    const dim3 g((n + 255) / 256), t(256);
    hipLaunchKernelGGL(k_adj_ion_local, g, t, 0, s, A);
    hipLaunchKernelGGL(k_adj_ion_gather, g, t, 0, s, A);
}

void launch_adj_lap_species(const LapAdj& L, hipStream_t s) {
    const int n = L.B.V();
    const dim3 g((n + 255) / 256), t(256);
    hipLaunchKernelGGL(k_adj_lapsp_coef, g, t, 0, s, L);
    hipLaunchKernelGGL(k_adj_lapsp_gather, g, t, 0, s, L);
}

void launch_adj_velocity(int n, const float* rho, const float* p[3], const float* vb[3],
                         float smallbit, float* rhob, float* pb[3], hipStream_t s) {
    hipLaunchKernelGGL(k_adj_velocity, dim3((n + 255) / 256), dim3(256), 0, s, n, rho, p[0],
                       p[1], p[2], vb[0], vb[1], vb[2], smallbit, rhob, pb[0], pb[1], pb[2]);
}

}  // namespace g3m3
