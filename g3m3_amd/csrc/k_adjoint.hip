// Discrete adjoints of the magnetic-field half of a box-step on one box --
// calcur, total B, fnd_evel, BandE with the uniform resistivity,
// push_bfld, bndry_outer's field rules, lap_bfld, fcsmooth -- and of
// trilinear craft sampling.
// Each kernel is the transpose of the linearisation of the forward kernel
// it names (k_fields.hip, k_bounds.hip, k_smooth.hip, k_fused.hip,
// k_misc.hip), about the forward state it is given,
// in exact arithmetic: a forward assignment y = f(x) becomes x_bar +=
// (df/dx)^T y_bar, and cells a forward stage overwrites give their adjoint
// to their sources and are zeroed, stages taken in reverse order.  The
// plasma (momenta, densities, electron pressure) is held fixed, so only
// the paths through b, J, ev, E and eta are differentiated.
#include "g3m3_craft.h"
#include "g3m3_dev.h"
#include "g3m3_launch.h"

namespace g3m3 {

namespace {

__device__ inline void add_to(float* a, int c, float v) { atomicAdd(a + c, v); }

// sum over a 256-thread block (every thread must call it)
__device__ inline double adj_block_sum(double v) {
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

// push_bfld (code/sub_push_bfld.f90:19-47) on interior cells c:
//   bx = ox - dt ((ez[c+Y] - ez[c-Y]) / dyt - (ey[c+Z] - ey[c-Z]) / dzt)
//   by = oy + dt ((ez[c+X] - ez[c-X]) / dxt - (ex[c+Z] - ex[c-Z]) / dzt)
//   bz = oz - dt ((ey[c+X] - ey[c-X]) / dxt - (ex[c+Y] - ex[c-Y]) / dyt)
// The output array held old before the push, so every cell of it depends
// on old at the same cell with weight 1 (interior: the old term; boundary:
// left as it was).  Adjoint: ob += bb everywhere; E_bar gathered at every
// cell m from the interior cells whose stencil reads m.  eb is overwritten.
__global__ void k_adj_push_bfld(Box B, const float* __restrict__ bxb,
                                const float* __restrict__ byb, const float* __restrict__ bzb,
                                float* oxb, float* oyb, float* ozb, float* exb, float* eyb,
                                float* ezb, float delt, float dxt, float dyt, float dzt) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const int m = B.IX(i, j, k), X = 1, Y = B.nx, Z = B.nx * B.ny;
    oxb[m] += bxb[m];
    oyb[m] += byb[m];
    ozb[m] += bzb[m];
    auto in = [&](int di, int dj, int dk) { return B.interior(i + di, j + dj, k + dk); };
    const float cx = delt / dxt, cy = delt / dyt, cz = delt / dzt;
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    // ez: read by bx at c = m -+ Y (coefficient -+ dt/dyt), by at c = m -+ X (+- dt/dxt)
    if (in(0, -1, 0)) gz -= cy * bxb[m - Y];
    if (in(0, 1, 0)) gz += cy * bxb[m + Y];
    if (in(-1, 0, 0)) gz += cx * byb[m - X];
    if (in(1, 0, 0)) gz -= cx * byb[m + X];
    // ey: read by bx at c = m -+ Z (+- dt/dzt), by bz at c = m -+ X (-+ dt/dxt)
    if (in(0, 0, -1)) gy += cz * bxb[m - Z];
    if (in(0, 0, 1)) gy -= cz * bxb[m + Z];
    if (in(-1, 0, 0)) gy -= cx * bzb[m - X];
    if (in(1, 0, 0)) gy += cx * bzb[m + X];
    // ex: read by by at c = m -+ Z (-+ dt/dzt), by bz at c = m -+ Y (+- dt/dyt)
    if (in(0, 0, -1)) gx -= cz * byb[m - Z];
    if (in(0, 0, 1)) gx += cz * byb[m + Z];
    if (in(0, -1, 0)) gx += cy * bzb[m - Y];
    if (in(0, 1, 0)) gx -= cy * bzb[m + Y];
    exb[m] = gx;
    eyb[m] = gy;
    ezb[m] = gz;
}

// BandE's face copies (code/sub_BandE.f90:125-159): every boundary cell
// took the interior cell at the clamped index.  Scatter and zero.
__global__ void k_adj_clamp_copy(Box B, float* ax, float* ay, float* az) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k) || B.interior(i, j, k)) return;
    const int s = B.IX(min(max(i, 1), B.nx - 2), min(max(j, 1), B.ny - 2),
                       min(max(k, 1), B.nz - 2));
    const int c = B.IX(i, j, k);
    add_to(ax, s, ax[c]);
    add_to(ay, s, ay[c]);
    add_to(az, s, az[c]);
    ax[c] = 0.0f;
    ay[c] = 0.0f;
    az[c] = 0.0f;
}

// BandE's body (code/sub_BandE.f90:32-84 and the eta term), pointwise:
//   E = -(ev x bs) - grad pe / (n Re) [interior] + eta J / n [interior]
// with pe and n fixed: ev_bar = E_bar x bs, bs_bar += ev x E_bar,
// J_bar += eta E_bar / n and eta_bar += E_bar . J / n on the interior.
// The per-block eta_bar sums are added to eta_bar[0] in double.
__global__ void k_adj_bande(Box B, const float* __restrict__ exb, const float* __restrict__ eyb,
                            const float* __restrict__ ezb, const float* __restrict__ bsx,
                            const float* __restrict__ bsy, const float* __restrict__ bsz,
                            const float* __restrict__ evx, const float* __restrict__ evy,
                            const float* __restrict__ evz, const float* __restrict__ cx,
                            const float* __restrict__ cy, const float* __restrict__ cz,
                            const float* __restrict__ q, const float* __restrict__ h,
                            const float* __restrict__ o, float rmq, float rmh, float rmo,
                            float eta, float* bsxb, float* bsyb, float* bszb, float* evxb,
                            float* evyb, float* evzb, float* cxb, float* cyb, float* czb,
                            double* etab) {
    int i, j, k;
    double et = 0.0;
    if (cell_of(B, gtid(), i, j, k)) {
        const int c = B.IX(i, j, k);
        const float Ex = exb[c], Ey = eyb[c], Ez = ezb[c];
        const float bx = bsx[c], by = bsy[c], bz = bsz[c];
        const float vx = evx[c], vy = evy[c], vz = evz[c];
        evxb[c] = Ey * bz - Ez * by;
        evyb[c] = Ez * bx - Ex * bz;
        evzb[c] = Ex * by - Ey * bx;
        bsxb[c] += vy * Ez - vz * Ey;
        bsyb[c] += vz * Ex - vx * Ez;
        bszb[c] += vx * Ey - vy * Ex;
        if (B.interior(i, j, k)) {
            const float eden = q[c] / rmq + h[c] / rmh + o[c] / rmo;
            cxb[c] += eta * Ex / eden;
            cyb[c] += eta * Ey / eden;
            czb[c] += eta * Ez / eden;
            et = ((double)Ex * cx[c] + (double)Ey * cy[c] + (double)Ez * cz[c]) / eden;
        }
    }
    const double s = adj_block_sum(et);
    if (threadIdx.x == 0) atomicAdd(etab, s);
}

// fnd_evel (code/sub_fnd_evel.f90:318-346): ev = tv - J / tot / Re with the
// plasma fixed; J_bar -= ev_bar / tot / Re.
__global__ void k_adj_evel(int n, const float* __restrict__ evxb, const float* __restrict__ evyb,
                           const float* __restrict__ evzb, const float* __restrict__ q,
                           const float* __restrict__ h, const float* __restrict__ o, float rmq,
                           float rmh, float rmo, float reynolds, float* cxb, float* cyb,
                           float* czb) {
    const int c = gtid();
    if (c >= n) return;
    const float tot = (q[c] + 0.000001f) / rmq + (h[c] + 0.000001f) / rmh +
                      (o[c] + 0.000001f) / rmo;
    const float f = 1.0f / tot / reynolds;
    cxb[c] -= evxb[c] * f;
    cyb[c] -= evyb[c] * f;
    czb[c] -= evzb[c] * f;
}

// calcur's corners (code/sub_calcur.f90:264-289), one field: the eight
// corner formulas of k_calcur_corners transposed, in reverse order.
__device__ inline void adj_calcur_corners1(Box B, float* c) {
    const int nx = B.nx, ny = B.ny, nz = B.nz, n1 = nx - 2, m1 = ny - 2, l1 = nz - 2;
    auto at = [&](int i, int j, int k) -> float& { return c[B.IX(i, j, k)]; };
    auto take = [&](int i, int j, int k) {
        const float v = at(i, j, k);
        at(i, j, k) = 0.0f;
        return v;
    };
    float v = take(nx - 1, ny - 1, nz - 1) / 3.0f;
    at(nx - 1, ny - 1, l1) += v;
    at(nx - 1, m1, nz - 1) += v;
    at(n1, ny - 1, nz - 1) += v;
    v = take(nx - 1, 0, nz - 1) / 3.0f;
    at(nx - 1, 0, l1) += v;
    at(nx - 1, 1, nz - 1) += v;
    at(n1, 0, nz - 1) += v;
    v = take(nx - 1, ny - 1, 0) / 3.0f;
    at(nx - 1, ny - 1, 1) += v;
    at(nx - 1, m1, 0) += v;
    at(n1, ny - 1, 0) += v;
    v = take(nx - 1, 0, 0) / 3.0f;
    at(nx - 1, 0, 1) += v;
    at(nx - 1, 1, 0) += v;
    at(n1, 0, 0) += v;
    v = take(0, ny - 1, nz - 1) / 2.0f;
    at(0, ny - 1, l1) += v;
    at(0, m1, nz - 1) += v;
    v = take(0, 0, nz - 1) / 2.0f;
    at(0, 0, l1) += v;
    at(0, 1, nz - 1) += v;
    v = take(0, ny - 1, 0);
    at(0, ny - 1, 1) += v;
    v = take(0, 0, 0);
    at(0, 0, 1) += v;
}
__global__ void k_adj_calcur_corners(Box B, float* cx, float* cy, float* cz) {
    if (threadIdx.x == 0) adj_calcur_corners1(B, cx);
    if (threadIdx.x == 1) adj_calcur_corners1(B, cy);
    if (threadIdx.x == 2) adj_calcur_corners1(B, cz);
}

// calcur's edge lines (code/sub_calcur.f90:213-262): each edge cell took a
// face cell (the source rule of k_calcur_edges).  Scatter and zero.
__global__ void k_adj_calcur_edges(Box B, float* cx, float* cy, float* cz) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    const bool fi = (i == 0 || i == nx - 1), fj = (j == 0 || j == ny - 1),
               fk = (k == 0 || k == nz - 1);
    if ((int)fi + (int)fj + (int)fk != 2) return;
    int si = i, sj = j, sk = k;
    if (!fi) {
        if (j == ny - 1 && k == nz - 1) sj = ny - 2;
        else sk = (k == 0) ? 1 : nz - 2;
    } else if (!fj) {
        sk = (k == 0) ? 1 : nz - 2;
    } else {
        sj = (j == 0) ? 1 : ny - 2;
    }
    const int c = B.IX(i, j, k), s = B.IX(si, sj, sk);
    add_to(cx, s, cx[c]);
    add_to(cy, s, cy[c]);
    add_to(cz, s, cz[c]);
    cx[c] = 0.0f;
    cy[c] = 0.0f;
    cz[c] = 0.0f;
}

// calcur's faces (code/sub_calcur.f90:161-209): each face cell took its
// inner neighbour along the face normal.  Scatter and zero.
__global__ void k_adj_calcur_faces(Box B, float* cx, float* cy, float* cz) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const bool fi = (i == 0 || i == B.nx - 1), fj = (j == 0 || j == B.ny - 1),
               fk = (k == 0 || k == B.nz - 1);
    if ((int)fi + (int)fj + (int)fk != 1) return;
    int si = i, sj = j, sk = k;
    if (fk) sk = (k == 0) ? 1 : B.nz - 2;
    if (fi) si = (i == 0) ? 1 : B.nx - 2;
    if (fj) sj = (j == 0) ? 1 : B.ny - 2;
    const int c = B.IX(i, j, k), s = B.IX(si, sj, sk);
    add_to(cx, s, cx[c]);
    add_to(cy, s, cy[c]);
    add_to(cz, s, cz[c]);
    cx[c] = 0.0f;
    cy[c] = 0.0f;
    cz[c] = 0.0f;
}

// calcur's interior curl (code/sub_calcur.f90:126-149), transposed:
// b_bar gathered at every cell m from the interior cells c reading it.
//   cx[c] = (bz[c+Y] - bz[c-Y]) / dyt - (by[c+Z] - by[c-Z]) / dzt
//   cy[c] = (bx[c+Z] - bx[c-Z]) / dzt - (bz[c+X] - bz[c-X]) / dxt
//   cz[c] = (by[c+X] - by[c-X]) / dxt - (bx[c+Y] - bx[c-Y]) / dyt
__global__ void k_adj_calcur_int(Box B, const float* __restrict__ cxb,
                                 const float* __restrict__ cyb, const float* __restrict__ czb,
                                 float* bxb, float* byb, float* bzb, float dxt, float dyt,
                                 float dzt) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const int m = B.IX(i, j, k), X = 1, Y = B.nx, Z = B.nx * B.ny;
    auto in = [&](int di, int dj, int dk) { return B.interior(i + di, j + dj, k + dk); };
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    if (in(0, 0, -1)) gx += cyb[m - Z] / dzt;
    if (in(0, 0, 1)) gx -= cyb[m + Z] / dzt;
    if (in(0, -1, 0)) gx -= czb[m - Y] / dyt;
    if (in(0, 1, 0)) gx += czb[m + Y] / dyt;
    if (in(0, 0, -1)) gy -= cxb[m - Z] / dzt;
    if (in(0, 0, 1)) gy += cxb[m + Z] / dzt;
    if (in(-1, 0, 0)) gy += czb[m - X] / dxt;
    if (in(1, 0, 0)) gy -= czb[m + X] / dxt;
    if (in(0, -1, 0)) gz += cxb[m - Y] / dyt;
    if (in(0, 1, 0)) gz -= cxb[m + Y] / dyt;
    if (in(-1, 0, 0)) gz -= cyb[m - X] / dxt;
    if (in(1, 0, 0)) gz += cyb[m + X] / dxt;
    bxb[m] += gx;
    byb[m] += gy;
    bzb[m] += gz;
}

// bndry_outer (code/sub_bndry_outer.f90:50-728) on bx, by, bz, transposed.
// The forward bndry_outer runs eight ordered passes; within a pass
// the cells written and the cells read are disjoint, so the adjoint of a
// pass moves each written cell's adjoint to the cells it was formed from
// and zeroes it, and the passes run in reverse.  Pass 0 (the inflow face,
// b = sb - b0) sends its adjoint to sb: block sums in double into
// sbb[comp].  blockIdx.y = component (0 bx, 1 by, 2 bz).
__global__ void k_adj_outer_inflow(Box B, float* bxb, float* byb, float* bzb, double* sbb) {
    const int comp = blockIdx.y;
    float* q = comp == 0 ? bxb : (comp == 1 ? byb : bzb);
    const int t = gtid();
    double v = 0.0;
    if (t < B.ny * B.nz) {
        const int c = B.IX(0, t % B.ny, t / B.ny);
        v = q[c];
        q[c] = 0.0f;
    }
    const double sum = adj_block_sum(v);
    if (threadIdx.x == 0) atomicAdd(sbb + comp, sum);
}

__global__ void k_adj_outer_pass(Box B, float* bxb, float* byb, float* bzb, int pass) {
    const int comp = blockIdx.y;
    float* q = comp == 0 ? bxb : (comp == 1 ? byb : bzb);
    const int t = gtid();
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    auto take = [&](int i, int j, int k) {
        const int c = B.IX(i, j, k);
        const float v = q[c];
        q[c] = 0.0f;
        return v;
    };
    auto give = [&](int i, int j, int k, float v) { add_to(q, B.IX(i, j, k), v); };
    switch (pass) {
    case 1: {  // back wall i=nx <- nx-1
        if (t >= ny * nz) return;
        const int j = t % ny, k = t / ny;
        give(nx - 2, j, k, take(nx - 1, j, k));
        break;
    }
    case 2: {  // y faces, i=2..nx
        if (t >= (nx - 1) * nz) return;
        const int i = 1 + t % (nx - 1), k = t / (nx - 1);
        give(i, 1, k, take(i, 0, k));
        give(i, ny - 2, k, take(i, ny - 1, k));
        break;
    }
    case 3: {  // z top
        if (t >= (nx - 1) * ny) return;
        const int i = 1 + t % (nx - 1), j = t / (nx - 1);
        give(i, j, nz - 2, take(i, j, nz - 1));
        break;
    }
    case 4: {  // z bottom
        if (t >= (nx - 1) * ny) return;
        const int i = 1 + t % (nx - 1), j = t / (nx - 1);
        give(i, j, 1, take(i, j, 0));
        break;
    }
    case 5: {  // corner lines of the back wall, k=2..nz-1: bx copies, by/bz average
        if (t >= nz - 2) return;
        const int k = 1 + t;
        const float lo = take(nx - 1, 0, k), hi = take(nx - 1, ny - 1, k);
        if (comp == 0) {
            give(nx - 2, 0, k, lo);
            give(nx - 2, ny - 1, k, hi);
        } else {
            give(nx - 2, 0, k, 0.5f * lo);
            give(nx - 1, 1, k, 0.5f * lo);
            give(nx - 2, ny - 1, k, 0.5f * hi);
            give(nx - 1, ny - 2, k, 0.5f * hi);
        }
        break;
    }
    case 6: {  // top/bottom lines of the back wall, j=2..ny-1
        if (t >= ny - 2) return;
        const int j = 1 + t;
        give(nx - 1, j, 1, take(nx - 1, j, 0));
        give(nx - 2, j, nz - 1, take(nx - 1, j, nz - 1));
        break;
    }
    case 7: {  // x-directed edge lines, i=2..nx
        if (t >= nx - 1) return;
        const int i = 1 + t;
        give(i, 0, 1, take(i, 0, 0));
        give(i, ny - 1, 1, take(i, ny - 1, 0));
        give(i, 0, nz - 2, take(i, 0, nz - 1));
        give(i, ny - 1, nz - 2, take(i, ny - 1, nz - 1));
        break;
    }
    }
}

// lap_bfld (code/sub_lap_bfld.f90:21-61) with the velocity v fixed:
//   w[c] = q[c] + chi (dt (uxp (q[c+X] - q[c]) - uxm (q[c] - q[c-X])) + ... )
// on interior cells c, uxp = |vx[c+X] - vx[c]|, uxm = |vx[c] - vx[c-X]|.
// Gathered at every cell m: qb[m] = the adjoint from every interior c whose
// stencil reads m (qb is assigned).  The forward leaves w's boundary cells
// as they were (the stage-1 field after its boundary pass); their adjoint
// goes to that field's adjoint, w1b.
__global__ void k_adj_lap(Box B, const float* __restrict__ wxb, const float* __restrict__ wyb,
                          const float* __restrict__ wzb, const float* __restrict__ vx,
                          const float* __restrict__ vy, const float* __restrict__ vz, float chi,
                          float delt, float* qxb, float* qyb, float* qzb, float* w1xb,
                          float* w1yb, float* w1zb) {
    int i, j, k;
    if (!cell_of(B, gtid(), i, j, k)) return;
    const int m = B.IX(i, j, k), X = 1, Y = B.nx, Z = B.nx * B.ny;
    const float* wb[3] = {wxb, wyb, wzb};
    float* qb[3] = {qxb, qyb, qzb};
    float* w1b[3] = {w1xb, w1yb, w1zb};
    const float g = chi * delt;
    auto in = [&](int di, int dj, int dk) { return B.interior(i + di, j + dj, k + dk); };
    float cs = 0.0f;   // coefficient of w[m] in q[m]'s own update
    float cn[6] = {};  // coefficient of w[n] for n = m-X, m+X, m-Y, m+Y, m-Z, m+Z
    const bool im = B.interior(i, j, k);
    if (im) {
        const float su = fabsf(vx[m + X] - vx[m]) + fabsf(vx[m] - vx[m - X]) +
                         fabsf(vy[m + Y] - vy[m]) + fabsf(vy[m] - vy[m - Y]) +
                         fabsf(vz[m + Z] - vz[m]) + fabsf(vz[m] - vz[m - Z]);
        cs = 1.0f - g * su;
    }
    // neighbour n = m - X reads q[n + X] = q[m] with uxp(n) = |vx[m] - vx[m-X]|
    if (in(-1, 0, 0)) cn[0] = g * fabsf(vx[m] - vx[m - X]);
    if (in(1, 0, 0)) cn[1] = g * fabsf(vx[m + X] - vx[m]);
    if (in(0, -1, 0)) cn[2] = g * fabsf(vy[m] - vy[m - Y]);
    if (in(0, 1, 0)) cn[3] = g * fabsf(vy[m + Y] - vy[m]);
    if (in(0, 0, -1)) cn[4] = g * fabsf(vz[m] - vz[m - Z]);
    if (in(0, 0, 1)) cn[5] = g * fabsf(vz[m + Z] - vz[m]);
    const int off[6] = {-X, X, -Y, Y, -Z, Z};
    for (int f = 0; f < 3; ++f) {
        const float* w = wb[f];
        float acc = im ? cs * w[m] : 0.0f;
        for (int d = 0; d < 6; ++d)
            if (cn[d] != 0.0f) acc += cn[d] * w[m + off[d]];
        qb[f][m] = acc;
        if (!im) w1b[f][m] += w[m];
    }
}

// fcsmooth (code/sub_fcsmooth.f90:7-242), transposed about (old, wrk) with
// the limiter's choices recomputed from them.  With D the diffused field
// (wrk + chi * 7-point Laplacian of old on the interior, wrk elsewhere),
// d = differences of wrk and d1 = differences of D along an axis, the
// forward forms at each interior index f the flux
//   fn(f) = s max(0, min(s d1(f-1), chi |d(f)|, s d1(f+1))),  s = sign d(f),
// copies fn(0) = fn(1), fn(n-1) = fn(n-2) and d1(n-1) = d1(n-2), and
// subtracts the flux differences axis by axis; the output keeps wrk on the
// k faces.  fn is linear in d1 or d with the branch fixed: 1 (d1(f-1)),
// 2 (chi d(f)), 3 (d1(f+1)) or 0.  Three fields at once, blockIdx.y.
struct FctAdj {
    Box B;
    const float* old[3];
    const float* wrk[3];
    float* ob[3];        // the output's adjoint (read)
    float* D[3];         // scratch: the diffused field
    float* Db[3];        // scratch: D's adjoint
    float* wb[3];        // wrk's adjoint (assigned)
    float* oldb[3];      // old's adjoint (accumulated, or assigned with assign_old)
    float chi;
    bool assign_old;
};

__global__ void k_adj_fct_diffuse(FctAdj A) {
    int i, j, k;
    if (!cell_of(A.B, gtid(), i, j, k)) return;
    const int f = blockIdx.y, c = A.B.IX(i, j, k), Y = A.B.nx, Z = A.B.nx * A.B.ny;
    const float* O = A.old[f];
    const float* W = A.wrk[f];
    A.D[f][c] = A.B.interior(i, j, k)
                    ? W[c] + A.chi * (O[c + 1] + O[c - 1] + O[c + Y] + O[c - Y] + O[c + Z] +
                                      O[c - Z] - 6.0f * O[c])
                    : W[c];
}

// the limiter's branch at interior index f of a line (stride st, length n)
// through base cell index c0 (index 0 of the line)
__device__ inline int fct_branch(const float* D, const float* W, int c0, int st, int n, int f,
                                 float chi) {
    auto d1 = [&](int t) {
        t = min(t, n - 2);
        return D[c0 + (t + 1) * st] - D[c0 + t * st];
    };
    const float d = W[c0 + (f + 1) * st] - W[c0 + f * st];
    const float s = copysignf(1.0f, d);
    float mn = s * d1(f - 1);
    int br = 1;
    const float v2 = chi * fabsf(d);
    if (v2 < mn) {
        mn = v2;
        br = 2;
    }
    const float v3 = s * d1(f + 1);
    if (v3 < mn) {
        mn = v3;
        br = 3;
    }
    return mn > 0.0f ? br : 0;
}

// per cell: D's and wrk's adjoints (the limiter's branches recomputed from
// D and wrk where they are used).  Along each axis the flux fn(u) of
// interior index u carries g(u) = -U(u) + U(u+1) (U: the output's adjoint
// where the axis's flux differences act; the copy fn(0) = fn(1) folded in)
// to d1(u-1) (branch 1), chi d(u) (branch 2) or d1(u+1) (branch 3, which
// is d1(n-2) for u = n-2); d1(t) = D(t+1) - D(t), d(t) = wrk(t+1) - wrk(t).
// The output's own terms: the x pass starts from D on interior i and wrk on
// the i faces, interior k; the k faces keep wrk.
__global__ void k_adj_fct_cells(FctAdj A) {
    int i, j, k;
    if (!cell_of(A.B, gtid(), i, j, k)) return;
    const int f = blockIdx.y, c = A.B.IX(i, j, k);
    const int nx = A.B.nx, ny = A.B.ny, nz = A.B.nz;
    const int pos[3] = {i, j, k}, len[3] = {nx, ny, nz}, st[3] = {1, nx, nx * ny};
    const float* ob = A.ob[f];
    const float* D = A.D[f];
    const float* W = A.wrk[f];
    const float chi = A.chi;
    const bool kin = k > 0 && k < nz - 1;
    float db = 0.0f, wbar = 0.0f;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        if (ax != 2 && !kin) continue;       // the x and y fluxes act on interior k only
        const int t = pos[ax], s = st[ax], n = len[ax];
        // U(u), u = t-2 .. t+2; g(u) and the branch br(u), u = t-2 .. t+1
        float U[5], G[4];
        int br[4];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const int u = t - 2 + q;
            const bool ok = ax == 2 ? (u > 0 && u < n - 1) : (u >= 0 && u < n);
            U[q] = ok ? ob[c + (q - 2) * s] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int u = t - 2 + q;
            G[q] = (u >= 2 ? -U[q] : 0.0f) + (u + 1 <= n - 2 ? U[q + 1] : 0.0f);
            br[q] = (u >= 1 && u <= n - 2) ? fct_branch(D, W, c - t * s, s, n, u, chi) : 0;
        }
        // adjoints of d1(t-1), d1(t) (p) and of d(t-1), d(t) (q); index q = u - t + 2
        const float p0 = (br[2] == 1 ? G[2] : 0.0f) + (br[0] == 3 ? G[0] : 0.0f) +
                         (t - 1 == n - 2 && br[1] == 3 ? G[1] : 0.0f);
        const float p1 = (br[3] == 1 ? G[3] : 0.0f) + (br[1] == 3 ? G[1] : 0.0f) +
                         (t == n - 2 && br[2] == 3 ? G[2] : 0.0f);
        const float q0 = br[1] == 2 ? chi * G[1] : 0.0f;
        const float q1 = br[2] == 2 ? chi * G[2] : 0.0f;
        if (t >= 1) {
            db += p0;
            wbar += q0;
        }
        if (t <= n - 2) {
            db -= p1;
            wbar -= q1;
        }
    }
    const float o = ob[c];
    if (kin) {
        if (i > 0 && i < nx - 1) db += o;     // the x pass's D term
        else wbar += o;                       // its i-face cells keep wrk
    } else {
        wbar += o;                            // k faces keep wrk
    }
    A.Db[f][c] = A.B.interior(i, j, k) ? db : 0.0f;
    A.wb[f][c] = wbar + db;                   // D = wrk + ...
}

__global__ void k_adj_fct_old(FctAdj A) {
    int i, j, k;
    if (!cell_of(A.B, gtid(), i, j, k)) return;
    const int f = blockIdx.y, c = A.B.IX(i, j, k);
    const int X = 1, Y = A.B.nx, Z = A.B.nx * A.B.ny;
    const float* Db = A.Db[f];
    float acc = 0.0f;
    if (i > 0) acc += Db[c - X];
    if (i < A.B.nx - 1) acc += Db[c + X];
    if (j > 0) acc += Db[c - Y];
    if (j < A.B.ny - 1) acc += Db[c + Y];
    if (k > 0) acc += Db[c - Z];
    if (k < A.B.nz - 1) acc += Db[c + Z];
    acc -= 6.0f * Db[c];
    A.oldb[f][c] = (A.assign_old ? 0.0f : A.oldb[f][c]) + A.chi * acc;
}

// y += x on three fields of n cells
__global__ void k_add3(int n, const float* __restrict__ x0, const float* __restrict__ x1,
                       const float* __restrict__ x2, float* y0, float* y1, float* y2) {
    const int c = gtid();
    if (c >= n) return;
    y0[c] += x0[c];
    y1[c] += x1[c];
    y2[c] += x2[c];
}

// trilinear craft sampling (k_sample), transposed: each sampled field's
// adjoint receives the sample's adjoint times the trilinear weight of each
// of the 8 cube corners.  Points outside a box (gridpt 0) give nothing.
__global__ void k_adj_sample(int npts, const float* __restrict__ sxyz, CraftGrid G, int nx,
                             int ny, int nz, int nq, FieldPtrs Fb, const float* __restrict__ sb) {
    const int p = gtid();
    if (p >= npts) return;
    const float s[3] = {sxyz[3 * p], sxyz[3 * p + 1], sxyz[3 * p + 2]};
    int gp[4];
    craft_findgrid(G, s, gp);
    if (!(gp[0] >= 1 && gp[0] <= nx - 1 && gp[1] >= 1 && gp[1] <= ny - 1 && gp[2] >= 1 &&
          gp[2] <= nz - 1))
        return;
    float v0[3], v1[3];
    craft_cube(G, gp, v0, v1);
    const float xd = (s[0] - v0[0]) / (v1[0] - v0[0]);
    const float yd = (s[1] - v0[1]) / (v1[1] - v0[1]);
    const float zd = (s[2] - v0[2]) / (v1[2] - v0[2]);
    const size_t V = (size_t)nx * ny * nz;
    const size_t c = (size_t)(gp[3] - 1) * V + (size_t)(gp[2] - 1) * nx * ny +
                     (size_t)(gp[1] - 1) * nx + (gp[0] - 1);
    const float wx[2] = {1.0f - xd, xd}, wy[2] = {1.0f - yd, yd}, wz[2] = {1.0f - zd, zd};
    for (int q = 0; q < nq; ++q) {
        const float g = sb[q + nq * p];
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx)
                    atomicAdd(Fb.p[q] + c + dx + (size_t)nx * dy + (size_t)nx * ny * dz,
                              g * (wx[dx] * wy[dy] * wz[dz]));
    }
}

// --- launches --------------------------------------------------------------

void launch_adj_push_bfld(const Box& B, const float* bb[3], float* ob[3], float* eb[3],
                          float delt, float rx, float ry, float rz, hipStream_t s) {
    const int n = B.V();
    hipLaunchKernelGGL(k_adj_push_bfld, dim3((n + 255) / 256), dim3(256), 0, s, B, bb[0], bb[1],
                       bb[2], ob[0], ob[1], ob[2], eb[0], eb[1], eb[2], delt, 2.0f * rx,
                       2.0f * ry, 2.0f * rz);
}

void launch_adj_bande(const Box& B, float* eb[3], const float* bs[3], const float* ev[3],
                      const float* cur[3], const float* rho[3], float rmq, float rmh, float rmo,
                      float eta, float* bsb[3], float* evb[3], float* curb[3], double* etab,
                      hipStream_t s) {
    const int n = B.V();
    const dim3 g((n + 255) / 256), t(256);
    hipLaunchKernelGGL(k_adj_clamp_copy, g, t, 0, s, B, eb[0], eb[1], eb[2]);
    hipLaunchKernelGGL(k_adj_bande, g, t, 0, s, B, eb[0], eb[1], eb[2], bs[0], bs[1], bs[2],
                       ev[0], ev[1], ev[2], cur[0], cur[1], cur[2], rho[0], rho[1], rho[2], rmq,
                       rmh, rmo, eta, bsb[0], bsb[1], bsb[2], evb[0], evb[1], evb[2], curb[0],
                       curb[1], curb[2], etab);
}

void launch_adj_bande_faces(const Box& B, float* eb[3], hipStream_t s) {
    const int n = B.V();
    hipLaunchKernelGGL(k_adj_clamp_copy, dim3((n + 255) / 256), dim3(256), 0, s, B, eb[0], eb[1],
                       eb[2]);
}

void launch_adj_evel(const Box& B, const float* evb[3], const float* rho[3], float rmq, float rmh,
                     float rmo, float reynolds, float* curb[3], hipStream_t s) {
    const int n = B.V();
    hipLaunchKernelGGL(k_adj_evel, dim3((n + 255) / 256), dim3(256), 0, s, n, evb[0], evb[1],
                       evb[2], rho[0], rho[1], rho[2], rmq, rmh, rmo, reynolds, curb[0], curb[1],
                       curb[2]);
}

void launch_adj_calcur(const Box& B, float* curb[3], float* bb[3], float rx, float ry, float rz,
                       hipStream_t s) {
    const int n = B.V();
    const dim3 g((n + 255) / 256), t(256);
    hipLaunchKernelGGL(k_adj_calcur_corners, dim3(1), dim3(64), 0, s, B, curb[0], curb[1],
                       curb[2]);
    hipLaunchKernelGGL(k_adj_calcur_edges, g, t, 0, s, B, curb[0], curb[1], curb[2]);
    hipLaunchKernelGGL(k_adj_calcur_faces, g, t, 0, s, B, curb[0], curb[1], curb[2]);
    hipLaunchKernelGGL(k_adj_calcur_int, g, t, 0, s, B, curb[0], curb[1], curb[2], bb[0], bb[1],
                       bb[2], 2.0f * rx, 2.0f * ry, 2.0f * rz);
}

void launch_adj_outer_b(const Box& B, float* bb[3], double* sbb, hipStream_t s) {
    const int nx = B.nx, ny = B.ny, nz = B.nz;
    const int ext[8] = {ny * nz, ny * nz, (nx - 1) * nz, (nx - 1) * ny, (nx - 1) * ny,
                        nz - 2, ny - 2, nx - 1};
    for (int p = 7; p >= 1; --p)
        hipLaunchKernelGGL(k_adj_outer_pass, dim3((ext[p] + 255) / 256, 3), dim3(256), 0, s, B,
                           bb[0], bb[1], bb[2], p);
    hipLaunchKernelGGL(k_adj_outer_inflow, dim3((ext[0] + 255) / 256, 3), dim3(256), 0, s, B,
                       bb[0], bb[1], bb[2], sbb);
}

void launch_adj_lap(const Box& B, const float* wb[3], const float* v[3], float chi, float delt,
                    float* qb[3], float* w1b[3], hipStream_t s) {
    const int n = B.V();
    hipLaunchKernelGGL(k_adj_lap, dim3((n + 255) / 256), dim3(256), 0, s, B, wb[0], wb[1], wb[2],
                       v[0], v[1], v[2], chi, delt, qb[0], qb[1], qb[2], w1b[0], w1b[1], w1b[2]);
}

size_t adj_fct_scratch_floats(const Box& B) { return (size_t)3 * 2 * B.V(); }

void launch_adj_fct(const Box& B, const float* old[3], const float* wrk[3], float* ob[3],
                    float* wb[3], float* oldb[3], float chi, float* scratch, hipStream_t s,
                    int nf, bool assign_old) {
    const int n = B.V();
    FctAdj A;
    A.B = B;
    A.chi = chi;
    A.assign_old = assign_old;
    for (int f = 0; f < nf; ++f) {
        float* base = scratch + (size_t)2 * n * f;
        A.old[f] = old[f];
        A.wrk[f] = wrk[f];
        A.ob[f] = ob[f];
        A.D[f] = base;
        A.Db[f] = base + (size_t)n;
        A.wb[f] = wb[f];
        A.oldb[f] = oldb[f];
    }
    const dim3 g((n + 255) / 256, nf), t(256);
    hipLaunchKernelGGL(k_adj_fct_diffuse, g, t, 0, s, A);
    hipLaunchKernelGGL(k_adj_fct_cells, g, t, 0, s, A);
    hipLaunchKernelGGL(k_adj_fct_old, g, t, 0, s, A);
}

void launch_add3(int n, const float* x[3], float* y[3], hipStream_t s) {
    hipLaunchKernelGGL(k_add3, dim3((n + 255) / 256), dim3(256), 0, s, n, x[0], x[1], x[2], y[0],
                       y[1], y[2]);
}

void launch_adj_sample(int npts, const float* sxyz, const CraftGrid& G, int nx, int ny, int nz,
                       int nq, const FieldPtrs& Fb, const float* sb, hipStream_t s) {
    if (npts <= 0) return;
    hipLaunchKernelGGL(k_adj_sample, dim3((npts + 255) / 256), dim3(256), 0, s, npts, sxyz, G,
                       nx, ny, nz, nq, Fb, sb);
}

}  // namespace g3m3
