// Spacecraft sampling geometry shared by host and device code:
// code/cft_findgrid.f90:4-60 (smallest box containing the point, lower
// grid corner), the cube vertices of code/cft_find_aux_cube.f90:62-65 and
// the craft position between trajectory points (code/cft_linterp.f90).
#pragma once
#include <hip/hip_runtime.h>

namespace g3m3 {

constexpr int kCraftMaxGrids = 8;

struct CraftGrid {
    int n_grids;
    float mn[kCraftMaxGrids][3], mx[kCraftMaxGrids][3], xspac[kCraftMaxGrids];
    float re_equiv;
};

// gridpt[0..2] = 1-based lower corner, gridpt[3] = box; returns false when
// the point lies in no box (it is then clamped into the outermost one, as
// the reference does after printing its warning)
__host__ __device__ inline bool craft_findgrid(const CraftGrid& G, const float xyz[3],
                                               int gridpt[4]) {
    const float abit = 0.001f;
    float xa[3] = {xyz[0] / G.re_equiv, xyz[1] / G.re_equiv, xyz[2] / G.re_equiv};
    int box = 0;
    for (int b = 0; b < G.n_grids; ++b) {
        if (xa[0] > G.mn[b][0] && xa[1] > G.mn[b][1] && xa[2] > G.mn[b][2] &&
            xa[0] < G.mx[b][0] && xa[1] < G.mx[b][1] && xa[2] < G.mx[b][2]) {
            box = b + 1;
            break;
        }
    }
    const bool inside = box != 0;
    if (!inside) {
        box = G.n_grids;
        for (int a = 0; a < 3; ++a) {
            xa[a] = fmaxf(xa[a], G.mn[box - 1][a] + abit);
            xa[a] = fminf(xa[a], G.mx[box - 1][a] - abit);
        }
    }
    gridpt[3] = box;
    for (int a = 0; a < 3; ++a) {
        const float delta = xa[a] - G.mn[box - 1][a];
        gridpt[a] = 1 + (int)(delta / G.xspac[box - 1]);
    }
    return inside;
}

// cube_vertices(3,2): v0[a] low corner, v1[a] high corner (physical units)
__host__ __device__ inline void craft_cube(const CraftGrid& G, const int gridpt[4], float v0[3],
                                           float v1[3]) {
    const int b = gridpt[3] - 1;
    for (int a = 0; a < 3; ++a) {
        v0[a] = (G.mn[b][a] + ((float)gridpt[a] - 1.0f) * G.xspac[b]) * G.re_equiv;
        v1[a] = v0[a] + G.xspac[b] * G.re_equiv;
    }
}

// The craft position at time ut (code/cft_find_aux_cube.f90:43-47 with
// code/cft_linterp.f90:17-27): the trajectory point xc(1:3) when the two
// points share a time or ut has reached the second, otherwise the straight
// line from xc to zc at ut; xc, zc are (x, y, z, t)
__host__ __device__ inline void craft_position(const float xc[4], const float zc[4], float ut,
                                               float s[3]) {
    if (xc[3] == zc[3] || ut >= zc[3]) {
        for (int a = 0; a < 3; ++a) s[a] = xc[a];
    } else {
        for (int a = 0; a < 3; ++a) s[a] = xc[a] + (zc[a] - xc[a]) / (zc[3] - xc[3]) * (ut - xc[3]);
    }
}

}  // namespace g3m3
