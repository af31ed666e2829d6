// Device-side helpers shared by the G3M^3 hot-path kernels (gfx950).
//
// Layout: a field is the reference's Fortran array q(nx,ny,nz,n_grids),
// column-major, i fastest.  A kernel always receives the base address of
// one box slice, q + (box-1)*V, and indexes it with IX(i,j,k) (0-based).
//
// All arithmetic is IEEE single precision with no contraction into fused
// multiply-adds (the library is built with -ffp-contract=off), and each
// expression keeps the reference's left-to-right order, so a kernel gives
// the same bits as a serial build of the reference subroutine it replaces.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace g3m3 {

constexpr int kMaxFields = 34;

struct Box {
    int nx, ny, nz;
    __host__ __device__ int V() const { return nx * ny * nz; }
    __host__ __device__ int IX(int i, int j, int k) const { return i + nx * (j + ny * k); }
    __device__ bool interior(int i, int j, int k) const {
        return i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && k > 0 && k < nz - 1;
    }
};

// Flat thread id over a box -> (i,j,k); false past the end.
__device__ inline bool cell_of(const Box& b, int gid, int& i, int& j, int& k) {
    if (gid >= b.V()) return false;
    i = gid % b.nx;
    int t = gid / b.nx;
    j = t % b.ny;
    k = t / b.ny;
    return true;
}

__device__ inline int gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// Largest box volume for cell_of_fast (host launchers check it).
constexpr int kFastDivMax = 1 << 23;

// n / d for 0 <= n < 2^23, d >= 1: a float-reciprocal estimate (within 1
// of the quotient at these sizes) and two exact integer corrections --
// a few instructions instead of a software integer division.
__device__ inline int udiv_small(int n, int d) {
    int q = (int)((float)n * __builtin_amdgcn_rcpf((float)d));
    int r = n - q * d;
    q += (r >= d) - (r < 0);
    r = n - q * d;
    q += (r >= d) - (r < 0);
    return q;
}

// cell_of for boxes of fewer than 2^23 cells
__device__ inline bool cell_of_fast(const Box& b, int gid, int& i, int& j, int& k) {
    if (gid >= b.V()) return false;
    const int t = udiv_small(gid, b.nx);
    i = gid - t * b.nx;
    k = udiv_small(t, b.ny);
    j = t - k * b.ny;
    return true;
}

// A table of field pointers passed by value (kernel argument space).
struct FieldPtrs {
    float* p[kMaxFields];
};

__device__ inline float sgn1(float x) { return copysignf(1.0f, x); }


}  // namespace g3m3
