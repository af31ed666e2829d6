// Parameter blocks of the fused box-step kernels (passed by value).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include <cmath>

#include <cstring>
#include <vector>

#include "g3m3_dev.h"
#include "g3m3_launch.h"

namespace g3m3 {

// one ion species: 0 rho, 1..3 presx/y/z, 4..6 px/py/pz, 7..9 presxy/xz/yz
struct SpeciesSet {
    float* f[10];
};

struct EParams {
    Box B;
    const float* b[3];
    const float* b0[3];
    const float* mom[9];     // q px,py,pz; h px,py,pz; o px,py,pz
    const float* rho[3];
    const float* ep;
    const float* rst;        // rst on BandE's list points (box <= mbndry), else null
    float* e[3];
    float* ev[3];
    float* tv[3];
    float* bs[3];
    float rmq, rmh, rmo, reynolds, dxt, dyt, dzt, eta;
    // reciprocals of the above, used when pow2: each of them is a power of
    // two, so x * (1/d) is x / d exactly
    float rrmq, rrmh, rrmo, rreynolds, rdxt, rdyt, rdzt;
    int pow2;
};

struct PushParams {
    Box B;
    SpeciesSet src[3], old[3], out[3];
    const float* bs[3];
    const float* tv[3];
    const float* e[3];
    const float* ev[3];
    float* bout[3];
    const float* bold[3];
    const float* wrkepres;
    const float* oldepres;
    float* epres;
    float rmass[3];
    float rrmass[3];            // 1/rmass (used when pow2)
    float bmax[3];              // 0.20*rmass/delt/reynolds per species
    float delt, reynolds, smallbit;
    float dxt, dyt, dzt;        // ions: 2 * grid spacing of the box extents
    float fdxt, fdyt, fdzt;     // fields and electrons: 2 * xspac
    float rdxt, rdyt, rdzt, rfdxt, rfdyt, rfdzt;   // reciprocals (used when pow2)
    int pow2;                   // every step and mass above is a power of two
    float xmin, ymin, zmin, ddx, ddy, ddz, re_equiv, grav, gamma, gamma1;
    int do_elec;
    int copy_shell;             // copy non-interior cells old -> out
    int march;                  // species by k_push_march; k_push does B, electrons, shells
    int nspecies;               // species k_push_march pushes (src/old/out[0 .. nspecies-1])
    int march_kc;               // k-planes per k_push_march block (0: 16)
    const float* btot;          // non-null: |B| of the species push read from here (the
                                // reference's btot) instead of formed from bs
    // time step from device memory (captured steps): delt = dt_scale *
    // *dt_dev, and bmax is formed from it; null: delt and bmax above
    const float* dt_dev;
    float dt_scale;
};

struct ValEntry {
    int cell;
    int species;
    float v[10];
};

struct ValParams {
    Box B;
    SpeciesSet sp[3];
    float* ep;
    const float* b[3];
    const float* b0[3];
    ValEntry* list;
    int* count;              // entries of this pass (zero on entry)
    int* count_next;         // the next pass's counter, cleared by this pass
    float* partials;         // per-block maxima (null: not needed this pass)
    float aniso, smallbit, vlim, cslim, iso_flag;
    int isotropic;
    float svlim[3], sclim[3];
};

struct LapParams {
    Box B;
    SpeciesSet src[3], dst[3];
    const float* ep_src;
    float* ep_dst;
    const float* b_src[3];
    float* b_dst[3];
    float chirho, chipxyz, chierg, delt, smallbit;
    const float* dt_dev;     // non-null: delt = *dt_dev (captured steps)
    int march;               // species by k_lap_march; k_lap does epres and B
};

struct FctParams {
    Box B;
    float* out[kMaxFields];
    const float* old[kMaxFields];
    const float* wrk[kMaxFields];
    float chi[kMaxFields];
    int nf;                  // fields in use (k_fct_fused)
    int kc_len;              // k-planes per block (0: 16)
    int no_remap;            // 1: plain block order instead of the XCD-contiguous one
};

// One cell of the 2-layer flank shell of a fine box with its precomputed
// trilinear weights (code/sub_flanks.f90:25-45): c = the fine cell, cb =
// the coarse cell (ib-1, jb-1, kb-1), dx/dy/dz = the fractional offsets,
// flags = bits 0-5 panels bottom, top, front, back, left, right containing
// the cell, bits 8-13 the same panels when the cell is on their inner layer.
// cb, sy and sz index the array the coarse values are read from: the coarse
// box itself (sy = nx, sz = nx*ny) or, after build_shell_slab, the gathered
// slab (cb a slab index, sy/sz the strides of the cell's face block).
struct ShellPt {
    int c, cb;
    float dx, dy, dz;
    int flags;
    int sy, sz;
};

// the shell of a box in panel-slab order, weights computed as the kernel
// of the reference computes them (host, IEEE single precision)
void build_shell(const Box& B, const FlankGeom& G, ShellPt* out);
int shell_count(const Box& B);
// The coarse cells a shell reads, as six dense blocks (one per face of the
// shell: the bounding box of its cells' trilinear corners, x fastest):
// gather[n] = the coarse cell of slab entry n; the points' cb/sy/sz are
// rewritten to index the slab.  False if a corner lies outside the box.
bool build_shell_slab(const Box& B, ShellPt* pts, int npts, std::vector<int>& gather);
// slab[(2f) n + t] = old[f][idx[t]], slab[(2f+1) n + t] = cur[f][idx[t]]
void launch_slab_gather(const int* idx, int n, const FlankPtrs& Fp, int nf, float* slab,
                        hipStream_t s);
void launch_flanks_table(const Box& B, const ShellPt* tab, int npts, const FlankPtrs& Fp, int nf,
                         float t1, float t2, hipStream_t s);

// ---------------------------------------------------------------------------
// Several runs in one launch.  Run r owns the logical blocks [start[r],
// start[r + 1]) of the launch (start[nruns] = total); the grid is the total
// padded to a multiple of 8, taken in the XCD-aware order, so each XCD
// sweeps a contiguous run of the concatenated blocks.  Per-run arguments
// live in device memory, one struct per run.
struct MultiRun {
    const int* start;
    int nruns, total;
};
// atomicAdd on a counter in device memory, issued as a global atomic.  A
// pointer read from device memory (a multi-run launch's per-run arguments)
// is not known to the compiler to point there, and it would otherwise
// issue a flat atomic, which runs many times slower here on a contended
// counter.
__device__ inline int atomic_add_global(int* p, int v) {
    return __hip_atomic_fetch_add((__attribute__((address_space(1))) int*)p, v, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
}

// The run of this block of a multi-run launch and the block's logical
// index within that run's blocks; false for the grid's padding blocks
__device__ inline bool multi_block(const MultiRun& M, int& r, int& lb) {
    const int b = blockIdx.x, nb = gridDim.x;
    const int g = (b & 7) * (nb >> 3) + (b >> 3);   // the XCD-aware order
    if (g >= M.total) return false;
    int lo = 0, hi = M.nruns - 1;   // the last run whose first block is <= g
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (M.start[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    r = lo;
    lb = g - M.start[lo];
    return true;
}
// the arguments of the stage kernels that take more than their parameter
// struct (as one struct, the form a multi-run launch reads per run)
struct MarchArgs {
    PushParams P;
    int ntx, nty, nkc;
};
struct LapMarchArgs {
    LapParams L;
    int ntx, nty, nkc;
};
struct FctArgs {
    FctParams F;
    int ntx, nty, nkc;
};
struct ApplyArgs {
    ValParams Q;
    float* maxima;
    int nblocks;
};
struct OuterArgs {
    Box B;
    OuterPtrs O;
};
// which = -1: the three lists of Ls (blockIdx.y = the list); 0..2 one list
struct InnerArgs {
    Box B;
    InnerPtrs I;
    InnerLists Ls;
    int which, mbndry, b;
    float xmin, ymin, dx, dy, re_equiv, v_rot;
};
struct NextDtArgs {
    const float* m;
    float stepsz, xspac, t_step_max;
    float* dt;
    double* t;
};
// the kernels a single-grid box-step launches, as recorded launches
enum KernelId {
    kKEfield = 0, kKEfieldP2, kKPushMarch, kKPushMarchP2, kKPush, kKPushP2, kKValidate,
    kKValidateApply, kKLapMarch, kKLap, kKFct, kKOuter, kKInner, kKNextDt, kKernelIds
};
// One launch as the one-run path would issue it: the kernel, its grid and
// block, and its arguments (the struct above, or the parameter struct)
struct LaunchRec {
    int id;
    dim3 grid, block;
    std::vector<unsigned char> args;
};
struct Recorder {
    std::vector<LaunchRec> recs;
    template <class A>
    void add(int id, dim3 grid, dim3 block, const A& a) {
        LaunchRec r{id, grid, block, {}};
        r.args.resize(sizeof(A));
        std::memcpy(r.args.data(), &a, sizeof(A));
        recs.push_back(std::move(r));
    }
};
// Non-null: the launchers of the stage kernels append to it instead of
// launching (the host thread's own recorder)
extern thread_local Recorder* g_rec;
size_t kernel_args_size(int id);
// every run's arguments for kernel id (nruns structs of kernel_args_size)
void launch_multi(int id, const void* args, const MultiRun& M, dim3 block, unsigned gy,
                  hipStream_t s);
void launch_multi_bounds(int id, const void* args, const MultiRun& M, dim3 block, unsigned gy,
                         hipStream_t s);

void launch_efield(const EParams& P, hipStream_t s);
void launch_push(const PushParams& P, hipStream_t s);
// the species of P (nspecies of them) by k_push_march alone
void launch_push_species(const PushParams& P, hipStream_t s);

// x is a finite nonzero power of two (then 1/x is exact)
inline bool is_pow2(float x) {
    int e = 0;
    return x != 0.0f && std::isfinite(x) && std::fabs(std::frexp(x, &e)) == 0.5f;
}
void launch_validate(const ValParams& Q, float* maxima, hipStream_t s);
int validate_blocks(const Box& B);
void launch_lap(const LapParams& L, hipStream_t s);
void launch_fct_fused(const FctParams& F, int nf, hipStream_t s);


}  // namespace g3m3
