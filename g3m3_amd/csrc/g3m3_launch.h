// Host-side launchers of the G3M^3 kernels and the pointer tables they take.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "g3m3_dev.h"

namespace g3m3 {

// one ion species: index 0 rho, 1..3 presx/y/z, 4..6 px/py/pz, 7..9 presxy/xz/yz
struct IonPtrs {
    float* out[10];
    const float* old[10];
    const float* wrk[10];
    const float* bs[3];
    const float* btot;
    const float* e[3];
    float* qv[3];
    const float* tv[3];
};

struct SpeciesPtrs {
    float* f[3][10];   // [species q,h,o][same index as IonPtrs]
};

struct SpeedParams {
    float aniso, smallbit, vlim, cslim, iso_flag;
    int isotropic;
    float svlim[3], sclim[3];
};

struct LapPtrs {
    const float* src[10];
    float* dst[10];
    float chi[10];
};

enum OuterKind { kRHO = 0, kPRES, kPX, kPY, kPZ, kOFF, kEP, kBX, kBY, kBZ };
struct OuterPtrs {
    float* f[kMaxFields];
    int kind[kMaxFields];
    float val[kMaxFields];
    const float* b0[3];
};

// q(10), h(10), o(10), epres, bx, by, bz -- IonPtrs order within a species
struct InnerPtrs {
    float* f[kMaxFields];
};

struct InnerLists {
    const int* ij[3];
    const float* parm[3];
    int num[3];
};

struct FlankPtrs {
    float* tgt[kMaxFields];
    const float* cur[kMaxFields];
    const float* old[kMaxFields];
};

// fine box (s*) and coarse box (b*) origins and spacings
struct FlankGeom {
    float sxmin, sx, bxmin, delx;
    float symin, sy, bymin, dely;
    float szmin, sz, bzmin, delz;
};

enum { kDivbRef = 0, kDivbF64 = 1, kDivbCG = 2, kDivbRefine = 3 };

// one box's origin and cell size (code/3d_main.f90:1247-1252)
struct InitGeom {
    float xmin, ymin, zmin, dx, dy, dz;
};
// common /rotation/'s dipole centre, tilt and strength (code/3d_main.f90:
// 633-634) with the sine and cosine of its rotation angle
struct RotationGeom {
    float xdip, ydip, zdip, sin_rot, cos_rot, sin_tilt, cos_tilt, b0;
};
constexpr int kMaxInitGrids = 32;
// the initialisation block (code/3d_main.f90:1238-1633): fields in the
// order q, h, o x (rho, presx, presy, presz, px, py, pz, presxy, presxz,
// presyz), epres, bx, by, bz, each (nx,ny,nz,n_grids); the point lists of
// boxes 1..mbndry as the reference lays them out, ij(mbndry,3,m) and
// parm(mbndry,7,m); cls = a class byte per cell of boxes 1..mbndry;
// counts = (srf, mid, zero) per box, cut at the capacity, then an
// overflow flag
struct InitPlasma {
    Box B;
    float* f[34];
    float* b0f[3];
    unsigned char* cls;
    int* counts;
    int *ijsrf, *ijmid, *ijzero;
    float *parm_srf, *parm_mid, *parm_zero;
    int mbndry, msrf, mmid, mzero;
    InitGeom geom[kMaxInitGrids];
    RotationGeom rot;
    float re_equiv, r_rot, v_rot, zero_bndry, mid_bndry, srf_bndry, r_inner, alpha_e, erho, d_min,
        cs_inner;
    float rmassq, rmassh, rmasso, ti_te;
    float srho, spx, spy, spz, spress, cs_wind, gamma, rho_frac, wind_bnd;
};


// Optional per-kernel timing: when the kernel named by g3m3_prof_select()
// is launched, HIP events are recorded around it on its stream.
void prof_begin(const char* kernel, hipStream_t s);
void prof_end(const char* kernel, hipStream_t s);
struct ProfScope {
    const char* k;
    hipStream_t s;
    ProfScope(const char* kernel, hipStream_t st) : k(kernel), s(st) { prof_begin(k, s); }
    ~ProfScope() { prof_end(k, s); }
};
enum { kScAnum = 0, kScGam = 1, kScTmp = 4 };

void launch_calcur(const Box& B, const float* bx, const float* by, const float* bz, float* cx,
                   float* cy, float* cz, float rx, float ry, float rz, hipStream_t s);
void launch_bfield_evel(const Box& B, const float* b0[3], const float* b[3], float* bs[3],
                        float* btot, const float* mom[9], const float* rho[3], const float* cur[3],
                        float* ev[3], float* tv[3], float rmq, float rmh, float rmo,
                        float reynolds, float smallbit, int do_b, int do_v, hipStream_t s);
void launch_bande(const Box& B, float* e[3], const float* bs[3], const float* cur[3],
                  const float* ev[3], const float* ep, const float* rho[3], const float* rst,
                  float resist, float reynolds, int box, float rmq, float rmh, float rmo,
                  const int* ijmid, int nummid, const int* ijzero, int numzero, int mbndry,
                  float rx, float ry, float rz, float eta, hipStream_t s);
void launch_push_elec(const Box& B, const float* w, const float* old, float* ep, const float* ev[3],
                      float gamma, float gamma1, float delt, float rx, float ry, float rz,
                      hipStream_t s);
void launch_push_bfld(const Box& B, float* b[3], const float* old[3], const float* e[3], float delt,
                      float rx, float ry, float rz, hipStream_t s);
void launch_copy_fields(int n, int nf, const FieldPtrs& dst, const FieldPtrs& src, hipStream_t s);
// adjoints of the field half of a box-step (k_adjoint.hip); box-local arrays
void launch_adj_push_bfld(const Box& B, const float* bb[3], float* ob[3], float* eb[3],
                          float delt, float rx, float ry, float rz, hipStream_t s);
void launch_adj_bande(const Box& B, float* eb[3], const float* bs[3], const float* ev[3],
                      const float* cur[3], const float* rho[3], float rmq, float rmh, float rmo,
                      float eta, float* bsb[3], float* evb[3], float* curb[3], double* etab,
                      hipStream_t s);
void launch_adj_evel(const Box& B, const float* evb[3], const float* rho[3], float rmq, float rmh,
                     float rmo, float reynolds, float* curb[3], hipStream_t s);
// BandE's face copies, transposed (scatter to the clamped interior cell, zero)
void launch_adj_bande_faces(const Box& B, float* eb[3], hipStream_t s);
void launch_adj_calcur(const Box& B, float* curb[3], float* bb[3], float rx, float ry, float rz,
                       hipStream_t s);
void launch_adj_outer_b(const Box& B, float* bb[3], double* sbb, hipStream_t s);
void launch_adj_lap(const Box& B, const float* wb[3], const float* v[3], float chi, float delt,
                    float* qb[3], float* w1b[3], hipStream_t s);
size_t adj_fct_scratch_floats(const Box& B);
// old's adjoint is accumulated, or with assign_old assigned (its first write)
void launch_adj_fct(const Box& B, const float* old[3], const float* wrk[3], float* ob[3],
                    float* wb[3], float* oldb[3], float chi, float* scratch, hipStream_t s,
                    int nf = 3, bool assign_old = false);
void launch_add3(int n, const float* x[3], float* y[3], hipStream_t s);
struct CraftGrid;
void launch_adj_sample(int npts, const float* sxyz, const CraftGrid& G, int nx, int ny, int nz,
                       int nq, const FieldPtrs& Fb, const float* sb, hipStream_t s);

// adjoints of push_ion (one species) and of lap_plasma (k_adjoint_ion.hip);
// box-local arrays, fields in IonPtrs order; all adjoints accumulate
struct IonAdj {
    Box B;
    const float* src[10];
    const float* bs[3];
    const float* btot;
    const float* e[3];
    const float* tv[3];
    const float* ob[10];     // adjoint of the output
    float* sb[10];           // adjoint of the source (wrk) fields
    float* oldb[10];         // adjoint of the old fields
    float* bsb[3];
    float* btotb;
    float* eb[3];
    float* tvb[3];
    float* scr;              // adj_ion_scratch_floats
    float rmass, delt, grav, re_equiv, reynolds;
    float xmin, ymin, zmin, ddx, ddy, ddz, dxt, dyt, dzt, bmax;
};
struct LapAdj {
    Box B;
    const float* q[10];
    const float* v[3];
    const float* wb[10];     // adjoint of the smoothed fields
    float* qb[10];
    float* vb[3];
    float chi[10];
    float delt;
    float* scr;              // adj_lapsp_scratch_floats
    int nf = 10;             // fields in use
    bool assign = false;     // qb assigned (=) instead of accumulated
};
size_t adj_ion_scratch_floats(const Box& B);
size_t adj_lapsp_scratch_floats(const Box& B);
void launch_adj_push_ion(const IonAdj& A, hipStream_t s);
void launch_adj_lap_species(const LapAdj& L, hipStream_t s);
void launch_adj_velocity(int n, const float* rho, const float* p[3], const float* vb[3],
                         float smallbit, float* rhob, float* pb[3], hipStream_t s);

// adjoints of the plasma couplings of a box-step (k_adjoint_full.hip)
struct EfAdj {
    Box B;
    const float* eb[3];      // E's adjoint after BandE's face copies are transposed
    const float* bs[3];
    const float* ev[3];
    const float* cur[3];
    const float* rho[3];
    const float* mom[9];     // q px,py,pz, h ..., o ...
    const float* ep;
    const float* tvb[3];     // tv's adjoint from its other uses (the ion push)
    float* evb[3];           // scratch: ev's adjoint
    float* bsb[3];
    float* curb[3];
    float* rhob[3];
    float* momb[9];
    float* epb;
    double* etab;
    float* scr;              // 3 * V
    float eta, rmq, rmh, rmo, reynolds, dxt, dyt, dzt;
};
struct PeAdj {
    Box B;
    const float* w;          // push_elec's source pressure
    const float* ev[3];
    const float* ob;         // adjoint of the output pressure
    float* oldb;
    float* wb;
    float* evb[3];
    float delt, gamma, gamma1, dxt, dyt, dzt;
};
struct VtotAdj {
    const float* rho[3];
    const float* mom[9];
    const float* vb[3];
    float* rhob[3];
    float* momb[9];
};
struct OuterAdj {
    float* f[kMaxFields];          // adjoint arrays (transposed in place)
    const float* val[kMaxFields];   // the field after the forward pass (signs of |.|)
    int kind[kMaxFields];
    int ctl[kMaxFields];            // control index of the inflow value, -1 for none
    float coef[kMaxFields];         // d(inflow value)/d(control)
};
// BandE's body with its plasma terms (assigns ev's adjoint), then fnd_evel's
// transpose of ev's adjoint plus tv's; between them a caller adds
// push_elec's part of ev's adjoint
void launch_adj_bande_plasma(const EfAdj& A, hipStream_t s);
void launch_adj_evel_plasma(const EfAdj& A, hipStream_t s);
void launch_adj_btot(int n, const float* b[3], const float* btb, float smallbit, float* bb[3],
                     hipStream_t s);
void launch_adj_vtot(int n, const VtotAdj& A, hipStream_t s);
void launch_adj_push_elec(const PeAdj& A, hipStream_t s);
void launch_adj_outer(const Box& B, const OuterAdj& O, int nf, double* ctlb, hipStream_t s);

// write_graphing_data's cut-plane quantities (k_graph_cuts in k_misc.hip)
struct GraphCuts {
    Box B;
    int cut[3];                 // 0-based: k of the xy plane, j of xz, i of yz
    const float* sp[3][7];      // per species rho, px, py, pz, presx, presy, presz
    const float* epres;
    const float* b[3];
    const float* b0[3];
    const float* cur[3];
    const float* tv[3];
    const float* e[3];
    float rmass[3];
    float smallbit, v_equiv, b_equiv, rho_equiv, re_equiv;
    float temp_equiv, pres_equiv, cur_equiv, e_equiv, mu0, m_prot;
    float xmin, ymin, zmin, xs;
};
void launch_graph_cuts(const GraphCuts& G, float* out, hipStream_t s);
void launch_fnd_pres(int n, const float* const p[3], const float* const v[3],
                     const float* const b[3], float* para, float* perp, float* cross,
                     hipStream_t s);
struct CraftGrid;
void launch_sample(int npts, float* sxyz, const float* xcraft, const float* zcraft, float ut,
                   const CraftGrid& G, int nx, int ny, int nz, int nq, const FieldPtrs& F,
                   float* out, int* gridpt, hipStream_t s);
void launch_swap_fields(int n, int nf, const FieldPtrs& a, const FieldPtrs& b, hipStream_t s);
void launch_push_ion(const Box& B, const IonPtrs& P, float rmass, float delt, float grav,
                     float re_equiv, float reynolds, const float gmin[3], const float gmax[3],
                     hipStream_t s);
void launch_set_rho(int n, const SpeciesPtrs& S, float* ep, hipStream_t s);
// partials: speed_partials_bytes(V) of scratch for the per-block maxima
void launch_set_speed(const Box& B, const SpeciesPtrs& S, const SpeedParams& Q, const float* b0[3],
                      const float* b[3], float* btot, float* ep, float* maxima, void* list,
                      int* count, float* partials, hipStream_t s);
size_t speed_list_bytes(int ncells);
size_t speed_partials_bytes(int ncells);
void launch_velocity(int n, const float* r, const float* p[3], float* v[3], float dmin,
                     hipStream_t s);
void launch_vtot(int n, const float* mom[9], const float* rho[3], float* v[3], hipStream_t s);
void launch_lapidus(const Box& B, const LapPtrs& L, int nf, const float* v[3], float delt,
                    hipStream_t s);
void launch_bndry_outer(const Box& B, const OuterPtrs& O, int nf, hipStream_t s);
void launch_bndry_inner(const Box& B, const InnerPtrs& I, const int* ij[3], const float* parm[3],
                        const int num[3], int mbndry, int box, float xmin, float ymin, float dx,
                        float dy, float re_equiv, float v_rot, hipStream_t s, bool disjoint = false);
void launch_flanks(const Box& B, const FlankPtrs& Fp, int nf, const FlankGeom& G, float t1,
                   float t2, hipStream_t s);
void launch_corer(const Box& B, const FieldPtrs& fine, const FieldPtrs& coarse, int nf, int nskip,
                  int nbx1, int nby1, int nbz1, hipStream_t s);
void launch_flanks_synced(const Box& B, const FieldPtrs& fine, const FieldPtrs& coarse, int nf,
                          const FlankGeom& G, const int kbi[6], hipStream_t s);
int divb_correct(const Box& B, float* bx, float* by, float* bz, float rx, float errlim, int mode,
                 long max_iters, float* work, double* dwork, long* iters_out, double* res2_out,
                 hipStream_t s);
size_t divb_work_floats(int ncells);
int sparse_solve(int npts, int nbands, const int* pos, const float* M, const float* rhs,
                 float* soln, float errlim, bool recurse, int redo_lim, int mode, long max_iters,
                 float* work, double* dwork, long* iters_out, double* res2_out, hipStream_t s);
size_t sparse_work_floats(int npts);
size_t divb_work_doubles(int ncells);
void launch_banded_matvec(int npts, const int* pos, const float* M, const float* x, float* y,
                          hipStream_t s);
void launch_trilin(int npts, const float* sxyz, const float* gridpts, const float* qty,
                   float* out, hipStream_t s);
void dipole_host(float ax, float ay, float az, const RotationGeom& R, float& abx, float& aby,
                 float& abz);
void launch_mak_dip_grd(const Box& B, float* bx0, float* by0, float* bz0, const InitGeom& G,
                        const RotationGeom& R, float r_inner, const int* ijzero, int mbndry,
                        int box, int numzero, hipStream_t s);
// point counts (srf, mid, zero) from dcounts on the device, or from the
// host array hcounts when dcounts is null
void launch_set_resist(const Box& B, float* rst, int mbndry, float resist, float b0,
                       const int* ijzero, const int* ijmid, const int* ijsrf, const int* dcounts,
                       const int hcounts[3], int max_points, hipStream_t s);
void launch_init_plasma(const InitPlasma& P, int n_grids, hipStream_t s);
// the torus source on box 1 (code/3d_main.f90:3149-3253); sp[s].f = rho,
// presx, presy, presz, px, py, pz, presxy, presxz, presyz of species q, h, o
struct TorusParams {
    Box B;
    SpeciesPtrs sp;
    float* epres;
    const float *bx0, *by0, *bz0, *bx, *by, *bz;
    float* contrib;   // 3 * V scratch
    float xmin, ymin, zmin, dx, dy, dz, xdip, ydip, zdip, re_equiv, v_rot;
    float torus_inj_dist, torus_rad, srf_bndry, denq, denh, deno, rmassq, rmassh, rmasso;
    float t_torus, aniso_factor, reduct;
};
void launch_torus(const TorusParams& T, float* tot, hipStream_t s);
// P.p = q, h, o x (px, py, pz) of one box; sums[18] as k_flux_walls lays them out
void launch_flux_walls(const Box& B, const FieldPtrs& P, float* sums, hipStream_t s);

}  // namespace g3m3
