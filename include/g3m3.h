/*
 * g3m3.h -- C-ABI of the MI355X-native G3M^3 per-timestep hot path.
 *
 * Every entry point mirrors one reference subroutine: same name (prefixed
 * g3m3_), same dummy-argument order and meaning, every argument passed by
 * reference exactly as the reference's implicit interfaces receive them,
 * so a Fortran caller binds each one with an ISO_C_BINDING interface
 * (bind(c), no VALUE attributes) and keeps its call sites unchanged.
 *
 * Arrays:
 *   - a field q(nx,ny,nz,n_grids) is float[n_grids][nz][ny][nx] (Fortran
 *     column-major, i fastest); box-local scratch q(nx,ny,nz) likewise;
 *   - pointers may be device pointers (hipMalloc / torch tensors) or host
 *     addresses inside a range registered with g3m3_register(); the latter
 *     are translated to their device mirror (the mirror is the data the
 *     kernels read and write -- use g3m3_to_device / g3m3_to_host to move
 *     values between a host array and its mirror).
 *   - integer dummies are int32, real dummies float, real(dp) are double,
 *     LOGICAL dummies are int32 (nonzero = .true.).
 *
 * Execution: calls are ordered on the library stream (g3m3_set_stream,
 * default the null stream).  A call returns before its kernels finish,
 * except calls with host scalar outputs (g3m3_set_speed_agrd,
 * g3m3_divb_correct_ex), which synchronise.  No entry point returns a
 * status: HIP and argument errors are latched and read with
 * g3m3_last_error().  Not reentrant: one host thread per process.
 *
 * Arithmetic is IEEE single precision, evaluated in the reference's order
 * with no fused multiply-add, so results agree with a serial build of the
 * reference subroutine bit for bit (the parity tests check this), except
 * where noted (divB dot products accumulate in double).
 */
#ifndef G3M3_H
#define G3M3_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library control ---------------------------------------------------- */
int g3m3_init(void);                          /* 0 on success; selects device 0 */
int g3m3_last_error(void);                    /* 0 = none; clears the latch */
const char* g3m3_last_error_msg(void);
void g3m3_set_stream(void* hip_stream);       /* hipStream_t, NULL = null stream */
void g3m3_sync(void);                         /* wait for the library stream */
/* No reference counterpart: per-kernel timing for measurement harnesses.
 * Launches of the named kernel are bracketed by HIP events on the stream
 * they run on; g3m3_prof_read returns the summed time and launch count. */
void g3m3_prof_select(const char* kernel);
void g3m3_prof_read(double* ms, long long* launches);
/* device mirrors of host arrays (drop-in use from an unchanged Fortran driver) */
int g3m3_register(void* host_ptr, long long bytes);
int g3m3_unregister(void* host_ptr);
int g3m3_to_device(void* host_ptr);           /* host array -> its mirror */
int g3m3_to_host(void* host_ptr);             /* mirror -> host array */

/* ---- box-step pieces ------------------------------------------------------ */
/* code/sub_store_array.f90:1 -- bx(:,:,:,box) = bx0(:,:,:,box) */
void g3m3_store_array(float* bx, const float* bx0, const int* nx, const int* ny,
                      const int* nz, const int* n_grids, const int* box);

/* code/sub_calcur.f90:5-6 -- J = curl(b) with the reference's face rules */
void g3m3_calcur(const float* bx, const float* by, const float* bz, const int* nx,
                 const int* ny, const int* nz, const int* n_grids, const int* box,
                 float* curx, float* cury, float* curz, const float* rx, const float* ry,
                 const float* rz);

/* code/sub_fnd_evel.f90:4-7 -- ion (tv) and electron (ev) bulk velocity */
void g3m3_fnd_evel(const float* qpx, const float* qpy, const float* qpz, const float* qrho,
                   const float* hpx, const float* hpy, const float* hpz, const float* hrho,
                   const float* opx, const float* opy, const float* opz, const float* orho,
                   const float* curx, const float* cury, const float* curz, float* evx,
                   float* evy, float* evz, float* tvx, float* tvy, float* tvz, const int* nx,
                   const int* ny, const int* nz, const int* n_grids, const int* box,
                   const float* rmassq, const float* rmassh, const float* rmasso,
                   const float* reynolds);

/* code/sub_BandE.f90:5-10 -- Ohm's law E = -ve x B - grad pe/(n Re) + rst J/n */
void g3m3_bande(float* efldx, float* efldy, float* efldz, const float* bsx, const float* bsy,
                const float* bsz, const float* curx, const float* cury, const float* curz,
                const float* evx, const float* evy, const float* evz, const float* btot,
                const float* epres, const float* qrho, const float* hrho, const float* orho,
                const float* rst, const float* resist, const float* reynolds, const int* nx,
                const int* ny, const int* nz, const int* n_grids, const int* box,
                const float* rmassq, const float* rmassh, const float* rmasso, const int* ijmid,
                const int* nummid, const int* ijzero, const int* mbndry, const int* numzero,
                const int* mmid, const int* mzero, const float* rx, const float* ry,
                const float* rz);

/* No reference counterpart: g3m3_bande plus a uniform resistivity eta*J/n
 * on every interior cell (applied after the list term). */
void g3m3_bande_eta(float* efldx, float* efldy, float* efldz, const float* bsx, const float* bsy,
                    const float* bsz, const float* curx, const float* cury, const float* curz,
                    const float* evx, const float* evy, const float* evz, const float* btot,
                    const float* epres, const float* qrho, const float* hrho, const float* orho,
                    const float* rst, const float* resist, const float* reynolds, const int* nx,
                    const int* ny, const int* nz, const int* n_grids, const int* box,
                    const float* rmassq, const float* rmassh, const float* rmasso,
                    const int* ijmid, const int* nummid, const int* ijzero, const int* mbndry,
                    const int* numzero, const int* mmid, const int* mzero, const float* rx,
                    const float* ry, const float* rz, const float* eta);

/* code/3d_main.f90:2248-2254 -- bs = b0 + b, btot = max(|bs|, smallbit) */
void g3m3_bfield_total(const float* bx, const float* by, const float* bz, const float* bx0,
                       const float* by0, const float* bz0, float* bsx, float* bsy, float* bsz,
                       float* btot, const int* nx, const int* ny, const int* nz,
                       const int* n_grids, const int* box, const float* smallbit);

/* code/sub_push_elec.f90:4-5 */
void g3m3_push_elec(const float* wrkepres, const float* oldepres, float* epres, const float* evx,
                    const float* evy, const float* evz, const float* gamma, const float* gamma1,
                    const int* nx, const int* ny, const int* nz, const int* n_grids,
                    const int* box, const float* delt, const float* rx, const float* ry,
                    const float* rz);

/* code/sub_push_ion.f90:4-17 -- one ion species (qvx..qvz are outputs) */
void g3m3_push_ion(float* qrho, float* qpresx, float* qpresy, float* qpresz, float* qpx,
                   float* qpy, float* qpz, const float* oldqrho, const float* oldqpresx,
                   const float* oldqpresy, const float* oldqpresz, const float* oldqpx,
                   const float* oldqpy, const float* oldqpz, const float* wrkqrho,
                   const float* wrkqpresx, const float* wrkqpresy, const float* wrkqpresz,
                   const float* wrkqpx, const float* wrkqpy, const float* wrkqpz,
                   float* qpresxy, float* qpresxz, float* qpresyz, const float* oldqpresxy,
                   const float* oldqpresxz, const float* oldqpresyz, const float* wrkqpresxy,
                   const float* wrkqpresxz, const float* wrkqpresyz, const float* bsx,
                   const float* bsy, const float* bsz, const float* btot, const float* efldx,
                   const float* efldy, const float* efldz, const float* rmassq, float* qvx,
                   float* qvy, float* qvz, const float* tvx, const float* tvy, const float* tvz,
                   const float* gamma, const float* gamma1, const int* nx, const int* ny,
                   const int* nz, const int* n_grids, const int* box, const float* delt,
                   const float* grav, const float* re_equiv, const float* reynolds,
                   const float* grd_xmin, const float* grd_xmax, const float* grd_ymin,
                   const float* grd_ymax, const float* grd_zmin, const float* grd_zmax,
                   const int* isotropic);

/* code/sub_push_bfld.f90:4-6 -- induction */
void g3m3_push_bfld(float* bx, float* by, float* bz, const float* oldbx, const float* oldby,
                    const float* oldbz, const float* efldx, const float* efldy, const float* efldz,
                    const int* nx, const int* ny, const int* nz, const int* n_grids,
                    const int* box, const float* delt, const float* rx, const float* ry,
                    const float* rz);

/* code/sub_set_rho.f90:5-11 -- density floors and negative-pressure reset */
void g3m3_set_rho(float* qrho, float* qpresx, float* qpresy, float* qpresz, float* qpresxy,
                  float* qpresxz, float* qpresyz, const float* rmassq, float* hrho, float* hpresx,
                  float* hpresy, float* hpresz, float* hpresxy, float* hpresxz, float* hpresyz,
                  const float* rmassh, float* orho, float* opresx, float* opresy, float* opresz,
                  float* opresxy, float* opresxz, float* opresyz, const float* rmasso,
                  float* epres, const int* nx, const int* ny, const int* nz, const int* n_grids,
                  const int* box, const float* o_conc);

/* code/sub_set_speed_agrd.f90:6-22 -- speed limits, averaging, CFL maxima */
void g3m3_set_speed_agrd(const int* nx, const int* ny, const int* nz, const int* n_grids,
                         const int* box, const float* rmassq, const float* rmassh,
                         const float* rmasso, const int* isotropic, const float* smallbit,
                         const float* aniso_limit, const float* vlim, const float* qvlim,
                         const float* hvlim, const float* ovlim, const float* qclim,
                         const float* hclim, const float* oclim, const float* cslim,
                         const float* bx, const float* by, const float* bz, const float* bx0,
                         const float* by0, const float* bz0, float* qrho, float* qpx, float* qpy,
                         float* qpz, float* qpresx, float* qpresy, float* qpresz, float* qpresxy,
                         float* qpresxz, float* qpresyz, float* hrho, float* hpx, float* hpy,
                         float* hpz, float* hpresx, float* hpresy, float* hpresz, float* hpresxy,
                         float* hpresxz, float* hpresyz, float* orho, float* opx, float* opy,
                         float* opz, float* opresx, float* opresy, float* opresz,
                         float* opresxy, float* opresxz, float* opresyz, float* epres,
                         float* btot, float* pxmax, float* pymax, float* pzmax, float* csmax,
                         float* alfmax, float* fastest);

/* No counterpart in the reference.  g3m3_set_speed_agrd with the same
 * arguments and results, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): it calls no hipMalloc/hipFree.  Its scratch is a fixed
 * 18.9 MB device array per GPU, sized for boxes of at most 131072 cells
 * (nx*ny*nz); a larger box latches an error and nothing runs.  Every call on
 * one GPU shares that array, so calls that overlap on one GPU (several
 * streams or host threads) race. */
void g3m3_set_speed_agrd_noalloc(
    const int* nx, const int* ny, const int* nz, const int* n_grids, const int* box,
    const float* rmassq, const float* rmassh, const float* rmasso, const int* isotropic,
    const float* smallbit, const float* aniso_limit, const float* vlim, const float* qvlim,
    const float* hvlim, const float* ovlim, const float* qclim, const float* hclim,
    const float* oclim, const float* cslim, const float* bx, const float* by, const float* bz,
    const float* bx0, const float* by0, const float* bz0, float* qrho, float* qpx, float* qpy,
    float* qpz, float* qpresx, float* qpresy, float* qpresz, float* qpresxy, float* qpresxz,
    float* qpresyz, float* hrho, float* hpx, float* hpy, float* hpz, float* hpresx,
    float* hpresy, float* hpresz, float* hpresxy, float* hpresxz, float* hpresyz, float* orho,
    float* opx, float* opy, float* opz, float* opresx, float* opresy, float* opresz,
    float* opresxy, float* opresxz, float* opresyz, float* epres, float* btot, float* pxmax,
    float* pymax, float* pzmax, float* csmax, float* alfmax, float* fastest);

/* code/sub_lap_plasma.f90:4-10 -- Lapidus smoothing, one species */
void g3m3_lap_plasma(const float* rho, const float* px, const float* py, const float* pz,
                     const float* presx, const float* presy, const float* presz,
                     const float* presxy, const float* presxz, const float* presyz,
                     float* wrkrho, float* wrkpx, float* wrkpy, float* wrkpz, float* wrkpresx,
                     float* wrkpresy, float* wrkpresz, float* wrkpresxy, float* wrkpresxz,
                     float* wrkpresyz, const float* vx, const float* vy, const float* vz,
                     const int* nx, const int* ny, const int* nz, const int* n_grids,
                     const int* box, const float* chirho, const float* chipxyz,
                     const float* chierg, const float* delt, const int* isotropic);

/* code/sub_lap_elec.f90:4-5 */
void g3m3_lap_elec(const float* ppres, float* wrkppres, const float* vx, const float* vy,
                   const float* vz, const int* nx, const int* ny, const int* nz,
                   const int* n_grids, const int* box, const float* chirho,
                   const float* chipxyz, const float* chierg, const float* delt);

/* code/sub_lap_bfld.f90:4-5 */
void g3m3_lap_bfld(const float* bx, const float* by, const float* bz, float* wrkbx, float* wrkby,
                   float* wrkbz, const float* vx, const float* vy, const float* vz,
                   const int* nx, const int* ny, const int* nz, const int* n_grids,
                   const int* box, const float* chirho, const float* chipxyz,
                   const float* chierg, const float* delt);

/* code/sub_fnd_vtot.f90:4-6 */
void g3m3_fnd_vtot(const float* qpx, const float* qpy, const float* qpz, const float* qrho,
                   const float* hpx, const float* hpy, const float* hpz, const float* hrho,
                   const float* opx, const float* opy, const float* opz, const float* orho,
                   float* vx, float* vy, float* vz, const int* nx, const int* ny,
                   const int* nz, const int* n_grids, const int* box, const float* rmassq,
                   const float* rmassh, const float* rmasso);

/* code/3d_main.f90:2698-2748 -- v = p/max(rho, smallbit) for Lapidus */
void g3m3_species_velocity(const float* rho, const float* px, const float* py, const float* pz,
                           float* vx, float* vy, float* vz, const int* nx, const int* ny,
                           const int* nz, const int* n_grids, const int* box,
                           const float* smallbit);

/* code/sub_fcsmooth.f90:7-8 -- FCT on one field (tempx..tempz are workspace
 * in the reference and are not written here) */
void g3m3_fcsmooth(float* px, const float* oldpx, const float* wrkpx, const int* nx,
                   const int* ny, const int* nz, const int* n_grids, const int* box,
                   const float* chipx, float* tempx, float* tempy, float* tempz);

/* code/sub_flux_correct.f90:4-35 -- FCT on all 34 fields, one launch set */
void g3m3_flux_correct(
    float* qrho, float* qpresx, float* qpresy, float* qpresz, float* qpresxy, float* qpresxz,
    float* qpresyz, float* qpx, float* qpy, float* qpz, const float* wrkqrho,
    const float* wrkqpresx, const float* wrkqpresy, const float* wrkqpresz,
    const float* wrkqpresxy, const float* wrkqpresxz, const float* wrkqpresyz,
    const float* wrkqpx, const float* wrkqpy, const float* wrkqpz, const float* oldqrho,
    const float* oldqpresx, const float* oldqpresy, const float* oldqpresz,
    const float* oldqpresxy, const float* oldqpresxz, const float* oldqpresyz,
    const float* oldqpx, const float* oldqpy, const float* oldqpz, float* hrho, float* hpresx,
    float* hpresy, float* hpresz, float* hpresxy, float* hpresxz, float* hpresyz, float* hpx,
    float* hpy, float* hpz, const float* wrkhrho, const float* wrkhpresx,
    const float* wrkhpresy, const float* wrkhpresz, const float* wrkhpresxy,
    const float* wrkhpresxz, const float* wrkhpresyz, const float* wrkhpx, const float* wrkhpy,
    const float* wrkhpz, const float* oldhrho, const float* oldhpresx, const float* oldhpresy,
    const float* oldhpresz, const float* oldhpresxy, const float* oldhpresxz,
    const float* oldhpresyz, const float* oldhpx, const float* oldhpy, const float* oldhpz,
    float* orho, float* opresx, float* opresy, float* opresz, float* opresxy, float* opresxz,
    float* opresyz, float* opx, float* opy, float* opz, const float* wrkorho,
    const float* wrkopresx, const float* wrkopresy, const float* wrkopresz,
    const float* wrkopresxy, const float* wrkopresxz, const float* wrkopresyz,
    const float* wrkopx, const float* wrkopy, const float* wrkopz, const float* oldorho,
    const float* oldopresx, const float* oldopresy, const float* oldopresz,
    const float* oldopresxy, const float* oldopresxz, const float* oldopresyz,
    const float* oldopx, const float* oldopy, const float* oldopz, float* epres,
    const float* wrkepres, const float* oldepres, float* bx, float* by, float* bz,
    const float* wrkbx, const float* wrkby, const float* wrkbz, const float* oldbx,
    const float* oldby, const float* oldbz, float* vvx, float* vvy, float* vvz, const int* nx,
    const int* ny, const int* nz, const int* n_grids, const int* box, const float* difrho,
    const float* diferg, const float* xspac, const int* isotropic);

/* code/sub_bndry_outer.f90:5-16 -- outermost box (box = n_grids) */
void g3m3_bndry_outer(
    float* qrho, float* qpresx, float* qpresy, float* qpresz, float* qpx, float* qpy, float* qpz,
    float* hrho, float* hpresx, float* hpresy, float* hpresz, float* hpx, float* hpy, float* hpz,
    float* orho, float* opresx, float* opresy, float* opresz, float* opx, float* opy, float* opz,
    float* epres, float* qpresxy, float* qpresxz, float* qpresyz, float* hpresxy,
    float* hpresxz, float* hpresyz, float* opresxy, float* opresxz, float* opresyz,
    const float* rmassq, const float* rmassh, const float* rmasso, float* bx, float* by,
    float* bz, const float* bx0, const float* by0, const float* bz0, const int* nx,
    const int* ny, const int* nz, const int* n_grids, const float* srho, const float* rho_frac,
    const float* o_conc, const float* spress, const float* spx, const float* spy,
    const float* spz, const float* sbx_wind, const float* sby_wind, const float* sbz_wind,
    const float* ti_te, const int* isotropic);

/* code/sub_bndry_inner.f90:5-19 -- planet lists on box mbndry.  v_rot and
 * r_rot are the reference's common /rotation/ values (code/3d_main.f90:633). */
void g3m3_bndry_inner(
    float* qrho, float* qpresx, float* qpresy, float* qpresz, float* qpx, float* qpy, float* qpz,
    const float* rmassq, float* hrho, float* hpresx, float* hpresy, float* hpresz, float* hpx,
    float* hpy, float* hpz, const float* rmassh, float* orho, float* opresx, float* opresy,
    float* opresz, float* opx, float* opy, float* opz, const float* rmasso, float* epres,
    float* bx, float* by, float* bz, float* qpresxy, float* qpresxz, float* qpresyz,
    float* hpresxy, float* hpresxz, float* hpresyz, float* opresxy, float* opresxz,
    float* opresyz, const int* nx, const int* ny, const int* nz, const int* n_grids,
    const float* parm_srf, const float* parm_mid, const float* parm_zero, const int* ijsrf,
    const int* numsrf, const int* ijmid, const int* nummid, const int* ijzero,
    const int* numzero, const int* mbndry, const int* msrf, const int* mmid, const int* mzero,
    const float* erho, const float* epress, const float* alpha_e, const float* ti_te,
    const float* o_conc, const float* re_equiv, const float* r_inner, const float* sbx_wind,
    const float* spress, const float* grd_xmin, const float* grd_xmax, const float* grd_ymin,
    const float* grd_ymax, const float* grd_zmin, const float* grd_zmax, const float* v_rot,
    const float* r_rot);

/* code/sub_flanks.f90:5-7 -- coarse box+1 -> fine box shell, one field */
void g3m3_flanks(float* wrkrho, const float* rho, const float* oldrho, const int* nx,
                 const int* ny, const int* nz, const int* n_grids, const int* box,
                 const double* t_new, const double* t_old, const double* t, float* work,
                 const float* grd_xmin, const float* grd_xmax, const float* grd_ymin,
                 const float* grd_ymax, const float* grd_zmin, const float* grd_zmax);

/* code/sub_bndry_flanks.f90:5-35 -- flanks on all 34 fields, one launch.
 * tgt/cur/old are arrays of 34 field pointers in the reference's argument
 * order (q: rho,px,py,pz,presx,presy,presz,presxy,presxz,presyz; h; o;
 * epres, bx, by, bz). */
void g3m3_bndry_flanks_v(float* const* tgt, const float* const* cur, const float* const* old,
                         float* work, const int* nx, const int* ny, const int* nz,
                         const int* n_grids, const int* box, const double* t_old,
                         const double* t_new, const double* t, const float* grd_xmin,
                         const float* grd_xmax, const float* grd_ymin, const float* grd_ymax,
                         const float* grd_zmin, const float* grd_zmax);

/* code/sub_corer.f90:5-6 -- fine box -> coarse box+1 core, one field */
void g3m3_corer(float* rho, const int* nx, const int* ny, const int* nz, const int* n_grids,
                const int* box, const float* grd_xmin, const float* grd_xmax,
                const float* grd_ymin, const float* grd_ymax, const float* grd_zmin,
                const float* grd_zmax);

/* code/sub_bndry_corer.f90:5-14 -- corer on 34 fields given as a pointer
 * array in the reference's argument order */
void g3m3_bndry_corer_v(float* const* fields, const int* nx, const int* ny, const int* nz,
                        const int* n_grids, const int* box, const float* grd_xmin,
                        const float* grd_xmax, const float* grd_ymin, const float* grd_ymax,
                        const float* grd_zmin, const float* grd_zmax);

/* code/sub_bndry_flanks.f90:5-35 -- the reference's argument list: the 34
 * target arrays (wrk..., in the reference's order), the 34 current and 34
 * old arrays of the same fields, work, the box and times, the grid extents.
 * The target may be the current array itself (code/3d_main.f90:2945-2976).
 * The same kernel as g3m3_bndry_flanks_v. */
void g3m3_bndry_flanks(
    float* wrkqrho, float* wrkqpx, float* wrkqpy, float* wrkqpz, float* wrkqpresx,
    float* wrkqpresy, float* wrkqpresz, float* wrkqpresxy, float* wrkqpresxz, float* wrkqpresyz,
    float* wrkhrho, float* wrkhpx, float* wrkhpy, float* wrkhpz, float* wrkhpresx,
    float* wrkhpresy, float* wrkhpresz, float* wrkhpresxy, float* wrkhpresxz, float* wrkhpresyz,
    float* wrkorho, float* wrkopx, float* wrkopy, float* wrkopz, float* wrkopresx,
    float* wrkopresy, float* wrkopresz, float* wrkopresxy, float* wrkopresxz, float* wrkopresyz,
    float* wrkepres, float* wrkbx, float* wrkby, float* wrkbz, const float* qrho,
    const float* qpx, const float* qpy, const float* qpz, const float* qpresx,
    const float* qpresy, const float* qpresz, const float* qpresxy, const float* qpresxz,
    const float* qpresyz, const float* hrho, const float* hpx, const float* hpy,
    const float* hpz, const float* hpresx, const float* hpresy, const float* hpresz,
    const float* hpresxy, const float* hpresxz, const float* hpresyz, const float* orho,
    const float* opx, const float* opy, const float* opz, const float* opresx,
    const float* opresy, const float* opresz, const float* opresxy, const float* opresxz,
    const float* opresyz, const float* epres, const float* bx, const float* by, const float* bz,
    const float* oldqrho, const float* oldqpx, const float* oldqpy, const float* oldqpz,
    const float* oldqpresx, const float* oldqpresy, const float* oldqpresz,
    const float* oldqpresxy, const float* oldqpresxz, const float* oldqpresyz,
    const float* oldhrho, const float* oldhpx, const float* oldhpy, const float* oldhpz,
    const float* oldhpresx, const float* oldhpresy, const float* oldhpresz,
    const float* oldhpresxy, const float* oldhpresxz, const float* oldhpresyz,
    const float* oldorho, const float* oldopx, const float* oldopy, const float* oldopz,
    const float* oldopresx, const float* oldopresy, const float* oldopresz,
    const float* oldopresxy, const float* oldopresxz, const float* oldopresyz,
    const float* oldepres, const float* oldbx, const float* oldby, const float* oldbz,
    float* work, const int* nx, const int* ny, const int* nz, const int* n_grids, const int* box,
    const double* t_old, const double* t_new, const double* t, const float* grd_xmin,
    const float* grd_xmax, const float* grd_ymin, const float* grd_ymax, const float* grd_zmin,
    const float* grd_zmax);

/* code/sub_bndry_corer.f90:5-14 -- the reference's argument list: the 34
 * fields in its order, the box, the grid extents.  The same kernel as
 * g3m3_bndry_corer_v. */
void g3m3_bndry_corer(
    float* qrho, float* qpresx, float* qpresy, float* qpresz, float* qpx, float* qpy, float* qpz,
    float* hrho, float* hpresx, float* hpresy, float* hpresz, float* hpx, float* hpy, float* hpz,
    float* orho, float* opresx, float* opresy, float* opresz, float* opx, float* opy, float* opz,
    float* epres, float* bx, float* by, float* bz, float* qpresxy, float* qpresxz,
    float* qpresyz, float* hpresxy, float* hpresxz, float* hpresyz, float* opresxy,
    float* opresxz, float* opresyz, const int* nx, const int* ny, const int* nz,
    const int* n_grids, const int* box, const float* grd_xmin, const float* grd_xmax,
    const float* grd_ymin, const float* grd_ymax, const float* grd_zmin, const float* grd_zmax);

/* code/sub_flanks_synced.f90:5-6 -- coarse -> fine faces after divB */
void g3m3_flanks_synced(float* rho, const int* nx, const int* ny, const int* nz,
                        const int* n_grids, const int* box, const float* grd_xmin,
                        const float* grd_xmax, const float* grd_ymin, const float* grd_ymax,
                        const float* grd_zmin, const float* grd_zmax);

/* code/sub_banded_tpose.f90:6 -- y = M^T x for the 7-band matrix M(7,npts) */
void g3m3_banded_tpose(const int* nbands, const int* npts, const int* divB_band_pos,
                       const float* banded_matrix, const float* x_vec, float* tp_result);

/* code/sub_divB_correct.f90:25-27 -- divB clean-up by CGNR (matrix-free).
 * The reference starts from an uninitialised solution vector; this starts
 * from zero. */
void g3m3_divb_correct(const int* nx, const int* ny, const int* nz, const int* n_grids,
                       const int* box, const int* nbands, const int* divB_band_pos,
                       const float* xspac, const float* divB_errlim, float* bx, float* by,
                       float* bz);

/* No reference counterpart: g3m3_divb_correct with a solver mode
 *   0 = the reference's CGNR, single-precision scalars,
 *   1 = the reference's CGNR, double-precision scalars and updates,
 *   2 = conjugate gradients on the same (symmetric negative definite)
 *       system, FP32 vectors with FP64 dot products and scalars, stopping
 *       on the reference's criterion -- about sqrt of CGNR's iterations,
 *   3 = mode 2 inside FP64 residual refinement: the solution and the
 *       residual b - A x in FP64, corrections by the FP32-vector CG, until
 *       the FP64 residual meets the criterion (relative residuals well
 *       below FP32's reach, e.g. 1e-8),
 * an iteration cap (0 = the reference's 10*npts), and the iteration count
 * and final squared residual |A x - b|^2 returned.  Returns the solver
 * status: 1 converged, 2 stopped after restarts, 3 iteration cap, 4 a HIP call
 * failed (the error is latched). */
int g3m3_divb_correct_ex(const int* nx, const int* ny, const int* nz, const int* n_grids,
                         const int* box, const float* xspac, const float* divB_errlim,
                         float* bx, float* by, float* bz, const int* mode,
                         const long long* max_iters, long long* iters, double* residual2);

/* code/cft_trilin_interp.f90:10 -- one point (sxyz(3), gridpts(3,2), qty(2,2,2)) */
void g3m3_trilin_interp(const float* sxyz, const float* gridpts, const float* qty,
                        float* scdata);

/* code/sub_sparse_solver.f90:19-20 -- CGNR on an explicit banded matrix
 * banded_matrix(nbands,npts) (device or registered), band offsets
 * divB_band_pos(nbands) (host), soln_vec = initial guess in, solution out;
 * residual2 is a host scalar, written before return.  Dot products
 * accumulate in double (the reference: single). */
void g3m3_sparse_solver(const int* nbands, const int* npts, const int* divB_band_pos,
                        const float* divB_errlim, const int* recurse, const int* redo_lim,
                        const float* divB, float* soln_vec, float* residual2,
                        const float* banded_matrix);

/* code/sub_fnd_pres.f90:5-6 -- parallel / perpendicular / cross pressure of
 * one species (pxy, pxz, pyz are taken, as by the reference, but unused) */
void g3m3_fnd_pres(const float* px, const float* py, const float* pz, const float* pxy,
                   const float* pxz, const float* pyz, float* p_para, float* p_perp,
                   float* p_cross, const float* vvx, const float* vvy, const float* vvz,
                   const float* bxt, const float* byt, const float* bzt, const int* nx,
                   const int* ny, const int* nz, const int* n_grids, const int* box);

/* code/cft_findgrid.f90:4 -- host arrays; grid_minvals/maxvals are (3,n_grids);
 * cname_len is the CHARACTER*8 hidden length a Fortran caller appends.
 * Prints the reference's warning and clamps when the point is in no box. */
void g3m3_findgrid(const float* xyz, const int* n_grids, const float* grid_minvals,
                   const float* grid_maxvals, const float* xspac, int* gridpt,
                   const float* re_equiv, const char* cname, size_t cname_len);

/* Craft sampling at npts points in one launch: for each point sxyz(3,p),
 * findgrid + the cube of code/cft_find_aux_cube.f90:62-65 + trilin_interp
 * of each of nq fields q(nx,ny,nz,n_grids) (code/cft_craft_handling.f90:
 * 165-195, code/cft_record_aux_data.f90:27-30).  scdata(nq,npts),
 * gridpt(4,npts) are device (or registered) arrays; grid_minvals,
 * grid_maxvals, xspac are host arrays.  A point whose cube would leave its
 * box gets gridpt(1,p) = 0 and zero samples. */
void g3m3_sample_fields(const int* npts, const float* sxyz, const int* n_grids,
                        const float* grid_minvals, const float* grid_maxvals, const float* xspac,
                        const float* re_equiv, const int* nx, const int* ny, const int* nz,
                        const int* nq, const float* const* fields, float* scdata, int* gridpt);

/* code/cft_find_aux_cube.f90:4-8 -- host arrays: the craft's position at
 * time ut between its trajectory points xcraftn and zcraftn, (x, y, z, t)
 * each, then findgrid and the cube vertices cube_vertices(3,2).  As in the
 * reference, a gridding problem (gridpt(1) <= 0) clears recordingn and
 * prints a notice; closing the craft's output unit scout+i_craft is left to
 * the Fortran caller.  cname_len is the CHARACTER*8 hidden length. */
void g3m3_find_aux_cube(const char* cname, const int* i_craft, const int* scout,
                        const int* n_grids, const float* grid_minvals, const float* grid_maxvals,
                        const float* xspac, const float* re_equiv, const float* xcraftn,
                        const float* zcraftn, const float* ut, int* ntimesn, int* recordingn,
                        int* gridpt, float* sxyz, float* cube_vertices, size_t cname_len);

/* No single reference counterpart: the aux-craft branch of craft_handling
 * (code/cft_craft_handling.f90:127-200) for many craft in one launch.  Each
 * craft n is placed at time ut between xcraft(:,n) and zcraft(:,n) (as
 * find_aux_cube places it) and written to sxyz(3,ncraft); then findgrid,
 * the aux cube and trilin_interp of the nq fields give scdata(nq,ncraft)
 * (record_aux_data's samples, code/cft_record_aux_data.f90:27-30) and
 * gridpt(4,ncraft).  Device arrays except the grid extents. */
void g3m3_sample_craft(const int* ncraft, const float* xcraft, const float* zcraft,
                       const float* ut, const int* n_grids, const float* grid_minvals,
                       const float* grid_maxvals, const float* xspac, const float* re_equiv,
                       const int* nx, const int* ny, const int* nz, const int* nq,
                       const float* const* fields, float* sxyz, float* scdata, int* gridpt);

/* code/sub_flux_counter.f90:5-8 -- the species' particle fluxes through
 * the six walls one cell inside box `box`; the six results are host
 * scalars (the call synchronises).  Wall sums run in the reference's loop
 * order in single precision. */
void g3m3_flux_counter(const float* qpx, const float* qpy, const float* qpz, const float* hpx,
                       const float* hpy, const float* hpz, const float* opx, const float* opy,
                       const float* opz, const int* nx, const int* ny, const int* nz,
                       const int* n_grids, const int* box, const float* rmassq,
                       const float* rmassh, const float* rmasso, const float* grid_spacing,
                       float* qflux_in, float* qflux_out, float* hflux_in, float* hflux_out,
                       float* oflux_in, float* oflux_out);

/* code/3d_main.f90:3149-3253 -- the torus source, which has no subroutine
 * of its own: h, q and o plasma added to box 1 in the torus around the
 * planet.  Fields are (nx,ny,nz,n_grids); grid_minvals/grid_maxvals are
 * the driver's (3,n_grids) host arrays; scalars are the driver's variables
 * of the same names (denq/denh/deno_torus, t_torus, aniso_factor, reduct
 * from its namelists).  tot_q/h/o (host) are the injected masses summed
 * over the cells, sum of den*dx*dy*dz in the reference's loop order, before
 * the driver's scaling to kg/s (:3255-3276); the call synchronises.  The
 * Gaussian profiles' exp is evaluated in double and rounded once. */
void g3m3_torus_inject(
    float* qrho, float* qpresx, float* qpresy, float* qpresz, float* qpx, float* qpy, float* qpz,
    float* qpresxy, float* qpresxz, float* qpresyz, float* hrho, float* hpresx, float* hpresy,
    float* hpresz, float* hpx, float* hpy, float* hpz, float* hpresxy, float* hpresxz,
    float* hpresyz, float* orho, float* opresx, float* opresy, float* opresz, float* opx,
    float* opy, float* opz, float* opresxy, float* opresxz, float* opresyz, float* epres,
    const float* bx, const float* by, const float* bz, const float* bx0, const float* by0,
    const float* bz0, const int* nx, const int* ny, const int* nz, const int* n_grids,
    const float* grid_minvals, const float* grid_maxvals, const float* xdip, const float* ydip,
    const float* zdip, const float* re_equiv, const float* v_rot, const float* torus_inj_dist,
    const float* torus_rad, const float* srf_bndry, const float* denq_torus,
    const float* denh_torus, const float* deno_torus, const float* rmassq, const float* rmassh,
    const float* rmasso, const float* t_torus, const float* aniso_factor, const float* reduct,
    float* tot_q, float* tot_h, float* tot_o);

/* code/sub_write_graphing_data.f90:348-760 -- the quantities the reference
 * tabulates for plotting on three cut planes of box `box`: the xy plane
 * k = limit/2+1, the xz plane j = limit+1 and the yz plane i = limit+1
 * (limit+1-15 for box 5), points in the reference's loop order.  For each
 * point 53 values, in its physical units and column order: plas (log10
 * densities per cm^3, temperatures) 8, flow (species and net velocities,
 * Alfven Mach number) 13, pres (log10 parallel / perpendicular / cross and
 * electron pressures) 10, bande (B, E, J) 11, model (position, perturbation
 * and dipole B) 11.  `out` (device) holds (nx*ny + nx*nz + ny*nz) * 53
 * floats.  The file writing and the plotting call stay with the caller
 * (g3m3_amd.graphing writes the reference's files). */
void g3m3_graphing_cuts(
    const float* qrho, const float* qpx, const float* qpy, const float* qpz, const float* qpresx,
    const float* qpresy, const float* qpresz, const float* hrho, const float* hpx,
    const float* hpy, const float* hpz, const float* hpresx, const float* hpresy,
    const float* hpresz, const float* orho, const float* opx, const float* opy,
    const float* opz, const float* opresx, const float* opresy, const float* opresz,
    const float* bx, const float* by, const float* bz, const float* epres, const float* bx0,
    const float* by0, const float* bz0, const float* resistive, const int* ijzero,
    const int* numzero, const int* ijmid, const int* nummid, const int* nx, const int* ny,
    const int* nz, const int* n_grids, const int* box, const int* mbndry, const int* mmid,
    const int* mzero, const float* limit, const float* xspac, const float* grid_minvals,
    const float* rmassq, const float* rmassh, const float* rmasso, const float* reynolds,
    const float* resist, const float* smallbit, const float* re_equiv, const float* b_equiv,
    const float* v_equiv, const float* rho_equiv, const float* planet_rad, float* out);

/* ---- initialisation -----------------------------------------------------------
 * common /rotation/ (code/3d_main.f90:633-634), which the reference's
 * dipole and mak_dip_grd read, comes in as the trailing arguments rot_angle,
 * xdip, ydip, zdip, sin_tilt, cos_tilt, b0; sin and cos of rot_angle are
 * taken on the host with the C library's sinf / cosf, as a compiled
 * reference takes them. */

/* code/sub_dipole.f90:4 -- the dipole field at one point, on the host. */
void g3m3_dipole(float* abx, float* aby, float* abz, const float* ax, const float* ay,
                 const float* az, const float* rot_angle, const float* xdip, const float* ydip,
                 const float* zdip, const float* sin_tilt, const float* cos_tilt, const float* b0);

/* code/sub_mak_dip_grd.f90:4-6 -- the dipole on box `box`, zero inside
 * r_inner-1.5 and at the box's ijzero points; numzero is a host array. */
void g3m3_mak_dip_grd(float* bx0, float* by0, float* bz0, const int* nx, const int* ny,
                      const int* nz, const int* n_grids, const int* mbndry, const int* box,
                      const int* ijzero, const int* numzero, const int* mzero,
                      const float* r_inner, const float* grd_xmin, const float* grd_xmax,
                      const float* grd_ymin, const float* grd_ymax, const float* grd_zmin,
                      const float* grd_zmax, const float* rot_angle, const float* xdip,
                      const float* ydip, const float* zdip, const float* sin_tilt,
                      const float* cos_tilt, const float* b0);

/* code/sub_set_resist.f90:10-13 -- rst(nx,ny,nz,mbndry); the point counts
 * are host arrays, and (as in the reference) those of box mbndry are used
 * for every box. */
void g3m3_set_resist(float* rst, const int* nx, const int* ny, const int* nz, const int* mbndry,
                     const float* resist, const int* ijzero, const int* numzero,
                     const int* ijmid, const int* nummid, const int* ijsrf, const int* numsrf,
                     const int* msrf, const int* mmid, const int* mzero, const float* b0);

/* code/3d_main.f90:1238-1633 -- the main program's initialisation block,
 * which has no subroutine of its own: the dipole field b0 of every box, the
 * ambient ionospheric plasma, the inner-boundary point lists of boxes
 * 1..mbndry with their saved values, the solar wind beyond r_rot/re_equiv,
 * the speed check, and set_resist(resistive, ..., b0 = 1.) (:1631-1633).
 * Fields come in the order q, h, o x (rho, presx, presy, presz, px, py, pz,
 * presxy, presxz, presyz), then epres, bx, by, bz, bx0, by0, bz0, each
 * (nx,ny,nz,n_grids); resistive is (nx,ny,nz,mbndry).  grid_minvals and
 * grid_maxvals are the driver's (3,n_grids) host arrays.  The lists are
 * written on the device; their counts num*(mbndry) are host outputs (the
 * call synchronises).  A list longer than its capacity (msrf, mmid, mzero)
 * is cut there and the error latch is set.  Scalars are the driver's
 * variables of the same names; rot_angle is the one computed at :1218. */
void g3m3_init_plasma(
    float* qrho, float* qpresx, float* qpresy, float* qpresz, float* qpx, float* qpy, float* qpz,
    float* qpresxy, float* qpresxz, float* qpresyz, float* hrho, float* hpresx, float* hpresy,
    float* hpresz, float* hpx, float* hpy, float* hpz, float* hpresxy, float* hpresxz,
    float* hpresyz, float* orho, float* opresx, float* opresy, float* opresz, float* opx,
    float* opy, float* opz, float* opresxy, float* opresxz, float* opresyz, float* epres,
    float* bx, float* by, float* bz, float* bx0, float* by0, float* bz0, float* resistive,
    const int* nx, const int* ny, const int* nz, const int* n_grids, const int* mbndry,
    const int* msrf, const int* mmid, const int* mzero, const float* grid_minvals,
    const float* grid_maxvals, float* parm_srf, float* parm_mid, float* parm_zero, int* ijsrf,
    int* ijmid, int* ijzero, int* numsrf, int* nummid, int* numzero, const float* rmassq,
    const float* rmassh, const float* rmasso, const float* ti_te, const float* re_equiv,
    const float* r_inner, const float* zero_bndry, const float* mid_bndry,
    const float* srf_bndry, const float* alpha_e, const float* erho, const float* d_min,
    const float* cs_inner, const float* srho, const float* spx, const float* spy,
    const float* spz, const float* spress, const float* cs_wind, const float* gamma,
    const float* rho_frac, const float* resist, const float* v_rot, const float* r_rot,
    const float* rot_angle, const float* xdip, const float* ydip, const float* zdip,
    const float* sin_tilt, const float* cos_tilt, const float* b0);

/* ---- fluid (restart) files ------------------------------------------------
 * The reference's restart format (written at code/3d_main.f90:3310-3349,
 * read at :972-1011): Fortran unformatted sequential records with 4-byte
 * byte counts -- t (double), the 37 arrays q(nx,ny,nz,n_grids) in this
 * order: per species q, h, o: rho, px, py, pz, presx, presy, presz,
 * presxy, presxz, presyz; then bx, by, bz, epres, bx0, by0, bz0 -- and one
 * record of the planet lists (parm_srf(mbndry,7,msrf), parm_mid, parm_zero,
 * ijzero(mbndry,3,mzero), numzero(mbndry), ijmid, nummid, ijsrf, numsrf),
 * msrf/mmid/mzero being the lists' declared sizes.  Host arrays; returns
 * 0, 1 (cannot open) or 2 (write failed / record sizes do not match). */
#define G3M3_FLUID_ARRAYS 37
int g3m3_fluid_write(const char* path, const double* t, const int* nx, const int* ny,
                     const int* nz, const int* n_grids, const float* const* fields,
                     const int* mbndry, const int* msrf, const int* mmid, const int* mzero,
                     const float* parm_srf, const float* parm_mid, const float* parm_zero,
                     const int* ijsrf, const int* ijmid, const int* ijzero, const int* numsrf,
                     const int* nummid, const int* numzero);
int g3m3_fluid_read(const char* path, double* t, const int* nx, const int* ny, const int* nz,
                    const int* n_grids, float* const* fields, const int* mbndry, const int* msrf,
                    const int* mmid, const int* mzero, float* parm_srf, float* parm_mid,
                    float* parm_zero, int* ijsrf, int* ijmid, int* ijzero, int* numsrf,
                    int* nummid, int* numzero);

/* ---- native runtime ----------------------------------------------------------
 * No reference counterpart: the whole run resident in HBM, advanced a full
 * step (code/3d_main.f90:1925-3095) per call with fused kernels.  Results
 * equal the per-subroutine entry points called in the reference's order.
 *
 * Field pointer arrays use this order: for each ion species q, h, o --
 * rho, presx, presy, presz, px, py, pz, presxy, presxz, presyz -- then
 * epres, bx, by, bz; each a device array q(nx,ny,nz,n_grids).  wrk and
 * old must start zeroed (the reference's static arrays).  Index lists and
 * parm arrays are device arrays in the reference's layout; extents and
 * scalars are copied at creation.
 *
 * A box-step swaps the roles of the box's state and old slices instead of
 * copying the state into old, so between calls a box's state may live in
 * the old array.  g3m3_run_sync_arrays() puts every box's state back into
 * the state array and its old state into the old array (the reference's
 * arrays, bit for bit) and waits for the run's stream; call it before
 * reading or changing the arrays outside the runtime. */
#define G3M3_MAX_GRIDS 8
typedef struct {
    int nx, ny, nz, n_grids;
    float* state[34];
    float* old[34];
    float* wrk[34];
    const float* bx0;
    const float* by0;
    const float* bz0;
    const float* rst;                  /* rst(nx,ny,nz,mbndry) */
    const int* ijsrf;
    const int* ijmid;
    const int* ijzero;
    const float* parm_srf;
    const float* parm_mid;
    const float* parm_zero;
    int numsrf, nummid, numzero, mbndry;
    float grd_xmin[G3M3_MAX_GRIDS], grd_xmax[G3M3_MAX_GRIDS];
    float grd_ymin[G3M3_MAX_GRIDS], grd_ymax[G3M3_MAX_GRIDS];
    float grd_zmin[G3M3_MAX_GRIDS], grd_zmax[G3M3_MAX_GRIDS];
    float xspac[G3M3_MAX_GRIDS];
    float rmassq, rmassh, rmasso, gamma, gamma1, reynolds, resist, grav, re_equiv, smallbit;
    float aniso_limit, qvlim, hvlim, ovlim, qclim, hclim, oclim, cslim;
    float chirho, chipxyz, chierg, difrho, diferg, o_conc, rho_frac, ti_te;
    float stepsz, t_step_max, vlim_factor;
    float srho, spx, spy, spz, spress, sbx_wind, sby_wind, sbz_wind, v_rot, r_rot;
    float eta;                         /* uniform resistivity (0 = reference) */
    int isotropic;
    /* Dipole tilting during the run (code/3d_main.f90:1966-1977, 2173-2187).
     * With tilting set, every box-step first recomputes its box's b0 with
     * mak_dip_grd for the tilt and rotation at its half step, rewriting
     * bx0..bz0 in place; tilt, ut advance once per full step from the
     * values given here.  tilting = 0: b0 is left alone. */
    int tilting;
    float tilt, dtilt;                 /* tilt now (degrees), its rate per unit time */
    double ut;                         /* universal time now (hours) */
    float t_equiv, planet_per;         /* the time unit (s) and the planet's period (h) */
    float xdip, ydip, zdip, dip_b0;    /* common /rotation/: dipole offset and strength */
    float r_inner;
} g3m3_run_desc;

void* g3m3_run_create(const g3m3_run_desc* desc);   /* NULL on failure */
void g3m3_run_destroy(void* run);
void g3m3_run_set_stream(void* run, void* hip_stream);
int g3m3_run_init_dt(void* run);                     /* code/3d_main.f90:1861-1903 */
int g3m3_run_full_step(void* run, const double* dt_override);   /* NULL = own dt */
/* g3m3_run_full_step in two halves, so that many runs on different streams
 * can be in flight together: _async issues the full step's kernels on the
 * run's stream and returns; _finish waits for them and reads the CFL maxima
 * that set the next step (the one host synchronisation of a full step). */
int g3m3_run_full_step_async(void* run, const double* dt_override);
int g3m3_run_finish(void* run);
/* Single-grid runs (n_grids = 1, no tilting): nsteps full steps replayed from HIP
 * graphs captured on the first call (one per buffer-rotation parity).  The
 * time step lives on the device: a kernel forms each next step from the
 * step's CFL maxima exactly as g3m3_run_finish does on the host, so no step
 * waits for the host and results equal g3m3_run_full_step's bit for bit.
 * Returns once the steps are issued; g3m3_run_graph_finish waits and brings
 * the run's times (t, t_step, t_stepnew) up to date.  A run without a
 * stream gets one of its own. */
int g3m3_run_graph_steps(void* run, int nsteps);
int g3m3_run_graph_finish(void* run);
/* Many single-grid runs (g3m3_run_create handles, no tilting) stepped in
 * lockstep: one launch per stage for all of them, the concatenation of the
 * runs' own grids, each run's blocks reading that run's own arguments, and
 * one captured graph per step.  A run's results equal its own
 * g3m3_run_graph_steps / g3m3_run_full_step results bit for bit.  No
 * counterpart in the reference (one run per process there).  stream: the
 * HIP stream the batch runs on (NULL: its own).  g3m3_batch_steps issues
 * nsteps full steps of every run and returns; g3m3_batch_finish waits and
 * brings each run's time, time step and error back to its handle (call it
 * before using a run on its own again).  The runs must outlive the batch. */
void* g3m3_batch_create(void* const* runs, int nruns, void* stream);   /* NULL on failure */
int g3m3_batch_steps(void* batch, int nsteps);
int g3m3_batch_finish(void* batch);
int g3m3_batch_launches(void* batch);    /* multi-run launches per lockstep step */
void g3m3_batch_destroy(void* batch);
/* The FCT launch shape the run's autotuner chose at g3m3_run_create: the
 * k-planes per block (8, 16 or 32) and the block order (0 XCD-contiguous,
 * 1 plain), with the measured microseconds of the six candidates (depth
 * major, order minor).  Every shape gives the same results. */
int g3m3_run_fct_shape(void* run, int* kc_len, int* no_remap, float* cand_us);
/* Overrides that choice (kc_len >= 1); returns nonzero on a bad value. */
int g3m3_run_set_fct_shape(void* run, int kc_len, int no_remap);
/* The species push and Lapidus kernels the autotuner chose at
 * g3m3_run_create (1: the k-marching, LDS-tiled kernels; 0: one thread per
 * cell), with the measured microseconds of one push plus one Lapidus pass
 * by each (index = the value above).  Both give the same results
 * (code/sub_push_ion.f90, code/sub_lap_plasma.f90). */
int g3m3_run_species_kernels(void* run, int* march, float* cand_us);
/* Overrides that choice (0 or 1); returns nonzero on a bad value. */
int g3m3_run_set_species_kernels(void* run, int march);
int g3m3_run_box_step(void* run, int box);
void g3m3_run_set_time(void* run, double t);          /* after a restart read */
int g3m3_run_sync_arrays(void* run);
void g3m3_run_info(void* run, double* t, double* t_step, double* t_stepnew, double* t_old,
                   double* t_new, long long* cell_updates, long long* box_steps);
/* The box-steps the last full step took, in order: box number, t_old and
 * t_new of that box after its step (what code/3d_main.f90:2167-2168 writes
 * to speeds.dat).  Fills at most cap entries; returns how many there were. */
int g3m3_run_step_log(void* run, int cap, int* box, double* t_old, double* t_new);

/* ---- adjoint of the field update ------------------------------------------
 * No reference counterpart: the magnetic-field half of a box-step on one
 * outermost box (box = n_grids = 1, no planet lists) with the plasma held
 * fixed, and its discrete adjoint, for gradients of a misfit at virtual
 * spacecraft with respect to the upstream field and a uniform resistivity.
 * One step, from b = (bx, by, bz) (three contiguous device arrays of
 * nx*ny*nz cells), follows the field calls of code/3d_main.f90: calcur,
 * total B, fnd_evel and BandE (plus eta J / n on the interior, after the
 * list term) from b (:2243-2270), push_bfld with delt/2 from old = b
 * (:2324) and bndry_outer (:2331); the same from that field with delt
 * (:2486-2565, 2576); lap_bfld (:2765) over the stage-1 field and
 * bndry_outer (:2774); fcsmooth with diferg (code/sub_flux_correct.f90:183)
 * and bndry_outer (:2932).  The fixed plasma gives fnd_evel's ion velocity
 * and lap_bfld's velocity (fnd_vtot).  ctl = {eta, sbx_wind, sby_wind,
 * sbz_wind} (host).  Device arrays; the descriptor's pointers are kept. */
typedef struct {
    int nx, ny, nz;
    const float* rho[3];      /* qrho, hrho, orho */
    const float* mom[9];      /* qpx, qpy, qpz, hpx, hpy, hpz, opx, opy, opz */
    const float* epres;
    const float* b0[3];       /* bx0, by0, bz0 */
    float rx, delt, rmassq, rmassh, rmasso, reynolds, smallbit, chierg, diferg;
} g3m3_field_model_desc;
void* g3m3_field_model_create(const g3m3_field_model_desc* desc);
void g3m3_field_model_destroy(void* model);
void g3m3_field_model_set_stream(void* model, void* hip_stream);
/* floats of one step's tape (24 arrays of nx*ny*nz) */
long long g3m3_field_model_tape_floats(void* model);
/* b_out = one step from b_in (distinct arrays); tape (or NULL) receives
 * what the adjoint of the step needs.  Returns 0 on success. */
int g3m3_field_model_step(void* model, const float* b_in, float* b_out, float* tape,
                          const float* ctl);
/* From the step's tape and the adjoint of its output: bin_bar = the adjoint
 * of its input (assigned); ctl_bar[0..3] (device doubles) += the adjoints
 * of eta and sb.  bout_bar is not changed.  Returns 0 on success. */
int g3m3_field_model_adjoint(void* model, const float* tape, const float* bout_bar,
                             float* bin_bar, double* ctl_bar, const float* ctl);
/* ---- adjoint of the whole box-step ----------------------------------------
 * No reference counterpart: one box-step of the outermost box (box =
 * n_grids = 1, no planet lists, fixed delt, the validation calls set_rho
 * and set_speed_agrd left out) with every field evolving, and its discrete
 * adjoint, for gradients of a misfit at virtual spacecraft with respect to
 * the upstream plasma and field and a uniform resistivity.  The step
 * follows code/3d_main.f90:2163-3030 for that box: stage 1 (:2243-2340:
 * E-field, push_ion for q, h, o and push_bfld with delt/2, bndry_outer),
 * stage 2 (:2486-2580: the same from the stage-1 set with delt, plus
 * push_elec), the Lapidus smoothing (:2698-2778) and flux_correct on every
 * field (:2896-2935, code/sub_flux_correct.f90) with bndry_outer after each.
 * The state is 35 contiguous device arrays of nx*ny*nz cells: q, h, o (ten
 * each: rho, presx, presy, presz, px, py, pz, presxy, presxz, presyz),
 * epres, bx, by, bz, and the electron pressure of the last Lapidus pass,
 * which the reference keeps in wrkepres and reads as the source of the next
 * step's push_elec and stage-2 BandE.  ctl (host, 9 floats) = {eta,
 * sbx_wind, sby_wind, sbz_wind, srho, spx, spy, spz, spress}: the
 * resistivity and the upstream values of code/sub_bndry_outer.f90:50-103
 * (the h and o values follow from them through rho_frac, o_conc and the
 * masses, the electron pressure through ti_te).  b0 points at device
 * arrays kept by the model. */
typedef struct {
    int nx, ny, nz;
    const float* b0[3];       /* bx0, by0, bz0 of the box */
    float rx;                 /* xspac of the box */
    float xmin, xmax, ymin, ymax, zmin, zmax;   /* grd_* of the box */
    float delt, rmassq, rmassh, rmasso, reynolds, smallbit, gamma, gamma1, grav, re_equiv;
    float chirho, chipxyz, chierg, difrho, diferg, rho_frac, o_conc, ti_te;
} g3m3_full_model_desc;
void* g3m3_full_model_create(const g3m3_full_model_desc* desc);
void g3m3_full_model_destroy(void* model);
void g3m3_full_model_set_stream(void* model, void* hip_stream);
/* floats of one step's tape (203 arrays of nx*ny*nz) */
long long g3m3_full_model_tape_floats(void* model);
/* s_out = one step from s_in (35 arrays each, distinct); tape (or NULL)
 * receives what the adjoint of the step needs.  Returns 0 on success. */
int g3m3_full_model_step(void* model, const float* s_in, float* s_out, float* tape,
                         const float* ctl);
/* From the step's tape and the adjoint of its output: sin_bar = the
 * adjoint of its input (assigned, 35 arrays); ctl_bar[0..8] (device
 * doubles) += the adjoints of the controls.  sout_bar is not changed.
 * Returns 0 on success. */
int g3m3_full_model_adjoint(void* model, const float* tape, const float* sout_bar,
                            float* sin_bar, double* ctl_bar, const float* ctl);
/* The model as the inner box of a nested pair (box < n_grids of the
 * reference; code/sub_flanks.f90, code/sub_corer.f90): geom = this box's
 * xmin, xmax, ymin, ymax, zmin, zmax, then its parent's (the parent has
 * the same extents in cells and twice the spacing, or any whole multiple).
 * Its steps then bound each stage by the flanks from the parent's states
 * p_old and p_cur (34 arrays each, canonical order) with the time weights
 * t12 = (t1, t2) at step 1's t_grid and (t1, t2) at the later stages'
 * (t1 = (t_new - t_grid) / (t_new - t_old) of the parent,
 * code/sub_flanks.f90:21-22); their adjoint adds the flanks' adjoints into
 * p_old_bar and p_cur_bar.  g3m3_full_model_corer sets the parent's core
 * from this box's state; its adjoint moves the parent core's adjoint into
 * this box's (then zeroes it).  Returns 0 on success. */
int g3m3_full_model_set_parent(void* model, const float* geom);
int g3m3_full_model_step_inner(void* model, const float* s_in, float* s_out, float* tape,
                               const float* ctl, const float* p_old, const float* p_cur,
                               const float* t12);
int g3m3_full_model_adjoint_inner(void* model, const float* tape, const float* sout_bar,
                                  float* sin_bar, double* ctl_bar, const float* ctl,
                                  const float* t12, float* p_old_bar, float* p_cur_bar);
int g3m3_full_model_corer(void* model, const float* s, float* p_state);
int g3m3_full_model_corer_adjoint(void* model, float* s_bar, float* p_bar);

/* The adjoint of g3m3_sample_fields (code/cft_trilin_interp.f90:10-66 for
 * each point): fields_bar[q] += the trilinear weights of point p's cube
 * times scdata_bar(q,p).  Points outside every box add nothing. */
void g3m3_sample_fields_adjoint(const int* npts, const float* sxyz, const int* n_grids,
                                const float* grid_minvals, const float* grid_maxvals,
                                const float* xspac, const float* re_equiv, const int* nx,
                                const int* ny, const int* nz, const int* nq,
                                float* const* fields_bar, const float* scdata_bar);

/* No reference counterparts: the adjoints of single subroutines, for an
 * adjoint driver that runs a step's calls backwards.  Box-local device
 * arrays of nx*ny*nz cells; a species' ten fields are passed as host arrays
 * of ten device pointers in the order rho, presx, presy, presz, px, py, pz,
 * presxy, presxz, presyz.  Every adjoint ACCUMULATES (+=) into its array;
 * the output adjoint is not changed.
 *
 * g3m3_push_ion_adjoint -- code/sub_push_ion.f90:4-362 for one species,
 * transposed: from the adjoint of the ten outputs (interior cells; the
 * boundary cells push_ion does not write carry nothing), the adjoints of
 * the source (wrk) fields, of the old fields, of bs, btot, efld and tv.
 * Scalars and grid extents are the box's (grd_*min/max of this box).
 * work: 15*nx*ny*nz floats of scratch. */
void g3m3_push_ion_adjoint(const float* const* src, const float* bsx, const float* bsy,
                           const float* bsz, const float* btot, const float* efldx,
                           const float* efldy, const float* efldz, const float* tvx,
                           const float* tvy, const float* tvz, const float* const* out_bar,
                           float* const* src_bar, float* const* old_bar, float* bsx_bar,
                           float* bsy_bar, float* bsz_bar, float* btot_bar, float* efldx_bar,
                           float* efldy_bar, float* efldz_bar, float* tvx_bar, float* tvy_bar,
                           float* tvz_bar, const int* nx, const int* ny, const int* nz,
                           const float* rmass, const float* delt, const float* grav,
                           const float* re_equiv, const float* reynolds, const float* grd_xmin,
                           const float* grd_xmax, const float* grd_ymin, const float* grd_ymax,
                           const float* grd_zmin, const float* grd_zmax, float* work);
/* code/sub_lap_plasma.f90:4-147 for one species, transposed: from the
 * adjoint of the ten smoothed fields, the adjoints of the ten fields and of
 * the velocity v (chirho for rho, chipxyz for the rest).  work: 6*nx*ny*nz
 * floats. */
void g3m3_lap_plasma_adjoint(const float* const* q, const float* vx, const float* vy,
                             const float* vz, const float* const* w_bar, float* const* q_bar,
                             float* vx_bar, float* vy_bar, float* vz_bar, const int* nx,
                             const int* ny, const int* nz, const float* chirho,
                             const float* chipxyz, const float* delt, float* work);
/* v = p / max(rho, smallbit) (code/3d_main.f90:2698-2703), transposed over
 * n cells. */
void g3m3_species_velocity_adjoint(const float* rho, const float* px, const float* py,
                                   const float* pz, const float* vx_bar, const float* vy_bar,
                                   const float* vz_bar, float* rho_bar, float* px_bar,
                                   float* py_bar, float* pz_bar, const int* n,
                                   const float* smallbit);

/* ---- Fortran implicit-interface names ------------------------------------------
 * A Fortran call with no interface block (the reference's own style) names
 * g3m3_<name>_ and passes everything by reference.  Every subroutine entry
 * point above is also exported under that name (an alias); the control and
 * registry calls have by-reference forms: */
int g3m3_init_(void);
int g3m3_last_error_(void);
void g3m3_sync_(void);
int g3m3_register_(void* host_ptr, const long long* bytes);
int g3m3_unregister_(void* host_ptr);
int g3m3_to_device_(void* host_ptr);
int g3m3_to_host_(void* host_ptr);

#ifdef __cplusplus
}
#endif
#endif /* G3M3_H */
