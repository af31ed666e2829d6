// extern "C" entry points of include/g3m3.h: argument unpacking, pointer
// translation for registered host arrays, workspace, error latch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/g3m3.h"
#include "g3m3_craft.h"
#include "g3m3_fused.h"
#include "g3m3_launch.h"

using namespace g3m3;

namespace {

hipStream_t g_stream = nullptr;
int g_err = 0;
std::string g_msg;
float g_iso_flag = 1.0f;   // the reference's SAVEd iso_flag (sub_set_speed_agrd.f90:68)

void latch(int code, const char* what) {
    if (g_err == 0) {
        g_err = code;
        g_msg = what;
    }
}

// true when e is hipSuccess; otherwise the error is latched
bool hip_ok(hipError_t e, const char* where) {
    if (e == hipSuccess) return true;
    std::string m = std::string(where) + ": " + hipGetErrorString(e);
    latch((int)e, m.c_str());
    return false;
}

void check_launch(const char* where) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::string m = std::string(where) + ": " + hipGetErrorString(e);
        latch((int)e, m.c_str());
    }
}

// --- host-array registry -------------------------------------------------------
struct Mirror {
    char* dev;
    size_t bytes;
};
std::map<uintptr_t, Mirror> g_reg;   // keyed by host base address

template <typename T>
T* R(T* p) {
    if (!p || g_reg.empty()) return p;
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    auto it = g_reg.upper_bound(a);
    if (it == g_reg.begin()) return p;
    --it;
    if (a >= it->first && a < it->first + it->second.bytes)
        return reinterpret_cast<T*>(it->second.dev + (a - it->first));
    return p;
}

// --- grow-only device workspace ---------------------------------------------------
struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    void* get(size_t need) {
        if (need > bytes) {
            if (p) hip_ok(hipFree(p), "workspace free");
            p = nullptr;
            if (hipMalloc(&p, need) != hipSuccess) {
                latch(2, "workspace allocation failed");
                bytes = 0;
                return nullptr;
            }
            bytes = need;
        }
        return p;
    }
};
Buf g_speed_list, g_speed_part, g_small, g_divb_f, g_divb_d, g_tmp;

}  // namespace

// --- fixed scratch of the allocation-free entry point -----------------------------
// g3m3_set_speed_agrd_noalloc keeps its side-list, per-block maxima, maxima
// and count here instead of in grow-only workspace.  Layout (bytes): the 5
// maxima at 0 and the count at 64; per-block maxima from kNoAllocPartOff;
// the side-list (3 species x V entries of 48 bytes) from kNoAllocListOff.
// Each device holds its own copy (18.9 MB) until the process ends.
constexpr long long kNoAllocMaxCells = 131072;
constexpr size_t kNoAllocPartBytes = 5 * ((kNoAllocMaxCells + 255) / 256) * sizeof(float);
constexpr size_t kNoAllocListBytes = 3 * (size_t)kNoAllocMaxCells * 48;
constexpr size_t kNoAllocPartOff = 256;
constexpr size_t kNoAllocListOff = kNoAllocPartOff + (kNoAllocPartBytes + 255) / 256 * 256;
constexpr size_t kNoAllocBytes = kNoAllocListOff + kNoAllocListBytes;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;
unsigned char* g_scratch_dev[kMaxDevices] = {};   // scratch's address on each device

// The current device's address of `scratch`, looked up once per device
// (g3m3_init does it for device 0).  The array's host-side name is not a
// device address and is never passed to a kernel or a copy.
unsigned char* noalloc_scratch() {
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "noalloc scratch")) return nullptr;
    if (dev < 0 || dev >= kMaxDevices) {
        latch(4, "noalloc scratch: device ordinal out of range");
        return nullptr;
    }
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "noalloc scratch address"))
            return nullptr;
        g_scratch_dev[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_dev[dev];
}

inline Box box_of(const int* nx, const int* ny, const int* nz) { return Box{*nx, *ny, *nz}; }
inline size_t off(const Box& B, const int* box) { return (size_t)(*box - 1) * B.V(); }

template <typename T>
T* S(T* base, const Box& B, const int* box) {
    return R(base) + off(B, box);
}

// --- per-kernel event timing ------------------------------------------------------
std::string g_prof_name;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_prof_events;
hipEvent_t g_prof_open = nullptr;
double g_prof_ms = 0.0;
long long g_prof_count = 0;

void prof_drain() {
    for (auto& e : g_prof_events) {
        float ms = 0.f;
        if (hip_ok(hipEventSynchronize(e.second), "kernel timing") &&
            hip_ok(hipEventElapsedTime(&ms, e.first, e.second), "kernel timing")) {
            g_prof_ms += ms;
            ++g_prof_count;
        }
        hip_ok(hipEventDestroy(e.first), "kernel timing");
        hip_ok(hipEventDestroy(e.second), "kernel timing");
    }
    g_prof_events.clear();
}

}  // namespace

namespace g3m3 {
void prof_begin(const char* kernel, hipStream_t s) {
    if (g_prof_name.empty() || g_prof_name != kernel) return;
    if (!hip_ok(hipEventCreate(&g_prof_open), "kernel timing")) {
        g_prof_open = nullptr;
        return;
    }
    hip_ok(hipEventRecord(g_prof_open, s), "kernel timing");
}
void prof_end(const char* kernel, hipStream_t s) {
    if (g_prof_name.empty() || g_prof_name != kernel || !g_prof_open) return;
    hipEvent_t stop;
    if (!hip_ok(hipEventCreate(&stop), "kernel timing")) return;
    hip_ok(hipEventRecord(stop, s), "kernel timing");
    g_prof_events.emplace_back(g_prof_open, stop);
    g_prof_open = nullptr;
    if (g_prof_events.size() > 4096) prof_drain();
}
}  // namespace g3m3

extern "C" {

/* Select the kernel whose launches are timed with HIP events (NULL or "" =
 * none) and reset the totals. */
void g3m3_prof_select(const char* kernel) {
    prof_drain();
    g_prof_name = kernel ? kernel : "";
    g_prof_ms = 0.0;
    g_prof_count = 0;
}

/* Total milliseconds and launch count of the selected kernel so far. */
void g3m3_prof_read(double* ms, long long* launches) {
    prof_drain();
    if (ms) *ms = g_prof_ms;
    if (launches) *launches = g_prof_count;
}

int g3m3_init(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        latch(1, "no HIP device");
        return 1;
    }
    if (!hip_ok(hipSetDevice(0), "init")) return 1;
    if (!noalloc_scratch()) return 1;
    return 0;
}

int g3m3_last_error(void) {
    const int e = g_err;
    g_err = 0;
    return e;
}

const char* g3m3_last_error_msg(void) { return g_msg.c_str(); }

void g3m3_set_stream(void* s) { g_stream = (hipStream_t)s; }

void g3m3_sync(void) {
    if (hipStreamSynchronize(g_stream) != hipSuccess) latch(3, "stream synchronize failed");
    check_launch("sync");
}

int g3m3_register(void* host, long long bytes) {
    if (!host || bytes <= 0) return 1;
    void* d = nullptr;
    if (hipMalloc(&d, (size_t)bytes) != hipSuccess) {
        latch(2, "register: allocation failed");
        return 2;
    }
    g_reg[reinterpret_cast<uintptr_t>(host)] = Mirror{(char*)d, (size_t)bytes};
    return 0;
}

int g3m3_unregister(void* host) {
    auto it = g_reg.find(reinterpret_cast<uintptr_t>(host));
    if (it == g_reg.end()) return 1;
    hip_ok(hipStreamSynchronize(g_stream), "unregister");
    hip_ok(hipFree(it->second.dev), "unregister");
    g_reg.erase(it);
    return 0;
}

int g3m3_to_device(void* host) {
    auto it = g_reg.find(reinterpret_cast<uintptr_t>(host));
    if (it == g_reg.end()) return 1;
    return hip_ok(hipMemcpyAsync(it->second.dev, host, it->second.bytes, hipMemcpyHostToDevice,
                                 g_stream),
                  "to_device")
               ? 0
               : 2;
}

int g3m3_to_host(void* host) {
    auto it = g_reg.find(reinterpret_cast<uintptr_t>(host));
    if (it == g_reg.end()) return 1;
    if (!hip_ok(hipMemcpyAsync(host, it->second.dev, it->second.bytes, hipMemcpyDeviceToHost,
                               g_stream),
                "to_host") ||
        !hip_ok(hipStreamSynchronize(g_stream), "to_host"))
        return 2;
    return 0;
}

// -----------------------------------------------------------------------------
void g3m3_store_array(float* bx, const float* bx0, const int* nx, const int* ny, const int* nz,
                      const int* n_grids, const int* box) {
    const Box B = box_of(nx, ny, nz);
    FieldPtrs d{}, s{};
    d.p[0] = S(bx, B, box);
    s.p[0] = const_cast<float*>(S(bx0, B, box));
    launch_copy_fields(B.V(), 1, d, s, g_stream);
    check_launch("store_array");
}

void g3m3_calcur(const float* bx, const float* by, const float* bz, const int* nx, const int* ny,
                 const int* nz, const int* n_grids, const int* box, float* curx, float* cury,
                 float* curz, const float* rx, const float* ry, const float* rz) {
    const Box B = box_of(nx, ny, nz);
    launch_calcur(B, S(bx, B, box), S(by, B, box), S(bz, B, box), R(curx), R(cury), R(curz), *rx,
                  *ry, *rz, g_stream);
    check_launch("calcur");
}

void g3m3_fnd_evel(const float* qpx, const float* qpy, const float* qpz, const float* qrho,
                   const float* hpx, const float* hpy, const float* hpz, const float* hrho,
                   const float* opx, const float* opy, const float* opz, const float* orho,
                   const float* curx, const float* cury, const float* curz, float* evx,
                   float* evy, float* evz, float* tvx, float* tvy, float* tvz, const int* nx,
                   const int* ny, const int* nz, const int* n_grids, const int* box,
                   const float* rmassq, const float* rmassh, const float* rmasso,
                   const float* reynolds) {
    const Box B = box_of(nx, ny, nz);
    const float* mom[9] = {S(qpx, B, box), S(qpy, B, box), S(qpz, B, box),
                           S(hpx, B, box), S(hpy, B, box), S(hpz, B, box),
                           S(opx, B, box), S(opy, B, box), S(opz, B, box)};
    const float* rho[3] = {S(qrho, B, box), S(hrho, B, box), S(orho, B, box)};
    const float* cur[3] = {R(curx), R(cury), R(curz)};
    float* ev[3] = {R(evx), R(evy), R(evz)};
    float* tv[3] = {R(tvx), R(tvy), R(tvz)};
    const float* nb[3] = {nullptr, nullptr, nullptr};
    float* nbs[3] = {nullptr, nullptr, nullptr};
    launch_bfield_evel(B, nb, nb, nbs, nullptr, mom, rho, cur, ev, tv, *rmassq, *rmassh, *rmasso,
                       *reynolds, 0.0f, 0, 1, g_stream);
    check_launch("fnd_evel");
}

void g3m3_bfield_total(const float* bx, const float* by, const float* bz, const float* bx0,
                       const float* by0, const float* bz0, float* bsx, float* bsy, float* bsz,
                       float* btot, const int* nx, const int* ny, const int* nz,
                       const int* n_grids, const int* box, const float* smallbit) {
    const Box B = box_of(nx, ny, nz);
    const float* b0[3] = {S(bx0, B, box), S(by0, B, box), S(bz0, B, box)};
    const float* b[3] = {S(bx, B, box), S(by, B, box), S(bz, B, box)};
    float* bs[3] = {R(bsx), R(bsy), R(bsz)};
    const float* nm[9] = {};
    const float* nr[3] = {};
    float* nv[3] = {};
    launch_bfield_evel(B, b0, b, bs, R(btot), nm, nr, nr, nv, nv, 1.f, 1.f, 1.f, 1.f, *smallbit, 1,
                       0, g_stream);
    check_launch("bfield_total");
}

static void bande_impl(float* efldx, float* efldy, float* efldz, const float* bsx,
                       const float* bsy, const float* bsz, const float* curx, const float* cury,
                       const float* curz, const float* evx, const float* evy, const float* evz,
                       const float* epres, const float* qrho, const float* hrho,
                       const float* orho, const float* rst, const float* resist,
                       const float* reynolds, const int* nx, const int* ny, const int* nz,
                       const int* box, const float* rmassq, const float* rmassh,
                       const float* rmasso, const int* ijmid, const int* nummid,
                       const int* ijzero, const int* mbndry, const int* numzero, const float* rx,
                       const float* ry, const float* rz, float eta) {
    const Box B = box_of(nx, ny, nz);
    float* e[3] = {R(efldx), R(efldy), R(efldz)};
    const float* bs[3] = {R(bsx), R(bsy), R(bsz)};
    const float* cur[3] = {R(curx), R(cury), R(curz)};
    const float* ev[3] = {R(evx), R(evy), R(evz)};
    const float* rho[3] = {S(qrho, B, box), S(hrho, B, box), S(orho, B, box)};
    const int b = *box - 1;
    const int nm = (*box <= *mbndry) ? nummid[b] : 0;
    const int nz0 = (*box <= *mbndry) ? numzero[b] : 0;
    launch_bande(B, e, bs, cur, ev, S(epres, B, box), rho, R(rst), *resist, *reynolds, *box,
                 *rmassq, *rmassh, *rmasso, R(ijmid), nm, R(ijzero), nz0, *mbndry, *rx, *ry, *rz,
                 eta, g_stream);
    check_launch("bande");
}

void g3m3_bande(float* efldx, float* efldy, float* efldz, const float* bsx, const float* bsy,
                const float* bsz, const float* curx, const float* cury, const float* curz,
                const float* evx, const float* evy, const float* evz, const float* btot,
                const float* epres, const float* qrho, const float* hrho, const float* orho,
                const float* rst, const float* resist, const float* reynolds, const int* nx,
                const int* ny, const int* nz, const int* n_grids, const int* box,
                const float* rmassq, const float* rmassh, const float* rmasso, const int* ijmid,
                const int* nummid, const int* ijzero, const int* mbndry, const int* numzero,
                const int* mmid, const int* mzero, const float* rx, const float* ry,
                const float* rz) {
    bande_impl(efldx, efldy, efldz, bsx, bsy, bsz, curx, cury, curz, evx, evy, evz, epres, qrho,
               hrho, orho, rst, resist, reynolds, nx, ny, nz, box, rmassq, rmassh, rmasso, ijmid,
               nummid, ijzero, mbndry, numzero, rx, ry, rz, 0.0f);
}

void g3m3_bande_eta(float* efldx, float* efldy, float* efldz, const float* bsx, const float* bsy,
                    const float* bsz, const float* curx, const float* cury, const float* curz,
                    const float* evx, const float* evy, const float* evz, const float* btot,
                    const float* epres, const float* qrho, const float* hrho, const float* orho,
                    const float* rst, const float* resist, const float* reynolds, const int* nx,
                    const int* ny, const int* nz, const int* n_grids, const int* box,
                    const float* rmassq, const float* rmassh, const float* rmasso,
                    const int* ijmid, const int* nummid, const int* ijzero, const int* mbndry,
                    const int* numzero, const int* mmid, const int* mzero, const float* rx,
                    const float* ry, const float* rz, const float* eta) {
    bande_impl(efldx, efldy, efldz, bsx, bsy, bsz, curx, cury, curz, evx, evy, evz, epres, qrho,
               hrho, orho, rst, resist, reynolds, nx, ny, nz, box, rmassq, rmassh, rmasso, ijmid,
               nummid, ijzero, mbndry, numzero, rx, ry, rz, *eta);
}

void g3m3_push_elec(const float* wrkepres, const float* oldepres, float* epres, const float* evx,
                    const float* evy, const float* evz, const float* gamma, const float* gamma1,
                    const int* nx, const int* ny, const int* nz, const int* n_grids,
                    const int* box, const float* delt, const float* rx, const float* ry,
                    const float* rz) {
    const Box B = box_of(nx, ny, nz);
    const float* ev[3] = {R(evx), R(evy), R(evz)};
    launch_push_elec(B, S(wrkepres, B, box), S(oldepres, B, box), S(epres, B, box), ev, *gamma,
                     *gamma1, *delt, *rx, *ry, *rz, g_stream);
    check_launch("push_elec");
}

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
                   const int* isotropic) {
    const Box B = box_of(nx, ny, nz);
    IonPtrs P;
    float* out[10] = {qrho, qpresx, qpresy, qpresz, qpx, qpy, qpz, qpresxy, qpresxz, qpresyz};
    const float* old[10] = {oldqrho, oldqpresx, oldqpresy, oldqpresz, oldqpx, oldqpy, oldqpz,
                            oldqpresxy, oldqpresxz, oldqpresyz};
    const float* wrk[10] = {wrkqrho, wrkqpresx, wrkqpresy, wrkqpresz, wrkqpx, wrkqpy, wrkqpz,
                            wrkqpresxy, wrkqpresxz, wrkqpresyz};
    for (int f = 0; f < 10; ++f) {
        P.out[f] = S(out[f], B, box);
        P.old[f] = S(old[f], B, box);
        P.wrk[f] = S(wrk[f], B, box);
    }
    P.bs[0] = R(bsx);
    P.bs[1] = R(bsy);
    P.bs[2] = R(bsz);
    P.btot = R(btot);
    P.e[0] = R(efldx);
    P.e[1] = R(efldy);
    P.e[2] = R(efldz);
    P.qv[0] = R(qvx);
    P.qv[1] = R(qvy);
    P.qv[2] = R(qvz);
    P.tv[0] = R(tvx);
    P.tv[1] = R(tvy);
    P.tv[2] = R(tvz);
    const int b = *box - 1;
    const float gmin[3] = {grd_xmin[b], grd_ymin[b], grd_zmin[b]};
    const float gmax[3] = {grd_xmax[b], grd_ymax[b], grd_zmax[b]};
    launch_push_ion(B, P, *rmassq, *delt, *grav, *re_equiv, *reynolds, gmin, gmax, g_stream);
    check_launch("push_ion");
}

void g3m3_push_bfld(float* bx, float* by, float* bz, const float* oldbx, const float* oldby,
                    const float* oldbz, const float* efldx, const float* efldy, const float* efldz,
                    const int* nx, const int* ny, const int* nz, const int* n_grids,
                    const int* box, const float* delt, const float* rx, const float* ry,
                    const float* rz) {
    const Box B = box_of(nx, ny, nz);
    float* b[3] = {S(bx, B, box), S(by, B, box), S(bz, B, box)};
    const float* o[3] = {S(oldbx, B, box), S(oldby, B, box), S(oldbz, B, box)};
    const float* e[3] = {R(efldx), R(efldy), R(efldz)};
    launch_push_bfld(B, b, o, e, *delt, *rx, *ry, *rz, g_stream);
    check_launch("push_bfld");
}

void g3m3_set_rho(float* qrho, float* qpresx, float* qpresy, float* qpresz, float* qpresxy,
                  float* qpresxz, float* qpresyz, const float* rmassq, float* hrho, float* hpresx,
                  float* hpresy, float* hpresz, float* hpresxy, float* hpresxz, float* hpresyz,
                  const float* rmassh, float* orho, float* opresx, float* opresy, float* opresz,
                  float* opresxy, float* opresxz, float* opresyz, const float* rmasso,
                  float* epres, const int* nx, const int* ny, const int* nz, const int* n_grids,
                  const int* box, const float* o_conc) {
    const Box B = box_of(nx, ny, nz);
    SpeciesPtrs SP{};
    float* g[3][7] = {{qrho, qpresx, qpresy, qpresz, qpresxy, qpresxz, qpresyz},
                      {hrho, hpresx, hpresy, hpresz, hpresxy, hpresxz, hpresyz},
                      {orho, opresx, opresy, opresz, opresxy, opresxz, opresyz}};
    for (int s = 0; s < 3; ++s) {
        SP.f[s][0] = S(g[s][0], B, box);
        for (int d = 0; d < 3; ++d) SP.f[s][1 + d] = S(g[s][1 + d], B, box);
        for (int d = 0; d < 3; ++d) SP.f[s][7 + d] = S(g[s][4 + d], B, box);
    }
    launch_set_rho(B.V(), SP, S(epres, B, box), g_stream);
    check_launch("set_rho");
}

// The dummies of code/sub_set_speed_agrd.f90:6-22, in order, shared by the
// entry point and its allocation-free variant.
#define G3M3_SPEED_AGRD_PARAMS                                                                  \
    const int *nx, const int *ny, const int *nz, const int *n_grids, const int *box,            \
        const float *rmassq, const float *rmassh, const float *rmasso, const int *isotropic,    \
        const float *smallbit, const float *aniso_limit, const float *vlim, const float *qvlim, \
        const float *hvlim, const float *ovlim, const float *qclim, const float *hclim,         \
        const float *oclim, const float *cslim, const float *bx, const float *by,               \
        const float *bz, const float *bx0, const float *by0, const float *bz0, float *qrho,     \
        float *qpx, float *qpy, float *qpz, float *qpresx, float *qpresy, float *qpresz,        \
        float *qpresxy, float *qpresxz, float *qpresyz, float *hrho, float *hpx, float *hpy,    \
        float *hpz, float *hpresx, float *hpresy, float *hpresz, float *hpresxy,                \
        float *hpresxz, float *hpresyz, float *orho, float *opx, float *opy, float *opz,        \
        float *opresx, float *opresy, float *opresz, float *opresxy, float *opresxz,            \
        float *opresyz, float *epres, float *btot, float *pxmax, float *pymax, float *pzmax,    \
        float *csmax, float *alfmax, float *fastest
#define G3M3_SPEED_AGRD_ARGS                                                                    \
    nx, ny, nz, n_grids, box, rmassq, rmassh, rmasso, isotropic, smallbit, aniso_limit, vlim,   \
        qvlim, hvlim, ovlim, qclim, hclim, oclim, cslim, bx, by, bz, bx0, by0, bz0, qrho, qpx,  \
        qpy, qpz, qpresx, qpresy, qpresz, qpresxy, qpresxz, qpresyz, hrho, hpx, hpy, hpz,       \
        hpresx, hpresy, hpresz, hpresxy, hpresxz, hpresyz, orho, opx, opy, opz, opresx, opresy, \
        opresz, opresxy, opresxz, opresyz, epres, btot, pxmax, pymax, pzmax, csmax, alfmax,     \
        fastest

// list: the averaging side-list (speed_list_bytes), part: per-block maxima
// (speed_partials_bytes), small: the 5 maxima at byte 0 and the list count
// at byte 64.
static void speed_agrd_impl(G3M3_SPEED_AGRD_PARAMS, void* list, float* part, char* small) {
    const Box B = box_of(nx, ny, nz);
    SpeciesPtrs SP{};
    float* g[3][10] = {{qrho, qpresx, qpresy, qpresz, qpx, qpy, qpz, qpresxy, qpresxz, qpresyz},
                       {hrho, hpresx, hpresy, hpresz, hpx, hpy, hpz, hpresxy, hpresxz, hpresyz},
                       {orho, opresx, opresy, opresz, opx, opy, opz, opresxy, opresxz, opresyz}};
    for (int s = 0; s < 3; ++s)
        for (int f = 0; f < 10; ++f) SP.f[s][f] = S(g[s][f], B, box);
    if (*isotropic) g_iso_flag = 0.0f;
    SpeedParams Q{};
    Q.aniso = *aniso_limit;
    Q.smallbit = *smallbit;
    Q.vlim = *vlim;
    Q.cslim = *cslim;
    Q.iso_flag = g_iso_flag;
    Q.isotropic = *isotropic ? 1 : 0;
    Q.svlim[0] = *qvlim;
    Q.svlim[1] = *hvlim;
    Q.svlim[2] = *ovlim;
    Q.sclim[0] = *qclim;
    Q.sclim[1] = *hclim;
    Q.sclim[2] = *oclim;
    const float* b0[3] = {S(bx0, B, box), S(by0, B, box), S(bz0, B, box)};
    const float* b[3] = {S(bx, B, box), S(by, B, box), S(bz, B, box)};
    float* maxima = (float*)small;
    int* count = (int*)(small + 64);
    launch_set_speed(B, SP, Q, b0, b, R(btot), S(epres, B, box), maxima, list, count, part,
                     g_stream);
    check_launch("set_speed_agrd");
    float h[6];
    if (!hip_ok(hipMemcpyAsync(h, maxima, sizeof(h), hipMemcpyDeviceToHost, g_stream),
                "set_speed_agrd") ||
        !hip_ok(hipStreamSynchronize(g_stream), "set_speed_agrd"))
        return;
    *pxmax = h[0];
    *pymax = h[1];
    *pzmax = h[2];
    *csmax = h[3];
    *alfmax = h[4];
    const float f = std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2] + h[3] * h[3] + h[4] * h[4]);
    *fastest = std::max(0.0f, f);
}

void g3m3_set_speed_agrd(G3M3_SPEED_AGRD_PARAMS) {
    const Box B = box_of(nx, ny, nz);
    void* list = g_speed_list.get(speed_list_bytes(B.V()));
    float* part = (float*)g_speed_part.get(speed_partials_bytes(B.V()));
    char* small = (char*)g_small.get(256);
    if (!list || !part || !small) return;
    speed_agrd_impl(G3M3_SPEED_AGRD_ARGS, list, part, small);
}

// Allocation-free set_speed_agrd: the side-list, the per-block maxima, the
// maxima and the count live in `scratch`, a fixed device array sized for
// boxes of up to kNoAllocMaxCells cells.  Host code reaches the array only
// through the per-device address noalloc_scratch() returns.
void g3m3_set_speed_agrd_noalloc(G3M3_SPEED_AGRD_PARAMS) {
    const Box B = box_of(nx, ny, nz);
    const long long V = (long long)B.nx * B.ny * B.nz;
    if (B.nx < 1 || B.ny < 1 || B.nz < 1 || V > kNoAllocMaxCells ||
        speed_list_bytes((int)V) > kNoAllocListBytes ||
        speed_partials_bytes((int)V) > kNoAllocPartBytes) {
        latch(4, "set_speed_agrd_noalloc: the box exceeds the 131072-cell scratch (or has an "
                 "empty dimension); call g3m3_set_speed_agrd");
        return;
    }
    unsigned char* base = noalloc_scratch();
    if (!base) return;
    speed_agrd_impl(G3M3_SPEED_AGRD_ARGS, base + kNoAllocListOff,
                    (float*)(base + kNoAllocPartOff), (char*)base);
}

void g3m3_species_velocity(const float* rho, const float* px, const float* py, const float* pz,
                           float* vx, float* vy, float* vz, const int* nx, const int* ny,
                           const int* nz, const int* n_grids, const int* box,
                           const float* smallbit) {
    const Box B = box_of(nx, ny, nz);
    const float* p[3] = {S(px, B, box), S(py, B, box), S(pz, B, box)};
    float* v[3] = {R(vx), R(vy), R(vz)};
    launch_velocity(B.V(), S(rho, B, box), p, v, *smallbit, g_stream);
    check_launch("species_velocity");
}

static void lap_impl(const Box& B, const float* const* src, float* const* dst, const float* chi,
                     int nf, const float* vx, const float* vy, const float* vz, float delt,
                     const int* box) {
    LapPtrs L{};
    for (int f = 0; f < nf; ++f) {
        L.src[f] = S(src[f], B, box);
        L.dst[f] = S(dst[f], B, box);
        L.chi[f] = chi[f];
    }
    const float* v[3] = {R(vx), R(vy), R(vz)};
    launch_lapidus(B, L, nf, v, delt, g_stream);
    check_launch("lapidus");
}

void g3m3_lap_plasma(const float* rho, const float* px, const float* py, const float* pz,
                     const float* presx, const float* presy, const float* presz,
                     const float* presxy, const float* presxz, const float* presyz,
                     float* wrkrho, float* wrkpx, float* wrkpy, float* wrkpz, float* wrkpresx,
                     float* wrkpresy, float* wrkpresz, float* wrkpresxy, float* wrkpresxz,
                     float* wrkpresyz, const float* vx, const float* vy, const float* vz,
                     const int* nx, const int* ny, const int* nz, const int* n_grids,
                     const int* box, const float* chirho, const float* chipxyz,
                     const float* chierg, const float* delt, const int* isotropic) {
    const Box B = box_of(nx, ny, nz);
    const float* src[10] = {rho, px, py, pz, presx, presy, presz, presxy, presxz, presyz};
    float* dst[10] = {wrkrho, wrkpx, wrkpy, wrkpz, wrkpresx, wrkpresy, wrkpresz,
                      wrkpresxy, wrkpresxz, wrkpresyz};
    float chi[10];
    chi[0] = *chirho;
    for (int f = 1; f < 10; ++f) chi[f] = *chipxyz;
    lap_impl(B, src, dst, chi, 10, vx, vy, vz, *delt, box);
}

void g3m3_lap_elec(const float* ppres, float* wrkppres, const float* vx, const float* vy,
                   const float* vz, const int* nx, const int* ny, const int* nz,
                   const int* n_grids, const int* box, const float* chirho,
                   const float* chipxyz, const float* chierg, const float* delt) {
    const Box B = box_of(nx, ny, nz);
    const float* src[1] = {ppres};
    float* dst[1] = {wrkppres};
    lap_impl(B, src, dst, chipxyz, 1, vx, vy, vz, *delt, box);
}

void g3m3_lap_bfld(const float* bx, const float* by, const float* bz, float* wrkbx, float* wrkby,
                   float* wrkbz, const float* vx, const float* vy, const float* vz,
                   const int* nx, const int* ny, const int* nz, const int* n_grids,
                   const int* box, const float* chirho, const float* chipxyz,
                   const float* chierg, const float* delt) {
    const Box B = box_of(nx, ny, nz);
    const float* src[3] = {bx, by, bz};
    float* dst[3] = {wrkbx, wrkby, wrkbz};
    const float chi[3] = {*chierg, *chierg, *chierg};
    lap_impl(B, src, dst, chi, 3, vx, vy, vz, *delt, box);
}

void g3m3_fnd_vtot(const float* qpx, const float* qpy, const float* qpz, const float* qrho,
                   const float* hpx, const float* hpy, const float* hpz, const float* hrho,
                   const float* opx, const float* opy, const float* opz, const float* orho,
                   float* vx, float* vy, float* vz, const int* nx, const int* ny, const int* nz,
                   const int* n_grids, const int* box, const float* rmassq,
                   const float* rmassh, const float* rmasso) {
    const Box B = box_of(nx, ny, nz);
    const float* mom[9] = {S(qpx, B, box), S(qpy, B, box), S(qpz, B, box),
                           S(hpx, B, box), S(hpy, B, box), S(hpz, B, box),
                           S(opx, B, box), S(opy, B, box), S(opz, B, box)};
    const float* rho[3] = {S(qrho, B, box), S(hrho, B, box), S(orho, B, box)};
    float* v[3] = {R(vx), R(vy), R(vz)};
    launch_vtot(B.V(), mom, rho, v, g_stream);
    check_launch("fnd_vtot");
}

static void fct_impl(const Box& B, float* const* out, const float* const* old,
                     const float* const* wrk, const float* chi, int nf, const int* box) {
    FctParams F{};
    F.B = B;
    for (int f = 0; f < nf; ++f) {
        F.out[f] = S(out[f], B, box);
        F.old[f] = S(old[f], B, box);
        F.wrk[f] = S(wrk[f], B, box);
        F.chi[f] = chi[f];
    }
    launch_fct_fused(F, nf, g_stream);
    check_launch("fct");
}

void g3m3_fcsmooth(float* px, const float* oldpx, const float* wrkpx, const int* nx,
                   const int* ny, const int* nz, const int* n_grids, const int* box,
                   const float* chipx, float* tempx, float* tempy, float* tempz) {
    const Box B = box_of(nx, ny, nz);
    float* out[1] = {px};
    const float* old[1] = {oldpx};
    const float* wrk[1] = {wrkpx};
    fct_impl(B, out, old, wrk, chipx, 1, box);
}

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
    const float* diferg, const float* xspac, const int* isotropic) {
    const Box B = box_of(nx, ny, nz);
    float* out[34] = {qrho, qpresx, qpresy, qpresz, qpresxy, qpresxz, qpresyz, qpx, qpy, qpz,
                      hrho, hpresx, hpresy, hpresz, hpresxy, hpresxz, hpresyz, hpx, hpy, hpz,
                      orho, opresx, opresy, opresz, opresxy, opresxz, opresyz, opx, opy, opz,
                      epres, bx, by, bz};
    const float* wrk[34] = {wrkqrho, wrkqpresx, wrkqpresy, wrkqpresz, wrkqpresxy, wrkqpresxz,
                            wrkqpresyz, wrkqpx, wrkqpy, wrkqpz, wrkhrho, wrkhpresx, wrkhpresy,
                            wrkhpresz, wrkhpresxy, wrkhpresxz, wrkhpresyz, wrkhpx, wrkhpy, wrkhpz,
                            wrkorho, wrkopresx, wrkopresy, wrkopresz, wrkopresxy, wrkopresxz,
                            wrkopresyz, wrkopx, wrkopy, wrkopz, wrkepres, wrkbx, wrkby, wrkbz};
    const float* old[34] = {oldqrho, oldqpresx, oldqpresy, oldqpresz, oldqpresxy, oldqpresxz,
                            oldqpresyz, oldqpx, oldqpy, oldqpz, oldhrho, oldhpresx, oldhpresy,
                            oldhpresz, oldhpresxy, oldhpresxz, oldhpresyz, oldhpx, oldhpy, oldhpz,
                            oldorho, oldopresx, oldopresy, oldopresz, oldopresxy, oldopresxz,
                            oldopresyz, oldopx, oldopy, oldopz, oldepres, oldbx, oldby, oldbz};
    float chi[34];
    for (int f = 0; f < 31; ++f) chi[f] = *difrho;
    for (int f = 31; f < 34; ++f) chi[f] = *diferg;
    fct_impl(B, out, old, wrk, chi, 34, box);
}

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
    const float* ti_te, const int* isotropic) {
    const Box B = box_of(nx, ny, nz);
    const int* box = n_grids;
    const float ofrac = *rho_frac * *o_conc;
    const float frac_h = *rmassh / *rmassq;
    const float frac_o = *rmasso / *rmassq;
    const float pq = 0.5f * *spress;
    const float ph = 0.5f * *spress * *rho_frac;
    const float po = 0.5f * *spress * ofrac;
    struct E {
        float* f;
        int kind;
        float val;
    } fs[34] = {
        {qrho, kRHO, *srho}, {qpx, kPX, *spx}, {qpy, kPY, *spy}, {qpz, kPZ, *spz},
        {qpresx, kPRES, pq}, {qpresy, kPRES, pq}, {qpresz, kPRES, pq},
        {hrho, kRHO, *srho * *rho_frac * frac_h}, {hpx, kPX, *spx * *rho_frac * frac_h},
        {hpy, kPY, *spy * *rho_frac * frac_h}, {hpz, kPZ, *spz * *rho_frac * frac_h},
        {hpresx, kPRES, ph}, {hpresy, kPRES, ph}, {hpresz, kPRES, ph},
        {orho, kRHO, *srho * ofrac * frac_o}, {opx, kPX, *spx * ofrac * frac_o},
        {opy, kPY, *spy * ofrac * frac_o}, {opz, kPZ, *spz * ofrac * frac_o},
        {opresx, kPRES, po}, {opresy, kPRES, po}, {opresz, kPRES, po},
        {epres, kEP, pq / *ti_te},
        {qpresxy, kOFF, 0.f}, {qpresxz, kOFF, 0.f}, {qpresyz, kOFF, 0.f},
        {hpresxy, kOFF, 0.f}, {hpresxz, kOFF, 0.f}, {hpresyz, kOFF, 0.f},
        {opresxy, kOFF, 0.f}, {opresxz, kOFF, 0.f}, {opresyz, kOFF, 0.f},
        {bx, kBX, *sbx_wind}, {by, kBY, *sby_wind}, {bz, kBZ, *sbz_wind}};
    OuterPtrs O{};
    for (int f = 0; f < 34; ++f) {
        O.f[f] = S(fs[f].f, B, box);
        O.kind[f] = fs[f].kind;
        O.val[f] = fs[f].val;
    }
    O.b0[0] = S(bx0, B, box);
    O.b0[1] = S(by0, B, box);
    O.b0[2] = S(bz0, B, box);
    launch_bndry_outer(B, O, 34, g_stream);
    check_launch("bndry_outer");
}

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
    const float* r_rot) {
    const Box B = box_of(nx, ny, nz);
    const int* box = mbndry;
    float* g[34] = {qrho, qpresx, qpresy, qpresz, qpx, qpy, qpz, qpresxy, qpresxz, qpresyz,
                    hrho, hpresx, hpresy, hpresz, hpx, hpy, hpz, hpresxy, hpresxz, hpresyz,
                    orho, opresx, opresy, opresz, opx, opy, opz, opresxy, opresxz, opresyz,
                    epres, bx, by, bz};
    InnerPtrs I{};
    for (int f = 0; f < 34; ++f) I.f[f] = S(g[f], B, box);
    const int b = *mbndry - 1;
    const int* ij[3] = {R(ijsrf), R(ijmid), R(ijzero)};
    const float* parm[3] = {R(parm_srf), R(parm_mid), R(parm_zero)};
    const int num[3] = {numsrf[b], nummid[b], numzero[b]};
    const float dx = (grd_xmax[b] - grd_xmin[b]) / (float)(*nx - 1.0f);
    const float dy = (grd_ymax[b] - grd_ymin[b]) / (float)(*ny - 1.0f);
    launch_bndry_inner(B, I, ij, parm, num, *mbndry, *mbndry, grd_xmin[b], grd_ymin[b], dx, dy,
                       *re_equiv, *v_rot, g_stream);
    check_launch("bndry_inner");
}

static FlankGeom flank_geom(int nx, int ny, int nz, int box, const float* gx0, const float* gx1,
                            const float* gy0, const float* gy1, const float* gz0,
                            const float* gz1) {
    const int mb = box, ms = box - 1;   // 0-based coarse, fine
    FlankGeom G;
    G.delx = (gx1[mb] - gx0[mb]) / (nx - 1.0f);
    G.dely = (gy1[mb] - gy0[mb]) / (ny - 1.0f);
    G.delz = (gz1[mb] - gz0[mb]) / (nz - 1.0f);
    G.sx = (gx1[ms] - gx0[ms]) / (nx - 1.0f);
    G.sy = (gy1[ms] - gy0[ms]) / (ny - 1.0f);
    G.sz = (gz1[ms] - gz0[ms]) / (nz - 1.0f);
    G.sxmin = gx0[ms];
    G.symin = gy0[ms];
    G.szmin = gz0[ms];
    G.bxmin = gx0[mb];
    G.bymin = gy0[mb];
    G.bzmin = gz0[mb];
    return G;
}

void g3m3_bndry_flanks_v(float* const* tgt, const float* const* cur, const float* const* old,
                         float* work, const int* nx, const int* ny, const int* nz,
                         const int* n_grids, const int* box, const double* t_old,
                         const double* t_new, const double* t, const float* grd_xmin,
                         const float* grd_xmax, const float* grd_ymin, const float* grd_ymax,
                         const float* grd_zmin, const float* grd_zmax) {
    const Box B = box_of(nx, ny, nz);
    const int mb = *box;   // 0-based index of box+1
    const float t1 = (float)((t_new[mb] - *t) / (t_new[mb] - t_old[mb]));
    const float t2 = 1.0f - t1;
    const int cbox = *box + 1;
    FlankPtrs P{};
    for (int f = 0; f < 34; ++f) {
        P.tgt[f] = S(tgt[f], B, box);
        P.cur[f] = S(cur[f], B, &cbox);
        P.old[f] = S(old[f], B, &cbox);
    }
    const FlankGeom G = flank_geom(*nx, *ny, *nz, *box, grd_xmin, grd_xmax, grd_ymin, grd_ymax,
                                   grd_zmin, grd_zmax);
    launch_flanks(B, P, 34, G, t1, t2, g_stream);
    check_launch("bndry_flanks");
}

void g3m3_flanks(float* wrkrho, const float* rho, const float* oldrho, const int* nx,
                 const int* ny, const int* nz, const int* n_grids, const int* box,
                 const double* t_new, const double* t_old, const double* t, float* work,
                 const float* grd_xmin, const float* grd_xmax, const float* grd_ymin,
                 const float* grd_ymax, const float* grd_zmin, const float* grd_zmax) {
    const Box B = box_of(nx, ny, nz);
    const int mb = *box;
    const float t1 = (float)((t_new[mb] - *t) / (t_new[mb] - t_old[mb]));
    const float t2 = 1.0f - t1;
    const int cbox = *box + 1;
    FlankPtrs P{};
    P.tgt[0] = S(wrkrho, B, box);
    P.cur[0] = S(rho, B, &cbox);
    P.old[0] = S(oldrho, B, &cbox);
    const FlankGeom G = flank_geom(*nx, *ny, *nz, *box, grd_xmin, grd_xmax, grd_ymin, grd_ymax,
                                   grd_zmin, grd_zmax);
    launch_flanks(B, P, 1, G, t1, t2, g_stream);
    check_launch("flanks");
}

static void corer_impl(float* const* fields, int nf, const int* nx, const int* ny, const int* nz,
                       const int* box, const float* gx0, const float* gx1, const float* gy0,
                       const float* gy1, const float* gz0, const float* gz1) {
    const Box B = box_of(nx, ny, nz);
    const int mb = *box, ms = *box - 1;
    const float sx = (gx1[ms] - gx0[ms]) / (*nx - 1.0f);
    const float dx = (gx1[mb] - gx0[mb]) / (*nx - 1.0f);
    const float dy = (gy1[mb] - gy0[mb]) / (*ny - 1.0f);
    const float dz = (gz1[mb] - gz0[mb]) / (*nz - 1.0f);
    const int nskip = (int)(dx / sx);
    const int nbx1 = (int)(1.0f + (gx0[ms] - gx0[mb]) / dx + 0.5f);
    const int nby1 = (int)(1.0f + (gy0[ms] - gy0[mb]) / dy + 0.5f);
    const int nbz1 = (int)(1.0f + (gz0[ms] - gz0[mb]) / dz + 0.5f);
    const int cbox = *box + 1;
    FieldPtrs fine{}, coarse{};
    for (int f = 0; f < nf; ++f) {
        fine.p[f] = S(fields[f], B, box);
        coarse.p[f] = S(fields[f], B, &cbox);
    }
    launch_corer(B, fine, coarse, nf, nskip, nbx1, nby1, nbz1, g_stream);
    check_launch("corer");
}

void g3m3_corer(float* rho, const int* nx, const int* ny, const int* nz, const int* n_grids,
                const int* box, const float* grd_xmin, const float* grd_xmax,
                const float* grd_ymin, const float* grd_ymax, const float* grd_zmin,
                const float* grd_zmax) {
    float* f[1] = {rho};
    corer_impl(f, 1, nx, ny, nz, box, grd_xmin, grd_xmax, grd_ymin, grd_ymax, grd_zmin, grd_zmax);
}

void g3m3_bndry_corer_v(float* const* fields, const int* nx, const int* ny, const int* nz,
                        const int* n_grids, const int* box, const float* grd_xmin,
                        const float* grd_xmax, const float* grd_ymin, const float* grd_ymax,
                        const float* grd_zmin, const float* grd_zmax) {
    corer_impl(fields, 34, nx, ny, nz, box, grd_xmin, grd_xmax, grd_ymin, grd_ymax, grd_zmin,
               grd_zmax);
}

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
    const float* grd_zmax) {
    float* const tgt[34] = {wrkqrho, wrkqpx, wrkqpy, wrkqpz, wrkqpresx, wrkqpresy, wrkqpresz,
        wrkqpresxy, wrkqpresxz, wrkqpresyz, wrkhrho, wrkhpx, wrkhpy, wrkhpz, wrkhpresx,
        wrkhpresy, wrkhpresz, wrkhpresxy, wrkhpresxz, wrkhpresyz, wrkorho, wrkopx, wrkopy,
        wrkopz, wrkopresx, wrkopresy, wrkopresz, wrkopresxy, wrkopresxz, wrkopresyz, wrkepres,
        wrkbx, wrkby, wrkbz};
    const float* const cur[34] = {qrho, qpx, qpy, qpz, qpresx, qpresy, qpresz, qpresxy, qpresxz,
        qpresyz, hrho, hpx, hpy, hpz, hpresx, hpresy, hpresz, hpresxy, hpresxz, hpresyz, orho,
        opx, opy, opz, opresx, opresy, opresz, opresxy, opresxz, opresyz, epres, bx, by, bz};
    const float* const old[34] = {oldqrho, oldqpx, oldqpy, oldqpz, oldqpresx, oldqpresy,
        oldqpresz, oldqpresxy, oldqpresxz, oldqpresyz, oldhrho, oldhpx, oldhpy, oldhpz,
        oldhpresx, oldhpresy, oldhpresz, oldhpresxy, oldhpresxz, oldhpresyz, oldorho, oldopx,
        oldopy, oldopz, oldopresx, oldopresy, oldopresz, oldopresxy, oldopresxz, oldopresyz,
        oldepres, oldbx, oldby, oldbz};
    g3m3_bndry_flanks_v(tgt, cur, old, work, nx, ny, nz, n_grids, box, t_old, t_new, t,
                        grd_xmin, grd_xmax, grd_ymin, grd_ymax, grd_zmin, grd_zmax);
}

void g3m3_bndry_corer(
    float* qrho, float* qpresx, float* qpresy, float* qpresz, float* qpx, float* qpy, float* qpz,
    float* hrho, float* hpresx, float* hpresy, float* hpresz, float* hpx, float* hpy, float* hpz,
    float* orho, float* opresx, float* opresy, float* opresz, float* opx, float* opy, float* opz,
    float* epres, float* bx, float* by, float* bz, float* qpresxy, float* qpresxz,
    float* qpresyz, float* hpresxy, float* hpresxz, float* hpresyz, float* opresxy,
    float* opresxz, float* opresyz, const int* nx, const int* ny, const int* nz,
    const int* n_grids, const int* box, const float* grd_xmin, const float* grd_xmax,
    const float* grd_ymin, const float* grd_ymax, const float* grd_zmin, const float* grd_zmax) {
    float* const fields[34] = {qrho, qpresx, qpresy, qpresz, qpx, qpy, qpz, hrho, hpresx, hpresy,
        hpresz, hpx, hpy, hpz, orho, opresx, opresy, opresz, opx, opy, opz, epres, bx, by, bz,
        qpresxy, qpresxz, qpresyz, hpresxy, hpresxz, hpresyz, opresxy, opresxz, opresyz};
    g3m3_bndry_corer_v(fields, nx, ny, nz, n_grids, box, grd_xmin, grd_xmax, grd_ymin,
                       grd_ymax, grd_zmin, grd_zmax);
}

void g3m3_flanks_synced(float* rho, const int* nx, const int* ny, const int* nz,
                        const int* n_grids, const int* box, const float* gx0, const float* gx1,
                        const float* gy0, const float* gy1, const float* gz0, const float* gz1) {
    const Box B = box_of(nx, ny, nz);
    const int mb = *box, ms = *box - 1;
    const FlankGeom G = flank_geom(*nx, *ny, *nz, *box, gx0, gx1, gy0, gy1, gz0, gz1);
    const int kbi[6] = {(int)(1.0f + (gz1[ms] - gz0[mb]) / G.delz),
                        (int)(1.0f + (gz0[ms] - gz0[mb]) / G.delz),
                        (int)(1.0f + (gy1[ms] - gy0[mb]) / G.dely),
                        (int)(1.0f + (gy0[ms] - gy0[mb]) / G.dely),
                        (int)(1.0f + (gx1[ms] - gx0[mb]) / G.delx),
                        (int)(1.0f + (gx0[ms] - gx0[mb]) / G.delx)};
    const int cbox = *box + 1;
    FieldPtrs fine{}, coarse{};
    fine.p[0] = S(rho, B, box);
    coarse.p[0] = S(rho, B, &cbox);
    launch_flanks_synced(B, fine, coarse, 1, G, kbi, g_stream);
    check_launch("flanks_synced");
}

void g3m3_banded_tpose(const int* nbands, const int* npts, const int* divB_band_pos,
                       const float* banded_matrix, const float* x_vec, float* tp_result) {
    if (*nbands != 7) {
        latch(4, "banded_tpose: only the 7-band layout is supported");
        return;
    }
    int* pos = (int*)g_tmp.get(64);
    if (!pos) return;
    if (!hip_ok(hipMemcpyAsync(pos, divB_band_pos, 7 * sizeof(int), hipMemcpyHostToDevice,
                               g_stream),
                "banded_tpose"))
        return;
    launch_banded_matvec(*npts, pos, R(banded_matrix), R(x_vec), R(tp_result), g_stream);
    check_launch("banded_tpose");
}

int g3m3_divb_correct_ex(const int* nx, const int* ny, const int* nz, const int* n_grids,
                         const int* box, const float* xspac, const float* divB_errlim,
                         float* bx, float* by, float* bz, const int* mode,
                         const long long* max_iters, long long* iters, double* residual2) {
    const Box B = box_of(nx, ny, nz);
    float* w = (float*)g_divb_f.get(divb_work_floats(B.V()) * sizeof(float));
    double* d = (double*)g_divb_d.get(divb_work_doubles(B.V()) * sizeof(double));
    if (!w || !d) return -1;
    long it = 0;
    double r2 = 0.0;
    const int st = divb_correct(B, S(bx, B, box), S(by, B, box), S(bz, B, box), xspac[*box - 1],
                                *divB_errlim, mode ? *mode : kDivbRef,
                                max_iters ? (long)*max_iters : 0, w, d, &it, &r2, g_stream);
    check_launch("divb_correct");
    if (iters) *iters = it;
    if (residual2) *residual2 = r2;
    return st;
}

void g3m3_divb_correct(const int* nx, const int* ny, const int* nz, const int* n_grids,
                       const int* box, const int* nbands, const int* divB_band_pos,
                       const float* xspac, const float* divB_errlim, float* bx, float* by,
                       float* bz) {
    const int mode = kDivbRef;
    const long long cap = 0;
    g3m3_divb_correct_ex(nx, ny, nz, n_grids, box, xspac, divB_errlim, bx, by, bz, &mode, &cap,
                         nullptr, nullptr);
}

void g3m3_trilin_interp(const float* sxyz, const float* gridpts, const float* qty,
                        float* scdata) {
    float* d = (float*)g_tmp.get(256);
    if (!d) return;
    float h[3 + 6 + 8];
    std::memcpy(h, sxyz, 3 * sizeof(float));
    std::memcpy(h + 3, gridpts, 6 * sizeof(float));
    std::memcpy(h + 9, qty, 8 * sizeof(float));
    if (!hip_ok(hipMemcpyAsync(d, h, sizeof(h), hipMemcpyHostToDevice, g_stream),
                "trilin_interp"))
        return;
    launch_trilin(1, d, d + 3, d + 9, d + 17, g_stream);
    check_launch("trilin_interp");
    if (hip_ok(hipMemcpyAsync(scdata, d + 17, sizeof(float), hipMemcpyDeviceToHost, g_stream),
               "trilin_interp"))
        hip_ok(hipStreamSynchronize(g_stream), "trilin_interp");
}

void g3m3_sparse_solver(const int* nbands, const int* npts, const int* divB_band_pos,
                        const float* divB_errlim, const int* recurse, const int* redo_lim,
                        const float* divB, float* soln_vec, float* residual2,
                        const float* banded_matrix) {
    const int n = *npts;
    float* w = (float*)g_divb_f.get(sparse_work_floats(n) * sizeof(float));
    double* d = (double*)g_divb_d.get(divb_work_doubles(n) * sizeof(double));
    if (!w || !d) return;
    double r2 = 0.0;
    const int st = sparse_solve(n, *nbands, divB_band_pos, R(banded_matrix), R(divB),
                                R(soln_vec), *divB_errlim, *recurse != 0, *redo_lim, kDivbRef, 0,
                                w, d, nullptr, &r2, g_stream);
    if (st < 0) latch(4, "sparse_solver: nbands must be 1..8");
    check_launch("sparse_solver");
    if (residual2) *residual2 = (float)r2;
}

void g3m3_fnd_pres(const float* px, const float* py, const float* pz, const float* pxy,
                   const float* pxz, const float* pyz, float* p_para, float* p_perp,
                   float* p_cross, const float* vvx, const float* vvy, const float* vvz,
                   const float* bxt, const float* byt, const float* bzt, const int* nx,
                   const int* ny, const int* nz, const int* n_grids, const int* box) {
    (void)pxy;
    (void)pxz;
    (void)pyz;
    (void)n_grids;
    const Box B = box_of(nx, ny, nz);
    const float* p[3] = {S(px, B, box), S(py, B, box), S(pz, B, box)};
    const float* v[3] = {R(vvx), R(vvy), R(vvz)};
    const float* b[3] = {R(bxt), R(byt), R(bzt)};
    launch_fnd_pres(B.V(), p, v, b, S(p_para, B, box), S(p_perp, B, box), S(p_cross, B, box),
                    g_stream);
    check_launch("fnd_pres");
}

Buf g_graph;

// code/sub_write_graphing_data.f90:348-760: the whole-box calcs (calcur,
// total B, fnd_evel, BandE with the resistivity list term; species
// velocities and fnd_pres inside the cut kernel) and the tabulated
// quantities on the three cut planes of one box
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
    const float* v_equiv, const float* rho_equiv, const float* planet_rad, float* out) {
    const Box B = box_of(nx, ny, nz);
    const int b = *box - 1;
    // the cut planes (:386-395), 1-based; box 5 of the reference is offset
    const int x_adj[5] = {0, 0, 0, 0, -15};
    const int cv[3] = {(int)(*limit / 2.0f + 1.0f), (int)(*limit + 1.0f),
                       (int)(*limit + 1.0f + (float)(*box <= 5 ? x_adj[b] : 0))};
    if (cv[0] < 1 || cv[0] > B.nz || cv[1] < 1 || cv[1] > B.ny || cv[2] < 1 || cv[2] > B.nx) {
        latch(4, "graphing_cuts: a cut plane lies outside the box");
        return;
    }
    const size_t V = (size_t)B.V();
    float* w = (float*)g_graph.get(16 * V * sizeof(float));
    if (!w) return;
    float *cur = w, *bt = w + 3 * V, *btot = w + 6 * V, *ev = w + 7 * V, *tv = w + 10 * V,
          *e = w + 13 * V;
    const float rx = xspac[b];
    g3m3_calcur(bx, by, bz, nx, ny, nz, n_grids, box, cur, cur + V, cur + 2 * V, &rx, &rx, &rx);
    g3m3_bfield_total(bx, by, bz, bx0, by0, bz0, bt, bt + V, bt + 2 * V, btot, nx, ny, nz,
                      n_grids, box, smallbit);
    g3m3_fnd_evel(qpx, qpy, qpz, qrho, hpx, hpy, hpz, hrho, opx, opy, opz, orho, cur, cur + V,
                  cur + 2 * V, ev, ev + V, ev + 2 * V, tv, tv + V, tv + 2 * V, nx, ny, nz,
                  n_grids, box, rmassq, rmassh, rmasso, reynolds);
    g3m3_bande(e, e + V, e + 2 * V, bt, bt + V, bt + 2 * V, cur, cur + V, cur + 2 * V, ev,
               ev + V, ev + 2 * V, btot, epres, qrho, hrho, orho, resistive, resist, reynolds, nx,
               ny, nz, n_grids, box, rmassq, rmassh, rmasso, ijmid, nummid, ijzero, mbndry,
               numzero, mmid, mzero, &rx, &rx, &rx);
    GraphCuts G{};
    G.B = B;
    for (int a = 0; a < 3; ++a) G.cut[a] = cv[a] - 1;
    const float* sp[3][7] = {{qrho, qpx, qpy, qpz, qpresx, qpresy, qpresz},
                             {hrho, hpx, hpy, hpz, hpresx, hpresy, hpresz},
                             {orho, opx, opy, opz, opresx, opresy, opresz}};
    for (int q = 0; q < 3; ++q)
        for (int f = 0; f < 7; ++f) G.sp[q][f] = S(sp[q][f], B, box);
    G.epres = S(epres, B, box);
    const float* bb[3] = {bx, by, bz};
    const float* bb0[3] = {bx0, by0, bz0};
    for (int a = 0; a < 3; ++a) {
        G.b[a] = S(bb[a], B, box);
        G.b0[a] = S(bb0[a], B, box);
        G.cur[a] = cur + a * V;
        G.tv[a] = tv + a * V;
        G.e[a] = e + a * V;
    }
    G.rmass[0] = *rmassq;
    G.rmass[1] = *rmassh;
    G.rmass[2] = *rmasso;
    G.smallbit = *smallbit;
    G.v_equiv = *v_equiv;
    G.b_equiv = *b_equiv;
    G.rho_equiv = *rho_equiv;
    G.re_equiv = *re_equiv;
    // the unit factors (:205-209) and constants (code/mod_astrometry.f90:20-24),
    // single precision in the reference's order
    const float pi = 3.14159274f;
    const float mu0 = 4.e-7f * pi;
    const float m_prot = 1.67262e-27f, q_elec = 1.60218e-19f;
    const float vv = *v_equiv * 1.e3f;
    G.temp_equiv = m_prot * (vv * vv) / q_elec / 1.e3f;
    G.pres_equiv = *rho_equiv * 1.e6f * m_prot * (vv * vv);
    G.cur_equiv = *b_equiv / mu0 / (*planet_rad * 1.e3f * *re_equiv);
    G.e_equiv = *b_equiv * *v_equiv * 1.e3f / 1.e6f;
    G.mu0 = mu0;
    G.m_prot = m_prot;
    G.xmin = grid_minvals[3 * b];
    G.ymin = grid_minvals[3 * b + 1];
    G.zmin = grid_minvals[3 * b + 2];
    G.xs = xspac[b];
    launch_graph_cuts(G, R(out), g_stream);
    check_launch("graphing_cuts");
}

// ---- initialisation ------------------------------------------------------------

static RotationGeom rotation_geom(const float* rot_angle, const float* xdip, const float* ydip,
                                  const float* zdip, const float* sin_tilt,
                                  const float* cos_tilt, const float* b0) {
    RotationGeom g;
    g.xdip = *xdip;
    g.ydip = *ydip;
    g.zdip = *zdip;
    g.sin_rot = std::sin(*rot_angle);   // the C library's sinf / cosf
    g.cos_rot = std::cos(*rot_angle);
    g.sin_tilt = *sin_tilt;
    g.cos_tilt = *cos_tilt;
    g.b0 = *b0;
    return g;
}

// code/sub_mak_dip_grd.f90:22-24 and code/3d_main.f90:1247-1252
static InitGeom init_geom(int nx, int ny, int nz, float x0, float x1, float y0, float y1,
                          float z0, float z1) {
    InitGeom G;
    G.xmin = x0;
    G.ymin = y0;
    G.zmin = z0;
    G.dx = (x1 - x0) / (nx - 1.0f);
    G.dy = (y1 - y0) / (ny - 1.0f);
    G.dz = (z1 - z0) / (nz - 1.0f);
    return G;
}

void g3m3_dipole(float* abx, float* aby, float* abz, const float* ax, const float* ay,
                 const float* az, const float* rot_angle, const float* xdip, const float* ydip,
                 const float* zdip, const float* sin_tilt, const float* cos_tilt, const float* b0) {
    const RotationGeom rot = rotation_geom(rot_angle, xdip, ydip, zdip, sin_tilt, cos_tilt, b0);
    dipole_host(*ax, *ay, *az, rot, *abx, *aby, *abz);
}

void g3m3_mak_dip_grd(float* bx0, float* by0, float* bz0, const int* nx, const int* ny,
                      const int* nz, const int* n_grids, const int* mbndry, const int* box,
                      const int* ijzero, const int* numzero, const int* mzero,
                      const float* r_inner, const float* grd_xmin, const float* grd_xmax,
                      const float* grd_ymin, const float* grd_ymax, const float* grd_zmin,
                      const float* grd_zmax, const float* rot_angle, const float* xdip,
                      const float* ydip, const float* zdip, const float* sin_tilt,
                      const float* cos_tilt, const float* b0) {
    (void)n_grids;
    const Box B = box_of(nx, ny, nz);
    const int b = *box - 1;
    const InitGeom G = init_geom(*nx, *ny, *nz, grd_xmin[b], grd_xmax[b], grd_ymin[b],
                                 grd_ymax[b], grd_zmin[b], grd_zmax[b]);
    const RotationGeom rot = rotation_geom(rot_angle, xdip, ydip, zdip, sin_tilt, cos_tilt, b0);
    int nzero = 0;
    if (*box <= *mbndry) {
        nzero = numzero[b];
        if (nzero < 0 || nzero > *mzero) {
            latch(4, "mak_dip_grd: numzero(box) outside 0..mzero");
            return;
        }
    }
    launch_mak_dip_grd(B, S(bx0, B, box), S(by0, B, box), S(bz0, B, box), G, rot, *r_inner,
                       R(ijzero), *mbndry, *box, nzero, g_stream);
    check_launch("mak_dip_grd");
}

void g3m3_set_resist(float* rst, const int* nx, const int* ny, const int* nz, const int* mbndry,
                     const float* resist, const int* ijzero, const int* numzero,
                     const int* ijmid, const int* nummid, const int* ijsrf, const int* numsrf,
                     const int* msrf, const int* mmid, const int* mzero, const float* b0) {
    const Box B = box_of(nx, ny, nz);
    const int mb = *mbndry;
    if (mb < 1) {
        latch(4, "set_resist: mbndry < 1");
        return;
    }
    const int cnt[3] = {numsrf[mb - 1], nummid[mb - 1], numzero[mb - 1]};
    const int cap[3] = {*msrf, *mmid, *mzero};
    for (int q = 0; q < 3; ++q)
        if (cnt[q] < 0 || cnt[q] > cap[q]) {
            latch(4, "set_resist: a point count exceeds its list's size");
            return;
        }
    const int most = std::max(cnt[0], std::max(cnt[1], cnt[2]));
    launch_set_resist(B, R(rst), mb, *resist, *b0, R(ijzero), R(ijmid), R(ijsrf), nullptr,
                      cnt, most, g_stream);
    check_launch("set_resist");
}

Buf g_init_cls, g_init_cnt;

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
    const float* sin_tilt, const float* cos_tilt, const float* b0) {
    const int ng = *n_grids, mb = *mbndry;
    if (ng < 1 || ng > kMaxInitGrids || mb < 1 || mb > ng) {
        latch(4, "init_plasma: n_grids or mbndry out of range");
        return;
    }
    InitPlasma P{};
    P.B = box_of(nx, ny, nz);
    float* f[34] = {qrho, qpresx, qpresy, qpresz, qpx, qpy, qpz, qpresxy, qpresxz, qpresyz,
                    hrho, hpresx, hpresy, hpresz, hpx, hpy, hpz, hpresxy, hpresxz, hpresyz,
                    orho, opresx, opresy, opresz, opx, opy, opz, opresxy, opresxz, opresyz,
                    epres, bx, by, bz};
    for (int q = 0; q < 34; ++q) P.f[q] = R(f[q]);
    P.b0f[0] = R(bx0);
    P.b0f[1] = R(by0);
    P.b0f[2] = R(bz0);
    const size_t V = (size_t)P.B.V();
    P.cls = (unsigned char*)g_init_cls.get(V * mb);
    P.counts = (int*)g_init_cnt.get(sizeof(int) * (3 * mb + 1));
    if (!P.cls || !P.counts) return;
    if (!hip_ok(hipMemsetAsync(P.counts + 3 * mb, 0, sizeof(int), g_stream), "init_plasma"))
        return;
    P.ijsrf = R(ijsrf);
    P.ijmid = R(ijmid);
    P.ijzero = R(ijzero);
    P.parm_srf = R(parm_srf);
    P.parm_mid = R(parm_mid);
    P.parm_zero = R(parm_zero);
    P.mbndry = mb;
    P.msrf = *msrf;
    P.mmid = *mmid;
    P.mzero = *mzero;
    for (int b = 0; b < ng; ++b)
        P.geom[b] = init_geom(*nx, *ny, *nz, grid_minvals[3 * b], grid_maxvals[3 * b],
                              grid_minvals[3 * b + 1], grid_maxvals[3 * b + 1],
                              grid_minvals[3 * b + 2], grid_maxvals[3 * b + 2]);
    P.rot = rotation_geom(rot_angle, xdip, ydip, zdip, sin_tilt, cos_tilt, b0);
    P.re_equiv = *re_equiv;
    P.r_rot = *r_rot;
    P.v_rot = *v_rot;
    P.zero_bndry = *zero_bndry;
    P.mid_bndry = *mid_bndry;
    P.srf_bndry = *srf_bndry;
    P.r_inner = *r_inner;
    P.alpha_e = *alpha_e;
    P.erho = *erho;
    P.d_min = *d_min;
    P.cs_inner = *cs_inner;
    P.rmassq = *rmassq;
    P.rmassh = *rmassh;
    P.rmasso = *rmasso;
    P.ti_te = *ti_te;
    P.srho = *srho;
    P.spx = *spx;
    P.spy = *spy;
    P.spz = *spz;
    P.spress = *spress;
    P.cs_wind = *cs_wind;
    P.gamma = *gamma;
    P.rho_frac = *rho_frac;
    P.wind_bnd = *r_rot / *re_equiv;   // code/3d_main.f90:1467
    launch_init_plasma(P, ng, g_stream);
    // code/3d_main.f90:1631-1633
    const int most = std::max(*msrf, std::max(*mmid, *mzero));
    launch_set_resist(P.B, R(resistive), mb, *resist, 1.0f, P.ijzero, P.ijmid, P.ijsrf,
                      P.counts + 3 * (mb - 1), nullptr, most, g_stream);
    check_launch("init_plasma");
    std::vector<int> h(3 * mb + 1, 0);
    if (hipMemcpyAsync(h.data(), P.counts, sizeof(int) * h.size(), hipMemcpyDeviceToHost,
                       g_stream) != hipSuccess ||
        hipStreamSynchronize(g_stream) != hipSuccess) {
        latch(5, "init_plasma: reading the point counts failed");
        return;
    }
    for (int b = 0; b < mb; ++b) {
        numsrf[b] = h[3 * b];
        nummid[b] = h[3 * b + 1];
        numzero[b] = h[3 * b + 2];
    }
    if (h[3 * mb]) latch(4, "init_plasma: more boundary points than msrf/mmid/mzero; lists cut");
}

// code/sub_flux_counter.f90:5-158: the wall sums on the device, the
// in/out bookkeeping on the host in the reference's order
void g3m3_flux_counter(const float* qpx, const float* qpy, const float* qpz, const float* hpx,
                       const float* hpy, const float* hpz, const float* opx, const float* opy,
                       const float* opz, const int* nx, const int* ny, const int* nz,
                       const int* n_grids, const int* box, const float* rmassq,
                       const float* rmassh, const float* rmasso, const float* grid_spacing,
                       float* qflux_in, float* qflux_out, float* hflux_in, float* hflux_out,
                       float* oflux_in, float* oflux_out) {
    (void)n_grids;
    const Box B = box_of(nx, ny, nz);
    FieldPtrs P{};
    const float* mom[9] = {qpx, qpy, qpz, hpx, hpy, hpz, opx, opy, opz};
    for (int q = 0; q < 9; ++q) P.p[q] = const_cast<float*>(S(mom[q], B, box));
    float* d = (float*)g_tmp.get(256);
    if (!d) return;
    launch_flux_walls(B, P, d, g_stream);
    check_launch("flux_counter");
    float h[18];
    if (!hip_ok(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, g_stream),
                "flux_counter") ||
        !hip_ok(hipStreamSynchronize(g_stream), "flux_counter"))
        return;
    const float gs = *grid_spacing;
    const float rm[3] = {*rmassq, *rmassh, *rmasso};
    float in[3] = {0.0f, 0.0f, 0.0f}, out[3] = {0.0f, 0.0f, 0.0f};
    for (int axis = 0; axis < 3; ++axis) {
        const float* f1 = h + 6 * axis;       // q, h, o on the low wall
        const float* f2 = h + 6 * axis + 3;   // on the high wall
        // the h species' flux decides the direction for all three (:48-67)
        for (int s = 0; s < 3; ++s) {
            if (f1[1] > 0) in[s] = in[s] + f1[s] * gs / rm[s];
            else out[s] = out[s] - f1[s] * gs / rm[s];
        }
        for (int s = 0; s < 3; ++s) {
            if (f2[1] < 0) in[s] = in[s] - f2[s] * gs / rm[s];
            else out[s] = out[s] + f2[s] * gs / rm[s];
        }
    }
    *qflux_in = in[0];
    *qflux_out = out[0];
    *hflux_in = in[1];
    *hflux_out = out[1];
    *oflux_in = in[2];
    *oflux_out = out[2];
}

Buf g_torus;

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
    float* tot_q, float* tot_h, float* tot_o) {
    (void)n_grids;
    const Box B = box_of(nx, ny, nz);
    const int box1 = 1;
    TorusParams T{};
    T.B = B;
    float* f[30] = {qrho, qpresx, qpresy, qpresz, qpx, qpy, qpz, qpresxy, qpresxz, qpresyz,
                    hrho, hpresx, hpresy, hpresz, hpx, hpy, hpz, hpresxy, hpresxz, hpresyz,
                    orho, opresx, opresy, opresz, opx, opy, opz, opresxy, opresxz, opresyz};
    for (int sp = 0; sp < 3; ++sp)
        for (int m = 0; m < 10; ++m) T.sp.f[sp][m] = S(f[10 * sp + m], B, &box1);
    T.epres = S(epres, B, &box1);
    T.bx = S(bx, B, &box1);
    T.by = S(by, B, &box1);
    T.bz = S(bz, B, &box1);
    T.bx0 = S(bx0, B, &box1);
    T.by0 = S(by0, B, &box1);
    T.bz0 = S(bz0, B, &box1);
    float* scratch = (float*)g_torus.get(sizeof(float) * (3 * (size_t)B.V() + 64));
    if (!scratch) return;
    T.contrib = scratch + 64;
    // code/3d_main.f90:3155-3158
    T.dx = (grid_maxvals[0] - grid_minvals[0]) / (*nx - 1.0f);
    T.dy = (grid_maxvals[1] - grid_minvals[1]) / (*ny - 1.0f);
    T.dz = (grid_maxvals[2] - grid_minvals[2]) / (*nz - 1.0f);
    T.xmin = grid_minvals[0];
    T.ymin = grid_minvals[1];
    T.zmin = grid_minvals[2];
    T.xdip = *xdip;
    T.ydip = *ydip;
    T.zdip = *zdip;
    T.re_equiv = *re_equiv;
    T.v_rot = *v_rot;
    T.torus_inj_dist = *torus_inj_dist;
    T.torus_rad = *torus_rad;
    T.srf_bndry = *srf_bndry;
    T.denq = *denq_torus;
    T.denh = *denh_torus;
    T.deno = *deno_torus;
    T.rmassq = *rmassq;
    T.rmassh = *rmassh;
    T.rmasso = *rmasso;
    T.t_torus = *t_torus;
    T.aniso_factor = *aniso_factor;
    T.reduct = *reduct;
    launch_torus(T, scratch, g_stream);
    check_launch("torus_inject");
    float h[3];
    if (!hip_ok(hipMemcpyAsync(h, scratch, sizeof(h), hipMemcpyDeviceToHost, g_stream),
                "torus_inject") ||
        !hip_ok(hipStreamSynchronize(g_stream), "torus_inject"))
        return;
    *tot_q = h[0];
    *tot_h = h[1];
    *tot_o = h[2];
}

static bool craft_grid(CraftGrid& G, const int* n_grids, const float* grid_minvals,
                       const float* grid_maxvals, const float* xspac, const float* re_equiv) {
    if (*n_grids < 1 || *n_grids > kCraftMaxGrids) {
        latch(4, "craft sampling: n_grids out of range");
        return false;
    }
    G.n_grids = *n_grids;
    for (int b = 0; b < G.n_grids; ++b) {
        for (int a = 0; a < 3; ++a) {
            G.mn[b][a] = grid_minvals[a + 3 * b];
            G.mx[b][a] = grid_maxvals[a + 3 * b];
        }
        G.xspac[b] = xspac[b];
    }
    G.re_equiv = *re_equiv;
    return true;
}

void g3m3_findgrid(const float* xyz, const int* n_grids, const float* grid_minvals,
                   const float* grid_maxvals, const float* xspac, int* gridpt,
                   const float* re_equiv, const char* cname, size_t cname_len) {
    CraftGrid G;
    if (!craft_grid(G, n_grids, grid_minvals, grid_maxvals, xspac, re_equiv)) return;
    if (!craft_findgrid(G, xyz, gridpt)) {
        std::string name = cname ? std::string(cname, cname_len) : std::string();
        printf(" Craft outside of grid limits. Name, xyz: %s %g %g %g\n", name.c_str(),
               (double)xyz[0], (double)xyz[1], (double)xyz[2]);
    }
}

void g3m3_sample_fields(const int* npts, const float* sxyz, const int* n_grids,
                        const float* grid_minvals, const float* grid_maxvals, const float* xspac,
                        const float* re_equiv, const int* nx, const int* ny, const int* nz,
                        const int* nq, const float* const* fields, float* scdata, int* gridpt) {
    CraftGrid G;
    if (!craft_grid(G, n_grids, grid_minvals, grid_maxvals, xspac, re_equiv)) return;
    if (*nq < 0 || *nq > kMaxFields) {
        latch(4, "sample_fields: nq out of range");
        return;
    }
    FieldPtrs F{};
    for (int q = 0; q < *nq; ++q) F.p[q] = const_cast<float*>(R(fields[q]));
    // positions are only read here (no trajectories)
    launch_sample(*npts, const_cast<float*>(R(sxyz)), nullptr, nullptr, 0.0f, G, *nx, *ny, *nz,
                  *nq, F, R(scdata), R(gridpt), g_stream);
    check_launch("sample_fields");
}

void g3m3_sample_fields_adjoint(const int* npts, const float* sxyz, const int* n_grids,
                                const float* grid_minvals, const float* grid_maxvals,
                                const float* xspac, const float* re_equiv, const int* nx,
                                const int* ny, const int* nz, const int* nq,
                                float* const* fields_bar, const float* scdata_bar) {
    CraftGrid G;
    if (!craft_grid(G, n_grids, grid_minvals, grid_maxvals, xspac, re_equiv)) return;
    if (*nq < 0 || *nq > kMaxFields) {
        latch(4, "sample_fields_adjoint: nq out of range");
        return;
    }
    FieldPtrs F{};
    for (int q = 0; q < *nq; ++q) F.p[q] = R(fields_bar[q]);
    launch_adj_sample(*npts, R(sxyz), G, *nx, *ny, *nz, *nq, F, R(scdata_bar), g_stream);
    check_launch("sample_fields_adjoint");
}

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
                           const float* grd_zmin, const float* grd_zmax, float* work) {
    IonAdj A{};
    A.B = box_of(nx, ny, nz);
    for (int f = 0; f < 10; ++f) {
        A.src[f] = R(src[f]);
        A.ob[f] = R(out_bar[f]);
        A.sb[f] = R(src_bar[f]);
        A.oldb[f] = R(old_bar[f]);
    }
    const float* bs[3] = {bsx, bsy, bsz};
    const float* e[3] = {efldx, efldy, efldz};
    const float* tv[3] = {tvx, tvy, tvz};
    float* bsb[3] = {bsx_bar, bsy_bar, bsz_bar};
    float* eb[3] = {efldx_bar, efldy_bar, efldz_bar};
    float* tvb[3] = {tvx_bar, tvy_bar, tvz_bar};
    for (int a = 0; a < 3; ++a) {
        A.bs[a] = R(bs[a]);
        A.e[a] = R(e[a]);
        A.tv[a] = R(tv[a]);
        A.bsb[a] = R(bsb[a]);
        A.eb[a] = R(eb[a]);
        A.tvb[a] = R(tvb[a]);
    }
    A.btot = R(btot);
    A.btotb = R(btot_bar);
    A.scr = R(work);
    A.rmass = *rmass;
    A.delt = *delt;
    A.grav = *grav;
    A.re_equiv = *re_equiv;
    A.reynolds = *reynolds;
    A.xmin = *grd_xmin;
    A.ymin = *grd_ymin;
    A.zmin = *grd_zmin;
    A.ddx = (*grd_xmax - *grd_xmin) / (float)(*nx - 1);
    A.ddy = (*grd_ymax - *grd_ymin) / (float)(*ny - 1);
    A.ddz = (*grd_zmax - *grd_zmin) / (float)(*nz - 1);
    A.dxt = 2.0f * A.ddx;
    A.dyt = 2.0f * A.ddy;
    A.dzt = 2.0f * A.ddz;
    A.bmax = 0.20f * A.rmass / A.delt / A.reynolds;
    launch_adj_push_ion(A, g_stream);
    check_launch("push_ion_adjoint");
}

void g3m3_lap_plasma_adjoint(const float* const* q, const float* vx, const float* vy,
                             const float* vz, const float* const* w_bar, float* const* q_bar,
                             float* vx_bar, float* vy_bar, float* vz_bar, const int* nx,
                             const int* ny, const int* nz, const float* chirho,
                             const float* chipxyz, const float* delt, float* work) {
    LapAdj L{};
    L.B = box_of(nx, ny, nz);
    for (int f = 0; f < 10; ++f) {
        L.q[f] = R(q[f]);
        L.wb[f] = R(w_bar[f]);
        L.qb[f] = R(q_bar[f]);
        L.chi[f] = f == 0 ? *chirho : *chipxyz;
    }
    L.v[0] = R(vx);
    L.v[1] = R(vy);
    L.v[2] = R(vz);
    L.vb[0] = R(vx_bar);
    L.vb[1] = R(vy_bar);
    L.vb[2] = R(vz_bar);
    L.delt = *delt;
    L.scr = R(work);
    launch_adj_lap_species(L, g_stream);
    check_launch("lap_plasma_adjoint");
}

void g3m3_species_velocity_adjoint(const float* rho, const float* px, const float* py,
                                   const float* pz, const float* vx_bar, const float* vy_bar,
                                   const float* vz_bar, float* rho_bar, float* px_bar,
                                   float* py_bar, float* pz_bar, const int* n,
                                   const float* smallbit) {
    const float* p[3] = {R(px), R(py), R(pz)};
    const float* vb[3] = {R(vx_bar), R(vy_bar), R(vz_bar)};
    float* pb[3] = {R(px_bar), R(py_bar), R(pz_bar)};
    launch_adj_velocity(*n, R(rho), p, vb, *smallbit, R(rho_bar), pb, g_stream);
    check_launch("species_velocity_adjoint");
}

void g3m3_sample_craft(const int* ncraft, const float* xcraft, const float* zcraft,
                       const float* ut, const int* n_grids, const float* grid_minvals,
                       const float* grid_maxvals, const float* xspac, const float* re_equiv,
                       const int* nx, const int* ny, const int* nz, const int* nq,
                       const float* const* fields, float* sxyz, float* scdata, int* gridpt) {
    CraftGrid G;
    if (!craft_grid(G, n_grids, grid_minvals, grid_maxvals, xspac, re_equiv)) return;
    if (*nq < 0 || *nq > kMaxFields) {
        latch(4, "sample_craft: nq out of range");
        return;
    }
    FieldPtrs F{};
    for (int q = 0; q < *nq; ++q) F.p[q] = const_cast<float*>(R(fields[q]));
    launch_sample(*ncraft, R(sxyz), R(xcraft), R(zcraft), *ut, G, *nx, *ny, *nz, *nq, F,
                  R(scdata), R(gridpt), g_stream);
    check_launch("sample_craft");
}

void g3m3_find_aux_cube(const char* cname, const int* i_craft, const int* scout,
                        const int* n_grids, const float* grid_minvals, const float* grid_maxvals,
                        const float* xspac, const float* re_equiv, const float* xcraftn,
                        const float* zcraftn, const float* ut, int* ntimesn, int* recordingn,
                        int* gridpt, float* sxyz, float* cube_vertices, size_t cname_len) {
    (void)i_craft;
    (void)scout;
    (void)ntimesn;
    CraftGrid G;
    if (!craft_grid(G, n_grids, grid_minvals, grid_maxvals, xspac, re_equiv)) return;
    craft_position(xcraftn, zcraftn, *ut, sxyz);
    g3m3_findgrid(sxyz, n_grids, grid_minvals, grid_maxvals, xspac, gridpt, re_equiv, cname,
                  cname_len);
    if (gridpt[0] <= 0) {
        *recordingn = 0;
        const std::string name = cname ? std::string(cname, cname_len) : std::string();
        printf(" Gridding problem with craft %s at UT = %g Future points skipped.\n",
               name.c_str(), (double)*ut);
        return;
    }
    float v0[3], v1[3];
    craft_cube(G, gridpt, v0, v1);
    for (int a = 0; a < 3; ++a) {
        cube_vertices[a] = v0[a];       // cube_vertices(axis, 1)
        cube_vertices[3 + a] = v1[a];   // cube_vertices(axis, 2)
    }
}

}  // extern "C"

// -----------------------------------------------------------------------------
// Fortran implicit-interface names.  A Fortran caller that declares no
// interface (as the reference's own call sites do) calls g3m3_push_ion as
// the symbol g3m3_push_ion_ with every argument by reference; those names
// are aliases of the entry points above.  The registry and control calls,
// whose C forms take values, get by-reference forms.
extern "C" {
int g3m3_init_(void) { return g3m3_init(); }
int g3m3_last_error_(void) { return g3m3_last_error(); }
void g3m3_sync_(void) { g3m3_sync(); }
int g3m3_register_(void* host, const long long* bytes) { return g3m3_register(host, *bytes); }
int g3m3_unregister_(void* host) { return g3m3_unregister(host); }
int g3m3_to_device_(void* host) { return g3m3_to_device(host); }
int g3m3_to_host_(void* host) { return g3m3_to_host(host); }
}

#ifndef __HIP_DEVICE_COMPILE__
#define G3M3_FORTRAN_ALIAS(n) extern "C" void g3m3_##n##_() __attribute__((alias("g3m3_" #n)));
G3M3_FORTRAN_ALIAS(store_array)
G3M3_FORTRAN_ALIAS(calcur)
G3M3_FORTRAN_ALIAS(fnd_evel)
G3M3_FORTRAN_ALIAS(bfield_total)
G3M3_FORTRAN_ALIAS(bande)
G3M3_FORTRAN_ALIAS(bande_eta)
G3M3_FORTRAN_ALIAS(push_elec)
G3M3_FORTRAN_ALIAS(push_ion)
G3M3_FORTRAN_ALIAS(push_bfld)
G3M3_FORTRAN_ALIAS(set_rho)
G3M3_FORTRAN_ALIAS(set_speed_agrd)
G3M3_FORTRAN_ALIAS(species_velocity)
G3M3_FORTRAN_ALIAS(lap_plasma)
G3M3_FORTRAN_ALIAS(lap_elec)
G3M3_FORTRAN_ALIAS(lap_bfld)
G3M3_FORTRAN_ALIAS(fnd_vtot)
G3M3_FORTRAN_ALIAS(fcsmooth)
G3M3_FORTRAN_ALIAS(flux_correct)
G3M3_FORTRAN_ALIAS(bndry_outer)
G3M3_FORTRAN_ALIAS(bndry_inner)
G3M3_FORTRAN_ALIAS(bndry_flanks_v)
G3M3_FORTRAN_ALIAS(bndry_flanks)
G3M3_FORTRAN_ALIAS(flanks)
G3M3_FORTRAN_ALIAS(corer)
G3M3_FORTRAN_ALIAS(bndry_corer_v)
G3M3_FORTRAN_ALIAS(bndry_corer)
G3M3_FORTRAN_ALIAS(flanks_synced)
G3M3_FORTRAN_ALIAS(banded_tpose)
G3M3_FORTRAN_ALIAS(divb_correct)
G3M3_FORTRAN_ALIAS(divb_correct_ex)
G3M3_FORTRAN_ALIAS(trilin_interp)
G3M3_FORTRAN_ALIAS(fnd_pres)
G3M3_FORTRAN_ALIAS(sparse_solver)
G3M3_FORTRAN_ALIAS(findgrid)
G3M3_FORTRAN_ALIAS(sample_fields)
G3M3_FORTRAN_ALIAS(sample_fields_adjoint)
G3M3_FORTRAN_ALIAS(push_ion_adjoint)
G3M3_FORTRAN_ALIAS(lap_plasma_adjoint)
G3M3_FORTRAN_ALIAS(species_velocity_adjoint)
G3M3_FORTRAN_ALIAS(sample_craft)
G3M3_FORTRAN_ALIAS(find_aux_cube)
G3M3_FORTRAN_ALIAS(flux_counter)
G3M3_FORTRAN_ALIAS(torus_inject)
G3M3_FORTRAN_ALIAS(dipole)
G3M3_FORTRAN_ALIAS(mak_dip_grd)
G3M3_FORTRAN_ALIAS(set_resist)
G3M3_FORTRAN_ALIAS(init_plasma)
#undef G3M3_FORTRAN_ALIAS
#endif
