// Fluid (restart) files in the reference's format: Fortran unformatted
// sequential, each record framed by 4-byte little-endian byte counts, in
// the order written at code/3d_main.f90:3310-3349 and read at :972-1011:
//   t (real(dp));
//   37 field arrays q(nx,ny,nz,n_grids), float32, species q, h, o each as
//   rho, px, py, pz, presx, presy, presz, presxy, presxz, presyz, then
//   bx, by, bz, epres, bx0, by0, bz0;
//   one record: parm_srf(mbndry,7,msrf), parm_mid(mbndry,7,mmid),
//   parm_zero(mbndry,7,mzero), ijzero(mbndry,3,mzero), numzero(mbndry),
//   ijmid(mbndry,3,mmid), nummid(mbndry), ijsrf(mbndry,3,msrf),
//   numsrf(mbndry).
// Host arrays only; the runtime wrappers move device data.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/g3m3.h"

namespace {

bool put_record(FILE* f, const std::vector<std::pair<const void*, size_t>>& parts) {
    size_t n = 0;
    for (auto& p : parts) n += p.second;
    if (n > 0x7fffffffu) return false;   // one record of < 2 GiB (no subrecords)
    const int32_t m = (int32_t)n;
    if (fwrite(&m, 4, 1, f) != 1) return false;
    for (auto& p : parts)
        if (p.second && fwrite(p.first, 1, p.second, f) != p.second) return false;
    return fwrite(&m, 4, 1, f) == 1;
}

bool get_record(FILE* f, const std::vector<std::pair<void*, size_t>>& parts) {
    size_t n = 0;
    for (auto& p : parts) n += p.second;
    int32_t m = 0, m2 = 0;
    if (fread(&m, 4, 1, f) != 1 || m < 0 || (size_t)m != n) return false;
    for (auto& p : parts)
        if (p.second && fread(p.first, 1, p.second, f) != p.second) return false;
    return fread(&m2, 4, 1, f) == 1 && m2 == m;
}

struct ListSizes {
    size_t parm_srf, parm_mid, parm_zero, ij_srf, ij_mid, ij_zero, num;
};

ListSizes list_sizes(int mbndry, int msrf, int mmid, int mzero) {
    const size_t mb = (size_t)mbndry;
    return ListSizes{4 * mb * 7 * msrf, 4 * mb * 7 * mmid, 4 * mb * 7 * mzero, 4 * mb * 3 * msrf,
                     4 * mb * 3 * mmid, 4 * mb * 3 * mzero, 4 * mb};
}

}  // namespace

extern "C" {

int g3m3_fluid_write(const char* path, const double* t, const int* nx, const int* ny,
                     const int* nz, const int* n_grids, const float* const* fields,
                     const int* mbndry, const int* msrf, const int* mmid, const int* mzero,
                     const float* parm_srf, const float* parm_mid, const float* parm_zero,
                     const int* ijsrf, const int* ijmid, const int* ijzero, const int* numsrf,
                     const int* nummid, const int* numzero) {
    FILE* f = fopen(path, "wb");
    if (!f) return 1;
    const size_t bytes = sizeof(float) * (size_t)*nx * *ny * *nz * *n_grids;
    bool ok = put_record(f, {{t, sizeof(double)}});
    for (int a = 0; ok && a < G3M3_FLUID_ARRAYS; ++a) ok = put_record(f, {{fields[a], bytes}});
    const ListSizes L = list_sizes(*mbndry, *msrf, *mmid, *mzero);
    ok = ok && put_record(f, {{parm_srf, L.parm_srf},
                              {parm_mid, L.parm_mid},
                              {parm_zero, L.parm_zero},
                              {ijzero, L.ij_zero},
                              {numzero, L.num},
                              {ijmid, L.ij_mid},
                              {nummid, L.num},
                              {ijsrf, L.ij_srf},
                              {numsrf, L.num}});
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : 2;
}

int g3m3_fluid_read(const char* path, double* t, const int* nx, const int* ny, const int* nz,
                    const int* n_grids, float* const* fields, const int* mbndry, const int* msrf,
                    const int* mmid, const int* mzero, float* parm_srf, float* parm_mid,
                    float* parm_zero, int* ijsrf, int* ijmid, int* ijzero, int* numsrf,
                    int* nummid, int* numzero) {
    FILE* f = fopen(path, "rb");
    if (!f) return 1;
    const size_t bytes = sizeof(float) * (size_t)*nx * *ny * *nz * *n_grids;
    bool ok = get_record(f, {{t, sizeof(double)}});
    for (int a = 0; ok && a < G3M3_FLUID_ARRAYS; ++a) ok = get_record(f, {{fields[a], bytes}});
    const ListSizes L = list_sizes(*mbndry, *msrf, *mmid, *mzero);
    ok = ok && get_record(f, {{parm_srf, L.parm_srf},
                              {parm_mid, L.parm_mid},
                              {parm_zero, L.parm_zero},
                              {ijzero, L.ij_zero},
                              {numzero, L.num},
                              {ijmid, L.ij_mid},
                              {nummid, L.num},
                              {ijsrf, L.ij_srf},
                              {numsrf, L.num}});
    fclose(f);
    return ok ? 0 : 2;
}

}  // extern "C"
