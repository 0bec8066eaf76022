// wam_map_file.cpp -- the gridded integrated-parameter file (binary form) on the host.
//
// WRITE_INT_PAR_OUTPUT (wam_output_module.f90:786-911) writes, per output
// time, Fortran sequential unformatted records
//   CDTPRO, REAL(NX), REAL(NY), AMOWEP, AMOSOP, AMOEAP, AMONOP, CDATEA
//       (CHARACTER*14, 2 REAL, 4 INTEGER M_SEC, CHARACTER*14)
//   NLON_RG(1:NY), ZDELLO(1:NY)      (INTEGER)
//   FFLAG_P(1:70)                    (LOGICAL)
//   GRID(1:NX, 1:NY)                 one record per parameter IP with FFLAG_P(IP)
// which READ_GRID_FILE (print/read_grid_file.f90:62-140) reads back.  Land
// cells hold ZMISS = -9999999.
#include <cstring>

#include "../../include/wam_gpu.h"
#include "wam_restart.hpp"

namespace {

struct File {
    FILE *f = nullptr;
    ~File() { if (f) std::fclose(f); }
};

bool record(wam::RecordWriter &w, const void *p, uint64_t n) { return w.begin(n) && w.put(p, n) && w.end(); }

}  // namespace

extern "C" {

int wam_map_write(const char *path, int32_t append, const char *cdtpro, const char *cdatea, int32_t nx, int32_t ny,
                  const int32_t *corners, const int32_t *nlon_rg, const int32_t *zdello, const int32_t *fflag,
                  const float *grids)
{
    if (!path || !cdtpro || !cdatea || nx <= 0 || ny <= 0 || !corners || !nlon_rg || !zdello || !fflag)
        return WAM_E_ARG;
    int nsel = 0;
    for (int i = 0; i < WAM_MAP_NOUT; i++) nsel += fflag[i] != 0;
    if (nsel > 0 && !grids) return WAM_E_ARG;
    File fh;
    fh.f = std::fopen(path, append ? "ab" : "wb");
    if (!fh.f) return WAM_E_IO;
    wam::RecordWriter w(fh.f);
    char head[52];
    const float fnx = (float)nx, fny = (float)ny;
    std::memcpy(head, cdtpro, 14);
    std::memcpy(head + 14, &fnx, 4);
    std::memcpy(head + 18, &fny, 4);
    std::memcpy(head + 22, corners, 16);          // AMOWEP, AMOSOP, AMOEAP, AMONOP
    std::memcpy(head + 38, cdatea, 14);
    int32_t flags[WAM_MAP_NOUT];
    for (int i = 0; i < WAM_MAP_NOUT; i++) flags[i] = fflag[i] != 0 ? 1 : 0;     // .TRUE. as 1
    const uint64_t nrow = (uint64_t)ny * 4, ngrid = (uint64_t)nx * ny * 4;
    bool ok = record(w, head, 52) && w.begin(2 * nrow) && w.put(nlon_rg, nrow) && w.put(zdello, nrow) && w.end() &&
              record(w, flags, sizeof(flags));
    for (int i = 0; ok && i < nsel; i++) ok = record(w, grids + (size_t)i * nx * ny, ngrid);
    if (!ok) return WAM_E_IO;
    const int rc = std::fclose(fh.f);
    fh.f = nullptr;
    return rc == 0 ? WAM_OK : WAM_E_IO;
}

}  // extern "C"
