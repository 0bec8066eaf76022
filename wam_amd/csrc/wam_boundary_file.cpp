// wam_boundary_file.cpp -- the nest boundary value file (binary form) on the host.
//
// SAVE_BOUNDARY_FILE (wam_boundary_module.f90:709-808) opens the file with
// one header record of eight REALs
//   KL, ML, TH(1), FR(1), CO, NBOUNC, IDEL_B_OUT, IDEL_BO_FILE;
// BOUNDARY_OUTPUT (:364-378) then appends, per output time and per boundary
// point, two records
//   BLNGC, BLATC (INTEGER), CDTPRO (CHARACTER*14), EMEAN, THQ, FMEAN, DEPTH
//   FBC(1:KL, 1:ML)
// which READ_BOUNDARY_INPUT (chief/read_boundary_input.f90:66-195) reads back
// in the fine grid's run.  Fortran sequential unformatted records with 4-byte
// counts, as the restart file.
#include <cstring>

#include "../../include/wam_gpu.h"
#include "wam_restart.hpp"

namespace {

struct File {
    FILE *f = nullptr;
    ~File() { if (f) std::fclose(f); }
};

constexpr int POINT_BYTES = 4 + 4 + 14 + 4 * 4;     // the point record

}  // namespace

extern "C" {

int wam_boundary_write_header(const char *path, const float *hdr)
{
    if (!path || !hdr) return WAM_E_ARG;
    File fh;
    fh.f = std::fopen(path, "wb");
    if (!fh.f) return WAM_E_IO;
    wam::RecordWriter w(fh.f);
    if (!(w.begin(32) && w.put(hdr, 32) && w.end())) return WAM_E_IO;
    const int rc = std::fclose(fh.f);
    fh.f = nullptr;
    return rc == 0 ? WAM_OK : WAM_E_IO;
}

int wam_boundary_append(const char *path, int32_t nb, int32_t kl, int32_t ml, const int32_t *lon, const int32_t *lat,
                        const char *cdtpro, const float *mean, const float *depth, const float *fbc)
{
    if (!path || nb < 0 || kl <= 0 || ml <= 0) return WAM_E_ARG;
    if (nb > 0 && (!lon || !lat || !cdtpro || !mean || !depth || !fbc)) return WAM_E_ARG;
    File fh;
    fh.f = std::fopen(path, "ab");
    if (!fh.f) return WAM_E_IO;
    wam::RecordWriter w(fh.f);
    const uint64_t nspec = (uint64_t)kl * ml * 4;
    for (int32_t i = 0; i < nb; i++) {
        char rec[POINT_BYTES];
        const float v[4] = {mean[i], mean[nb + i], mean[2 * (size_t)nb + i], depth[i]};
        std::memcpy(rec, &lon[i], 4);
        std::memcpy(rec + 4, &lat[i], 4);
        std::memcpy(rec + 8, cdtpro, 14);
        std::memcpy(rec + 22, v, 16);
        if (!(w.begin(POINT_BYTES) && w.put(rec, POINT_BYTES) && w.end())) return WAM_E_IO;
        if (!(w.begin(nspec) && w.put(fbc + (size_t)i * kl * ml, nspec) && w.end())) return WAM_E_IO;
    }
    const int rc = std::fclose(fh.f);
    fh.f = nullptr;
    return rc == 0 ? WAM_OK : WAM_E_IO;
}

int wam_boundary_read_header(const char *path, float *hdr, int64_t *pos)
{
    if (!path || !hdr || !pos) return WAM_E_ARG;
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return WAM_E_IO;
    wam::RecordReader r(fh.f);
    if (!(r.begin(32) && r.get(hdr, 32) && r.end())) return WAM_E_IO;
    *pos = (int64_t)ftello(fh.f);
    return WAM_OK;
}

int wam_boundary_read(const char *path, int64_t *pos, int32_t nb, int32_t kl, int32_t ml, int32_t *lon, int32_t *lat,
                      char *dates, float *mean, float *depth, float *fbc)
{
    if (!path || !pos || *pos < 0 || nb <= 0 || kl <= 0 || ml <= 0) return WAM_E_ARG;
    if (!lon || !lat || !dates || !mean || !depth || !fbc) return WAM_E_ARG;
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return WAM_E_IO;
    if (fseeko(fh.f, (off_t)*pos, SEEK_SET) != 0) return WAM_E_IO;
    wam::RecordReader r(fh.f);
    const uint64_t nspec = (uint64_t)kl * ml * 4;
    for (int32_t i = 0; i < nb; i++) {
        char rec[POINT_BYTES];
        if (!(r.begin(POINT_BYTES) && r.get(rec, POINT_BYTES) && r.end())) return WAM_E_IO;
        float v[4];
        std::memcpy(&lon[i], rec, 4);
        std::memcpy(&lat[i], rec + 4, 4);
        std::memcpy(dates + 14 * (size_t)i, rec + 8, 14);
        std::memcpy(v, rec + 22, 16);
        mean[i] = v[0];
        mean[nb + i] = v[1];
        mean[2 * (size_t)nb + i] = v[2];
        depth[i] = v[3];
        if (!(r.begin(nspec) && r.get(fbc + (size_t)i * kl * ml, nspec) && r.end())) return WAM_E_IO;
    }
    *pos = (int64_t)ftello(fh.f);
    return WAM_OK;
}

}  // extern "C"
