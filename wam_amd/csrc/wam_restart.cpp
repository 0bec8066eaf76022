// wam_restart.cpp -- the WAM restart file (binary form) on the host.
//
// SAVE_RESTART_FILE (wam_restart_module.f90:431-600) writes, and
// CONNECT_RESTART (:109-327) reads, Fortran sequential unformatted records:
//   1. NSEA, CDTPRO, CDTSOU, CDA, CDTA, CDCA   (INTEGER*4, 5 x CHARACTER*14)
//   2. U10(NSEA)   3. UDIR(NSEA)   4. TAUW(NSEA)   5. FL3(NSEA,KL,ML)
//   6. DEPTH(NSEA)          only when CDTA is not blank
//   7. U(NSEA), V(NSEA)     only when CDCA is not blank (one record)
// in the sea-point order of one undivided domain.  Every record has a 4-byte
// little-endian byte count before and after its data.  A record longer than
// 2 147 483 639 bytes (FL3 from about 1.6 M sea points at 36 x 36) is split
// into subrecords, the scheme of the Intel and GNU Fortran runtimes: the
// leading count of every subrecord but the last and the trailing count of
// every subrecord but the first are negated.
#include "wam_restart.hpp"

#include <cstring>

namespace wam {

static bool put32(FILE *f, int32_t v) { return std::fwrite(&v, 4, 1, f) == 1; }

static bool get32(FILE *f, int32_t &v) { return std::fread(&v, 4, 1, f) == 1; }

bool RecordWriter::begin(uint64_t bytes)
{
    total_ = bytes;
    left_ = bytes;
    first_ = true;
    return open_sub();
}

bool RecordWriter::open_sub()
{
    sub_ = left_ > MAX_SUBRECORD ? MAX_SUBRECORD : left_;
    sub_left_ = sub_;
    const bool more = left_ > sub_;
    return put32(f_, more ? -(int32_t)sub_ : (int32_t)sub_);
}

bool RecordWriter::close_sub()
{
    const int32_t t = first_ ? (int32_t)sub_ : -(int32_t)sub_;
    first_ = false;
    return put32(f_, t);
}

bool RecordWriter::put(const void *p, uint64_t n)
{
    const char *c = static_cast<const char *>(p);
    while (n > 0) {
        if (sub_left_ == 0) {
            // the current subrecord is full and the record goes on
            if (!close_sub() || !open_sub()) return false;
        }
        const uint64_t k = n < sub_left_ ? n : sub_left_;
        if (std::fwrite(c, 1, k, f_) != k) return false;
        c += k;
        n -= k;
        sub_left_ -= k;
        left_ -= k;
    }
    return true;
}

bool RecordWriter::end()
{
    if (left_ != 0 || sub_left_ != 0) return false;
    return close_sub();
}

bool RecordReader::begin(uint64_t bytes)
{
    total_ = bytes;
    left_ = bytes;
    first_ = true;
    return open_sub();
}

bool RecordReader::open_sub()
{
    int32_t m;
    if (!get32(f_, m)) return false;
    more_ = m < 0;
    sub_ = (uint64_t)(m < 0 ? -(int64_t)m : m);
    sub_left_ = sub_;
    return sub_ <= left_ && (more_ || sub_ == left_);
}

bool RecordReader::close_sub()
{
    int32_t t;
    if (!get32(f_, t)) return false;
    const int32_t want = first_ ? (int32_t)sub_ : -(int32_t)sub_;
    first_ = false;
    return t == want;
}

bool RecordReader::get(void *p, uint64_t n)
{
    char *c = static_cast<char *>(p);
    while (n > 0) {
        if (sub_left_ == 0) {
            if (!more_ || !close_sub() || !open_sub()) return false;
        }
        const uint64_t k = n < sub_left_ ? n : sub_left_;
        if (std::fread(c, 1, k, f_) != k) return false;
        c += k;
        n -= k;
        sub_left_ -= k;
        left_ -= k;
    }
    return true;
}

bool RecordReader::skip(uint64_t n)
{
    while (n > 0) {
        if (sub_left_ == 0) {
            if (!more_ || !close_sub() || !open_sub()) return false;
        }
        const uint64_t k = n < sub_left_ ? n : sub_left_;
        if (fseeko(f_, (off_t)k, SEEK_CUR) != 0) return false;
        n -= k;
        sub_left_ -= k;
        left_ -= k;
    }
    return true;
}

bool RecordReader::end()
{
    if (left_ != 0 || sub_left_ != 0 || more_) return false;
    return close_sub();
}

bool date_blank(const char *d14)
{
    for (int i = 0; i < 14; i++)
        if (d14[i] != ' ') return false;
    return true;
}

}  // namespace wam

// ---------------------------------------------------------------------------
// The C-ABI host forms: arrays in host memory.

#include "../../include/wam_gpu.h"

namespace {

struct File {
    FILE *f = nullptr;
    ~File() { if (f) std::fclose(f); }
};

bool put_record(wam::RecordWriter &w, const void *p, uint64_t n) { return w.begin(n) && w.put(p, n) && w.end(); }

bool put_record2(wam::RecordWriter &w, const void *a, const void *b, uint64_t n)
{
    return w.begin(2 * n) && w.put(a, n) && w.put(b, n) && w.end();
}

bool get_record(wam::RecordReader &r, void *p, uint64_t n) { return r.begin(n) && r.get(p, n) && r.end(); }

}  // namespace

extern "C" {

int wam_restart_write(const char *path, int32_t nsea, int32_t kl, int32_t ml, const char *dates, const float *fl3,
                      const float *u10, const float *udir, const float *tauw, const float *depth, const float *u,
                      const float *v)
{
    if (!path || !dates || nsea <= 0 || kl <= 0 || ml <= 0 || !fl3 || !u10 || !udir || !tauw) return WAM_E_ARG;
    const bool with_depth = !wam::date_blank(dates + 42), with_cur = !wam::date_blank(dates + 56);
    if ((with_depth && !depth) || (with_cur && (!u || !v))) return WAM_E_ARG;
    File fh;
    fh.f = std::fopen(path, "wb");
    if (!fh.f) return WAM_E_IO;
    wam::RecordWriter w(fh.f);
    char head[74];
    std::memcpy(head, &nsea, 4);
    std::memcpy(head + 4, dates, 70);
    const uint64_t n = (uint64_t)nsea * 4, nb = n * (uint64_t)kl * (uint64_t)ml;
    bool ok = put_record(w, head, 74) && put_record(w, u10, n) && put_record(w, udir, n) && put_record(w, tauw, n) &&
              put_record(w, fl3, nb);
    if (ok && with_depth) ok = put_record(w, depth, n);
    if (ok && with_cur) ok = put_record2(w, u, v, n);
    if (!ok) return WAM_E_IO;
    if (std::fclose(fh.f) != 0) { fh.f = nullptr; return WAM_E_IO; }
    fh.f = nullptr;
    return WAM_OK;
}

int wam_restart_read(const char *path, int32_t nsea, int32_t kl, int32_t ml, char *dates, float *fl3, float *u10,
                     float *udir, float *tauw, float *depth, float *u, float *v, int32_t *present)
{
    if (!path || !dates || nsea <= 0 || kl <= 0 || ml <= 0 || !fl3 || !u10 || !udir || !tauw) return WAM_E_ARG;
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return WAM_E_IO;
    wam::RecordReader r(fh.f);
    char head[74];
    if (!get_record(r, head, 74)) return WAM_E_IO;
    int32_t nsea_r;
    std::memcpy(&nsea_r, head, 4);
    if (nsea_r != nsea) return WAM_E_IO;                // restart fields inconsistent with the grid
    std::memcpy(dates, head + 4, 70);
    const bool with_depth = !wam::date_blank(dates + 42), with_cur = !wam::date_blank(dates + 56);
    const uint64_t n = (uint64_t)nsea * 4, nb = n * (uint64_t)kl * (uint64_t)ml;
    bool ok = get_record(r, u10, n) && get_record(r, udir, n) && get_record(r, tauw, n) && get_record(r, fl3, nb);
    if (ok && with_depth) {
        if (depth) ok = get_record(r, depth, n);
        else ok = r.begin(n) && r.skip(n) && r.end();
    }
    if (ok && with_cur) {
        if (u && v) ok = r.begin(2 * n) && r.get(u, n) && r.get(v, n) && r.end();
        else ok = r.begin(2 * n) && r.skip(2 * n) && r.end();
    }
    if (!ok) return WAM_E_IO;
    if (present) *present = (with_depth ? 1 : 0) | (with_cur ? 2 : 0);
    return WAM_OK;
}

}  // extern "C"
