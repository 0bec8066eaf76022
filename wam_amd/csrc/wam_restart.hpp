// wam_restart.hpp -- Fortran sequential unformatted records for the restart file.
#pragma once

#include <cstdint>
#include <cstdio>
#include <sys/types.h>

namespace wam {

// Longest subrecord of the Intel and GNU Fortran runtimes, in bytes.
constexpr uint64_t MAX_SUBRECORD = 2147483639ull;

// Writes one logical record of a known byte count in pieces: begin(bytes),
// put(...) until bytes have been written, end().
class RecordWriter {
public:
    explicit RecordWriter(FILE *f) : f_(f) {}
    bool begin(uint64_t bytes);
    bool put(const void *p, uint64_t n);
    bool end();

private:
    bool open_sub();
    bool close_sub();
    FILE *f_;
    uint64_t total_ = 0, left_ = 0, sub_ = 0, sub_left_ = 0;
    bool first_ = true;
};

// Reads one logical record that must hold exactly `bytes` bytes.
class RecordReader {
public:
    explicit RecordReader(FILE *f) : f_(f) {}
    bool begin(uint64_t bytes);
    bool get(void *p, uint64_t n);
    bool skip(uint64_t n);
    bool end();

private:
    bool open_sub();
    bool close_sub();
    FILE *f_;
    uint64_t total_ = 0, left_ = 0, sub_ = 0, sub_left_ = 0;
    bool first_ = true, more_ = false;
};

// A CHARACTER*14 date that is all blanks (the optional record is absent).
bool date_blank(const char *d14);

}  // namespace wam
