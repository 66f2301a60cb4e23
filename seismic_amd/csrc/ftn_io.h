// ftn_io.h — the two Fortran file forms of the out-of-core L-BFGS protocol,
// read and written from C++ (host only):
//
//  * list-directed text records: libs/libnlinv/nlinv_lbfgs.f keeps the SAVE
//    scalars of LBFGSOC / MCSRCHOC as `write(20,*) a, b, ...` records and reads
//    them back with `read(20,*) a, b, ...` (nlinv_lbfgs.f:276-281, :755-760);
//    IFLAG and MYINFO travel as `write(30,"(I2)")` / `read(10,*)`
//    (progs/inversion/L_BFGS.f90:82-90, nlinv_read_array_params.f90:154-174).
//  * sequential unformatted records: W and DIAG are one `write(89) w` record
//    each (L_BFGS.f90:69-78), framed by a 4-byte little-endian byte count
//    before and after the data.
#pragma once
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"

namespace ftn {

// One list-directed READ statement's items, taken from a text file that holds
// several such statements (records).  A statement reads values from its first
// record on, continuing onto further records while items remain, and the rest
// of its last record is skipped, as Fortran's list-directed input does.  Value
// separators are blanks, tabs and commas; `r*c` repeats c r times, `r*` is r
// null values and `/` ends the statement (both leave the remaining items as
// they were).
class ListReader {
  public:
    explicit ListReader(const std::string &path) : path_(path) {
        FILE *f = fopen(path.c_str(), "r");
        if (!f) throw EngineError(SFD_EIO, "cannot open " + path);
        char buf[4096];
        std::string line;
        while (fgets(buf, sizeof buf, f)) {
            line += buf;
            if (!line.empty() && line.back() == '\n') {
                line.pop_back();
                if (!line.empty() && line.back() == '\r') line.pop_back();
                lines_.push_back(line);
                line.clear();
            }
        }
        if (!line.empty()) lines_.push_back(line);
        fclose(f);
    }

    // Starts the next READ statement: items come from the next record on.
    void statement() {
        pending_.clear();
        repeat_ = 0;
        ended_ = false;
        if (started_) line_++;
        started_ = true;
        pos_ = 0;
    }
    void real(double *v) {
        std::string t;
        if (!next(&t)) return;
        *v = parse_real(t);
    }
    void integer(int *v) {
        std::string t;
        if (!next(&t)) return;
        char *e = nullptr;
        errno = 0;
        const long x = strtol(t.c_str(), &e, 10);
        if (e == t.c_str() || *e || errno || x < INT32_MIN || x > INT32_MAX) bad(t, "an integer");
        *v = (int)x;
    }
    void logical(int *v) {
        std::string t;
        if (!next(&t)) return;
        size_t k = t[0] == '.' ? 1 : 0;
        const char c = k < t.size() ? (char)toupper((unsigned char)t[k]) : 0;
        if (c == 'T') *v = 1;
        else if (c == 'F') *v = 0;
        else bad(t, "a logical");
    }

  private:
    std::string path_;
    std::vector<std::string> lines_;
    size_t line_ = 0, pos_ = 0;
    bool started_ = false, ended_ = false;
    std::string pending_;   // value of an r*c repeat in progress
    long repeat_ = 0;

    [[noreturn]] void bad(const std::string &t, const char *what) {
        throw EngineError(SFD_EIO, path_ + ": '" + t + "' is not " + what);
    }
    static bool sep(char c) { return c == ' ' || c == '\t' || c == ','; }

    // Next value of the statement; false for a null value (the item keeps its
    // value).  Running out of records is an error, as Fortran's end of file.
    bool next(std::string *t) {
        if (ended_) return false;
        if (repeat_ > 0) {
            repeat_--;
            if (pending_.empty()) return false;
            *t = pending_;
            return true;
        }
        for (;;) {
            if (line_ >= lines_.size()) throw EngineError(SFD_EIO, path_ + ": end of file in a record");
            const std::string &L = lines_[line_];
            while (pos_ < L.size() && (L[pos_] == ' ' || L[pos_] == '\t')) pos_++;
            if (pos_ < L.size() && L[pos_] == ',') {   // an empty value between two commas is null
                pos_++;
                return false;
            }
            if (pos_ >= L.size()) {
                line_++;
                pos_ = 0;
                continue;
            }
            if (L[pos_] == '/') {
                ended_ = true;
                return false;
            }
            size_t e = pos_;
            while (e < L.size() && !sep(L[e]) && L[e] != '/') e++;
            std::string tok = L.substr(pos_, e - pos_);
            pos_ = e;
            while (pos_ < L.size() && (L[pos_] == ' ' || L[pos_] == '\t')) pos_++;
            if (pos_ < L.size() && L[pos_] == ',') pos_++;
            const size_t star = tok.find('*');
            if (star != std::string::npos) {
                char *end = nullptr;
                const long r = strtol(tok.c_str(), &end, 10);
                if (end != tok.c_str() + star || r < 1) bad(tok, "a repeat count");
                pending_ = tok.substr(star + 1);
                repeat_ = r - 1;
                if (pending_.empty()) return false;
                *t = pending_;
                return true;
            }
            *t = tok;
            return true;
        }
    }

    // A Fortran real: optional sign, digits with an optional point, and an
    // exponent written with E, D or Q (either case) or as a bare signed
    // number after the digits ("1.5+20").
    double parse_real(const std::string &t0) {
        std::string t = t0;
        for (char &c : t)
            if (c == 'd' || c == 'D' || c == 'q' || c == 'Q') c = 'E';
        size_t k = (t[0] == '+' || t[0] == '-') ? 1 : 0;
        size_t j = k;
        while (j < t.size() && (isdigit((unsigned char)t[j]) || t[j] == '.')) j++;
        if (j > k && j < t.size() && (t[j] == '+' || t[j] == '-')) t.insert(j, "E");
        char *e = nullptr;
        const double v = strtod(t.c_str(), &e);
        if (e == t.c_str() || *e) bad(t0, "a real");
        return v;
    }
};

// Writes list-directed records as a Fortran `write(u,*)` does: a leading
// blank, then the items separated by blanks.  Reals get the fewest significant
// digits (at most 17) that read back to the same double, so a text state file
// carries every scalar exactly.
class ListWriter {
  public:
    void real(double v) {
        char b[40];
        for (int p = 1; p <= 17; p++) {
            snprintf(b, sizeof b, "%.*E", p - 1, v);
            if (!std::isfinite(v) || strtod(b, nullptr) == v) break;
        }
        item(b);
    }
    void integer(int v) { item(std::to_string(v)); }
    void logical(int v) { item(v ? "T" : "F"); }
    void end_record() { text_ += '\n'; }
    // Writes the text to path through a temporary file and a rename, so a
    // reader never sees a partly written file.
    void save(const std::string &path) const { save_text(path, text_); }

    static void save_text(const std::string &path, const std::string &text) {
        const std::string tmp = path + ".tmp";
        FILE *f = fopen(tmp.c_str(), "w");
        if (!f) throw EngineError(SFD_EIO, "cannot create " + tmp);
        const bool ok = fwrite(text.data(), 1, text.size(), f) == text.size();
        if (fclose(f) != 0 || !ok) throw EngineError(SFD_EIO, "write failed: " + tmp);
        if (rename(tmp.c_str(), path.c_str()) != 0) throw EngineError(SFD_EIO, "cannot rename to " + path);
    }

  private:
    std::string text_;
    void item(const std::string &s) {
        text_ += ' ';
        text_ += s;
    }
};

// `(I2)` output of one integer (L_BFGS.f90:84,90): right-aligned in two
// columns, asterisks when it does not fit.
inline std::string i2_record(int v) {
    char b[16];
    snprintf(b, sizeof b, "%2d", v);
    return (strlen(b) > 2 ? std::string("**") : std::string(b)) + "\n";
}

// Whether a file exists (INQUIRE(FILE=..., EXIST=...)).
inline bool exists(const std::string &path) {
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return false;
    fclose(f);
    return true;
}

// One sequential unformatted record of doubles.  The reader takes the first
// `count` values of the file's first record (a Fortran READ may read fewer
// items than its record holds; a shorter record is an error) and accepts
// 4-byte or 8-byte length markers (the width whose trailing marker repeats
// the leading one).  The writer frames the record with 4-byte
// markers; a record of 2 GiB or more does not fit one 4-byte marker (Fortran
// runtimes split such records into subrecords) and is refused.
class RecordReader {
  public:
    RecordReader(const std::string &path, size_t count) : path_(path), count_(count) {
        f_ = fopen(path.c_str(), "rb");
        if (!f_) throw EngineError(SFD_EIO, "cannot open " + path);
        // the marker width is the one whose trailing marker repeats the length
        if (fseek(f_, 0, SEEK_END) != 0) fail("seek failed");
        const uint64_t size = (uint64_t)ftell(f_);
        const uint64_t need = (uint64_t)count * 8;
        bool found = false;
        for (size_t w : {(size_t)4, (size_t)8}) {
            uint64_t head = 0, tail = 0;
            if (size < 2 * w || !at(0, &head, w) || head >= size || size - 2 * w < head || !at(w + head, &tail, w) ||
                tail != head)
                continue;
            if (w == 4 && head >= 0x80000000u) continue;   // a subrecord marker
            len_ = head;
            marker_ = w;
            found = true;
            break;
        }
        if (!found) fail("not a Fortran sequential unformatted file (record markers do not match)");
        if (len_ < need) fail("its first record holds fewer than " + std::to_string(count) + " doubles");
        if (fseek(f_, (long)marker_, SEEK_SET) != 0) fail("seek failed");
    }
    ~RecordReader() {
        if (f_) fclose(f_);
    }
    // The next k values of the record.
    void read(double *dst, size_t k) {
        if (done_ + k > count_) fail("read past the expected values");
        if (k && fread(dst, sizeof(double), k, f_) != k) fail("truncated record");
        done_ += k;
    }
  private:
    std::string path_;
    size_t count_, done_ = 0, marker_ = 4;
    uint64_t len_ = 0;
    FILE *f_ = nullptr;
    [[noreturn]] void fail(const std::string &why) { throw EngineError(SFD_EIO, path_ + ": " + why); }
    bool at(uint64_t off, uint64_t *v, size_t w) {
        unsigned char m[8] = {0};
        if (fseek(f_, (long)off, SEEK_SET) != 0 || fread(m, 1, w, f_) != w) return false;
        *v = 0;
        memcpy(v, m, w);
        return true;
    }
};

class RecordWriter {
  public:
    RecordWriter(const std::string &path, size_t count) : path_(path), tmp_(path + ".tmp"), count_(count) {
        const uint64_t len = (uint64_t)count * 8;
        if (len >= 0x80000000ull)
            throw EngineError(SFD_EINVAL, path + ": a record of " + std::to_string(len) +
                                              " bytes exceeds one 4-byte record marker");
        len_ = (uint32_t)len;
        f_ = fopen(tmp_.c_str(), "wb");
        if (!f_) throw EngineError(SFD_EIO, "cannot create " + tmp_);
        put(&len_, 4);
    }
    ~RecordWriter() {
        if (f_) {
            fclose(f_);
            remove(tmp_.c_str());
        }
    }
    void write(const double *src, size_t k) {
        if (done_ + k > count_) throw EngineError(SFD_EINVAL, path_ + ": more values than the record holds");
        put(src, k * sizeof(double));
        done_ += k;
    }
    void close() {
        if (done_ != count_) throw EngineError(SFD_EINVAL, path_ + ": record not filled");
        put(&len_, 4);
        FILE *f = f_;
        f_ = nullptr;
        if (fclose(f) != 0) throw EngineError(SFD_EIO, "write failed: " + tmp_);
        if (rename(tmp_.c_str(), path_.c_str()) != 0) throw EngineError(SFD_EIO, "cannot rename to " + path_);
    }

  private:
    std::string path_, tmp_;
    size_t count_, done_ = 0;
    uint32_t len_ = 0;
    FILE *f_ = nullptr;
    void put(const void *p, size_t b) {
        if (b && fwrite(p, 1, b, f_) != b) throw EngineError(SFD_EIO, "write failed: " + tmp_);
    }
};

}  // namespace ftn
