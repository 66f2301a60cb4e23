// sep.hip — SEP files on disk (the I/O the reference does through SEPlib:
// from_aux / sreed / srite / to_history, libs/libwave/IO.f90:14-51,
// ReadSouVelRho.f90:59-297) and the shot geometry of a coordinate file
// (readgathercoords, number_of_shots, number_of_traces,
// ReadSouVelRho.f90:161-241, :828-890).  Host code only.
//
// A SEP history file is text: assignments name=value, several to a line,
// values optionally in double or single quotes; a later assignment of a name
// overrides an earlier one (each program appends its own lines).  `in` names
// the binary file; `esize` is the bytes per element (4 here); `data_format`
// is "xdr_float" (big-endian IEEE, SEPlib's default when absent) or
// "native_float".  n1..n6 / o1..o6 / d1..d6 describe the axes (absent: n 1,
// o 0, d 1).
#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/seisfd.h"
#include "common.h"

namespace sep {

using Dict = std::map<std::string, std::string>;

// The assignments of a history file, later ones winning.
Dict parse_history(const std::string &path) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) throw EngineError(SFD_EIO, "cannot open " + path);
    std::string text;
    char buf[65536];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    fclose(f);
    // a header followed by its data in the same file ends at the EOT bytes
    const size_t eot = text.find('\004');
    if (eot != std::string::npos) text.resize(eot);
    Dict d;
    size_t i = 0;
    const size_t n = text.size();
    auto is_name = [](char c) { return isalnum((unsigned char)c) || c == '_' || c == '.'; };
    while (i < n) {
        if (text[i] == '#') {   // comment to the end of the line
            while (i < n && text[i] != '\n') i++;
            continue;
        }
        if (!is_name(text[i])) {
            i++;
            continue;
        }
        size_t s = i;
        while (i < n && is_name(text[i])) i++;
        if (i >= n || text[i] != '=') continue;   // a word that is not an assignment
        const std::string name = text.substr(s, i - s);
        i++;
        std::string val;
        if (i < n && (text[i] == '"' || text[i] == '\'')) {
            const char q = text[i++];
            const size_t e = text.find(q, i);
            if (e == std::string::npos) throw EngineError(SFD_EIO, path + ": unterminated quote for " + name);
            val = text.substr(i, e - i);
            i = e + 1;
        } else {
            s = i;
            while (i < n && !isspace((unsigned char)text[i])) i++;
            val = text.substr(s, i - s);
        }
        d[name] = val;
    }
    return d;
}

bool get(const Dict &d, const std::string &k, std::string *v) {
    auto it = d.find(k);
    if (it == d.end()) return false;
    *v = it->second;
    return true;
}

// The directory of a path (for a relative `in`, resolved as SEPlib does:
// relative to the working directory, or next to the header when that fails).
std::string dir_of(const std::string &p) {
    const size_t s = p.find_last_of('/');
    return s == std::string::npos ? std::string() : p.substr(0, s + 1);
}

void read_header(const std::string &path, sfd_sep_header *h) {
    const Dict d = parse_history(path);
    memset(h, 0, sizeof *h);
    std::string v;
    h->ndim = 1;
    for (int a = 0; a < 6; a++) {
        const std::string k = std::to_string(a + 1);
        h->n[a] = 1;
        h->o[a] = 0.f;
        h->d[a] = 1.f;
        if (get(d, "n" + k, &v)) {
            char *e = nullptr;
            const long long x = strtoll(v.c_str(), &e, 10);
            if (e == v.c_str() || *e || x < 1) throw EngineError(SFD_EIO, path + ": bad n" + k + "=" + v);
            h->n[a] = x;
            if (x > 1) h->ndim = a + 1;
        }
        if (get(d, "o" + k, &v)) h->o[a] = strtof(v.c_str(), nullptr);
        if (get(d, "d" + k, &v)) h->d[a] = strtof(v.c_str(), nullptr);
    }
    h->esize = 4;
    if (get(d, "esize", &v)) h->esize = atoi(v.c_str());
    if (h->esize != 4) throw EngineError(SFD_EUNSUPPORTED, path + ": esize " + v + " (only 4-byte floats)");
    h->xdr = 1;
    if (get(d, "data_format", &v)) {
        if (v == "native_float") h->xdr = 0;
        else if (v == "xdr_float") h->xdr = 1;
        else throw EngineError(SFD_EUNSUPPORTED, path + ": data_format " + v);
    }
    if (!get(d, "in", &v) || v.empty()) throw EngineError(SFD_EIO, path + ": no in= (binary file)");
    if (v == "stdin" || v == "-") throw EngineError(SFD_EUNSUPPORTED, path + ": data on stdin");
    std::string p = v;
    if (p[0] != '/') {
        FILE *t = fopen(p.c_str(), "rb");
        if (t) fclose(t);
        else p = dir_of(path) + v;
    }
    if (p.size() >= sizeof h->in) throw EngineError(SFD_EINVAL, path + ": data path too long");
    memcpy(h->in, p.c_str(), p.size() + 1);
}

size_t elements(const sfd_sep_header &h) {
    size_t c = 1;
    for (int a = 0; a < 6; a++) c *= (size_t)h.n[a];
    return c;
}

void swap4(float *x, size_t n) {
    for (size_t i = 0; i < n; i++) {
        uint32_t u;
        memcpy(&u, x + i, 4);
        u = __builtin_bswap32(u);
        memcpy(x + i, &u, 4);
    }
}

void read_data(const sfd_sep_header &h, size_t offset, size_t count, float *out) {
    if (offset + count > elements(h)) throw EngineError(SFD_EINVAL, std::string(h.in) + ": read past the data");
    FILE *f = fopen(h.in, "rb");
    if (!f) throw EngineError(SFD_EIO, std::string("cannot open ") + h.in);
    const bool ok = fseeko(f, (off_t)(offset * 4), SEEK_SET) == 0 && fread(out, 4, count, f) == count;
    fclose(f);
    if (!ok) throw EngineError(SFD_EIO, std::string(h.in) + ": shorter than its header says");
    if (h.xdr) swap4(out, count);
}

// Writes the history file (one line of axes, then in / esize / data_format)
// and the binary data (native little-endian floats), each through a
// temporary name and a rename.
void write(const std::string &path, const sfd_sep_header &h, const float *data, size_t count) {
    if (count != elements(h)) throw EngineError(SFD_EINVAL, path + ": count differs from the header's n1*n2*...");
    std::string bin = h.in[0] ? std::string(h.in) : path + "@";
    {
        const std::string tmp = bin + ".tmp";
        FILE *f = fopen(tmp.c_str(), "wb");
        if (!f) throw EngineError(SFD_EIO, "cannot create " + tmp);
        const bool ok = fwrite(data, 4, count, f) == count;
        if (fclose(f) != 0 || !ok) throw EngineError(SFD_EIO, "write failed: " + tmp);
        if (rename(tmp.c_str(), bin.c_str()) != 0) throw EngineError(SFD_EIO, "cannot rename to " + bin);
    }
    std::string t;
    char b[160];
    for (int a = 0; a < 6; a++) {
        if (a >= std::max(1, h.ndim) && h.n[a] <= 1) break;
        snprintf(b, sizeof b, "%sn%d=%lld o%d=%.9g d%d=%.9g", a ? " " : "\t", a + 1, (long long)h.n[a], a + 1,
                 (double)h.o[a], a + 1, (double)h.d[a]);
        t += b;
    }
    t += "\n\tin=\"" + bin + "\" esize=4 data_format=\"native_float\"\n";
    const std::string tmp = path + ".tmp";
    FILE *f = fopen(tmp.c_str(), "w");
    if (!f) throw EngineError(SFD_EIO, "cannot create " + tmp);
    const bool ok = fwrite(t.data(), 1, t.size(), f) == t.size();
    if (fclose(f) != 0 || !ok) throw EngineError(SFD_EIO, "write failed: " + tmp);
    if (rename(tmp.c_str(), path.c_str()) != 0) throw EngineError(SFD_EIO, "cannot rename to " + path);
}

}  // namespace sep

extern "C" {

int sfd_sep_read_header(const char *hfile, sfd_sep_header *hdr) {
    if (!hfile || !hdr) return SFD_EINVAL;
    return abi_call(nullptr, [&]() { sep::read_header(hfile, hdr); });
}

int sfd_sep_read(const char *hfile, size_t offset, size_t count, float *data) {
    if (!hfile || (count && !data)) return SFD_EINVAL;
    return abi_call(nullptr, [&]() {
        sfd_sep_header h;
        sep::read_header(hfile, &h);
        sep::read_data(h, offset, count, data);
    });
}

int sfd_sep_write(const char *hfile, const sfd_sep_header *hdr, const float *data, size_t count) {
    if (!hfile || !hdr || (count && !data)) return SFD_EINVAL;
    return abi_call(nullptr, [&]() { sep::write(hfile, *hdr, data, count); });
}

int sfd_sep_gathers(const char *coordfile, const int keys[6], int *nshots, int *shot_first, int *shot_ntr,
                    float *sz, float *sx, float *rz, float *rx, int cap_shots, int cap_traces) {
    if (!coordfile || !keys || !nshots) return SFD_EINVAL;
    return abi_call(nullptr, [&]() {
        sfd_sep_header h;
        sep::read_header(coordfile, &h);
        const long nkeys = (long)h.n[0], ntr = (long)(sep::elements(h) / h.n[0]);
        for (int k = 0; k < 6; k++)
            if (keys[k] < 1 || keys[k] > nkeys)
                throw EngineError(SFD_EINVAL, "key index " + std::to_string(keys[k]) + " outside 1.." +
                                                  std::to_string(nkeys) + " of the coordinate file");
        std::vector<float> c((size_t)nkeys * ntr);
        sep::read_data(h, 0, c.size(), c.data());
        // keys[1] (gy) is checked but not used: the y coordinate is 0 in 2D
        const int kgx = keys[0] - 1, ksx = keys[2] - 1, ksy = keys[3] - 1, kse = keys[4] - 1, kge = keys[5] - 1;
        auto key = [&](long t, int k) { return c[(size_t)t * nkeys + k]; };
        // number_of_shots / number_of_traces: a shot is a run of consecutive
        // traces with the same (sx, sy)
        std::vector<long> first;
        for (long t = 0; t < ntr; t++)
            if (t == 0 || key(t, ksx) != key(t - 1, ksx) || key(t, ksy) != key(t - 1, ksy)) first.push_back(t);
        *nshots = (int)first.size();
        if (!shot_first) return;   // a query of the counts
        if ((long)first.size() > cap_shots || ntr > cap_traces)
            throw EngineError(SFD_EINVAL, "output arrays smaller than the shots or traces of the file");
        for (size_t s = 0; s < first.size(); s++) {
            const long b = first[s], e = s + 1 < first.size() ? first[s + 1] : ntr;
            shot_first[s] = (int)b;
            shot_ntr[s] = (int)(e - b);
            // the source's keys come from the shot's last trace (readgathercoords)
            if (sz) sz[s] = key(e - 1, kse);
            if (sx) sx[s] = key(e - 1, ksx);
        }
        for (long t = 0; t < ntr; t++) {
            if (rz) rz[t] = key(t, kge);
            if (rx) rx[t] = key(t, kgx);
        }
    });
}

}  // extern "C"
