// Dataset and variant files of the reference's callers, host only.
//
//   x,y CSV           src/c_test_prog.cpp:86-104   (fgets + sscanf "%[^,],%[^,]" + atof)
//   lon,lat[,TEC] CSV src/gpu/import_dataset.cpp:63-134 (latitude moved to [0, 180])
//   eps,minpts CSV    src/gpu/import_dataset.cpp:232-262 (atof, atoi)
//
// Fields are parsed with the C library's strtod/strtol, which is what atof and
// atoi are defined as, so every value equals the reference's. The reference
// copies each field into a 20-byte buffer (longer fields overflow it) and turns
// a line without a comma into a point built from the previous line's fields;
// here fields may have any length and lines without a comma are skipped.
#include "vdbscan.h"
#include "vdb_common.h"
#include "vdb_context.h"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace vdb;

namespace {

// Splits a line into its first `want` comma-separated fields (the last one
// runs to the end of the line). Returns the number found.
int split(const std::string &line, int want, std::string *out)
{
    size_t a = 0;
    for (int f = 0; f < want; f++) {
        if (a > line.size()) return f;
        size_t e = (f + 1 < want) ? line.find(',', a) : std::string::npos;
        if (e == std::string::npos) {
            if (f + 1 < want) {
                // fewer commas than fields asked for
                out[f] = line.substr(a);
                return f + 1;
            }
            e = line.size();
        }
        out[f] = line.substr(a, e - a);
        a = e + 1;
    }
    return want;
}

double atof_like(const std::string &s) { return strtod(s.c_str(), nullptr); }

int atoi_like(const std::string &s) { return (int)strtol(s.c_str(), nullptr, 10); }

// Reads every line of path; calls f(line) for each (without the newline).
template <class F> int for_lines(const char *path, F &&f)
{
    if (!path) {
        set_error("NULL path");
        return kErrInvalid;
    }
    FILE *in = fopen(path, "r");
    if (!in) {
        set_error((std::string("cannot open ") + path + ": " + strerror(errno)).c_str());
        return kErrInvalid;
    }
    std::string line;
    char buf[4096];
    bool ok = true;
    while (fgets(buf, sizeof buf, in)) {
        line += buf;
        if (!line.empty() && line.back() == '\n') {
            line.pop_back();
            if (!line.empty() && line.back() == '\r') line.pop_back();
            ok = f(line);
            line.clear();
            if (!ok) break;
        }
    }
    if (ok && !line.empty()) ok = f(line);
    fclose(in);
    if (!ok) {
        set_error("too many lines");
        return kErrInvalid;
    }
    return kOk;
}

template <class T> int give(const std::vector<T> &v, T **out)
{
    *out = (T *)malloc(sizeof(T) * (v.empty() ? 1 : v.size()));
    if (!*out) {
        set_error("host allocation failed");
        return kErrNoMemory;
    }
    if (!v.empty()) memcpy(*out, v.data(), sizeof(T) * v.size());
    return kOk;
}

// latitude in [-90, 90] to [0, 180] (90 -> 0, -90 -> 180), with the
// prototype's operations (src/gpu/import_dataset.cpp:84-97)
double lat_to_x(double lat)
{
    if (lat > 0.0) {
        double d = (90.0 - lat) / 90.0;
        return d * 90.0;
    }
    double d = (-lat) / 90.0;
    return (d * 90.0) + 90.0;
}

int read_points(const char *path, bool lonlat, double **x, double **y, uint32_t *n)
{
    if (!x || !y || !n) {
        set_error("NULL output argument");
        return kErrInvalid;
    }
    std::vector<double> vx, vy;
    std::string f[2];
    int s = for_lines(path, [&](const std::string &line) {
        if (split(line, 2, f) < 2 || line.find(',') == std::string::npos) return true;
        if (vx.size() >= 0xffffffffull) return false;
        if (lonlat) {
            vx.push_back(lat_to_x(atof_like(f[1].substr(0, f[1].find(',')))));
            vy.push_back(atof_like(f[0]));
        } else {
            vx.push_back(atof_like(f[0]));
            vy.push_back(atof_like(f[1]));
        }
        return true;
    });
    if (s != kOk) return s;
    VDB_TRY(give(vx, x));
    int t = give(vy, y);
    if (t != kOk) {
        free(*x);
        *x = nullptr;
        return t;
    }
    *n = (uint32_t)vx.size();
    return kOk;
}

} // namespace

extern "C" {

int vdb_read_xy_csv(const char *path, double **x, double **y, uint32_t *n)
{
    return read_points(path, false, x, y, n);
}

int vdb_read_lonlat_csv(const char *path, double **x, double **y, uint32_t *n)
{
    return read_points(path, true, x, y, n);
}

int vdb_read_variants_csv(const char *path, double **eps, uint32_t **minpts, uint32_t *nvar)
{
    if (!eps || !minpts || !nvar) {
        set_error("NULL output argument");
        return kErrInvalid;
    }
    std::vector<double> ve;
    std::vector<uint32_t> vm;
    std::string f[2];
    int s = for_lines(path, [&](const std::string &line) {
        if (split(line, 2, f) < 2 || line.find(',') == std::string::npos) return true;
        if (ve.size() >= 0xffffffffull) return false;
        ve.push_back(atof_like(f[0]));
        // the reference stores atoi's int in an unsigned field
        vm.push_back((uint32_t)atoi_like(f[1]));
        return true;
    });
    if (s != kOk) return s;
    VDB_TRY(give(ve, eps));
    int t = give(vm, minpts);
    if (t != kOk) {
        free(*eps);
        *eps = nullptr;
        return t;
    }
    *nvar = (uint32_t)ve.size();
    return kOk;
}

int vdb_write_labels_csv(const char *path, const double *x, const double *y, uint32_t n,
                         const uint32_t *labels, uint32_t nvar)
{
    if (!path || (n && (!x || !y || (nvar && !labels)))) {
        set_error("NULL argument");
        return kErrInvalid;
    }
    FILE *out = strcmp(path, "-") == 0 ? stdout : fopen(path, "w");
    if (!out) {
        set_error((std::string("cannot open ") + path + ": " + strerror(errno)).c_str());
        return kErrInvalid;
    }
    for (uint32_t i = 0; i < n; i++) {
        fprintf(out, "%.17g,%.17g", x[i], y[i]);
        for (uint32_t v = 0; v < nvar; v++) fprintf(out, ",%u", labels[(size_t)v * n + i]);
        fputc('\n', out);
    }
    int bad = ferror(out);
    if (out != stdout) bad |= fclose(out);
    else fflush(out);
    if (bad) {
        set_error("write failed");
        return kErrInvalid;
    }
    return kOk;
}

void vdb_free(void *p) { free(p); }

} // extern "C"
