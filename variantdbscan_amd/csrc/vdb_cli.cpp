// vdbscan: command-line front end of libSharedVDBSCAN.so.
//
// Reads a point file (x,y CSV like the reference's harness, src/c_test_prog.cpp,
// or lon,lat[,TEC] CSV like the GPU prototype, src/gpu/import_dataset.cpp) and
// the variants (an eps,minpts CSV or --eps/--minpts lists), clusters every
// variant and writes the labels: x,y,label_0,...,label_{V-1} per point, or
// with --harness the text the reference's harness prints
// (src/c_test_prog.cpp:144-153). --standard gives textbook DBSCAN labels
// instead (-1 = noise, clusters 0.., every border point in a cluster).
#include "vdbscan.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

void usage()
{
    fprintf(stderr,
            "usage: vdbscan --points FILE [--lonlat] (--variants FILE | --eps E1,E2,.. "
            "--minpts M1,M2,..)\n"
            "               [--out FILE] [--harness] [--path auto|fused|table|batch] "
            "[--fp32] [--standard] [--verbose]\n");
}

bool parse_list(const char *s, std::vector<double> &out)
{
    const char *p = s;
    while (*p) {
        char *e = nullptr;
        double v = strtod(p, &e);
        if (e == p) return false;
        out.push_back(v);
        p = e;
        if (*p == ',') p++;
        else if (*p) return false;
    }
    return !out.empty();
}

// labels as signed values (the standard border rule's -1 = noise)
int write_signed_csv(const char *path, const double *x, const double *y, uint32_t n,
                     const uint32_t *labels, uint32_t nvar)
{
    FILE *f = strcmp(path, "-") == 0 ? stdout : fopen(path, "w");
    if (!f) return 1;
    for (uint32_t i = 0; i < n; i++) {
        fprintf(f, "%.17g,%.17g", x[i], y[i]);
        for (uint32_t v = 0; v < nvar; v++) fprintf(f, ",%d", (int32_t)labels[(size_t)v * n + i]);
        fputc('\n', f);
    }
    int bad = ferror(f);
    if (f != stdout) bad |= fclose(f);
    else fflush(f);
    return bad ? 1 : 0;
}

int fail(int rc)
{
    fprintf(stderr, "vdbscan: %s (code %d)\n", vdb_last_error(), rc);
    return 1;
}

} // namespace

int main(int argc, char **argv)
{
    const char *points = nullptr, *variants = nullptr, *out = "-";
    std::vector<double> eps_l, mp_l;
    bool lonlat = false, harness = false, verbose = false;
    vdb_options opt;
    memset(&opt, 0, sizeof opt);
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
        if (a == "--points") points = next();
        else if (a == "--variants") variants = next();
        else if (a == "--out") out = next();
        else if (a == "--lonlat") lonlat = true;
        else if (a == "--harness") harness = true;
        else if (a == "--verbose") verbose = true;
        else if (a == "--fp32") opt.precision = VDB_PREC_FP32_RECHECK;
        else if (a == "--standard") opt.border = VDB_BORDER_STANDARD;
        else if (a == "--eps" || a == "--minpts") {
            const char *v = next();
            if (!v || !parse_list(v, a == "--eps" ? eps_l : mp_l)) {
                usage();
                return 2;
            }
        } else if (a == "--path") {
            const char *v = next();
            std::string p = v ? v : "";
            if (p == "auto") opt.path = VDB_PATH_AUTO;
            else if (p == "fused") opt.path = VDB_PATH_FUSED;
            else if (p == "table") opt.path = VDB_PATH_TABLE;
            else if (p == "batch") opt.path = VDB_PATH_BATCH;
            else {
                usage();
                return 2;
            }
        } else {
            usage();
            return 2;
        }
    }
    if (!points || !out || (!variants && (eps_l.empty() || eps_l.size() != mp_l.size()))) {
        usage();
        return 2;
    }
    opt.verbose = verbose;
    double *x = nullptr, *y = nullptr, *eps = nullptr;
    uint32_t *mp = nullptr, n = 0, nv = 0;
    int rc = lonlat ? vdb_read_lonlat_csv(points, &x, &y, &n) : vdb_read_xy_csv(points, &x, &y, &n);
    if (rc) return fail(rc);
    if (variants) {
        rc = vdb_read_variants_csv(variants, &eps, &mp, &nv);
        if (rc) return fail(rc);
    } else {
        nv = (uint32_t)eps_l.size();
        eps = (double *)malloc(sizeof(double) * nv);
        mp = (uint32_t *)malloc(sizeof(uint32_t) * nv);
        for (uint32_t v = 0; v < nv; v++) {
            eps[v] = eps_l[v];
            mp[v] = (uint32_t)mp_l[v];
        }
    }
    std::vector<uint32_t> labels((size_t)n * nv);
    vdb_stats st;
    rc = vdb_cluster(x, y, n, eps, mp, nv, labels.data(), 0, &opt, &st);
    if (rc) return fail(rc);
    if (verbose)
        fprintf(stderr, "vdbscan: %u points x %u variants in %.3f ms (device)\n", n, nv,
                st.ms_total);
    if (harness) {
        FILE *f = strcmp(out, "-") == 0 ? stdout : fopen(out, "w");
        if (!f) {
            fprintf(stderr, "vdbscan: cannot open %s\n", out);
            return 1;
        }
        for (uint32_t v = 0; v < nv; v++) {
            fprintf(f, "\n\nVariant: %u, Epsilon: %f, minpts: %u", v, eps[v], mp[v]);
            for (uint32_t j = 0; j < n; j++) {
                uint32_t l = labels[(size_t)v * n + j];
                if (opt.border == VDB_BORDER_STANDARD)
                    fprintf(f, "\n%u, %f, %f, %d", j, x[j], y[j], (int32_t)l);
                else
                    fprintf(f, "\n%u, %f, %f, %u", j, x[j], y[j], l);
            }
        }
        fprintf(f, "\n");
        if (f != stdout) fclose(f);
    } else if (opt.border == VDB_BORDER_STANDARD) {
        if (write_signed_csv(out, x, y, n, labels.data(), nv)) {
            fprintf(stderr, "vdbscan: cannot write %s\n", out);
            return 1;
        }
    } else {
        rc = vdb_write_labels_csv(out, x, y, n, labels.data(), nv);
        if (rc) return fail(rc);
    }
    vdb_free(x);
    vdb_free(y);
    free(eps);
    free(mp);
    return 0;
}
