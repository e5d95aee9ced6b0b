// Harness that times one libVDBSCAN call of the compiled reference library.
//
// TEST INFRASTRUCTURE ONLY (oracle side). The reference library keeps global
// per-thread buffers and calls omp_set_num_threads (src/DBScan.cpp:51,
// src/main.cpp:58), so every call runs in a fresh process of this program.
//
// usage: ref_runner <lib.so> <input.bin> <output.bin>
//   input : u32 n, u32 nvar, i32 mbb, u32 verbose, f64 x[n], f64 y[n],
//           f64 eps[nvar], u32 minpts[nvar]
//   output: u32 labels[nvar*n]  (retArr layout of src/main.cpp:273-284)
// prints one JSON line: {"seconds": wall time of the libVDBSCAN call}
#include <dlfcn.h>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

typedef int (*lib_fn)(double *, double *, unsigned int, double *, unsigned int *,
                      unsigned int, int, unsigned int *, bool);

static bool read_all(FILE *f, void *dst, size_t bytes)
{
    return bytes == 0 || fread(dst, 1, bytes, f) == bytes;
}

int main(int argc, char **argv)
{
    if (argc != 4) {
        fprintf(stderr, "usage: %s lib.so in.bin out.bin\n", argv[0]);
        return 2;
    }
    void *h = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
    if (!h) { fprintf(stderr, "dlopen: %s\n", dlerror()); return 3; }
    lib_fn fn = (lib_fn)dlsym(h, "libVDBSCAN");
    if (!fn) { fprintf(stderr, "dlsym: %s\n", dlerror()); return 3; }
    FILE *f = fopen(argv[2], "rb");
    if (!f) { perror("input"); return 4; }
    uint32_t n = 0, nvar = 0, verbose = 0;
    int32_t mbb = 70;
    if (!read_all(f, &n, 4) || !read_all(f, &nvar, 4) || !read_all(f, &mbb, 4) ||
        !read_all(f, &verbose, 4)) { fprintf(stderr, "short header\n"); return 4; }
    std::vector<double> x(n), y(n), eps(nvar);
    std::vector<uint32_t> mp(nvar);
    std::vector<uint32_t> ret((size_t)n * nvar, 0xFFFFFFFFu);
    if (!read_all(f, x.data(), 8ull * n) || !read_all(f, y.data(), 8ull * n) ||
        !read_all(f, eps.data(), 8ull * nvar) || !read_all(f, mp.data(), 4ull * nvar)) {
        fprintf(stderr, "short body\n");
        return 4;
    }
    fclose(f);
    auto t0 = std::chrono::steady_clock::now();
    int rc = fn(x.data(), y.data(), n, eps.data(), mp.data(), nvar, mbb, ret.data(),
                verbose != 0);
    auto t1 = std::chrono::steady_clock::now();
    double sec = std::chrono::duration<double>(t1 - t0).count();
    FILE *o = fopen(argv[3], "wb");
    if (!o) { perror("output"); return 5; }
    fwrite(ret.data(), 4, ret.size(), o);
    fclose(o);
    printf("{\"seconds\": %.9f, \"rc\": %d}\n", sec, rc);
    return 0;
}
