// Drives the C++ entry points (include/vdbscan.hpp) the way the reference's
// libVDBSCAN drives its own classes (src/main.cpp:65-284): sort the variants,
// bin the data, one DBScan per variant, a Schedule handing out the jobs, then
// labels mapped back to caller order. Prints, per caller variant, a checksum
// line "v <id> <clusters> <sum of labels> <noise>" and writes all labels to
// the file given as argv[2] for comparison with vdb_cluster.
//
// usage: api_check <points.bin> <labels.out> [scratch]
//   points.bin: u32 n, u32 nvar, f64 x[n], f64 y[n], f64 eps[nvar], u32 minpts[nvar]
//   scratch: every variant through algDBScanParallel (no reuse)
#include "vdbscan.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdint>
#include <string>
#include <vector>

int main(int argc, char **argv)
{
    if (argc != 3 && argc != 4) return 2;
    bool scratch = argc == 4 && std::string(argv[3]) == "scratch";
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 3;
    uint32_t n = 0, nv = 0;
    if (fread(&n, 4, 1, f) != 1 || fread(&nv, 4, 1, f) != 1) return 3;
    std::vector<double> x(n), y(n), eps(nv);
    std::vector<uint32_t> mp(nv);
    if (fread(x.data(), 8, n, f) != n || fread(y.data(), 8, n, f) != n ||
        fread(eps.data(), 8, nv, f) != nv || fread(mp.data(), 4, nv, f) != nv)
        return 3;
    fclose(f);

    std::vector<experiment> ex;
    for (uint32_t v = 0; v < nv; v++) ex.push_back({eps[v], (int)mp[v], v});
    std::sort(ex.begin(), ex.end(), compareByEpsilon);
    std::vector<dataElem> data(n);
    for (uint32_t i = 0; i < n; i++) data[i] = {x[i], y[i]};
    vdbscan::PointIndex index(&data);
    std::vector<DBScan *> inst;
    for (auto &e : ex) inst.push_back(new DBScan(&data, e.epsilon, e.minpts, &index));
    for (auto *d : inst) d->instanceVect = inst;
    Schedule sched(&ex);
    for (size_t i = 0; i < 2 * ex.size(); i++) {
        int reuse = 0, clus = 0;
        if (!sched.determineReuse((int)i, &reuse, &clus)) continue;
        if (reuse == clus || scratch) inst[clus]->algDBScanParallel();
        else inst[clus]->assignPointsToPredefinedCluster(reuse);
        sched.finishedList[clus] = true;
    }
    std::vector<uint32_t> ret((size_t)n * nv);
    const std::vector<int> &m = index.mapping();
    for (size_t i = 0; i < ex.size(); i++) {
        uint32_t *o = ret.data() + (size_t)ex[i].variantID * n;
        for (uint32_t k = 0; k < n; k++) o[m[k]] = (uint32_t)inst[i]->clusterIDs[k];
    }
    for (uint32_t v = 0; v < nv; v++) {
        const uint32_t *o = ret.data() + (size_t)v * n;
        uint64_t s = 0, z = 0, mx = 0;
        for (uint32_t k = 0; k < n; k++) {
            s += o[k];
            z += o[k] == 0;
            mx = o[k] > mx ? o[k] : mx;
        }
        printf("v %u %llu %llu %llu\n", v, (unsigned long long)mx, (unsigned long long)s,
               (unsigned long long)z);
    }
    FILE *g = fopen(argv[2], "wb");
    if (!g) return 4;
    fwrite(ret.data(), 4, ret.size(), g);
    fclose(g);
    for (auto *d : inst) delete d;
    return 0;
}
