// C++ entry points (include/vdbscan.hpp) over the C ABI of this library.
#include "vdbscan.h"
#include "vdbscan.hpp"

#include <mutex>
#include <stdexcept>
#include <string>

namespace vdbscan {

PointIndex::PointIndex(std::vector<dataElem> *ptrData) : data_(ptrData)
{
    size_t n = ptrData->size();
    std::vector<double> x(n), y(n);
    for (size_t i = 0; i < n; i++) {
        x[i] = (*ptrData)[i].x;
        y[i] = (*ptrData)[i].y;
    }
    std::vector<uint32_t> m(n);
    int rc = vdb_rank_order(x.data(), y.data(), (uint32_t)n, m.data(), 0);
    if (rc) throw std::runtime_error(std::string("vdb_rank_order: ") + vdb_last_error());
    // reorder the data in place, like binDataset (src/main.cpp:573-595)
    mapping_.resize(n);
    for (size_t k = 0; k < n; k++) {
        mapping_[k] = (int)m[k];
        (*ptrData)[k].x = x[m[k]];
        (*ptrData)[k].y = y[m[k]];
    }
}

} // namespace vdbscan

DBScan::DBScan(std::vector<dataElem> *ptrData, double epsilon, int minimumPts,
               vdbscan::PointIndex *idx)
    : clusterIDs(ptrData->size(), 0), dataPoints(ptrData), distance(epsilon),
      minPts(minimumPts), index(idx)
{
}

void DBScan::algDBScanParallel()
{
    size_t n = dataPoints->size();
    std::vector<double> x(n), y(n);
    for (size_t i = 0; i < n; i++) {
        x[i] = (*dataPoints)[i].x;
        y[i] = (*dataPoints)[i].y;
    }
    // the data are already in binned order, and binning is idempotent on
    // binned data only up to ties; cluster in the caller order of the index
    // and map back to binned positions
    std::vector<double> ox(n), oy(n);
    const std::vector<int> &m = index->mapping();
    for (size_t k = 0; k < n; k++) {
        ox[m[k]] = x[k];
        oy[m[k]] = y[k];
    }
    std::vector<uint32_t> lab(n);
    double e = distance;
    uint32_t mp = (uint32_t)minPts; // an int below 0 wraps above INT_MAX: never core
    int rc = vdb_cluster(ox.data(), oy.data(), (uint32_t)n, &e, &mp, 1, lab.data(), 0, nullptr,
                         nullptr);
    if (rc) throw std::runtime_error(std::string("vdb_cluster: ") + vdb_last_error());
    uint32_t mx = 0;
    for (size_t k = 0; k < n; k++) {
        clusterIDs[k] = (int)lab[m[k]];
        if (lab[m[k]] > mx) mx = lab[m[k]];
    }
    clusterCnt = (int)mx + 1;
    numClustersForStats = clusterCnt;
}

void DBScan::assignPointsToPredefinedCluster(int finishedInstanceID)
{
    (void)finishedInstanceID;
    algDBScanParallel();
}

int DBScan::getDBScanNumClusters() { return clusterCnt; }

Schedule::Schedule(std::vector<experiment> *experimentList) : expPtr(experimentList)
{
    finishedList.assign(experimentList->size(), false);
    for (size_t i = 0; i < experimentList->size(); i++) jobQueue.push((int)i);
}

static std::mutex g_sched_mu;

bool Schedule::determineReuse(int inExperID, int *outIDReuse, int *outInstanceCluster)
{
    (void)inExperID;
    int expID;
    {
        std::lock_guard<std::mutex> g(g_sched_mu);
        if (jobQueue.empty()) return false;
        expID = jobQueue.front();
        jobQueue.pop();
    }
    *outInstanceCluster = expID;
    *outIDReuse = expID;
    std::lock_guard<std::mutex> g(g_sched_mu);
    for (int i = expID - 1; i >= 0; i--) {
        if (finishedList[i] && (*expPtr)[i].minpts >= (*expPtr)[expID].minpts) {
            *outIDReuse = i;
            return true;
        }
    }
    return true;
}

bool compareByEpsilon(const experiment &a, const experiment &b)
{
    if (a.epsilon == b.epsilon) return a.minpts > b.minpts;
    return a.epsilon < b.epsilon;
}
