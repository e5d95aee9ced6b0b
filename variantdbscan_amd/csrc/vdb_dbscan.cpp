// C++ entry points (include/vdbscan.hpp) over the library.
#include "vdbscan.h"
#include "vdbscan.hpp"

#include "vdb_reuse.h"

#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

namespace {

void check(int rc, const char *what)
{
    if (rc) throw std::runtime_error(std::string(what) + ": " + vdb_last_error());
}

void check_hip(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

} // namespace

namespace vdbscan {

PointIndex::PointIndex(std::vector<dataElem> *ptrData) : data_(ptrData)
{
    size_t n = ptrData->size();
    n_ = (uint32_t)n;
    std::vector<double> x(n), y(n);
    for (size_t i = 0; i < n; i++) {
        x[i] = (*ptrData)[i].x;
        y[i] = (*ptrData)[i].y;
    }
    // the points stay on the device in the order handed in: every variant
    // over this index is clustered from them without another upload
    if (n) {
        check_hip(hipMalloc(&dx_, 8 * n), "hipMalloc");
        check_hip(hipMalloc(&dy_, 8 * n), "hipMalloc");
        check_hip(hipMemcpy(dx_, x.data(), 8 * n, hipMemcpyHostToDevice), "hipMemcpy");
        check_hip(hipMemcpy(dy_, y.data(), 8 * n, hipMemcpyHostToDevice), "hipMemcpy");
    }
    std::vector<uint32_t> m(n);
    check(vdb_rank_order(x.data(), y.data(), (uint32_t)n, m.data(), 0), "vdb_rank_order");
    // reorder the data in place, like binDataset (src/main.cpp:573-595)
    mapping_.resize(n);
    for (size_t k = 0; k < n; k++) {
        mapping_[k] = (int)m[k];
        (*ptrData)[k].x = x[m[k]];
        (*ptrData)[k].y = y[m[k]];
    }
}

PointIndex::~PointIndex()
{
    if (dx_) (void)hipFree(dx_);
    if (dy_) (void)hipFree(dy_);
    if (dlab_) (void)hipFree(dlab_);
}

void PointIndex::cluster(const std::vector<DBScan *> &batch)
{
    uint32_t nv = (uint32_t)batch.size();
    if (nv == 0) return;
    std::vector<double> eps(nv);
    std::vector<uint32_t> mp(nv);
    for (uint32_t v = 0; v < nv; v++) {
        eps[v] = batch[v]->distance;
        mp[v] = (uint32_t)batch[v]->minPts; // an int below 0 wraps above INT_MAX: never core
    }
    std::vector<uint32_t> lab((size_t)n_ * nv);
    if (n_) {
        size_t need = 4 * (size_t)n_ * nv;
        if (need > dlab_bytes_) {
            if (dlab_) check_hip(hipFree(dlab_), "hipFree");
            dlab_ = nullptr;
            check_hip(hipMalloc(&dlab_, need), "hipMalloc");
            dlab_bytes_ = need;
        }
        // all variants in one call: the batched engine when it applies
        check(vdb_cluster(dx_, dy_, n_, eps.data(), mp.data(), nv, (uint32_t *)dlab_,
                          VDB_DEVICE_PTRS, nullptr, nullptr),
              "vdb_cluster");
        check_hip(hipMemcpy(lab.data(), dlab_, need, hipMemcpyDeviceToHost), "hipMemcpy");
    }
    for (uint32_t v = 0; v < nv; v++) {
        DBScan &d = *batch[v];
        const uint32_t *l = lab.data() + (size_t)v * n_;
        uint32_t mx = 0;
        for (uint32_t k = 0; k < n_; k++) {
            d.clusterIDs[k] = (int)l[mapping_[k]]; // by binned position
            mx = l[mapping_[k]] > mx ? l[mapping_[k]] : mx;
        }
        d.clusterCnt = (int)mx + 1;
        d.numClustersForStats = d.clusterCnt;
        d.state_ = DBScan::kScratchReady;
    }
}

void PointIndex::reuse(DBScan &d, const DBScan &old)
{
    std::vector<uint32_t> o(n_), lab(n_);
    for (uint32_t k = 0; k < n_; k++) o[k] = (uint32_t)old.clusterIDs[k];
    uint32_t ncl = 1;
    check(vdb::reuse_labels(dx_, dy_, n_, d.distance, (uint32_t)d.minPts, o.data(),
                            (uint32_t)old.clusterCnt, lab.data(), &ncl),
          "reuse");
    for (uint32_t k = 0; k < n_; k++) d.clusterIDs[k] = (int)lab[k];
    d.clusterCnt = (int)ncl;
    d.numClustersForStats = d.clusterCnt;
}

} // namespace vdbscan

DBScan::DBScan(std::vector<dataElem> *ptrData, double epsilon, int minimumPts,
               vdbscan::PointIndex *idx)
    : clusterIDs(ptrData->size(), 0), dataPoints(ptrData), distance(epsilon),
      minPts(minimumPts), index(idx)
{
    // labels are computed over the index's copy of the points: the data
    // handed in must be the binned data the index was built from
    if (!idx || idx->data() != dataPoints)
        throw std::invalid_argument("DBScan: the data must be the index's own (binned) data");
}

void DBScan::algDBScanParallel()
{
    if (state_ == kScratchReady) {
        // computed with the other variants of its index in one batched call
        state_ = kDone;
        return;
    }
    // this variant and every variant over the same index not run yet
    std::vector<DBScan *> batch{this};
    for (DBScan *d : instanceVect)
        if (d && d != this && d->index == index && d->state_ == kNotRun) batch.push_back(d);
    index->cluster(batch);
    state_ = kDone;
}

void DBScan::assignPointsToPredefinedCluster(int finishedInstanceID)
{
    if (finishedInstanceID < 0 || (size_t)finishedInstanceID >= instanceVect.size() ||
        !instanceVect[finishedInstanceID] || instanceVect[finishedInstanceID]->index != index)
        throw std::runtime_error("assignPointsToPredefinedCluster: no finished variant " +
                                 std::to_string(finishedInstanceID) + " over this index");
    index->reuse(*this, *instanceVect[finishedInstanceID]);
    state_ = kDone;
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
