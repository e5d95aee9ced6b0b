// C++ entry points of the reference, kept by name, on top of the HIP engine.
//
//   experiment, dataElem   src/structs.h:35-39, :51-54 (same members)
//   DBScan                 src/DBScan.h:34-145  (constructor, algDBScanParallel,
//                          assignPointsToPredefinedCluster, getDBScanNumClusters,
//                          clusterIDs, instanceVect, numClustersForStats)
//   Schedule               src/schedule.h:36-67 (constructor, determineReuse,
//                          finishedList, expPtr)
//
// The reference's two R-tree pointers and the box lookup table become one
// PointIndex handle: the dataset in the reference's binned visit order
// (src/main.cpp:107), which every DBScan object over the same data shares.
// clusterIDs are indexed like the reference's: by binned position (use
// PointIndex::mapping() to go back to caller order, as src/main.cpp:255-262).
#ifndef VDBSCAN_HPP
#define VDBSCAN_HPP

#include <cstddef>
#include <cstdint>
#include <queue>
#include <vector>

struct experiment {
    double epsilon;
    int minpts;
    unsigned int variantID;
};

struct dataElem {
    double x;
    double y;
};

class DBScan;

namespace vdbscan {

// The data in binned visit order; replaces the MPB R-tree, the high-resolution
// R-tree and the MPB lookup of src/main.cpp:110-165. A copy of the points
// stays in device memory for the index's lifetime, so the variants over it
// are clustered without uploading them again.
class PointIndex {
public:
    // Bins *ptrData in place into the reference's visit order (src/main.cpp:459-605),
    // as binDataset does. Throws std::runtime_error on device failure.
    explicit PointIndex(std::vector<dataElem> *ptrData);
    ~PointIndex();
    PointIndex(const PointIndex &) = delete;
    PointIndex &operator=(const PointIndex &) = delete;
    // mapping()[k] = index, in the data as passed in, of binned point k
    const std::vector<int> &mapping() const { return mapping_; }
    std::vector<dataElem> *data() const { return data_; }

    // Scratch labels of several variants over this index in one library call
    // (the batched engine when it applies); fills each one's clusterIDs.
    void cluster(const std::vector<DBScan *> &batch);
    // d's labels reusing the finished variant old (the reference's
    // assignPointsToPredefinedCluster with one search thread).
    void reuse(DBScan &d, const DBScan &old);

private:
    std::vector<dataElem> *data_;
    std::vector<int> mapping_;
    uint32_t n_ = 0;
    double *dx_ = nullptr, *dy_ = nullptr; // points as handed in, on the device
    void *dlab_ = nullptr;                 // label staging of cluster()
    size_t dlab_bytes_ = 0;
};

} // namespace vdbscan

class DBScan {
public:
    // src/DBScan.h:50-51; the index replaces (indexPtr, ptr_MPB_lookup, highResIndex).
    // ptrData must be the data the index binned (index->data()); throws
    // std::invalid_argument otherwise.
    DBScan(std::vector<dataElem> *ptrData, double epsilon, int minimumPts,
           vdbscan::PointIndex *index);

    // Clusters from scratch (src/DBScan.cpp:523-609) on the GPU. The first
    // call over an index clusters, in one batched call, this variant and
    // every variant of instanceVect over the same index that has not run yet;
    // their own calls then return at once.
    void algDBScanParallel();
    // Seeds this variant with the clusters of the finished variant
    // instanceVect[finishedInstanceID] (src/DBScan.cpp:272-416, with the
    // helpers of :157-267 and the remainder pass of :429-518): the labels the
    // reference gives for that pair. Driven by a one-thread Schedule, the
    // variants get the reference's multi-variant labels with
    // NSEARCHTHREADS=1. Throws std::runtime_error without such a variant.
    void assignPointsToPredefinedCluster(int finishedInstanceID);
    // clusters + 1 (cluster 0 is noise), as src/DBScan.cpp:604
    int getDBScanNumClusters();

    std::vector<int> clusterIDs;          // by binned position, 0 = noise
    std::vector<DBScan *> instanceVect;   // other variants over the same data
    int numClustersForStats = 0;

private:
    friend class vdbscan::PointIndex;
    enum State { kNotRun, kScratchReady, kDone };
    std::vector<dataElem> *dataPoints;
    double distance;
    int minPts;
    int clusterCnt = 0;
    vdbscan::PointIndex *index;
    State state_ = kNotRun;
};

class Schedule {
public:
    // src/schedule.cpp:32-46: job queue in list order, nothing finished yet
    explicit Schedule(std::vector<experiment> *experimentList);
    // src/schedule.cpp:54-116 (greedy): next variant from the queue; the reuse
    // candidate is the closest earlier finished variant with minpts >= its own
    // (outIDReuse == outInstanceCluster: cluster from scratch). False when the
    // queue is empty. Thread-safe.
    bool determineReuse(int inExperID, int *outIDReuse, int *outInstanceCluster);

    std::vector<bool> finishedList;
    std::vector<experiment> *expPtr;

private:
    std::queue<int> jobQueue;
};

// src/main.cpp:612-620: ascending eps, descending minpts for equal eps
bool compareByEpsilon(const experiment &a, const experiment &b);

#endif
