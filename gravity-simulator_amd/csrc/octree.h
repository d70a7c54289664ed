// octree.h -- device octree records shared by the tree kernels and hosts.
#pragma once

#include <hip/hip_runtime.h>

namespace gs {

// One node, 64 bytes, read by the walk with wave-uniform (scalar) loads.
struct NodeRec {
    double cx, cy, cz;  // centre of mass
    double l2;          // squared cell side, w^2 4^-depth (w: box width)
    long long prefix;   // Morton prefix at `depth` (3*depth bits)
    int depth;
    int count;          // particles in the node
    int first;          // first particle, sorted position
    int child0;         // first child (children are contiguous), -1 for leaves
    int nchild;         // number of children, 0 for leaves
    float fmass;        // mass rounded to FP32 (the walk's monopole weight);
                        // the FP64 mass is mom[node].x
};
static_assert(sizeof(NodeRec) == 64, "NodeRec must stay 64 bytes");

struct DeviceTree {
    int n = 0;
    int leaf_cap = 1;
    int num_nodes = 0;
    int num_levels = 0;
    int node_cap = 0;
    double center[3] = {0, 0, 0};
    double width = 0.0;
    int level_begin[MORTON_MAX_LEVEL + 2] = {0};
    int level_end[MORTON_MAX_LEVEL + 2] = {0};
    double *box = nullptr;          // device: centre[3], width
    long long *skeys = nullptr;     // device: sorted keys
    int *sidx = nullptr;            // device: sorted original indices
    double4 *P = nullptr;           // device: (x, y, z, m) in sorted order
    NodeRec *nodes = nullptr;       // device: level-ordered nodes
    double4 *mom = nullptr;         // device: (M, sum m x, sum m y, sum m z)
    double *quad = nullptr;         // device: 8 per node (xx yy zz xy xz yz, 0, 0),
                                    // FP64, formed only with the third moments
    float *quadf = nullptr;         // device: 8 per node, the walks' FP32 moments
                                    // (xx yy zz xy xz yz, 1.5 tr, 0)
    bool has_quad = false;          // quadf holds this tree's second moments
    double *oct = nullptr;          // device: 12 per node (xxx yyy zzz xxy xxz xyy yyz
                                    // xzz yzz xyz, 0, 0): third central moments
    float *octf = nullptr;          // device: 12 per node, the walk's FP32 copy
    int oct_cap = 0;                // nodes oct / octf hold room for
    bool has_oct = false;           // oct holds this tree's third moments
    unsigned *gflag = nullptr;      // device: per node, 1 for the root and the nodes of
                                    // more than 64 particles (the group walk's parents)
    unsigned *tmp = nullptr;        // device: per-parent child counts
    unsigned *scan_tmp = nullptr;   // device: scan scratch
};

// Stable LSD radix sort of n (key, index) pairs on the low 8 * passes bits of
// the (non-negative) keys, alternating between (k0, i0) and (k1, i1); the
// sorted pairs end in (*ko, *io).  hist holds radix_hist_elements(n) counters.
size_t radix_hist_elements(int n);
ErrorStatus radix_sort_pairs(Context *c, long long *k0, int *i0, long long *k1, int *i1, int n,
                             int passes, unsigned *hist, long long **ko, int **io);
// Bounding box, Morton keys and the stable sort.  With m and packed not
// NULL the key kernel also writes packed[i] = (x_i, y_i, z_i, m_i) in the
// input order, for the gather into sorted order.
ErrorStatus tree_sort(Context *c, const double *x, int n, const double *box_center,
                      double box_width, DeviceTree *t, const double *m = nullptr,
                      double4 *packed = nullptr);

}  // namespace gs
