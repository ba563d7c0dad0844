// Binary BVH builders: binned object SAH (mesh BLAS) and centroid-median
// split (instance TLAS). Reference behaviour followed step by step, because
// the element order left by the in-place partition decides every later split
// and the packed triangle order that the GPU arrays inherit:
//   thirdparty/RadeonRays/split_bvh.cpp:60-218  (node split, right child first)
//   thirdparty/RadeonRays/split_bvh.cpp:220-337 (binned SAH search)
//   thirdparty/RadeonRays/bvh.cpp:68-243        (median split, left child first)
// The builders are iterative (explicit stack) so that degenerate inputs cannot
// overflow the host call stack; the visiting order equals a recursive build.
#include "bvh_build.h"

#include <cmath>
#include <limits>
#include <utility>

namespace glslpt {
namespace {

struct Request {
    int start, count;
    BBox bounds, cbounds;
    int level;
    int parent;     // index of the parent node, -1 for the root
    bool is_left;   // which child slot of the parent this request fills
};

struct Split {
    int dim = 0;
    float pos = std::numeric_limits<float>::quiet_NaN();
};

// Binned SAH over centroids; returns NaN position when no split is usable.
Split FindSahSplit(const Request& rq, const int* order, const BBox* boxes, const Vec3* cents,
                   float traversal_cost, int nb, std::vector<BBox>& bin_box,
                   std::vector<int>& bin_cnt, std::vector<BBox>& right_acc) {
    Split s;
    Vec3 ext = rq.cbounds.extents();
    if (Vec3::Dot(ext, ext) == 0.f) return s;

    float best = std::numeric_limits<float>::max();
    int best_bin = -1;
    const float invarea = 1.f / rq.bounds.surface_area();
    const Vec3 rootmin = rq.cbounds.pmin;

    for (int axis = 0; axis < 3; ++axis) {
        const float lo = rootmin[axis];
        const float rng = ext[axis];
        const float invrng = 1.f / rng;
        if (rng == 0.f) continue;

        for (int b = 0; b < nb; ++b) { bin_box[b] = BBox(); bin_cnt[b] = 0; }
        for (int i = rq.start; i < rq.start + rq.count; ++i) {
            const int p = order[i];
            float f = static_cast<float>(nb) * ((cents[p][axis] - lo) * invrng);
            int b = static_cast<int>(std::min<float>(f, static_cast<float>(nb - 1)));
            bin_cnt[b] += 1;
            bin_box[b].grow(boxes[p]);
        }
        // right_acc[i] = union of bins i+1 .. nb-1
        BBox acc;
        for (int b = nb - 1; b > 0; --b) { acc.grow(bin_box[b]); right_acc[b - 1] = acc; }

        BBox left;
        int nl = 0, nr = rq.count;
        for (int b = 0; b < nb - 1; ++b) {
            left.grow(bin_box[b]);
            nl += bin_cnt[b];
            nr -= bin_cnt[b];
            float cost = traversal_cost +
                         (nl * left.surface_area() + nr * right_acc[b].surface_area()) * invarea;
            if (cost < best) { best = cost; best_bin = b; s.dim = axis; }
        }
    }
    if (best_bin != -1) s.pos = rootmin[s.dim] + (best_bin + 1) * (ext[s.dim] / nb);
    return s;
}

// In-place two-sided partition of order[start, start+count) around `border`
// on `axis`. Elements classified "left" by the first comparison move to the
// front. Grows the child bounds while scanning, as the reference does, which
// matters when the fallback below keeps them. Returns the split position.
int Partition(const Request& rq, int* order, const BBox* boxes, const Vec3* cents, int axis,
              float border, bool near2far, BBox& lb, BBox& rb, BBox& lc, BBox& rc) {
    auto goes_left = [&](int p) {
        float c = cents[p][axis];
        return near2far ? (c < border) : (c >= border);
    };
    auto goes_right = [&](int p) {
        float c = cents[p][axis];
        return near2far ? (c >= border) : (c < border);
    };
    int first = rq.start, last = rq.start + rq.count;
    for (;;) {
        while (first != last && goes_left(order[first])) {
            lb.grow(boxes[order[first]]); lc.grow(cents[order[first]]); ++first;
        }
        bool done = (first == last);
        --last;
        if (done) break;
        rb.grow(boxes[order[first]]); rc.grow(cents[order[first]]);
        while (first != last && goes_right(order[last])) {
            rb.grow(boxes[order[last]]); rc.grow(cents[order[last]]); --last;
        }
        if (first == last) break;
        lb.grow(boxes[order[last]]); lc.grow(cents[order[last]]);
        std::swap(order[first], order[last]);
        ++first;
    }
    return first;
}

enum class Policy { kSah, kMedian };

BvhTree Build(const BBox* prims, int n, Policy policy, float traversal_cost, int nb) {
    BvhTree t;
    if (n <= 0) return t;
    const int leaf_below = policy == Policy::kSah ? 4 : 2;

    std::vector<Vec3> cents(n);
    std::vector<int> order(n);
    BBox cb;
    for (int i = 0; i < n; ++i) {
        t.bounds.grow(prims[i]);
    }
    for (int i = 0; i < n; ++i) {
        cents[i] = prims[i].center();
        cb.grow(cents[i]);
        order[i] = i;
    }
    t.nodes.reserve(2 * static_cast<size_t>(n));
    t.packed.reserve(n);

    std::vector<BBox> bin_box(nb), right_acc(nb > 1 ? nb - 1 : 1);
    std::vector<int> bin_cnt(nb);

    std::vector<Request> stack;
    stack.push_back(Request{0, n, t.bounds, cb, 0, -1, false});
    while (!stack.empty()) {
        Request rq = stack.back();
        stack.pop_back();
        t.height = std::max(t.height, rq.level);

        const int id = static_cast<int>(t.nodes.size());
        t.nodes.emplace_back();
        if (rq.parent >= 0) {
            if (rq.is_left) t.nodes[rq.parent].lc = id; else t.nodes[rq.parent].rc = id;
        }
        BvhBuildNode& node = t.nodes[id];
        node.bounds = rq.bounds;

        if (rq.count < leaf_below) {
            node.leaf = true;
            node.start = static_cast<int>(t.packed.size());
            node.count = rq.count;
            for (int i = rq.start; i < rq.start + rq.count; ++i) t.packed.push_back(order[i]);
            continue;
        }

        int axis = rq.cbounds.maxdim();
        float border = rq.cbounds.center()[axis];
        if (policy == Policy::kSah) {
            Split s = FindSahSplit(rq, order.data(), prims, cents.data(), traversal_cost, nb,
                                   bin_box, bin_cnt, right_acc);
            if (!std::isnan(s.pos)) { border = s.pos; axis = s.dim; }
        }

        BBox lb, rb, lc, rc;
        int split = rq.start;
        const bool near2far = ((rq.count + rq.start) & 1) != 0;
        if (rq.cbounds.extents()[axis] > 0.f)
            split = Partition(rq, order.data(), prims, cents.data(), axis, border, near2far,
                              lb, rb, lc, rc);
        if (split == rq.start || split == rq.start + rq.count) {
            split = rq.start + (rq.count >> 1);
            for (int i = rq.start; i < split; ++i) { lb.grow(prims[order[i]]); lc.grow(cents[order[i]]); }
            for (int i = split; i < rq.start + rq.count; ++i) { rb.grow(prims[order[i]]); rc.grow(cents[order[i]]); }
        }
        Request left{rq.start, split - rq.start, lb, lc, rq.level + 1, id, true};
        Request right{split, rq.count - (split - rq.start), rb, rc, rq.level + 1, id, false};
        // LIFO: the child pushed last is built next.
        if (policy == Policy::kSah) { stack.push_back(left); stack.push_back(right); }
        else { stack.push_back(right); stack.push_back(left); }
    }
    return t;
}

}  // namespace

BvhTree BuildSahBvh(const BBox* prims, int n, float traversal_cost, int num_bins) {
    return Build(prims, n, Policy::kSah, traversal_cost, num_bins);
}

BvhTree BuildMedianBvh(const BBox* prims, int n) {
    return Build(prims, n, Policy::kMedian, 0.f, 1);
}

}  // namespace glslpt
