// Host implementation of the acceleration-structure build (the algorithm is described in
// bvh.h): recursive binned-SAH BVH2, collapse to 4-wide nodes, quantisation.
#include "bvh.h"

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <stdexcept>

namespace vlra {
namespace {

struct HBox {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const HBox &b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    bool empty() const { return lo[0] > hi[0]; }
    float area() const { return bvhArea(lo, hi); }
};

int32_t leafCode(uint32_t first, uint32_t count) {
    return static_cast<int32_t>(~((first << kLeafShift) | (count - 1)));
}

class Bvh2Builder {
public:
    Bvh2Builder(const float *boxes, uint32_t n, uint32_t maxLeaf) : maxLeaf_(maxLeaf) {
        boxes_.resize(n);
        cent_.resize(size_t(n) * 3);
        order.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            for (int a = 0; a < 3; ++a) {
                boxes_[i].lo[a] = bvhNoNegZero(boxes[6 * size_t(i) + a]);
                boxes_[i].hi[a] = bvhNoNegZero(boxes[6 * size_t(i) + 3 + a]);
                cent_[3 * size_t(i) + a] = 0.5f * (boxes_[i].lo[a] + boxes_[i].hi[a]);
            }
            order[i] = i;
        }
    }
    std::vector<BVHNode> nodes;
    std::vector<uint32_t> order;

    void build() {
        nodes.clear();
        nodes.reserve(boxes_.size() + 1);
        nodes.push_back(BVHNode());
        const uint32_t n = static_cast<uint32_t>(boxes_.size());
        if (n <= maxLeaf_) {
            // root with a single leaf child; child code 0 marks an empty slot
            setChild(0, 0, rangeBox(0, n), n ? leafCode(0, n) : 0);
            setChild(0, 1, HBox(), 0);
            return;
        }
        splitInto(0, 0, n, 1);
    }

private:
    std::vector<HBox> boxes_;
    std::vector<float> cent_;
    uint32_t maxLeaf_;

    void setChild(uint32_t node, int slot, const HBox &b, int32_t code) {
        BVHNode &N = nodes[node];
        float *lo = slot == 0 ? N.lo0 : N.lo1;
        float *hi = slot == 0 ? N.hi0 : N.hi1;
        for (int a = 0; a < 3; ++a) {
            lo[a] = b.empty() ? 0.0f : b.lo[a];
            hi[a] = b.empty() ? 0.0f : b.hi[a];
        }
        N.child[slot] = code;
        N.pad[0] = N.pad[1] = 0;
    }
    HBox rangeBox(uint32_t b, uint32_t e) const {
        HBox r;
        for (uint32_t i = b; i < e; ++i) r.grow(boxes_[order[i]]);
        return r;
    }
    float cent(uint32_t p, int a) const { return cent_[3 * size_t(p) + a]; }

    uint32_t partition(uint32_t b, uint32_t e, uint32_t depth) {
        const uint32_t n = e - b;
        float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (uint32_t i = b; i < e; ++i)
            for (int a = 0; a < 3; ++a) {
                clo[a] = std::min(clo[a], cent(order[i], a));
                chi[a] = std::max(chi[a], cent(order[i], a));
            }
        constexpr int NB = kBvhBins;
        if (depth < kBvhSahMaxDepth) {
            float best = FLT_MAX;
            int bestAxis = -1, bestSplit = -1;
            for (int a = 0; a < 3; ++a) {
                const float ext = chi[a] - clo[a];
                if (!(ext > 0.0f)) continue;
                HBox bb[NB];
                uint32_t bc[NB] = {};
                for (uint32_t i = b; i < e; ++i) {
                    const int k = bvhBin(cent(order[i], a), clo[a], ext);
                    bb[k].grow(boxes_[order[i]]);
                    ++bc[k];
                }
                HBox left[NB];
                uint32_t nl[NB];
                HBox acc;
                uint32_t cnt = 0;
                for (int k = 0; k < NB; ++k) {
                    if (bc[k]) acc.grow(bb[k]);
                    cnt += bc[k];
                    left[k] = acc;
                    nl[k] = cnt;
                }
                HBox racc;
                uint32_t rc = 0;
                for (int k = NB - 1; k >= 1; --k) {
                    if (bc[k]) racc.grow(bb[k]);
                    rc += bc[k];
                    const uint32_t lcount = nl[k - 1];
                    if (!lcount || !rc) continue;
                    const float c = left[k - 1].area() * static_cast<float>(lcount) + racc.area() * static_cast<float>(rc);
                    if (c < best) {
                        best = c;
                        bestAxis = a;
                        bestSplit = k;
                    }
                }
            }
            if (bestAxis >= 0) {
                const float ext = chi[bestAxis] - clo[bestAxis];
                auto it = std::stable_partition(order.begin() + b, order.begin() + e, [&](uint32_t p) {
                    return bvhBin(cent(p, bestAxis), clo[bestAxis], ext) < bestSplit;
                });
                return static_cast<uint32_t>(it - order.begin());
            }
            return b + n / 2;   // coincident centroids
        }
        if (n > kBvhSmallPrims) return b + n / 2;
        int a = 0;
        for (int k = 1; k < 3; ++k)
            if (chi[k] - clo[k] > chi[a] - clo[a]) a = k;
        std::stable_sort(order.begin() + b, order.begin() + e,
                         [&](uint32_t x, uint32_t y) { return cent(x, a) < cent(y, a); });
        return b + n / 2;
    }
    // node `idx` gets children for [b, m) and [m, e)
    void splitInto(uint32_t idx, uint32_t b, uint32_t e, uint32_t depth) {
        const uint32_t m = partition(b, e, depth);
        const uint32_t ranges[2][2] = {{b, m}, {m, e}};
        for (int s = 0; s < 2; ++s) {
            const uint32_t rb = ranges[s][0], re = ranges[s][1];
            const HBox box = rangeBox(rb, re);
            if (re - rb <= maxLeaf_) {
                setChild(idx, s, box, leafCode(rb, re - rb));
            } else {
                const uint32_t child = static_cast<uint32_t>(nodes.size());
                nodes.push_back(BVHNode());
                setChild(idx, s, box, static_cast<int32_t>(child));
                splitInto(child, rb, re, depth + 1);
            }
        }
    }
};

class Bvh4Collapser {
public:
    Bvh4Collapser(const std::vector<BVHNode> &bin, const uint32_t *leafRemap, const std::vector<uint32_t> &order)
        : bin_(bin), remap_(leafRemap), order_(order) {}
    std::vector<BVH4Node> nodes;
    uint32_t depth = 0;
    void run() {
        nodes.clear();
        depth = 0;
        if (!bin_.empty()) emit(0, 1);
    }

private:
    struct Slot {
        float lo[3], hi[3];
        int32_t code;
        float area() const { return bvhArea(lo, hi); }
    };
    const std::vector<BVHNode> &bin_;
    const uint32_t *remap_;
    const std::vector<uint32_t> &order_;

    Slot slotOf(uint32_t n, int s) const {
        const BVHNode &N = bin_[n];
        Slot r;
        const float *lo = s == 0 ? N.lo0 : N.lo1, *hi = s == 0 ? N.hi0 : N.hi1;
        for (int a = 0; a < 3; ++a) {
            r.lo[a] = lo[a];
            r.hi[a] = hi[a];
        }
        r.code = N.child[s];
        return r;
    }
    int32_t leaf(int32_t code) const {
        if (!remap_) return code;
        const uint32_t first = static_cast<uint32_t>(~code) >> kLeafShift;
        return static_cast<int32_t>(~(remap_[order_[first]] << kLeafShift));
    }
    uint32_t emit(uint32_t binNode, uint32_t level) {
        depth = std::max(depth, level);
        const uint32_t idx = static_cast<uint32_t>(nodes.size());
        nodes.push_back(BVH4Node());
        Slot slots[4];
        int n = 0;
        for (int s = 0; s < 2; ++s) {
            const Slot c = slotOf(binNode, s);
            if (c.code != 0) slots[n++] = c;
        }
        while (n < 4) {
            int best = -1;
            float bestArea = -1.0f;
            for (int k = 0; k < n; ++k)
                if (slots[k].code > 0 && slots[k].area() > bestArea) {
                    best = k;
                    bestArea = slots[k].area();
                }
            if (best < 0) break;
            const uint32_t inner = static_cast<uint32_t>(slots[best].code);
            const Slot a = slotOf(inner, 0), b = slotOf(inner, 1);
            // the replaced slot's children take its place
            Slot tmp[4];
            int m = 0;
            for (int k = 0; k < n; ++k) {
                if (k == best) {
                    if (a.code != 0) tmp[m++] = a;
                    if (b.code != 0) tmp[m++] = b;
                } else {
                    tmp[m++] = slots[k];
                }
            }
            for (int k = 0; k < m; ++k) slots[k] = tmp[k];
            n = m;
        }
        int32_t codes[4] = {0, 0, 0, 0};
        for (int k = 0; k < n; ++k)
            codes[k] = slots[k].code > 0 ? static_cast<int32_t>(emit(static_cast<uint32_t>(slots[k].code), level + 1))
                                         : leaf(slots[k].code);
        BVH4Node &N = nodes[idx];
        for (int k = 0; k < 4; ++k) {
            const bool used = k < n;
            N.loX[k] = used ? slots[k].lo[0] : 0.0f;
            N.loY[k] = used ? slots[k].lo[1] : 0.0f;
            N.loZ[k] = used ? slots[k].lo[2] : 0.0f;
            N.hiX[k] = used ? slots[k].hi[0] : 0.0f;
            N.hiY[k] = used ? slots[k].hi[1] : 0.0f;
            N.hiZ[k] = used ? slots[k].hi[2] : 0.0f;
            N.child[k] = codes[k];
            N.pad[k] = 0;
        }
        return idx;
    }
};

} // namespace

void hostBuildBvh4(const float *boxes, uint32_t numPrims, uint32_t maxLeaf, const uint32_t *leafRemap,
                   HostBvh4 *out) {
    if (numPrims >= kBvhMaxPrims) throw std::runtime_error("BVH build: too many primitives");
    if (maxLeaf < 1 || maxLeaf > kMaxLeafPrims) throw std::runtime_error("BVH build: leaf size out of range");
    if (leafRemap && maxLeaf != 1) throw std::runtime_error("BVH build: leaf remapping needs one primitive per leaf");
    out->nodes.clear();
    out->order.clear();
    if (numPrims == 0) return;
    Bvh2Builder b2(boxes, numPrims, maxLeaf);
    b2.build();
    Bvh4Collapser c(b2.nodes, leafRemap, b2.order);
    c.run();
    out->nodes.resize(c.nodes.size());
    for (size_t i = 0; i < c.nodes.size(); ++i) out->nodes[i] = quantizeBvh4Node(c.nodes[i]);
    out->order = std::move(b2.order);
    out->depth = c.depth;
}

} // namespace vlra
