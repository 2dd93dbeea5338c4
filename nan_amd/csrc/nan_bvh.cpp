// Binned-SAH build, 4-wide collapse and 8-bit child-box quantization.
// See nan_bvh.h for the node format.  Compiled with -ffp-contract=off so the
// float arithmetic of the quantization check equals the kernels' decode.
#include "nan_bvh.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace nanbvh {
namespace {

struct Aabb {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const Aabb &b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float *p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    bool empty() const { return lo[0] > hi[0]; }
    double area() const {
        if (empty()) return 0.0;
        double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct BNode {
    Aabb box;
    uint32_t a = 0, b = 0;  // internal: children; leaf: first, count
    bool leaf = false;
};

struct Task {
    uint32_t node, begin, end;
};

class Builder {
  public:
    Builder(const float *world, uint32_t n, const BuildOptions &opt) : world_(world), n_(n), opt_(opt) {}

    void run() {
        idx_.resize(n_);
        boxes_.resize(n_);
        cent_.resize(3 * (size_t)n_);
        for (uint32_t i = 0; i < n_; ++i) {
            idx_[i] = i;
            const float *t = world_ + 9 * (size_t)i;
            Aabb b;
            b.grow(t);
            b.grow(t + 3);
            b.grow(t + 6);
            boxes_[i] = b;
            for (int k = 0; k < 3; ++k) cent_[3 * (size_t)i + k] = 0.5f * (b.lo[k] + b.hi[k]);
        }
        nodes_.reserve(2 * (size_t)n_);
        nodes_.push_back(BNode());
        std::vector<Task> stack;
        stack.push_back({0, 0, n_});
        uint32_t bins = std::max(4u, std::min(opt_.bins ? opt_.bins : 16u, 64u));
        uint32_t max_leaf = std::max(1u, std::min(opt_.max_leaf ? opt_.max_leaf : 4u, kMaxLeafSize));
        std::vector<Aabb> bin_box(bins);
        std::vector<uint32_t> bin_cnt(bins);
        std::vector<double> right_area(bins);
        std::vector<uint32_t> right_cnt(bins);
        while (!stack.empty()) {
            Task t = stack.back();
            stack.pop_back();
            Aabb box, cbox;
            for (uint32_t i = t.begin; i < t.end; ++i) {
                box.grow(boxes_[idx_[i]]);
                cbox.grow(&cent_[3 * (size_t)idx_[i]]);
            }
            nodes_[t.node].box = box;
            uint32_t count = t.end - t.begin;
            if (count <= 1) {
                make_leaf(t);
                continue;
            }
            double parent_area = box.area();
            double best_cost = INFINITY;
            int best_axis = -1;
            uint32_t best_split = 0;
            for (int axis = 0; axis < 3; ++axis) {
                float cmin = cbox.lo[axis], cmax = cbox.hi[axis];
                if (!(cmax > cmin)) continue;
                double scale = (double)bins / ((double)cmax - (double)cmin);
                for (uint32_t k = 0; k < bins; ++k) {
                    bin_box[k] = Aabb();
                    bin_cnt[k] = 0;
                }
                for (uint32_t i = t.begin; i < t.end; ++i) {
                    uint32_t k = bin_of(cent_[3 * (size_t)idx_[i] + axis], cmin, scale, bins);
                    bin_box[k].grow(boxes_[idx_[i]]);
                    bin_cnt[k]++;
                }
                Aabb acc;
                uint32_t acc_n = 0;
                for (int k = (int)bins - 1; k >= 1; --k) {
                    acc.grow(bin_box[k]);
                    acc_n += bin_cnt[k];
                    right_area[k] = acc.area();
                    right_cnt[k] = acc_n;
                }
                acc = Aabb();
                acc_n = 0;
                for (uint32_t k = 1; k < bins; ++k) {
                    acc.grow(bin_box[k - 1]);
                    acc_n += bin_cnt[k - 1];
                    if (acc_n == 0 || right_cnt[k] == 0) continue;
                    double cost = acc.area() * acc_n + right_area[k] * right_cnt[k];
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = axis;
                        best_split = k;
                    }
                }
            }
            double split_sah = parent_area > 0 ? 1.0 + best_cost / parent_area : INFINITY;
            if (count <= max_leaf && (best_axis < 0 || (double)count <= split_sah)) {
                make_leaf(t);
                continue;
            }
            uint32_t mid;
            if (best_axis >= 0) {
                float cmin = cbox.lo[best_axis];
                double scale = (double)bins / ((double)cbox.hi[best_axis] - (double)cmin);
                auto it = std::partition(idx_.begin() + t.begin, idx_.begin() + t.end, [&](uint32_t id) {
                    return bin_of(cent_[3 * (size_t)id + best_axis], cmin, scale, bins) < best_split;
                });
                mid = (uint32_t)(it - idx_.begin());
                if (mid == t.begin || mid == t.end) mid = t.begin + count / 2;
            } else {
                mid = t.begin + count / 2;
            }
            uint32_t left = (uint32_t)nodes_.size();
            nodes_.push_back(BNode());
            nodes_.push_back(BNode());
            nodes_[t.node].a = left;
            nodes_[t.node].b = left + 1;
            stack.push_back({left + 1, mid, t.end});
            stack.push_back({left, t.begin, mid});
        }
    }

    const std::vector<BNode> &nodes() const { return nodes_; }
    const std::vector<uint32_t> &order() const { return idx_; }

  private:
    static uint32_t bin_of(float c, float cmin, double scale, uint32_t bins) {
        double f = ((double)c - (double)cmin) * scale;
        int64_t k = (int64_t)f;
        if (k < 0) k = 0;
        if (k >= (int64_t)bins) k = bins - 1;
        return (uint32_t)k;
    }
    void make_leaf(const Task &t) {
        BNode &nd = nodes_[t.node];
        nd.leaf = true;
        nd.a = t.begin;
        nd.b = t.end - t.begin;
    }

    const float *world_;
    uint32_t n_;
    BuildOptions opt_;
    std::vector<uint32_t> idx_;
    std::vector<Aabb> boxes_;
    std::vector<float> cent_;
    std::vector<BNode> nodes_;
};

// Leaves larger than kMaxLeafSize cannot be encoded; they only arise when
// many centroids coincide, and are split by index here.
void split_oversized_leaves(std::vector<BNode> &nodes, const std::vector<uint32_t> &order, const float *world) {
    for (size_t i = 0; i < nodes.size(); ++i) {
        if (!nodes[i].leaf || nodes[i].b <= kMaxLeafSize) continue;
        uint32_t first = nodes[i].a, count = nodes[i].b, half = count / 2;
        BNode l, r;
        l.leaf = r.leaf = true;
        l.a = first;
        l.b = half;
        r.a = first + half;
        r.b = count - half;
        for (uint32_t k = 0; k < l.b; ++k) {
            const float *t = world + 9 * (size_t)order[l.a + k];
            l.box.grow(t); l.box.grow(t + 3); l.box.grow(t + 6);
        }
        for (uint32_t k = 0; k < r.b; ++k) {
            const float *t = world + 9 * (size_t)order[r.a + k];
            r.box.grow(t); r.box.grow(t + 3); r.box.grow(t + 6);
        }
        uint32_t li = (uint32_t)nodes.size();
        nodes.push_back(l);
        nodes.push_back(r);
        nodes[i].leaf = false;
        nodes[i].a = li;
        nodes[i].b = li + 1;
    }
}

// (float)q * scale is exact (q < 256, scale a power of two), so this equals the
// kernels' single-rounding fmaf(q, scale, origin) bit for bit.
inline float decode(float origin, uint32_t q, float scale) { return origin + (float)q * scale; }

inline float exp2i(int e) {
    uint32_t bits = (uint32_t)(e + 127) << 23;
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

bool quantize_axis(float origin, float max_hi, const float *lo, const float *hi, int n, int8_t *exp_out,
                   uint8_t *qlo, uint8_t *qhi) {
    double extent = (double)max_hi - (double)origin;
    int e = -100;
    if (extent > 0.0) {
        e = (int)std::ceil(std::log2(extent / 255.0));
        if (e < -100) e = -100;
    }
    while (e <= 127 && decode(origin, 255u, exp2i(e)) < max_hi) ++e;
    if (e > 127) return false;
    float s = exp2i(e);
    for (int c = 0; c < n; ++c) {
        double ql = std::floor(((double)lo[c] - (double)origin) / (double)s);
        uint32_t q = (uint32_t)std::max(0.0, std::min(255.0, ql));
        while (q > 0 && decode(origin, q, s) > lo[c]) --q;
        if (decode(origin, q, s) > lo[c]) return false;
        qlo[c] = (uint8_t)q;
        double qh = std::ceil(((double)hi[c] - (double)origin) / (double)s);
        uint32_t p = (uint32_t)std::max(0.0, std::min(255.0, qh));
        while (p < 255 && decode(origin, p, s) < hi[c]) ++p;
        if (decode(origin, p, s) < hi[c]) return false;
        if (p < q) p = q;
        qhi[c] = (uint8_t)p;
    }
    *exp_out = (int8_t)e;
    return true;
}

class Collapser {
  public:
    Collapser(const std::vector<BNode> &bn, Bvh4 *out) : bn_(bn), out_(out) {}

    // Returns the child entry for binary node `b`.
    uint32_t entry_of(uint32_t b) {
        const BNode &nd = bn_[b];
        if (nd.leaf) {
            out_->leaf_count++;
            return kLeafBit | ((nd.b - 1) << kLeafCountShift) | nd.a;
        }
        return emit(b);
    }

    bool ok = true;

  private:
    uint32_t emit(uint32_t b) {
        uint32_t self = out_->node_count++;
        out_->nodes.resize((size_t)out_->node_count * kNodeWords, 0u);
        // gather up to 4 children by opening the largest internal child
        uint32_t kids[4];
        int n = 0;
        kids[n++] = bn_[b].a;
        kids[n++] = bn_[b].b;
        while (n < 4) {
            int best = -1;
            double best_area = -1.0;
            for (int i = 0; i < n; ++i) {
                if (bn_[kids[i]].leaf) continue;
                double a = bn_[kids[i]].box.area();
                if (a > best_area) {
                    best_area = a;
                    best = i;
                }
            }
            if (best < 0) break;
            uint32_t open = kids[best];
            kids[best] = bn_[open].a;
            kids[n++] = bn_[open].b;
        }
        Aabb parent;
        for (int i = 0; i < n; ++i) parent.grow(bn_[kids[i]].box);
        uint32_t *w = &out_->nodes[(size_t)self * kNodeWords];
        std::memcpy(&w[0], &parent.lo[0], 4);
        std::memcpy(&w[1], &parent.lo[1], 4);
        std::memcpy(&w[2], &parent.lo[2], 4);
        int8_t ex[3];
        uint8_t qlo[3][4] = {}, qhi[3][4] = {};
        for (int axis = 0; axis < 3; ++axis) {
            float lo[4], hi[4];
            for (int i = 0; i < n; ++i) {
                lo[i] = bn_[kids[i]].box.lo[axis];
                hi[i] = bn_[kids[i]].box.hi[axis];
            }
            if (!quantize_axis(parent.lo[axis], parent.hi[axis], lo, hi, n, &ex[axis], qlo[axis], qhi[axis]))
                ok = false;
        }
        uint32_t valid = (1u << n) - 1u;
        w[3] = (uint32_t)(uint8_t)ex[0] | ((uint32_t)(uint8_t)ex[1] << 8) | ((uint32_t)(uint8_t)ex[2] << 16) |
               (valid << 24);
        for (int axis = 0; axis < 3; ++axis) {
            uint32_t l = 0, h = 0;
            for (int i = 0; i < 4; ++i) {
                l |= (uint32_t)qlo[axis][i] << (8 * i);
                h |= (uint32_t)qhi[axis][i] << (8 * i);
            }
            w[4 + axis] = l;
            w[7 + axis] = h;
        }
        for (int i = 0; i < 4; ++i) w[10 + i] = 0xFFFFFFFFu;
        // children emitted after the parent (pre-order)
        uint32_t entries[4];
        for (int i = 0; i < n; ++i) entries[i] = entry_of(kids[i]);
        w = &out_->nodes[(size_t)self * kNodeWords];
        for (int i = 0; i < n; ++i) w[10 + i] = entries[i];
        return self;
    }

    const std::vector<BNode> &bn_;
    Bvh4 *out_;
};

uint32_t stack_need(const Bvh4 &b, uint32_t entry) {
    if (entry & kLeafBit) return 0;
    const uint32_t *w = &b.nodes[(size_t)entry * kNodeWords];
    uint32_t valid = w[3] >> 24;
    uint32_t n = (uint32_t)__builtin_popcount(valid);
    uint32_t deeper = 0;
    for (uint32_t i = 0; i < n; ++i) deeper = std::max(deeper, stack_need(b, w[10 + i]));
    return (n - 1) + deeper;
}

}  // namespace

void decode_child(const uint32_t *node, int child, float lo[3], float hi[3]) {
    for (int axis = 0; axis < 3; ++axis) {
        float origin;
        std::memcpy(&origin, &node[axis], 4);
        int8_t e = (int8_t)((node[3] >> (8 * axis)) & 0xFFu);
        float s = exp2i(e);
        uint32_t ql = (node[4 + axis] >> (8 * child)) & 0xFFu;
        uint32_t qh = (node[7 + axis] >> (8 * child)) & 0xFFu;
        lo[axis] = decode(origin, ql, s);
        hi[axis] = decode(origin, qh, s);
    }
}

bool build_bvh4(const float *world, uint32_t triangle_count, const BuildOptions &opt, Bvh4 *out,
                const char **error) {
    *out = Bvh4();
    if (triangle_count == 0) {
        *error = "scene has no triangles";
        return false;
    }
    if (triangle_count > kLeafFirstMask) {
        *error = "too many triangles for 27-bit leaf offsets";
        return false;
    }
    Builder builder(world, triangle_count, opt);
    builder.run();
    std::vector<BNode> nodes = builder.nodes();
    const std::vector<uint32_t> &order = builder.order();
    split_oversized_leaves(nodes, order, world);

    Collapser col(nodes, out);
    out->root_entry = col.entry_of(0);
    if (!col.ok) {
        *error = "BVH quantization failed (scene extent outside float range)";
        return false;
    }
    out->tris.assign((size_t)triangle_count * kTriWords, 0.0f);
    for (uint32_t i = 0; i < triangle_count; ++i) {
        uint32_t id = order[i];
        float *r = &out->tris[(size_t)i * kTriWords];
        std::memcpy(r, world + 9 * (size_t)id, 9 * sizeof(float));
        std::memcpy(&r[9], &id, 4);
    }
    out->max_stack = stack_need(*out, out->root_entry);
    // SAH cost relative to the root box: internal nodes cost 1, triangles 1.
    double root_area = nodes[0].box.area();
    double cost = 0.0;
    if (root_area > 0) {
        for (const BNode &nd : nodes)
            cost += nd.leaf ? nd.box.area() / root_area * nd.b : 0.0;
        for (uint32_t i = 0; i < out->node_count; ++i) {
            float lo[3], hi[3];
            Aabb a;
            const uint32_t *w = &out->nodes[(size_t)i * kNodeWords];
            for (int k = 0; k < 3; ++k) std::memcpy(&a.lo[k], &w[k], 4);
            uint32_t n = (uint32_t)__builtin_popcount(w[3] >> 24);
            for (uint32_t c = 0; c < n; ++c) {
                decode_child(w, (int)c, lo, hi);
                for (int k = 0; k < 3; ++k) a.hi[k] = std::max(a.hi[k], hi[k]);
            }
            cost += a.area() / root_area;
        }
    }
    out->sah_cost = (float)cost;
    return true;
}

}  // namespace nanbvh
