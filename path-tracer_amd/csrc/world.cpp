// world.cpp - host scene preparation: the producers of the kernel's inputs.
//
// Restates, in C++, what the reference's World does between loading a scene
// and uploading it (src/render/world.rs):
//   * binned-SAH BVH with 8 bins over centroid bounds, leaf at <= 2 triangles
//     or when splitting costs more than the leaf, Hoare partition, children
//     allocated as an adjacent pair, recursion left then right
//     (world.rs:1339-1345, 1430-1591);
//   * triangle reorder into BVH order (world.rs:1348-1355);
//   * Vose alias tables in f64 (world.rs:539-575), emissive-triangle weights
//     2*pi*area*luminance (world.rs:1691-1735), sky weights per row solid
//     angle (world.rs:584-587, 1959-1975), fluxes and category probabilities
//     (world.rs:484-498, 1977-1987, 1738-1741);
//   * sky normalisation by solid-angle-weighted mean luminance (world.rs:1872-1957);
//   * camera (world.rs:2159-2193) and the headless option handling
//     (headless.rs:180-204).
// Floating-point steps follow the reference's types (f32 with mul_add where
// it uses mul_add, f64 where it widens), so the arrays come out bit-identical.
#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <limits>
#include <new>
#include <stdexcept>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "host_parallel.h"
#include "pt_internal.h"

namespace {

using ptcore::host_threads;
using ptcore::run_parallel;

// fmin/fmax as the reference's f32::min/max: a NaN operand gives the other
// one; of equal operands (+0 and -0) the second, as glibc's fminf/fmaxf give.
// Written so the compiler emits minss/maxss behind a NaN test that is never
// taken on real data: as library calls, or as compare-and-branch on the
// values, they were most of the BVH build's time.
inline float min_f(float x, float y) { return y != y ? x : (x < y ? x : y); }
inline float max_f(float x, float y) { return y != y ? x : (x > y ? x : y); }

struct Aabb {
    float mn[3] = {INFINITY, INFINITY, INFINITY};
    float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Aabb& o) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = min_f(mn[k], o.mn[k]);
            mx[k] = max_f(mx[k], o.mx[k]);
        }
    }
};

// e.x*e.y + e.y*e.z + e.z*e.x with the reference's mul_add nesting (world.rs:412-416)
inline float half_area(const float mn[3], const float mx[3]) {
    float ex = mx[0] - mn[0], ey = mx[1] - mn[1], ez = mx[2] - mn[2];
    return std::fmaf(ez, ex, std::fmaf(ex, ey, ey * ez));
}

struct BvhTri {
    Aabb box;
    float centroid[3];
    uint32_t original;
};

// Rust `f32 as usize`: saturating, NaN -> 0
inline size_t f32_as_usize(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 1.8446744e19f) return std::numeric_limits<size_t>::max();
    return (size_t)v;
}

// Triangles below which a BVH pass is not cut into chunks, and twice which a
// subtree is not split further before the workers take over:
// PT_BVH_TASK_GRAIN when set, else 4096.
size_t bvh_task_grain() {
    if (const char* v = std::getenv("PT_BVH_TASK_GRAIN")) {
        long g = std::strtol(v, nullptr, 10);
        if (g >= 1) return (size_t)g;
    }
    return 4096;
}

// Raised when a split sends every triangle to one side. The child would then
// repeat its parent's range and box, and the reference's recursion never
// ends (it overflows its stack): non-finite SAH costs from NaN or
// overflowing vertex coordinates lead there.
struct BvhNoProgress {};

// The BVH builder of world.rs:1430-1591: binned SAH over centroid bounds with
// PT_BVH_BINS bins, leaf at <= 2 triangles or when the best split costs at
// least the leaf, modified Hoare partition, children allocated as an adjacent
// pair, and allocation in the order of the reference's recursion (a node's
// pair, then the left subtree, then the right).
//
// Each node's decision depends only on its own range of `tris`, which only
// its ancestors' partitions rearranged, so nodes may be split in any order
// that puts parents first. The parallel build uses that: the largest nodes
// are split first (their passes over the triangles cut into chunks, one per
// thread), the remaining subtrees are built concurrently into local arrays,
// and a final walk in the reference's allocation order numbers every node.
// Per-chunk bounds and bins are folded in chunk order with the same fmin/fmax
// calls as the one-thread fold, so the nodes come out bit-identical.
class BvhBuilder {
public:
    BvhBuilder(std::vector<BvhTri>& tris, unsigned threads, size_t grain)
        : tris_(tris), threads_(threads ? threads : 1u), grain_(grain) {}

    void build(std::vector<PtBvhNode>& out);

private:
    // one node's decision: a leaf, or a split at `mid` after the partition
    struct Decision {
        bool leaf;
        size_t mid;
    };
    struct Task {
        size_t start, end;
        Aabb box;
        std::vector<PtBvhNode> nodes;  // local numbering, root at 0
        std::vector<uint8_t> inner;    // 1 where left_first is a local node index
    };
    // a node of the top levels, split before the subtrees are handed out
    struct Top {
        size_t start, end;
        Aabb box;
        int kind;      // 0 leaf, 1 split (children at `left`, `left` + 1), 2 subtree task
        size_t left;   // index into top_ (kind 1) or tasks_ (kind 2)
    };

    Aabb bounds(size_t start, size_t end, unsigned threads) const;
    Decision decide(const Aabb& box, size_t start, size_t end, unsigned threads);
    template <class GetBin>
    void partition_parallel(size_t start, size_t cut, size_t end, unsigned threads, const GetBin& get_bin, int axis,
                            size_t split_bin);
    void build_task(Task& t);
    unsigned chunks_for(size_t count) const {
        if (threads_ <= 1 || count < 2 * grain_) return 1u;
        return (unsigned)std::min<size_t>(threads_, count / grain_);
    }

    std::vector<BvhTri>& tris_;
    unsigned threads_;
    size_t grain_;
    std::vector<Top> top_;
    std::vector<Task> tasks_;
};

// union of the triangle boxes over [start, end), in index order
Aabb BvhBuilder::bounds(size_t start, size_t end, unsigned threads) const {
    const size_t n = end - start;
    if (threads <= 1) {
        Aabb b;
        for (size_t i = start; i < end; ++i) b.grow(tris_[i].box);
        return b;
    }
    std::vector<Aabb> part(threads);
    run_parallel(threads, [&](unsigned t) {
        const size_t a = start + n * t / threads, e = start + n * (t + 1) / threads;
        Aabb b;
        for (size_t i = a; i < e; ++i) b.grow(tris_[i].box);
        part[t] = b;
    });
    Aabb b;
    for (const Aabb& p : part) b.grow(p);
    return b;
}

BvhBuilder::Decision BvhBuilder::decide(const Aabb& node_box, size_t start, size_t end, unsigned threads) {
    const size_t tri_count = end - start;
    if (tri_count <= 2) return {true, 0};
    const size_t n = tri_count;
    // centroid bounds
    struct Bounds {
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    };
    // per-chunk partials; the one-chunk case (every node below the top levels) stays on the stack
    Bounds c_one;
    std::vector<Bounds> c_many(threads > 1 ? threads : 0);
    Bounds* cpart = threads > 1 ? c_many.data() : &c_one;
    run_parallel(threads, [&](unsigned t) {
        const size_t a = start + n * t / threads, e = start + n * (t + 1) / threads;
        Bounds c;
        for (size_t i = a; i < e; ++i)
            for (int k = 0; k < 3; ++k) {
                c.mn[k] = min_f(c.mn[k], tris_[i].centroid[k]);
                c.mx[k] = max_f(c.mx[k], tris_[i].centroid[k]);
            }
        cpart[t] = c;
    });
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (unsigned t = 0; t < threads; ++t)
        for (int k = 0; k < 3; ++k) {
            cmin[k] = min_f(cmin[k], cpart[t].mn[k]);
            cmax[k] = max_f(cmax[k], cpart[t].mx[k]);
        }
    // bins of the three axes in one pass; an axis with a flat centroid range is skipped
    bool live[3];
    float scale[3];
    for (int k = 0; k < 3; ++k) {
        live[k] = !(cmin[k] == cmax[k]);
        scale[k] = (float)PT_BVH_BINS / (cmax[k] - cmin[k]);
    }
    struct Bins {
        Aabb box[3][PT_BVH_BINS];
        uint32_t count[3][PT_BVH_BINS] = {};
    };
    Bins b_one;
    std::vector<Bins> b_many(threads > 1 ? threads : 0);
    Bins* bpart = threads > 1 ? b_many.data() : &b_one;
    run_parallel(threads, [&](unsigned t) {
        const size_t a = start + n * t / threads, e = start + n * (t + 1) / threads;
        Bins& bn = bpart[t];
        for (size_t i = a; i < e; ++i) {
            const BvhTri& tri = tris_[i];
            for (int k = 0; k < 3; ++k) {
                if (!live[k]) continue;
                size_t b = std::min(f32_as_usize((tri.centroid[k] - cmin[k]) * scale[k]), (size_t)PT_BVH_BINS - 1);
                bn.count[k][b] += 1;
                bn.box[k][b].grow(tri.box);
            }
        }
    });
    int best_axis = 0;
    size_t best_split = 0, best_left = 0;
    float best_cost = std::numeric_limits<float>::max();
    for (int axis = 0; axis < 3; ++axis) {
        if (!live[axis]) continue;
        Aabb bins[PT_BVH_BINS];
        uint32_t counts[PT_BVH_BINS] = {0};
        for (unsigned t = 0; t < threads; ++t)
            for (int b = 0; b < PT_BVH_BINS; ++b) {
                counts[b] += bpart[t].count[axis][b];
                bins[b].grow(bpart[t].box[axis][b]);
            }
        float left_area[PT_BVH_BINS - 1], right_area[PT_BVH_BINS - 1];
        uint32_t left_count[PT_BVH_BINS - 1], right_count[PT_BVH_BINS - 1];
        Aabb lb;
        uint32_t ls = 0;
        for (int i = 0; i < PT_BVH_BINS - 1; ++i) {
            ls += counts[i];
            lb.grow(bins[i]);
            left_count[i] = ls;
            left_area[i] = half_area(lb.mn, lb.mx);
        }
        Aabb rb;
        uint32_t rs = 0;
        for (int i = PT_BVH_BINS - 1; i >= 1; --i) {
            rs += counts[i];
            rb.grow(bins[i]);
            right_count[i - 1] = rs;
            right_area[i - 1] = half_area(rb.mn, rb.mx);
        }
        for (int i = 0; i < PT_BVH_BINS - 1; ++i) {
            float cost = std::fmaf((float)left_count[i], left_area[i], (float)right_count[i] * right_area[i]);
            if (cost < best_cost) {
                best_cost = cost;
                best_axis = axis;
                best_split = (size_t)i;
                best_left = left_count[i];
            }
        }
    }
    const float leaf_cost = (float)tri_count * half_area(node_box.mn, node_box.mx);
    if (best_cost >= leaf_cost) return {true, 0};
    const float amin = cmin[best_axis];
    const float sc = scale[best_axis];
    auto get_bin = [&](float c) { return std::min(f32_as_usize((c - amin) * sc), (size_t)PT_BVH_BINS - 1); };
    size_t mid;
    if (threads <= 1) {
        // modified Hoare partition (world.rs:1554-1573)
        int64_t left = (int64_t)start, right = (int64_t)end - 1;
        for (;;) {
            while (left <= right && get_bin(tris_[left].centroid[best_axis]) <= best_split) ++left;
            while (right > left && get_bin(tris_[right].centroid[best_axis]) > best_split) --right;
            if (left >= right) break;
            std::swap(tris_[left], tris_[right]);
            ++left;
            --right;
        }
        mid = (size_t)left;
    } else {
        mid = start + best_left;
        partition_parallel(start, mid, end, threads, get_bin, best_axis, best_split);
    }
    if (mid == start || mid == end) throw BvhNoProgress{};
    return {false, mid};
}

// The Hoare partition above swaps the k-th triangle of [start, cut) that
// belongs right (in ascending order) with the k-th triangle of [cut, end) that
// belongs left (in descending order), for every k, where cut - start is the
// count of triangles that belong left. The pairs are disjoint, so they are
// found by chunked scans and swapped concurrently, with the same result.
template <class GetBin>
void BvhBuilder::partition_parallel(size_t start, size_t cut, size_t end, unsigned threads, const GetBin& get_bin,
                                    int axis, size_t split_bin) {
    const size_t nl = cut - start, nr = end - cut;
    auto goes_right = [&](size_t i) { return get_bin(tris_[i].centroid[axis]) > split_bin; };
    // chunk t of the left part is [start + nl*t/T, start + nl*(t+1)/T); chunk t
    // of the right part is the t-th slice counted down from end
    auto lchunk = [&](unsigned t, size_t& a, size_t& e) {
        a = start + nl * t / threads;
        e = start + nl * (t + 1) / threads;
    };
    auto rchunk = [&](unsigned t, size_t& hi, size_t& lo) {
        hi = end - nr * t / threads;
        lo = end - nr * (t + 1) / threads;
    };
    std::vector<size_t> lcount(threads + 1, 0), rcount(threads + 1, 0);
    run_parallel(threads, [&](unsigned t) {
        size_t a, e, hi, lo, c = 0;
        lchunk(t, a, e);
        for (size_t i = a; i < e; ++i) c += goes_right(i);
        lcount[t + 1] = c;
        c = 0;
        rchunk(t, hi, lo);
        for (size_t i = lo; i < hi; ++i) c += !goes_right(i);
        rcount[t + 1] = c;
    });
    for (unsigned t = 0; t < threads; ++t) {
        lcount[t + 1] += lcount[t];
        rcount[t + 1] += rcount[t];
    }
    const size_t swaps = lcount[threads];
    if (swaps != rcount[threads]) throw std::logic_error("BVH partition counts disagree with the bins");
    std::vector<size_t> lpos(swaps), rpos(swaps);
    run_parallel(threads, [&](unsigned t) {
        size_t a, e, hi, lo, k = lcount[t];
        lchunk(t, a, e);
        for (size_t i = a; i < e; ++i)
            if (goes_right(i)) lpos[k++] = i;
        k = rcount[t];
        rchunk(t, hi, lo);
        for (size_t i = hi; i-- > lo;)
            if (!goes_right(i)) rpos[k++] = i;
    });
    run_parallel(threads, [&](unsigned t) {
        const size_t a = swaps * t / threads, e = swaps * (t + 1) / threads;
        for (size_t k = a; k < e; ++k) std::swap(tris_[lpos[k]], tris_[rpos[k]]);
    });
}

// one subtree, depth first in the reference's order, into the task's local arrays
void BvhBuilder::build_task(Task& t) {
    struct Item {
        size_t node, start, end;
    };
    auto set_box = [](PtBvhNode& nd, const Aabb& b) {
        for (int k = 0; k < 3; ++k) {
            nd.aabb_min[k] = b.mn[k];
            nd.aabb_max[k] = b.mx[k];
        }
    };
    t.nodes.clear();
    t.inner.clear();
    t.nodes.reserve(2 * (t.end - t.start));
    t.inner.reserve(2 * (t.end - t.start));
    t.nodes.push_back(PtBvhNode{});
    t.inner.push_back(0);
    set_box(t.nodes[0], t.box);
    std::vector<Item> stack;
    stack.push_back({0, t.start, t.end});
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        Aabb box;
        for (int k = 0; k < 3; ++k) {
            box.mn[k] = t.nodes[it.node].aabb_min[k];
            box.mx[k] = t.nodes[it.node].aabb_max[k];
        }
        Decision d = decide(box, it.start, it.end, 1u);
        if (d.leaf) {
            t.nodes[it.node].left_first = (uint32_t)it.start;
            t.nodes[it.node].tri_count = (uint32_t)(it.end - it.start);
            continue;
        }
        const size_t lc = t.nodes.size();
        t.nodes.push_back(PtBvhNode{});
        t.nodes.push_back(PtBvhNode{});
        t.inner.push_back(0);
        t.inner.push_back(0);
        set_box(t.nodes[lc], bounds(it.start, d.mid, 1u));
        set_box(t.nodes[lc + 1], bounds(d.mid, it.end, 1u));
        t.nodes[it.node].left_first = (uint32_t)lc;
        t.nodes[it.node].tri_count = 0;
        t.inner[it.node] = 1;
        stack.push_back({lc + 1, d.mid, it.end});
        stack.push_back({lc, it.start, d.mid});
    }
}

void BvhBuilder::build(std::vector<PtBvhNode>& out) {
    const size_t n = tris_.size();
    top_.clear();
    tasks_.clear();
    top_.push_back(Top{0, n, bounds(0, n, chunks_for(n)), 2, 0});
    // split the largest pending node while it is worth more than one task's share
    const size_t target_tasks = threads_ <= 1 ? 1 : 8 * (size_t)threads_;
    auto larger = [&](size_t a, size_t b) { return top_[a].end - top_[a].start < top_[b].end - top_[b].start; };
    std::vector<size_t> pending{0};
    while (!pending.empty()) {
        std::pop_heap(pending.begin(), pending.end(), larger);
        const size_t x = pending.back();
        const size_t count = top_[x].end - top_[x].start;
        if (threads_ <= 1 || count < 2 * grain_ || pending.size() >= target_tasks) break;
        pending.pop_back();
        const Decision d = decide(top_[x].box, top_[x].start, top_[x].end, chunks_for(count));
        if (d.leaf) {
            top_[x].kind = 0;
            continue;
        }
        const size_t s = top_[x].start, e = top_[x].end;
        Top l{s, d.mid, bounds(s, d.mid, chunks_for(d.mid - s)), 2, 0};
        Top r{d.mid, e, bounds(d.mid, e, chunks_for(e - d.mid)), 2, 0};
        top_[x].kind = 1;
        top_[x].left = top_.size();
        top_.push_back(l);
        top_.push_back(r);
        pending.push_back(top_[x].left);
        std::push_heap(pending.begin(), pending.end(), larger);
        pending.push_back(top_[x].left + 1);
        std::push_heap(pending.begin(), pending.end(), larger);
    }
    // the rest are subtree tasks, handed out largest first
    std::sort(pending.begin(), pending.end(), [&](size_t a, size_t b) { return larger(b, a); });
    tasks_.resize(pending.size());
    for (size_t i = 0; i < pending.size(); ++i) {
        Top& tp = top_[pending[i]];
        tasks_[i].start = tp.start;
        tasks_[i].end = tp.end;
        tasks_[i].box = tp.box;
        tp.left = i;
    }
    std::atomic<size_t> next_task{0};
    run_parallel(std::min<size_t>(threads_, tasks_.size()), [&](unsigned) {
        for (size_t i; (i = next_task.fetch_add(1)) < tasks_.size();) build_task(tasks_[i]);
    });
    // number the nodes in the reference's allocation order: a split node takes
    // the next pair, a subtree takes its local count minus its root
    size_t total = 1;
    for (const Top& tp : top_)
        if (tp.kind == 1) total += 2;
    for (const Task& t : tasks_) total += t.nodes.size() - 1;
    if (total > 0xFFFFFFFFull) throw std::bad_alloc();
    out.assign(total, PtBvhNode{});
    std::vector<size_t> task_root(tasks_.size()), task_base(tasks_.size());
    struct Item {
        size_t top, global;
    };
    std::vector<Item> stack{{0, 0}};
    size_t next = 1;
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        const Top& tp = top_[it.top];
        PtBvhNode& nd = out[it.global];
        for (int k = 0; k < 3; ++k) {
            nd.aabb_min[k] = tp.box.mn[k];
            nd.aabb_max[k] = tp.box.mx[k];
        }
        if (tp.kind == 0) {
            nd.left_first = (uint32_t)tp.start;
            nd.tri_count = (uint32_t)(tp.end - tp.start);
        } else if (tp.kind == 1) {
            nd.left_first = (uint32_t)next;
            nd.tri_count = 0;
            stack.push_back({tp.left + 1, next + 1});
            stack.push_back({tp.left, next});
            next += 2;
        } else {
            task_root[tp.left] = it.global;
            task_base[tp.left] = next;
            next += tasks_[tp.left].nodes.size() - 1;
        }
    }
    // copy each subtree into its block, local index i >= 1 at base + i - 1
    std::atomic<size_t> next_copy{0};
    run_parallel(std::min<size_t>(threads_, tasks_.size()), [&](unsigned) {
        for (size_t i; (i = next_copy.fetch_add(1)) < tasks_.size();) {
            const Task& t = tasks_[i];
            const size_t base = task_base[i];
            auto place = [&](size_t local) { return local == 0 ? task_root[i] : base + local - 1; };
            for (size_t j = 0; j < t.nodes.size(); ++j) {
                PtBvhNode nd = t.nodes[j];
                if (t.inner[j]) nd.left_first = (uint32_t)place(nd.left_first);
                out[place(j)] = nd;
            }
        }
    });
}

// luminance of max(rgb, 0) in f64 (world.rs:578-581)
inline double luminance(const float* rgb) {
    double r = std::fmax((double)rgb[0], 0.0), g = std::fmax((double)rgb[1], 0.0), b = std::fmax((double)rgb[2], 0.0);
    return (r * 0.2126 + g * 0.7152) + b * 0.0722;
}
// exact equirectangular row solid angle without the longitude factor (world.rs:584-587)
inline double row_weight(uint32_t y, uint32_t height) {
    const double step = M_PI / (double)height;
    return std::cos((double)y * step) - std::cos((double)(y + 1) * step);
}

}  // namespace

namespace ptcore {

// Vose's alias method in f64 with LIFO small/large worklists (world.rs:539-575)
std::vector<PtSkyboxAliasEntry> alias_table(const std::vector<double>& weights) {
    double total = 0.0;
    for (double w : weights) total += w;
    std::vector<PtSkyboxAliasEntry> entries;
    if (!(total > 0.0)) return entries;
    const size_t n = weights.size();
    entries.resize(n);
    std::vector<double> scaled(n);
    std::vector<size_t> small, large;
    for (size_t i = 0; i < n; ++i) {
        entries[i].threshold = 1.0f;
        entries[i].alias_index = (uint32_t)i;
        entries[i].selection_probability = (float)(weights[i] / total);
        scaled[i] = weights[i] / total * (double)n;
    }
    for (size_t i = 0; i < n; ++i) (scaled[i] < 1.0 ? small : large).push_back(i);
    while (!small.empty() && !large.empty()) {
        size_t low = small.back(), high = large.back();
        small.pop_back();
        large.pop_back();
        entries[low].threshold = (float)std::min(std::max(scaled[low], 0.0), 1.0);
        entries[low].alias_index = (uint32_t)high;
        scaled[high] = (scaled[high] + scaled[low]) - 1.0;
        (scaled[high] < 1.0 ? small : large).push_back(high);
    }
    return entries;
}

// GpuSceneUniforms::update_light_probabilities (world.rs:484-498)
void update_light_probabilities(PtSceneUniforms& u, float skybox_flux) {
    float geometry_flux = u.emissive_triangles_count > 0 ? u.total_luminous_flux : 0.0f;
    float scale = std::fmax(geometry_flux, skybox_flux);
    if (scale > 0.0f) {
        float g = geometry_flux / scale, s = skybox_flux / scale;
        u.light_probabilities[0] = g / (g + s);
        u.light_probabilities[1] = s / (g + s);
    } else {
        u.light_probabilities[0] = u.light_probabilities[1] = 0.0f;
    }
}

// Scene::uniforms_and_alias_table (world.rs:1691-1735)
void scene_lights(const World& w, PtSceneUniforms& u, std::vector<PtSceneAliasEntry>& out) {
    std::vector<uint32_t> indices;
    std::vector<double> weights;
    for (size_t i = 0; i < w.geometries.size(); ++i) {
        const PtMaterial& m = w.materials[w.attributes[i].material];
        double emission = luminance(m.emission_color_and_luminance) *
                          (double)std::fmax(m.emission_color_and_luminance[3], 0.0f) *
                          (double)std::fmax(m.emission_weight, 0.0f);
        if (emission == 0.0) continue;
        const float* v = w.geometries[i].v;
        double ax = (double)v[3] - v[0], ay = (double)v[4] - v[1], az = (double)v[5] - v[2];
        double bx = (double)v[6] - v[0], by = (double)v[7] - v[1], bz = (double)v[8] - v[2];
        double cx = ay * bz - by * az, cy = az * bx - bz * ax, cz = ax * by - bx * ay;
        double area = std::sqrt((cx * cx + cy * cy) + cz * cz) * 0.5;
        double flux = 2.0 * M_PI * area * emission;
        if (flux > 0.0 && std::isfinite(flux)) {
            indices.push_back((uint32_t)i);
            weights.push_back(flux);
        }
    }
    double total = 0.0;
    for (double x : weights) total += x;
    u.total_luminous_flux = (float)std::min(total, (double)std::numeric_limits<float>::max());
    u.emissive_triangles_count = (uint32_t)indices.size();
    u.light_probabilities[0] = u.light_probabilities[1] = 0.0f;
    std::vector<PtSkyboxAliasEntry> t = alias_table(weights);
    out.resize(t.size());
    for (size_t i = 0; i < t.size(); ++i)
        out[i] = PtSceneAliasEntry{t[i].threshold, t[i].alias_index, indices[i], t[i].selection_probability};
}

// Camera::gpu_camera (world.rs:2159-2173) with glam's Mat3 products
PtCamera gpu_camera(const float pos[3], float yaw, float pitch, float fov, float aspect) {
    float sy = sinf(yaw), cy = cosf(yaw), sp = sinf(pitch), cp = cosf(pitch);
    // columns of Ry and Rx
    float ry[3][3] = {{cy, 0.0f, -sy}, {0.0f, 1.0f, 0.0f}, {sy, 0.0f, cy}};
    float rx[3][3] = {{1.0f, 0.0f, 0.0f}, {0.0f, cp, sp}, {0.0f, -sp, cp}};
    auto mulv = [](const float m[3][3], const float v[3], float out[3]) {
        for (int k = 0; k < 3; ++k) out[k] = (m[0][k] * v[0] + m[1][k] * v[1]) + m[2][k] * v[2];
    };
    float rot[3][3];
    for (int c = 0; c < 3; ++c) mulv(ry, rx[c], rot[c]);
    const float X[3] = {1.0f, 0.0f, 0.0f}, NY[3] = {0.0f, -1.0f, 0.0f}, NZ[3] = {0.0f, 0.0f, -1.0f};
    float height = 2.0f * tanf(fov * 0.5f);
    float rxv[3], rny[3], rnz[3];
    mulv(rot, X, rxv);
    mulv(rot, NY, rny);
    mulv(rot, NZ, rnz);
    PtCamera c{};
    for (int k = 0; k < 3; ++k) {
        float h = rxv[k] * aspect * height;
        float v = rny[k] * height;
        c.position[k] = pos[k];
        c.horizontal[k] = h;
        c.vertical[k] = v;
    }
    for (int k = 0; k < 3; ++k)
        c.lower_left_corner_minus_position[k] = rnz[k] - (c.horizontal[k] + c.vertical[k]) * 0.5f;
    return c;
}

// Scene::pick_material (world.rs:1595-1688): its own traversal (boxes tested
// at pop, left child popped first, no distance ordering) and its own
// triangle-distance expression, so the pick equals the reference's.
int32_t pick_material(const World& w, const float pos[3], float yaw, float pitch) {
    float sy = sinf(yaw), cy = cosf(yaw), sp = sinf(pitch), cp = cosf(pitch);
    float ry[3][3] = {{cy, 0.0f, -sy}, {0.0f, 1.0f, 0.0f}, {sy, 0.0f, cy}};
    float rx[3][3] = {{1.0f, 0.0f, 0.0f}, {0.0f, cp, sp}, {0.0f, -sp, cp}};
    auto mulv = [](const float m[3][3], const float v[3], float out[3]) {
        for (int k = 0; k < 3; ++k) out[k] = (m[0][k] * v[0] + m[1][k] * v[1]) + m[2][k] * v[2];
    };
    float rot[3][3];
    for (int c = 0; c < 3; ++c) mulv(ry, rx[c], rot[c]);
    const float NZ[3] = {0.0f, 0.0f, -1.0f};
    float d[3], rcp[3], ad[3];
    mulv(rot, NZ, d);
    for (int k = 0; k < 3; ++k) {
        rcp[k] = 1.0f / d[k];
        ad[k] = std::fabs(d[k]);
    }
    const int kxy = ad[0] <= ad[1] ? 1 : 0;
    const int kz = ad[kxy] > ad[2] ? kxy : 2;
    const int kx_val = kz == 2 ? 0 : kz + 1;
    const int ky_val = kx_val == 2 ? 0 : kx_val + 1;
    const bool is_neg = d[kz] < 0.0f;
    const int kx = is_neg ? ky_val : kx_val, ky = is_neg ? kx_val : ky_val;
    const float sz = rcp[kz], sx = -d[kx] * sz, syy = -d[ky] * sz;
    std::vector<uint32_t> stack{0u};
    float closest_t = std::numeric_limits<float>::max();
    int64_t hit = -1;
    while (!stack.empty()) {
        const PtBvhNode& node = w.bvh_nodes[stack.back()];
        stack.pop_back();
        float t_near = -INFINITY, t_far = INFINITY;
        for (int k = 0; k < 3; ++k) {
            float t0 = (node.aabb_min[k] - pos[k]) * rcp[k], t1 = (node.aabb_max[k] - pos[k]) * rcp[k];
            float lo = std::fmin(t0, t1), hi = std::fmax(t0, t1);
            t_near = k == 0 ? lo : std::fmax(t_near, lo);
            t_far = k == 0 ? hi : std::fmin(t_far, hi);
        }
        if (!(t_near <= t_far && t_far > 0.0f && t_near < closest_t)) continue;
        if (node.tri_count > 0) {
            for (uint32_t i = 0; i < node.tri_count; ++i) {
                const uint32_t ti = node.left_first + i;
                const float* v = w.geometries[ti].v;
                float a[3], b[3], c[3];
                for (int k = 0; k < 3; ++k) {
                    a[k] = v[k] - pos[k];
                    b[k] = v[3 + k] - pos[k];
                    c[k] = v[6 + k] - pos[k];
                }
                float ax = std::fmaf(sx, a[kz], a[kx]), ay = std::fmaf(syy, a[kz], a[ky]);
                float bx = std::fmaf(sx, b[kz], b[kx]), by = std::fmaf(syy, b[kz], b[ky]);
                float cx = std::fmaf(sx, c[kz], c[kx]), cyy = std::fmaf(syy, c[kz], c[ky]);
                float u = std::fmaf(cyy, -bx, cx * by);
                float vv = std::fmaf(ay, -cx, ax * cyy);
                float ww = std::fmaf(by, -ax, bx * ay);
                if (std::fmin(u, std::fmin(vv, ww)) < 0.0f && std::fmax(u, std::fmax(vv, ww)) > 0.0f) continue;
                float det = u + vv + ww;
                if (det == 0.0f) continue;
                float t = std::fmaf(ww * sz, c[kz], std::fmaf(u * sz, a[kz], vv * sz * b[kz])) / det;
                if (t > 0.0f && t < closest_t) {
                    closest_t = t;
                    hit = ti;
                }
            }
        } else {
            stack.push_back(node.left_first + 1u);
            stack.push_back(node.left_first);
        }
    }
    return hit < 0 ? -1 : (int32_t)w.attributes[(size_t)hit].material;
}

// Skybox::update_total_luminous_flux (world.rs:1977-1987)
void update_sky_flux(World& w) {
    w.skybox.total_luminous_flux =
        w.sky_alias.empty() ? 0.0f
                            : (float)std::min(4.0 * (M_PI * M_PI) * w.radius_squared * (double)w.skybox.average_luminance,
                                              (double)std::numeric_limits<float>::max());
}

}  // namespace ptcore

using namespace ptcore;

extern "C" {

pt_status pt_world_build(const pt_world_desc* d, pt_world** out) {
    if (!d || !out) return set_error(PT_ERR_INVALID, "pt_world_build: null argument");
    *out = nullptr;
    if (d->triangle_count == 0) return set_error(PT_ERR_INVALID, "scene has no triangles");
    if (!d->geometries || !d->attributes) return set_error(PT_ERR_INVALID, "missing triangle arrays");
    if (d->material_count == 0 || !d->materials) return set_error(PT_ERR_INVALID, "scene has no materials");
    for (uint32_t i = 0; i < d->triangle_count; ++i)
        if (d->attributes[i].material >= d->material_count)
            return set_error(PT_ERR_INVALID, "triangle %u uses material %u of %u", i, d->attributes[i].material,
                             d->material_count);
    if ((d->srgb_atlas_count || d->linear_atlas_count) && d->atlas_size == 0)
        return set_error(PT_ERR_INVALID, "atlas_size is 0");
    for (uint32_t i = 0; i < d->material_count; ++i) {
        const PtMaterial& m = d->materials[i];
        int32_t srgb_layers[3] = {m.base_color_tex_layer, m.emission_tex_layer, m.specular_color_tex_layer};
        int32_t lin_layers[3] = {m.normal_tex_layer, m.metallic_roughness_tex_layer, m.specular_weight_tex_layer};
        for (int k = 0; k < 3; ++k) {
            if (srgb_layers[k] >= (int32_t)d->srgb_atlas_count)
                return set_error(PT_ERR_INVALID, "material %u: sRGB atlas layer %d out of range", i, srgb_layers[k]);
            if (lin_layers[k] >= (int32_t)d->linear_atlas_count)
                return set_error(PT_ERR_INVALID, "material %u: linear atlas layer %d out of range", i, lin_layers[k]);
        }
    }
    if (!d->skybox_rgba || d->skybox_width == 0 || d->skybox_height == 0)
        return set_error(PT_ERR_INVALID, "missing skybox");
    pt_world* w = new (std::nothrow) pt_world();
    if (!w) return set_error(PT_ERR_MEMORY, "out of host memory");
    try {
        const uint32_t n = d->triangle_count;
        // BVH triangles in input order (world.rs:1290-1297)
        const unsigned threads = host_threads();
        const size_t grain = bvh_task_grain();
        // element-wise passes: one chunk per thread once there is enough work
        const unsigned pass_threads = (unsigned)std::max<size_t>(1, std::min<size_t>(threads, n / (4 * grain)));
        auto each = [&](auto&& f) {
            run_parallel(pass_threads, [&](unsigned t) {
                const size_t a = (size_t)n * t / pass_threads, e = (size_t)n * (t + 1) / pass_threads;
                for (size_t i = a; i < e; ++i) f(i);
            });
        };
        std::vector<BvhTri> tris(n);
        each([&](size_t i) {
            const float* v = d->geometries[i].v;
            BvhTri& t = tris[i];
            for (int k = 0; k < 3; ++k) {
                t.box.mn[k] = min_f(min_f(v[k], v[3 + k]), v[6 + k]);
                t.box.mx[k] = max_f(max_f(v[k], v[3 + k]), v[6 + k]);
                t.centroid[k] = ((v[k] + v[3 + k]) + v[6 + k]) / 3.0f;
            }
            t.original = (uint32_t)i;
        });
        BvhBuilder(tris, threads, grain).build(w->bvh_nodes);
        w->geometries.resize(n);
        w->attributes.resize(n);
        each([&](size_t i) {
            w->geometries[i] = d->geometries[tris[i].original];
            w->attributes[i] = d->attributes[tris[i].original];
        });
        w->materials.assign(d->materials, d->materials + d->material_count);
        w->atlas_size = d->atlas_size;
        const size_t layer_bytes = (size_t)d->atlas_size * d->atlas_size * 4;
        if (d->srgb_atlas_count) w->srgb.assign(d->srgb_atlases, d->srgb_atlases + layer_bytes * d->srgb_atlas_count);
        if (d->linear_atlas_count)
            w->linear.assign(d->linear_atlases, d->linear_atlases + layer_bytes * d->linear_atlas_count);
        w->srgb_count = d->srgb_atlas_count;
        w->linear_count = d->linear_atlas_count;

        // Skybox::new (world.rs:1872-1957)
        const uint32_t W = d->skybox_width, H = d->skybox_height;
        std::vector<float> px(d->skybox_rgba, d->skybox_rgba + (size_t)W * H * 4);
        double total_luma = 0.0;
        const double total_weight = 2.0 * (double)W;
        for (uint32_t y = 0; y < H; ++y) {
            double wgt = row_weight(y, H);
            for (uint32_t x = 0; x < W; ++x) total_luma = std::fma(luminance(&px[((size_t)y * W + x) * 4]), wgt, total_luma);
        }
        if (total_luma == 0.0) {
            w->sky.assign(4, 0.0f);
            w->skybox = PtSkyboxUniforms{0.0f, 0.0f, 1u, 1u};
        } else {
            double avg = total_luma / total_weight;
            double scale = 1.0 / avg;
            for (size_t i = 0; i < px.size(); i += 4)
                for (int c = 0; c < 3; ++c) px[i + c] = (float)((double)px[i + c] * scale);
            w->sky = std::move(px);
            w->skybox = PtSkyboxUniforms{0.0f, (float)avg, W, H};
        }
        // World::edit_skybox_average_luminance (world.rs:646-648)
        if (d->skybox_luminance >= 0.0f) w->skybox.average_luminance = std::fmax(d->skybox_luminance, 0.0f);

        // World::take_updates ordering (world.rs:833-848)
        scene_lights(*w, w->scene_uniforms, w->scene_alias);
        if (total_luma != 0.0) {
            std::vector<double> weights;
            weights.reserve((size_t)W * H);
            for (uint32_t y = 0; y < H; ++y) {
                double wgt = row_weight(y, H);
                for (uint32_t x = 0; x < W; ++x) weights.push_back(luminance(&w->sky[((size_t)y * W + x) * 4]) * wgt);
            }
            w->sky_alias = alias_table(weights);
        }
        const PtBvhNode& root = w->bvh_nodes[0];
        double dx = (double)root.aabb_max[0] - root.aabb_min[0], dy = (double)root.aabb_max[1] - root.aabb_min[1],
               dz = (double)root.aabb_max[2] - root.aabb_min[2];
        w->radius_squared = ((dx * dx + dy * dy) + dz * dz) * 0.25;
        update_sky_flux(*w);
        update_light_probabilities(w->scene_uniforms, w->skybox.total_luminous_flux);

        // camera: Camera::new from the root box, then the headless options (headless.rs:193-202)
        const float* mn = root.aabb_min;
        const float* mx = root.aabb_max;
        float pos[3] = {(mn[0] + mx[0]) * 0.5f, (mn[1] + mx[1]) * 0.5f,
                        std::fmaf(std::fmax(mx[0] - mn[0], mx[1] - mn[1]), 0.6f, mx[2])};
        for (int k = 0; k < 3; ++k) w->initial_cam.position[k] = pos[k];
        if (!std::isnan(d->camera_position[0]))
            for (int k = 0; k < 3; ++k) pos[k] = d->camera_position[k];
        const float rads_per_deg = 3.14159265358979323846f / 180.0f;
        float yaw = std::isnan(d->camera_yaw_degrees) ? 0.0f : d->camera_yaw_degrees * rads_per_deg;
        float pitch = std::isnan(d->camera_pitch_degrees) ? 0.0f : d->camera_pitch_degrees * rads_per_deg;
        const float half_pi = 1.57079632679489661923f;
        pitch = std::fmin(std::fmax(pitch, -half_pi), half_pi);
        float fov = half_pi;
        if (!std::isnan(d->camera_fov_degrees)) {
            fov = d->camera_fov_degrees * rads_per_deg;
            fov = std::fmin(std::fmax(fov, 0.0f), 3.14159265358979323846f - std::numeric_limits<float>::epsilon());
        }
        float aspect = (std::isfinite(d->aspect_ratio) && d->aspect_ratio > 0.0f) ? d->aspect_ratio : 1.0f;
        // the initial camera is Camera::new resized to the aspect ratio (world.rs:596-597)
        w->initial_cam.yaw = 0.0f;
        w->initial_cam.pitch = 0.0f;
        w->initial_cam.fov = half_pi;
        w->initial_cam.aspect_ratio = aspect;
        w->camera = gpu_camera(pos, yaw, pitch, fov, aspect);
        for (int k = 0; k < 3; ++k) w->cam.position[k] = pos[k];
        w->cam.yaw = yaw;
        w->cam.pitch = pitch;
        w->cam.fov = fov;
        w->cam.aspect_ratio = aspect;
        w->initial_materials = w->materials;
        w->selected_material = pick_material(*w, pos, yaw, pitch);
    } catch (const std::bad_alloc&) {
        delete w;
        return set_error(PT_ERR_MEMORY, "out of host memory while building the world");
    } catch (const BvhNoProgress&) {
        delete w;
        return set_error(PT_ERR_INVALID,
                         "BVH split made no progress: vertex positions are non-finite or too large for the "
                         "surface-area costs");
    } catch (const std::system_error& e) {
        delete w;
        return set_error(PT_ERR_MEMORY, "could not start host threads: %s", e.what());
    } catch (const std::exception& e) {
        delete w;
        return set_error(PT_ERR_INVALID, "world build failed: %s", e.what());
    }
    *out = w;
    return PT_OK;
}

pt_status pt_world_destroy(pt_world* w) {
    delete w;
    return PT_OK;
}

pt_status pt_world_scene_desc(const pt_world* wp, pt_scene_desc* s) {
    if (!wp || !s) return set_error(PT_ERR_INVALID, "pt_world_scene_desc: null argument");
    const World& w = *wp;
    std::memset(s, 0, sizeof(*s));
    s->geometries = w.geometries.data();
    s->attributes = w.attributes.data();
    s->triangle_count = (uint32_t)w.geometries.size();
    s->materials = w.materials.data();
    s->material_count = (uint32_t)w.materials.size();
    s->bvh_nodes = w.bvh_nodes.data();
    s->bvh_node_count = (uint32_t)w.bvh_nodes.size();
    s->srgb_atlases = w.srgb.empty() ? nullptr : w.srgb.data();
    s->srgb_atlas_count = w.srgb_count;
    s->linear_atlases = w.linear.empty() ? nullptr : w.linear.data();
    s->linear_atlas_count = w.linear_count;
    s->atlas_size = w.atlas_size;
    s->camera = w.camera;
    s->skybox_rgba = w.sky.data();
    s->skybox = w.skybox;
    s->skybox_alias = w.sky_alias.empty() ? nullptr : w.sky_alias.data();
    s->skybox_alias_count = (uint32_t)w.sky_alias.size();
    s->scene_uniforms = w.scene_uniforms;
    s->scene_alias = w.scene_alias.empty() ? nullptr : w.scene_alias.data();
    s->scene_alias_count = (uint32_t)w.scene_alias.size();
    return PT_OK;
}

pt_status pt_world_camera(const pt_world* w, int initial, pt_camera_params* out) {
    if (!w || !out) return set_error(PT_ERR_INVALID, "pt_world_camera: null argument");
    *out = initial ? w->initial_cam : w->cam;
    return PT_OK;
}

pt_status pt_world_edit_camera(pt_world* w, const pt_camera_params* c) {
    if (!w || !c) return set_error(PT_ERR_INVALID, "pt_world_edit_camera: null argument");
    w->pending_camera = true;
    w->pending_cam = *c;
    return PT_OK;
}

pt_status pt_world_material(const pt_world* w, uint32_t index, PtMaterial* current, PtMaterial* initial) {
    if (!w) return set_error(PT_ERR_INVALID, "pt_world_material: null argument");
    if (index >= w->materials.size()) return set_error(PT_ERR_INVALID, "material %u out of range", index);
    if (current) *current = w->materials[index];
    if (initial) *initial = w->initial_materials[index];
    return PT_OK;
}

pt_status pt_world_edit_material(pt_world* w, uint32_t index, const PtMaterial* candidate, int* accepted) {
    if (!w || !candidate) return set_error(PT_ERR_INVALID, "pt_world_edit_material: null argument");
    if (accepted) *accepted = 0;
    if (index >= w->materials.size()) return PT_OK;  // the reference ignores unknown materials
    auto it = w->pending_materials.find(index);
    const PtMaterial original = it != w->pending_materials.end() ? it->second : w->materials[index];
    // the editable fields are the f32 prefix before alpha_mode
    constexpr size_t editable = offsetof(PtMaterial, alpha_mode);
    float vals[editable / sizeof(float)];
    std::memcpy(vals, candidate, editable);
    for (float v : vals)
        if (!std::isfinite(v)) return PT_OK;
    PtMaterial acc = original;
    std::memcpy(&acc, candidate, editable);
    if (std::memcmp(&acc, &w->materials[index], sizeof(PtMaterial)) == 0) w->pending_materials.erase(index);
    else w->pending_materials[index] = acc;
    if (accepted) *accepted = 1;
    return PT_OK;
}

pt_status pt_world_edit_emission_scale(pt_world* w, float scale) {
    if (!w) return set_error(PT_ERR_INVALID, "pt_world_edit_emission_scale: null argument");
    scale = std::fmax(scale, 0.0f);
    w->emission_scale = scale;
    for (uint32_t i = 0; i < w->materials.size(); ++i) {
        auto it = w->pending_materials.find(i);
        PtMaterial e = it != w->pending_materials.end() ? it->second : w->materials[i];
        // emissive: luminance(emission colour) * emission weight > 0 (world.rs:676-681)
        if (luminance(e.emission_color_and_luminance) * (double)e.emission_weight <= 0.0) continue;
        e.emission_color_and_luminance[3] = w->initial_materials[i].emission_color_and_luminance[3] * scale;
        if (std::memcmp(&e, &w->materials[i], sizeof(PtMaterial)) == 0) w->pending_materials.erase(i);
        else w->pending_materials[i] = e;
    }
    return PT_OK;
}

pt_status pt_world_emission_scale(const pt_world* w, float* scale) {
    if (!w || !scale) return set_error(PT_ERR_INVALID, "pt_world_emission_scale: null argument");
    *scale = w->emission_scale;
    return PT_OK;
}

pt_status pt_world_edit_skybox_luminance(pt_world* w, float luminance) {
    if (!w) return set_error(PT_ERR_INVALID, "pt_world_edit_skybox_luminance: null argument");
    w->pending_luminance = true;
    w->pending_lum = std::fmax(luminance, 0.0f);
    return PT_OK;
}

pt_status pt_world_pick_material(const pt_world* w, const float position[3], float yaw, float pitch,
                                 int32_t* material) {
    if (!w || !position || !material) return set_error(PT_ERR_INVALID, "pt_world_pick_material: null argument");
    *material = pick_material(*w, position, yaw, pitch);
    return PT_OK;
}

pt_status pt_world_take_updates(pt_world* w, pt_world_updates* out) {
    if (!w || !out) return set_error(PT_ERR_INVALID, "pt_world_take_updates: null argument");
    std::memset(out, 0, sizeof(*out));
    bool camera_changed = false, selection_changed = false;
    if (w->pending_camera) {
        const pt_camera_params& c = w->pending_cam;
        selection_changed = c.position[0] != w->cam.position[0] || c.position[1] != w->cam.position[1] ||
                            c.position[2] != w->cam.position[2] || c.yaw != w->cam.yaw || c.pitch != w->cam.pitch;
        camera_changed = selection_changed || c.fov != w->cam.fov || c.aspect_ratio != w->cam.aspect_ratio;
        w->cam = c;
        w->pending_camera = false;
    }
    bool luminance_changed = false;
    if (w->pending_luminance) {
        luminance_changed = w->pending_lum != w->skybox.average_luminance;
        w->skybox.average_luminance = w->pending_lum;
        w->pending_luminance = false;
    }
    bool lighting_changed = false;
    w->update_indices.clear();
    w->update_materials.clear();
    for (const auto& [index, m] : w->pending_materials) {
        PtMaterial& old = w->materials[index];
        lighting_changed = lighting_changed || old.emission_weight != m.emission_weight ||
                           std::memcmp(old.emission_color_and_luminance, m.emission_color_and_luminance,
                                       sizeof(m.emission_color_and_luminance)) != 0;
        old = m;
        w->update_indices.push_back(index);
        w->update_materials.push_back(m);
    }
    w->pending_materials.clear();
    if (lighting_changed) scene_lights(*w, w->scene_uniforms, w->scene_alias);
    const bool skybox_uniforms_changed = luminance_changed;
    if (skybox_uniforms_changed) update_sky_flux(*w);
    const bool scene_uniforms_changed = lighting_changed || skybox_uniforms_changed;
    if (scene_uniforms_changed) update_light_probabilities(w->scene_uniforms, w->skybox.total_luminous_flux);
    if (camera_changed) w->camera = gpu_camera(w->cam.position, w->cam.yaw, w->cam.pitch, w->cam.fov, w->cam.aspect_ratio);
    if (selection_changed) w->selected_material = pick_material(*w, w->cam.position, w->cam.yaw, w->cam.pitch);

    out->camera_changed = camera_changed;
    out->camera = w->camera;
    out->material_count = (uint32_t)w->update_indices.size();
    out->material_indices = w->update_indices.empty() ? nullptr : w->update_indices.data();
    out->materials = w->update_materials.empty() ? nullptr : w->update_materials.data();
    out->scene_uniforms_changed = scene_uniforms_changed;
    out->scene_uniforms = w->scene_uniforms;
    out->scene_alias_changed = lighting_changed;
    out->scene_alias = w->scene_alias.empty() ? nullptr : w->scene_alias.data();
    out->scene_alias_count = (uint32_t)w->scene_alias.size();
    out->skybox_uniforms_changed = skybox_uniforms_changed;
    out->skybox_uniforms = w->skybox;
    out->reset_accumulation = camera_changed || luminance_changed || !w->update_indices.empty();
    out->selected_material = w->selected_material;
    return PT_OK;
}

}  // extern "C"
