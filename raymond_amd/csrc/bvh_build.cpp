// The scene BVH built natively on the host: the tree of src/BVH.mjs:57-234
// (binned SAH over 32 centroid bins per axis, leaves of at most
// max_leaf_size primitives unless splitting pays, the in-place two-pointer
// partition, a median fallback), emitted in the same 32-byte node layout
// and the same node numbering, bit for bit.
//
// Every split decision depends only on the primitives of the node's own
// range, and partitions stay inside that range, so disjoint subtrees are
// built on separate threads. The numbering of the sequential builder — the
// two children of a split are appended together, in the order a LIFO stack
// pops the nodes it splits — is then recovered by replaying that stack over
// the finished tree, which is cheap (one pass over the nodes).
//
// Arithmetic is double precision, in the order the JS builder writes it
// (compiled without FMA contraction), and minima and maxima follow
// Math.min / Math.max: a NaN wins, and -0 is below +0. Bins are filled in
// one pass for all three axes; each bin's count and box are the same
// whatever order its primitives arrive in.
#include <raymond.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

namespace rmi {
int fail(int code, const std::string& message);
}

namespace {

constexpr int BINS = 32;
constexpr double INF = std::numeric_limits<double>::infinity();
// Subtrees of fewer primitives are built by one thread from start to end.
constexpr uint32_t GRAIN = 4096;
// Nodes of this many primitives or more are binned by all threads at once.
constexpr uint32_t PARALLEL_BINS = 1u << 18;

inline uint64_t bits(double x) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return u;
}

inline double from_bits(uint64_t u) {
    double x;
    std::memcpy(&x, &u, 8);
    return x;
}

// Selects rather than branches: bin boxes grow in data-dependent order. Of
// two equal values, OR-ing their bits picks -0 over +0 (and is the value
// itself otherwise); AND-ing picks +0.
inline double js_min(double a, double b) {
    const double r = a == b ? from_bits(bits(a) | bits(b)) : (a < b ? a : b);
    return a != a || b != b ? a + b : r;
}

inline double js_max(double a, double b) {
    const double r = a == b ? from_bits(bits(a) & bits(b)) : (a > b ? a : b);
    return a != a || b != b ? a + b : r;
}

// Without NaN or -0 among the values, Math.min and Math.max are the plain
// comparisons (equal values then have equal bits). Boxes only ever hold
// input values, and a centroid is -0 only when both its inputs are, so a
// scan of the input decides which the whole build may use.
template <bool Exact> inline double min_of(double a, double b) { return Exact ? js_min(a, b) : (b < a ? b : a); }
template <bool Exact> inline double max_of(double a, double b) { return Exact ? js_max(a, b) : (b > a ? b : a); }

bool needs_exact(const double* v, size_t n) {
    bool odd = false;
    for (size_t i = 0; i < n; ++i) odd |= v[i] != v[i] || (v[i] == 0.0 && std::signbit(v[i]));
    return odd;
}

struct Box {
    double lo[3], hi[3];
    void clear() {
        lo[0] = lo[1] = lo[2] = INF;
        hi[0] = hi[1] = hi[2] = -INF;
    }
    template <bool E> void grow(const double* mn, const double* mx) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = min_of<E>(lo[a], mn[a]);
            hi[a] = max_of<E>(hi[a], mx[a]);
        }
    }
    template <bool E> void grow(const Box& b) { grow<E>(b.lo, b.hi); }
    double half_area() const {
        const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.0 && dy >= 0.0 && dz >= 0.0)) return 0.0;
        return dx * dy + dy * dz + dz * dx;
    }
};

// left: the first of the two children (the second is left + 1) once the
// node is split; 0 for a leaf, since the root is nobody's child.
struct Node {
    Box box;
    uint32_t first, count, left;
};

// A primitive's bounds travel with it through the partitions, so that every
// pass over a node's range reads memory in order; id is its original index.
struct Prim {
    double lo[3], hi[3];
    uint32_t id;
    double centroid(int a) const { return (lo[a] + hi[a]) * 0.5; }
};

// The bounds of a node's centroids, which its bins divide.
struct Centroids {
    double lo[3], hi[3];
    void clear() {
        lo[0] = lo[1] = lo[2] = INF;
        hi[0] = hi[1] = hi[2] = -INF;
    }
    template <bool E> void grow(const Prim& p) {
        for (int a = 0; a < 3; ++a) {
            const double c = p.centroid(a);
            lo[a] = min_of<E>(lo[a], c);
            hi[a] = max_of<E>(hi[a], c);
        }
    }
};

// A node to split, with its centroids' bounds.
struct Task {
    uint32_t node;
    Centroids c;
};

// A node's bins on each axis. Only occupied bins hold a box and count: an
// empty bin leaves both sweeps' boxes and counts as they were, so the cost
// of a split before bin b is the same for every b after one occupied bin up
// to the next, and the first of them — just after the occupied bin — is the
// one a sweep over all 32 keeps.
struct Bins {
    Box box[3][BINS];
    uint32_t count[3][BINS];
    uint32_t occupied[3];

    template <bool E> void add(int a, int b, const Box& x, uint32_t n) {
        if (occupied[a] >> b & 1u) {
            count[a][b] += n;
            box[a][b].grow<E>(x);
        } else {
            occupied[a] |= 1u << b;
            count[a][b] = n;
            box[a][b] = x;
        }
    }
};

template <bool E> struct Builder {
    Prim* prims;
    Node* nodes;
    uint32_t max_leaf;
    unsigned threads;   // for binning nodes of PARALLEL_BINS primitives or more
    std::atomic<uint32_t> used{1};

    void bin(uint32_t first, uint32_t end, const Centroids& c, const bool live[3], const double scale[3],
             Bins& out) const {
        out.occupied[0] = out.occupied[1] = out.occupied[2] = 0;
        for (uint32_t i = first; i < end; ++i) {
            const Prim& p = prims[i];
            for (int a = 0; a < 3; ++a) {
                if (!live[a]) continue;
                const double f = std::floor((p.centroid(a) - c.lo[a]) * scale[a]);
                const int b = f < (double)(BINS - 1) ? (int)f : BINS - 1;
                if (out.occupied[a] >> b & 1u) {
                    ++out.count[a][b];
                } else {
                    out.occupied[a] |= 1u << b;
                    out.count[a][b] = 1;
                    out.box[a][b].clear();
                }
                out.box[a][b].template grow<E>(p.lo, p.hi);
            }
        }
    }

    // The same bins from `threads` slices of the range, merged: counts add,
    // and boxes grow the same in any order.
    void bin_parallel(uint32_t first, uint32_t end, const Centroids& c, const bool live[3], const double scale[3],
                      Bins& out) const {
        const unsigned t = threads;
        std::vector<Bins> part(t);
        std::vector<std::thread> pool;
        const uint32_t n = end - first;
        auto slice = [&](unsigned k) { return first + (uint32_t)((uint64_t)n * k / t); };
        for (unsigned k = 1; k < t; ++k)
            pool.emplace_back([&, k] { bin(slice(k), slice(k + 1), c, live, scale, part[k]); });
        bin(slice(0), slice(1), c, live, scale, part[0]);
        for (auto& th : pool) th.join();
        out = part[0];
        for (unsigned k = 1; k < t; ++k)
            for (int a = 0; a < 3; ++a)
                for (uint32_t m = part[k].occupied[a]; m; m &= m - 1) {
                    const int b = __builtin_ctz(m);
                    out.add<E>(a, b, part[k].box[a][b], part[k].count[a][b]);
                }
    }

    void measure(uint32_t first, uint32_t n, Box& box, Centroids& c) const {
        box.clear();
        c.clear();
        for (uint32_t i = first; i < first + n; ++i) {
            box.grow<E>(prims[i].lo, prims[i].hi);
            c.template grow<E>(prims[i]);
        }
    }

    // src/BVH.mjs:99-209 for one popped node: leaves it a leaf (false), or
    // splits it and gives its children's tasks.
    bool split(const Task& t, Task out[2]) {
        Node& node = nodes[t.node];
        const uint32_t first = node.first, n = node.count, end = first + n;

        bool live[3];
        double scale[3];
        for (int a = 0; a < 3; ++a) {
            const double extent = t.c.hi[a] - t.c.lo[a];
            live[a] = extent > 0.0;
            scale[a] = live[a] ? BINS / extent : 0.0;
        }

        Bins bins;
        bins.occupied[0] = bins.occupied[1] = bins.occupied[2] = 0;
        if (live[0] || live[1] || live[2]) {
            if (threads > 1 && n >= PARALLEL_BINS)
                bin_parallel(first, end, t.c, live, scale, bins);
            else
                bin(first, end, t.c, live, scale, bins);
        }

        double best_cost = INF, best_split = 0.0;
        int best_axis = -1;
        for (int a = 0; a < 3; ++a) {
            if (!live[a]) continue;
            int at[BINS], k = 0;
            for (uint32_t m = bins.occupied[a]; m; m &= m - 1) at[k++] = __builtin_ctz(m);
            double right_area[BINS];
            uint32_t right_count[BINS];
            Box box;
            box.clear();
            uint32_t total = 0;
            for (int i = k - 1; i > 0; --i) {
                box.grow<E>(bins.box[a][at[i]]);
                total += bins.count[a][at[i]];
                right_area[i] = box.half_area();
                right_count[i] = total;
            }
            box.clear();
            uint32_t left_count = 0;
            for (int i = 0; i + 1 < k; ++i) {
                box.grow<E>(bins.box[a][at[i]]);
                left_count += bins.count[a][at[i]];
                const double cost = (double)left_count * box.half_area() + (double)right_count[i + 1] * right_area[i + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_split = t.c.lo[a] + (double)(at[i] + 1) / scale[a];
                }
            }
        }

        const double parent_area = node.box.half_area();
        uint32_t mid;
        Box lbox, rbox;
        bool measured = false;
        if (best_axis < 0) {
            if (n <= max_leaf) return false;
            mid = first + (n >> 1);
        } else {
            if (n <= max_leaf && best_cost + parent_area >= (double)n * parent_area) return false;
            // The two-pointer partition; each primitive is final once it is
            // passed on the left or swapped to the right, so the children's
            // bounds are gathered on the way.
            lbox.clear();
            rbox.clear();
            out[0].c.clear();
            out[1].c.clear();
            int64_t i = first, j = (int64_t)end - 1;
            while (i <= j) {
                if (prims[i].centroid(best_axis) < best_split) {
                    lbox.grow<E>(prims[i].lo, prims[i].hi);
                    out[0].c.template grow<E>(prims[i]);
                    ++i;
                } else {
                    std::swap(prims[i], prims[j]);
                    rbox.grow<E>(prims[j].lo, prims[j].hi);
                    out[1].c.template grow<E>(prims[j]);
                    --j;
                }
            }
            mid = (uint32_t)i;
            measured = true;
            if (mid == first || mid == end) {
                mid = first + (n >> 1);
                measured = false;
            }
        }
        if (!measured) {
            measure(first, mid - first, lbox, out[0].c);
            measure(mid, end - mid, rbox, out[1].c);
        }

        const uint32_t left = used.fetch_add(2, std::memory_order_relaxed);
        nodes[left] = Node{lbox, first, mid - first, 0};
        nodes[left + 1] = Node{rbox, mid, end - mid, 0};
        node.left = left;
        out[0].node = left;
        out[1].node = left + 1;
        return true;
    }

    // The whole subtree under t, on this thread.
    void subtree(const Task& t, std::vector<Task>& stack) {
        stack.clear();
        stack.push_back(t);
        while (!stack.empty()) {
            const Task k = stack.back();
            stack.pop_back();
            Task out[2];
            if (!split(k, out)) continue;
            for (const Task& c : out)
                if (nodes[c.node].count > 1) stack.push_back(c);
        }
    }
};

// Splits big nodes as tasks, shared out to `threads` threads; subtrees
// smaller than GRAIN are finished by the thread that reaches them.
template <bool E> void build_parallel(Builder<E>& B, const Task& root, unsigned threads) {
    std::mutex lock;
    std::condition_variable wake;
    std::vector<Task> queue{root};
    size_t pending = 1;

    auto worker = [&]() {
        std::vector<Task> stack;
        for (;;) {
            Task t;
            {
                std::unique_lock<std::mutex> g(lock);
                wake.wait(g, [&] { return !queue.empty() || pending == 0; });
                if (queue.empty()) return;
                t = queue.back();
                queue.pop_back();
            }
            Task out[2];
            if (B.split(t, out)) {
                for (const Task& c : out) {
                    const uint32_t n = B.nodes[c.node].count;
                    if (n <= 1) continue;
                    if (n < GRAIN) {
                        B.subtree(c, stack);
                    } else {
                        std::lock_guard<std::mutex> g(lock);
                        queue.push_back(c);
                        ++pending;
                        wake.notify_one();
                    }
                }
            }
            std::lock_guard<std::mutex> g(lock);
            if (--pending == 0) wake.notify_all();
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < threads; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
}

// Builds the tree into nodes (root at 0) with prims left in leaf order;
// *used receives the number of nodes.
template <bool E>
int build(const double* mins, const double* maxs, uint32_t count, uint32_t max_leaf_size, uint32_t threads,
          Node* nodes, Prim* prims, uint32_t* used) {
    Builder<E> B;
    B.prims = prims;
    B.nodes = nodes;
    B.max_leaf = max_leaf_size;
    B.threads = 1;
    for (uint32_t i = 0; i < count; ++i) {
        Prim& p = prims[i];
        for (int a = 0; a < 3; ++a) {
            p.lo[a] = mins[(size_t)i * 3 + a];
            p.hi[a] = maxs[(size_t)i * 3 + a];
        }
        p.id = i;
    }

    Task root;
    root.node = 0;
    nodes[0] = Node{Box{}, 0, count, 0};
    B.measure(0, count, nodes[0].box, root.c);
    if (count > max_leaf_size) {
        unsigned t = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
        t = std::min(t, 64u);
        if (t == 1 || count < 2 * GRAIN) {
            std::vector<Task> stack;
            B.subtree(root, stack);
        } else {
            try {
                B.threads = t;
                build_parallel(B, root, t);
            } catch (const std::system_error&) {
                return rmi::fail(RM_ERR_MEMORY, "Could not start the BVH build's threads.");
            }
        }
    }
    *used = B.used.load();
    return RM_OK;
}

}  // namespace

extern "C" int rm_bvh_build(const double* mins, const double* maxs, uint32_t count, uint32_t max_leaf_size,
                            uint32_t threads, void* nodes_out, uint32_t node_capacity, uint32_t* node_count,
                            uint32_t* order) {
    if ((count && (!mins || !maxs || !order)) || !nodes_out || !node_count)
        return rmi::fail(RM_ERR_INVALID, "rm_bvh_build needs bounds, room for nodes and the order, and a count.");
    if (count >= (1u << 31)) return rmi::fail(RM_ERR_INVALID, "rm_bvh_build takes fewer than 2^31 primitives.");
    const uint32_t most = count > 1 ? 2 * count - 1 : 1;
    if (node_capacity < most)
        return rmi::fail(RM_ERR_INVALID, "rm_bvh_build needs room for max(1, 2·count − 1) nodes.");
    uint32_t* out = static_cast<uint32_t*>(nodes_out);
    float* outf = static_cast<float*>(nodes_out);
    if (count == 0) {
        // No primitives: one empty node at the origin (src/BVH.mjs:224-233).
        std::memset(out, 0, 32);
        *node_count = 1;
        return RM_OK;
    }

    std::unique_ptr<Node, decltype(&std::free)> nodes(static_cast<Node*>(std::malloc(sizeof(Node) * most)), &std::free);
    std::unique_ptr<Prim, decltype(&std::free)> prims(static_cast<Prim*>(std::malloc(sizeof(Prim) * count)), &std::free);
    if (!nodes || !prims) return rmi::fail(RM_ERR_MEMORY, "Out of host memory for the BVH build.");
    uint32_t used = 0;
    const int r = needs_exact(mins, (size_t)count * 3) || needs_exact(maxs, (size_t)count * 3)
                      ? build<true>(mins, maxs, count, max_leaf_size, threads, nodes.get(), prims.get(), &used)
                      : build<false>(mins, maxs, count, max_leaf_size, threads, nodes.get(), prims.get(), &used);
    if (r != RM_OK) return r;
    for (uint32_t i = 0; i < count; ++i) order[i] = prims.get()[i].id;
    const Node* built = nodes.get();

    // The sequential builder's numbering (src/BVH.mjs:89-90, 200-209): its
    // stack pops nodes with more than one primitive, and each split appends
    // the two children at the next free index.
    std::vector<uint32_t> renum(used);
    renum[0] = 0;
    uint32_t next = 1;
    std::vector<uint32_t> stack;
    if (count > max_leaf_size) stack.push_back(0);
    while (!stack.empty()) {
        const uint32_t k = stack.back();
        stack.pop_back();
        const uint32_t l = built[k].left;
        if (l == 0) continue;
        renum[l] = next;
        renum[l + 1] = next + 1;
        next += 2;
        if (built[l].count > 1) stack.push_back(l);
        if (built[l + 1].count > 1) stack.push_back(l + 1);
    }

    // Packed as src/BVH.mjs:212-223 writes them: f32 min[3], u32 first,
    // f32 max[3], u32 count (0 for an inner node, whose children are
    // first and first + 1).
    for (uint32_t k = 0; k < used; ++k) {
        const Node& n = built[k];
        const size_t o = (size_t)renum[k] * 8;
        for (int a = 0; a < 3; ++a) {
            outf[o + a] = (float)n.box.lo[a];
            outf[o + 4 + a] = (float)n.box.hi[a];
        }
        out[o + 3] = n.left ? renum[n.left] : n.first;
        out[o + 7] = n.left ? 0u : n.count;
    }
    *node_count = used;
    return RM_OK;
}

// ------------------------------------------------------------ quantized tree
//
// The 8-bit child-quantized form of a packed tree, one 32-byte record per
// node, indexed like the nodes:
//
//   inner node p, children c and c + 1 (both boxes relative to p's box):
//     f32 origin[3]   p's min corner
//     u32 meta        step exponents + 127 (bits 0-7, 8-15, 16-23: the
//                     step is 2^e per axis), bit 24: child c is a leaf,
//                     bit 25: child c + 1 is a leaf
//     u8  box[12]     c's lo xyz, hi xyz, then c + 1's lo xyz, hi xyz
//     u32 c
//   leaf: u32 first, u32 count, then zeros.
//
// A coordinate decodes as origin + q·step in FP32 (q·step is exact), the
// very expression the device evaluates; each lo is the largest q whose
// decoded value is at most the child's FP32 lo, each hi the smallest whose
// decoded value is at least its hi. Decoded boxes therefore enclose the
// FP32 ones, and since a slab test's entry distance only falls and its exit
// only grows as a box grows, every ray that enters a child's FP32 box
// enters its decoded box: no leaf the FP32 tree reaches is skipped.
namespace {

inline float decode(float origin, uint32_t q, float step) { return origin + (float)q * step; }

// The smallest exponent in [-126, 127] whose 255 steps reach hi from lo.
int step_exponent(float lo, float hi) {
    if (!(hi > lo)) return -126;
    int e = (int)std::ceil(std::log2(((double)hi - (double)lo) / 255.0)) - 1;
    e = std::max(-126, std::min(127, e));
    while (e < 127 && !(decode(lo, 255, std::ldexp(1.0f, e)) >= hi)) ++e;
    return e;
}

bool quantize_axis(float origin, float step, float lo, float hi, uint32_t& qlo, uint32_t& qhi) {
    double g = std::floor(((double)lo - origin) / step);
    uint32_t a = g < 0.0 ? 0u : g > 255.0 ? 255u : (uint32_t)g;
    while (a > 0 && decode(origin, a, step) > lo) --a;
    while (a < 255 && decode(origin, a + 1, step) <= lo) ++a;
    g = std::ceil(((double)hi - origin) / step);
    uint32_t b = g < 0.0 ? 0u : g > 255.0 ? 255u : (uint32_t)g;
    while (b < 255 && decode(origin, b, step) < hi) ++b;
    while (b > 0 && decode(origin, b - 1, step) >= hi) --b;
    qlo = a;
    qhi = b;
    return decode(origin, a, step) <= lo && decode(origin, b, step) >= hi;
}

}  // namespace

extern "C" int rm_bvh_quantize(const void* nodes, uint32_t node_count, void* qnodes) {
    if (!nodes || !qnodes || node_count == 0)
        return rmi::fail(RM_ERR_INVALID, "rm_bvh_quantize needs nodes and room for as many quantized nodes.");
    const uint32_t* nu = static_cast<const uint32_t*>(nodes);
    const float* nf = static_cast<const float*>(nodes);
    uint32_t* qu = static_cast<uint32_t*>(qnodes);
    float* qf = static_cast<float*>(qnodes);
    for (uint32_t p = 0; p < node_count; ++p) {
        const size_t o = (size_t)p * 8;
        const uint32_t first = nu[o + 3], count = nu[o + 7];
        std::memset(qu + o, 0, 32);
        if (count > 0) {
            qu[o] = first;
            qu[o + 1] = count;
            continue;
        }
        if ((uint64_t)first + 1 >= node_count)
            return rmi::fail(RM_ERR_INVALID, "BVH node " + std::to_string(p) + " points outside the tree.");
        uint32_t meta = 0, q[2][6];
        for (int a = 0; a < 3; ++a) {
            const float lo = nf[o + a], hi = nf[o + 4 + a];
            if (!std::isfinite(lo) || !std::isfinite(hi))
                return rmi::fail(RM_ERR_INVALID, "BVH node " + std::to_string(p) + " has a box that isn't finite.");
            const int e = step_exponent(lo, hi);
            const float step = std::ldexp(1.0f, e);
            meta |= (uint32_t)(e + 127) << (8 * a);
            for (uint32_t k = 0; k < 2; ++k) {
                const size_t co = (size_t)(first + k) * 8;
                const float clo = nf[co + a], chi = nf[co + 4 + a];
                if (!(clo >= lo && chi <= hi) || !quantize_axis(lo, step, clo, chi, q[k][a], q[k][3 + a]))
                    return rmi::fail(RM_ERR_INVALID,
                                     "BVH node " + std::to_string(first + k) + " lies outside its parent's box.");
            }
        }
        for (uint32_t k = 0; k < 2; ++k)
            if (nu[(size_t)(first + k) * 8 + 7] > 0) meta |= 1u << (24 + k);
        qf[o] = nf[o];
        qf[o + 1] = nf[o + 1];
        qf[o + 2] = nf[o + 2];
        qu[o + 3] = meta;
        const uint32_t* b = &q[0][0];   // c lo xyz, hi xyz, c + 1 lo xyz, hi xyz
        for (int w = 0; w < 3; ++w)
            qu[o + 4 + w] = b[4 * w] | b[4 * w + 1] << 8 | b[4 * w + 2] << 16 | b[4 * w + 3] << 24;
        qu[o + 7] = first;
    }
    return RM_OK;
}

// ------------------------------------------------------ 4-wide quantized tree
//
// Every other level of the binary tree folded into its parent: a wide node
// made from binary node p has p's children as its children, each inner one
// replaced by its own two children — two to four in all. One 64-byte record
// per wide node (node 0 from the root), the children's boxes quantized
// against p's box exactly as rm_bvh_quantize quantizes them:
//
//   f32 origin[3]    p's min corner
//   u32 meta         step exponents + 127 (bits 0-7, 8-15, 16-23),
//                    bits 24-27: child k is a leaf, bits 28-30: children
//   u8  box[24]      child k's lo xyz, hi xyz at bytes 6k .. 6k + 5
//   u32 ref[4]       a leaf's first primitive, or an inner child's wide node
//   u8  count[4]     a leaf child's primitive count (at most 255)
//   u32 zero
//
// A walk pushes up to three children per wide node; the producer refuses a
// tree where that could pass the 64-entry traversal stack on any path.
extern "C" int rm_bvh_quantize_wide(const void* nodes, uint32_t node_count, void* wide, uint32_t capacity,
                                    uint32_t* wide_count) {
    if (!nodes || !wide || !wide_count || node_count == 0)
        return rmi::fail(RM_ERR_INVALID, "rm_bvh_quantize_wide needs nodes and room for the wide nodes.");
    const uint32_t* nu = static_cast<const uint32_t*>(nodes);
    const float* nf = static_cast<const float*>(nodes);
    uint32_t* wu = static_cast<uint32_t*>(wide);
    float* wf = static_cast<float*>(wide);
    *wide_count = 0;
    if (nu[7] > 0) return RM_OK;   // the root is a leaf: nothing to fold
    struct Item { uint32_t binary, wide, pushes; };
    std::vector<Item> todo{{0u, 0u, 0u}};
    uint32_t used = 1;
    while (!todo.empty()) {
        const Item it = todo.back();
        todo.pop_back();
        const uint32_t p = it.binary;
        const size_t po = (size_t)p * 8;
        const uint32_t c = nu[po + 3];
        if ((uint64_t)c + 1 >= node_count)
            return rmi::fail(RM_ERR_INVALID, "BVH node " + std::to_string(p) + " points outside the tree.");
        uint32_t kids[4], n = 0;
        for (uint32_t k = c; k < c + 2; ++k) {
            if (nu[(size_t)k * 8 + 7] > 0) {
                kids[n++] = k;
            } else {
                const uint32_t g = nu[(size_t)k * 8 + 3];
                if ((uint64_t)g + 1 >= node_count)
                    return rmi::fail(RM_ERR_INVALID, "BVH node " + std::to_string(k) + " points outside the tree.");
                kids[n++] = g;
                kids[n++] = g + 1;
            }
        }
        const uint32_t pushes = it.pushes + (n - 1);
        if (pushes > 64)
            return rmi::fail(RM_ERR_INVALID, "The tree is too deep for the wide walk's 64-entry stack.");
        uint32_t* r = wu + (size_t)it.wide * 16;
        std::memset(r, 0, 64);
        uint32_t meta = (n << 28);
        uint8_t* box = reinterpret_cast<uint8_t*>(r + 4);
        uint8_t* counts = reinterpret_cast<uint8_t*>(r + 14);
        for (int a = 0; a < 3; ++a) {
            const float lo = nf[po + a], hi = nf[po + 4 + a];
            if (!std::isfinite(lo) || !std::isfinite(hi))
                return rmi::fail(RM_ERR_INVALID, "BVH node " + std::to_string(p) + " has a box that isn't finite.");
            const int e = step_exponent(lo, hi);
            const float step = std::ldexp(1.0f, e);
            meta |= (uint32_t)(e + 127) << (8 * a);
            wf[(size_t)it.wide * 16 + a] = lo;
            for (uint32_t k = 0; k < n; ++k) {
                const size_t ko = (size_t)kids[k] * 8;
                uint32_t qlo, qhi;
                if (!(nf[ko + a] >= lo && nf[ko + 4 + a] <= hi) ||
                    !quantize_axis(lo, step, nf[ko + a], nf[ko + 4 + a], qlo, qhi))
                    return rmi::fail(RM_ERR_INVALID,
                                     "BVH node " + std::to_string(kids[k]) + " lies outside its parent's box.");
                box[6 * k + a] = (uint8_t)qlo;
                box[6 * k + 3 + a] = (uint8_t)qhi;
            }
        }
        for (uint32_t k = 0; k < n; ++k) {
            const size_t ko = (size_t)kids[k] * 8;
            if (nu[ko + 7] > 0) {
                if (nu[ko + 7] > 255)
                    return rmi::fail(RM_ERR_INVALID, "A leaf of more than 255 primitives can't be folded.");
                meta |= 1u << (24 + k);
                r[10 + k] = nu[ko + 3];
                counts[k] = (uint8_t)nu[ko + 7];
            } else {
                if (used >= capacity) return rmi::fail(RM_ERR_INVALID, "rm_bvh_quantize_wide ran out of room.");
                r[10 + k] = used;
                todo.push_back({kids[k], used, pushes});
                ++used;
            }
        }
        r[3] = meta;
    }
    *wide_count = used;
    return RM_OK;
}
