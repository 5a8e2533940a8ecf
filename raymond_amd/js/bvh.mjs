// Binned SAH BVH builder with the reference's decisions and packed layout
// (src/BVH.mjs:57-234, docs/bvh.md:34-131): 32 centroid bins per axis, a
// right-then-left sweep costing N_L·A_L + N_R·A_R, leaves of at most
// maxLeafSize unless the SAH says splitting pays (cost + parent area < n ·
// parent area), the in-place two-pointer partition, a median fallback, and a
// LIFO stack of nodes with more than one primitive. Children are appended
// together, so the right child is always first + 1.
//
// Node layout, 32 bytes: f32 min[3], u32 first, f32 max[3], u32 count
// (count 0: inner node whose children are first and first + 1).
//
// Written over flat typed arrays so that two million triangles build in
// seconds under Node 12. The addon's native builder (rm_bvh_build) makes the
// same tree, node for node, an order of magnitude faster; BVH.build uses it
// whenever the addon loads, unless told otherwise.

import { loadAddon } from "./addon.mjs";

const BINS = 32;
const BUILDERS = ["auto", "native", "js"];
let nativeBuild;   // undefined until tried; null when the addon can't load

function native(required) {
    if (nativeBuild === undefined) {
        try {
            nativeBuild = loadAddon().bvhBuild || null;
        }
        catch (error) {
            nativeBuild = null;
        }
    }
    if (!nativeBuild && required)
        throw new Error("The native BVH builder needs the Raymond addon (raymond_amd/lib/raymond.node).");
    return nativeBuild;
}

const f64 = (a) => (a instanceof Float64Array ? a : Float64Array.from(a));

function halfArea(x0, y0, z0, x1, y1, z1) {
    const dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
    if (!(dx >= 0.0 && dy >= 0.0 && dz >= 0.0))
        return 0.0;
    return dx * dy + dy * dz + dz * dx;
}

export class BVH {

    #nodes; #nodeCount; #order;

    constructor(nodes, nodeCount, order) {
        this.#nodes = nodes;
        this.#nodeCount = nodeCount;
        this.#order = order;
    }

    static get NODE_BYTES() { return 32; }

    /** The packed nodes, 32 bytes each. */
    get nodes() { return this.#nodes; }
    get nodeCount() { return this.#nodeCount; }
    /** order[i] is the original index of the i-th primitive in leaf order. */
    get order() { return this.#order; }

    /**
     * @param {Float64Array|Float32Array|number[]} mins 3 per primitive
     * @param {Float64Array|Float32Array|number[]} maxs 3 per primitive
     * builder: "auto" (native when the addon loads, else JS), "native" or
     * "js" — the same tree either way. threads: the native builder's
     * threads (0: one per core).
     */
    static build(mins, maxs, { maxLeafSize = 4, builder = "auto", threads = 0 } = {}) {
        if (!BUILDERS.includes(builder))
            throw new RangeError(`builder is one of ${BUILDERS.join(", ")}.`);
        const fn = builder == "js" ? null : native(builder == "native");
        if (fn) {
            const b = fn(f64(mins), f64(maxs), maxLeafSize, threads);
            return new BVH(b.nodes, b.nodeCount, b.order);
        }
        return buildJs(mins, maxs, maxLeafSize);
    }
}

// The builder in JS (src/BVH.mjs:57-234).
function buildJs(mins, maxs, maxLeafSize) {

    const count = mins.length / 3;
    const order = new Uint32Array(count);
    for (let i = 0; i < count; ++i)
        order[i] = i;

    const cen = new Float64Array(count * 3);
    for (let i = 0; i < count * 3; ++i)
        cen[i] = (mins[i] + maxs[i]) * 0.5;

    // Growable node storage (FP64 boxes until packing).
    let cap = Math.max(16, 2 * count + 1);
    let nBox = new Float64Array(cap * 6);
    let nFirst = new Uint32Array(cap);
    let nCount = new Uint32Array(cap);
    let used = 0;

    const addNode = (first, n) => {
        if (used == cap) {
            cap *= 2;
            const b = new Float64Array(cap * 6); b.set(nBox); nBox = b;
            const f = new Uint32Array(cap); f.set(nFirst); nFirst = f;
            const c = new Uint32Array(cap); c.set(nCount); nCount = c;
        }
        let x0 = Infinity, y0 = Infinity, z0 = Infinity, x1 = -Infinity, y1 = -Infinity, z1 = -Infinity;
        for (let i = first; i < first + n; ++i) {
            const p = order[i] * 3;
            x0 = Math.min(x0, mins[p]); y0 = Math.min(y0, mins[p + 1]); z0 = Math.min(z0, mins[p + 2]);
            x1 = Math.max(x1, maxs[p]); y1 = Math.max(y1, maxs[p + 1]); z1 = Math.max(z1, maxs[p + 2]);
        }
        const o = used * 6;
        nBox[o] = x0; nBox[o + 1] = y0; nBox[o + 2] = z0;
        nBox[o + 3] = x1; nBox[o + 4] = y1; nBox[o + 5] = z1;
        nFirst[used] = first;
        nCount[used] = n;
        return used++;
    };

    addNode(0, count);
    const stack = count > maxLeafSize ? [0] : [];

    const binBox = new Float64Array(BINS * 6);
    const binCount = new Uint32Array(BINS);
    const rightArea = new Float64Array(BINS);
    const rightCount = new Uint32Array(BINS);

    while (stack.length > 0) {

        const node = stack.pop();
        const first = nFirst[node];
        const n = nCount[node];

        const cmin = [Infinity, Infinity, Infinity];
        const cmax = [-Infinity, -Infinity, -Infinity];
        for (let i = first; i < first + n; ++i) {
            const p = order[i] * 3;
            for (let a = 0; a < 3; ++a) {
                const c = cen[p + a];
                if (c < cmin[a]) cmin[a] = c;
                if (c > cmax[a]) cmax[a] = c;
            }
        }

        let bestCost = Infinity, bestAxis = -1, bestSplit = 0;

        for (let axis = 0; axis < 3; ++axis) {
            const extent = cmax[axis] - cmin[axis];
            if (!(extent > 0.0))
                continue;

            for (let b = 0; b < BINS; ++b) {
                binBox[b * 6] = binBox[b * 6 + 1] = binBox[b * 6 + 2] = Infinity;
                binBox[b * 6 + 3] = binBox[b * 6 + 4] = binBox[b * 6 + 5] = -Infinity;
                binCount[b] = 0;
            }

            const scale = BINS / extent;
            for (let i = first; i < first + n; ++i) {
                const p = order[i] * 3;
                const b = Math.min(BINS - 1, Math.floor((cen[p + axis] - cmin[axis]) * scale));
                binCount[b]++;
                const o = b * 6;
                binBox[o] = Math.min(binBox[o], mins[p]);
                binBox[o + 1] = Math.min(binBox[o + 1], mins[p + 1]);
                binBox[o + 2] = Math.min(binBox[o + 2], mins[p + 2]);
                binBox[o + 3] = Math.max(binBox[o + 3], maxs[p]);
                binBox[o + 4] = Math.max(binBox[o + 4], maxs[p + 1]);
                binBox[o + 5] = Math.max(binBox[o + 5], maxs[p + 2]);
            }

            // Suffix areas and counts, sweeping in from the right.
            let x0 = Infinity, y0 = Infinity, z0 = Infinity, x1 = -Infinity, y1 = -Infinity, z1 = -Infinity;
            let total = 0;
            for (let b = BINS - 1; b > 0; --b) {
                if (binCount[b] > 0) {
                    const o = b * 6;
                    x0 = Math.min(x0, binBox[o]); y0 = Math.min(y0, binBox[o + 1]); z0 = Math.min(z0, binBox[o + 2]);
                    x1 = Math.max(x1, binBox[o + 3]); y1 = Math.max(y1, binBox[o + 4]); z1 = Math.max(z1, binBox[o + 5]);
                }
                total += binCount[b];
                rightArea[b] = halfArea(x0, y0, z0, x1, y1, z1);
                rightCount[b] = total;
            }

            // Then from the left, costing a split before each bin.
            x0 = Infinity; y0 = Infinity; z0 = Infinity; x1 = -Infinity; y1 = -Infinity; z1 = -Infinity;
            let leftCount = 0;
            for (let b = 1; b < BINS; ++b) {
                if (binCount[b - 1] > 0) {
                    const o = (b - 1) * 6;
                    x0 = Math.min(x0, binBox[o]); y0 = Math.min(y0, binBox[o + 1]); z0 = Math.min(z0, binBox[o + 2]);
                    x1 = Math.max(x1, binBox[o + 3]); y1 = Math.max(y1, binBox[o + 4]); z1 = Math.max(z1, binBox[o + 5]);
                }
                leftCount += binCount[b - 1];
                if (leftCount == 0 || rightCount[b] == 0)
                    continue;
                const cost = leftCount * halfArea(x0, y0, z0, x1, y1, z1) + rightCount[b] * rightArea[b];
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = axis;
                    bestSplit = cmin[axis] + b / scale;
                }
            }
        }

        const o = node * 6;
        const parentArea = halfArea(nBox[o], nBox[o + 1], nBox[o + 2], nBox[o + 3], nBox[o + 4], nBox[o + 5]);

        let mid;
        if (bestAxis < 0) {
            if (n <= maxLeafSize)
                continue;
            mid = first + (n >> 1);
        }
        else {
            if (n <= maxLeafSize && bestCost + parentArea >= n * parentArea)
                continue;
            let i = first, j = first + n - 1;
            while (i <= j) {
                if (cen[order[i] * 3 + bestAxis] < bestSplit)
                    ++i;
                else {
                    const t = order[i]; order[i] = order[j]; order[j] = t;
                    --j;
                }
            }
            mid = i;
            if (mid == first || mid == first + n)
                mid = first + (n >> 1);
        }

        const left = addNode(first, mid - first);
        addNode(mid, first + n - mid);
        nFirst[node] = left;
        nCount[node] = 0;
        if (nCount[left] > 1) stack.push(left);
        if (nCount[left + 1] > 1) stack.push(left + 1);
    }

    const nodeCount = Math.max(used, 1);
    const buffer = new ArrayBuffer(nodeCount * 32);
    const f = new Float32Array(buffer);
    const u = new Uint32Array(buffer);
    for (let i = 0; i < used; ++i) {
        const o = i * 8, b = i * 6;
        f[o] = nBox[b]; f[o + 1] = nBox[b + 1]; f[o + 2] = nBox[b + 2];
        u[o + 3] = nFirst[i];
        f[o + 4] = nBox[b + 3]; f[o + 5] = nBox[b + 4]; f[o + 6] = nBox[b + 5];
        u[o + 7] = nCount[i];
    }
    // No primitives: one empty box at the origin (the scene packer makes
    // it a leaf over a primitive that nothing hits).
    if (count == 0) {
        f.fill(0.0);
        u[3] = 0;
        u[7] = 0;
    }
    return new BVH(buffer, nodeCount, order);
}
