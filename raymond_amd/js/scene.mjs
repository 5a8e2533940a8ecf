// Scene: the object list, the packed buffers the kernels read, and the CPU
// raycast. Packed formats follow src/Scene.mjs:15-60 and :294-539:
//
//   primitive, 64 B:  u32 kind, u32 material, u32 light, u32 attributes,
//                     vec4 p0, vec4 p1, vec4 p2   (BVH leaf order, then volumes)
//     sphere: p0 = centre, radius; p1 = globe x axis; p2 = north pole
//     quad:   p0 = corner; p1 = u; p2 = v
//     box:    p0..p2 = rows of the world-to-box matrix
//     triangle: p0..p2 = vertices
//   attribute block, 4 × vec4: (n0, u0), (n1, u1), (n2, u2), (v0, v1, v2, flags)
//     flags: 1 smooth normals, 2 texture coordinates
//   material, 96 B: vec3 color, u32 type, vec3 emission, f32 roughness,
//     vec3 checker, f32 checkerSize (0: none), f32 ior, u32 texture,
//     vec2 textureScale, u32 normalMap, f32 normalScale, u32 emissionTexture,
//     f32 dispersion, f32 density, f32 anisotropy, 2 × f32 unused
//
// The packer writes flat typed arrays directly, so meshes of millions of
// triangles pack without a per-primitive allocation.
import { BVH } from "./bvh.mjs";
import { Box, Mesh, Quad, Sphere } from "./geometry.mjs";
import { MATERIAL_TYPES } from "./material.mjs";
import { Vec2, Vec3 } from "./math.mjs";
import { Sky } from "./sky.mjs";

export const PRIMITIVE_KINDS = Object.freeze({ sphere: 0, quad: 1, box: 2, triangle: 3 });

/** Marks a missing index in packed data. */
export const NONE = 0xffffffff;

const PRIM_FLOATS = 16;
const MAT_FLOATS = 24;
const SMOOTH = 1;
const TEXTURED = 2;

export class Scene {

    #objects = [];
    #sky;
    #version = 0;
    #packed = null;

    constructor({ sky = new Sky() } = {}) {
        this.#sky = sky;
    }

    get sky() { return this.#sky; }
    set sky(value) { this.#sky = value; }
    get objects() { return this.#objects.slice(); }
    get version() { return this.#version; }

    add(...objects) {
        for (const o of objects) {
            if (!(o instanceof Sphere || o instanceof Quad || o instanceof Box || o instanceof Mesh))
                throw new TypeError("Scenes hold Spheres, Quads, Boxes and Meshes.");
        }
        this.#objects.push(...objects);
        this.#version++;
        this.#packed = null;
        return this;
    }

    remove(object) {
        const i = this.#objects.indexOf(object);
        if (i < 0)
            return false;
        this.#objects.splice(i, 1);
        this.#version++;
        this.#packed = null;
        return true;
    }

    clear() {
        this.#objects = [];
        this.#version++;
        this.#packed = null;
    }

    get primitiveCount() {
        let n = 0;
        for (const o of this.#objects)
            n += o instanceof Mesh ? o.triangleCount : 1;
        return n;
    }

    /** The root box (padded primitive bounds), or null for an empty scene. */
    bounds() {
        const packed = this.pack();
        if (packed.primitiveCount == 0)
            return null;
        const f = new Float32Array(packed.nodes, 0, 8);
        return { min: new Vec3(f[0], f[1], f[2]), max: new Vec3(f[4], f[5], f[6]) };
    }

    pack() {
        if (!this.#packed)
            this.#packed = packObjects(this.#objects);
        return this.#packed;
    }

    /**
     * The first hit of a ray on the CPU, in double precision over the packed
     * FP32 data (src/Scene.mjs:202-283), or null. distance is in units of
     * the direction's length.
     */
    raycast(origin, direction, maxDistance = Infinity) {
        const packed = this.pack();
        if (packed.primitiveCount == 0)
            return null;
        const pf = new Float32Array(packed.primitives);
        const pu = new Uint32Array(packed.primitives);
        const nf = new Float32Array(packed.nodes);
        const nu = new Uint32Array(packed.nodes);
        const o = [origin.x, origin.y, origin.z];
        const d = [direction.x, direction.y, direction.z];
        const inv = [1.0 / d[0], 1.0 / d[1], 1.0 / d[2]];

        let bestT = maxDistance, best = -1, bu = 0.0, bv = 0.0;
        const stack = [0];
        while (stack.length > 0) {
            const n = stack.pop();
            const nb = n * 8;
            if (slabs64(nf, nb, o, inv) >= bestT)
                continue;
            const count = nu[nb + 7], first = nu[nb + 3];
            if (count == 0) {
                const tl = slabs64(nf, first * 8, o, inv);
                const tr = slabs64(nf, (first + 1) * 8, o, inv);
                if (tl <= tr) stack.push(first + 1, first);
                else stack.push(first, first + 1);
                continue;
            }
            for (let p = first; p < first + count; ++p) {
                const h = intersect64(pf, pu, p, o, d, bestT);
                if (h) {
                    bestT = h[0]; bu = h[1]; bv = h[2]; best = p;
                }
            }
        }
        if (best < 0)
            return null;

        const point = origin.add(direction.scale(bestT));
        const material = packed.materialList[pu[best * PRIM_FLOATS + 1]];
        const { geometric, shading } = normals64(packed, pf, pu, best, point, bu, bv);
        const uv = uv64(packed, pf, pu, best, point, bu, bv);
        let n = shading;
        if (material.normalMap) {
            const { tangent, up } = tangents64(packed, pf, pu, best, point, bu, bv);
            const s = material.textureScale;
            n = bump(shading, geometric, tangent, up, material.normalMap.sample(uv.x * s.x, uv.y * s.y, false), material.normalScale);
        }
        const front = geometric.dot(direction) < 0.0;
        return {
            distance: bestT,
            point,
            normal: front ? n : n.negate(),
            frontFace: front,
            uv,
            object: packed.objects[best],
            material,
            primitive: best,
        };
    }
}

function packObjects(objects) {

    const materialIndex = new Map();
    const materialList = [];
    const materialOf = (m) => {
        let i = materialIndex.get(m);
        if (i === undefined) {
            i = materialList.length;
            materialList.push(m);
            materialIndex.set(m, i);
        }
        return i;
    };
    const textureIndex = new Map(), textureList = [];
    const normalIndex = new Map(), normalList = [];
    const textureOf = (t, index, list) => {
        if (!t) return NONE;
        let i = index.get(t);
        if (i === undefined) {
            i = list.length;
            list.push(t);
            index.set(t, i);
        }
        return i;
    };

    let total = 0;
    for (const o of objects) {
        if (o.material.type == "volume" && !(o instanceof Sphere || o instanceof Box))
            throw new TypeError("A volume must be a Sphere or a Box.");
        total += o instanceof Mesh ? o.triangleCount : 1;
    }

    // Records in object order, their bounds, owners and attribute blocks.
    const rec = new Float32Array(total * PRIM_FLOATS);
    const recU = new Uint32Array(rec.buffer);
    const bounds = new Float64Array(total * 6);
    const owners = new Array(total);
    const blockOf = new Int32Array(total).fill(-1);
    let blockStore = [];
    let k = 0;

    const begin = (owner, kind, material) => {
        const b = k * PRIM_FLOATS;
        recU[b] = kind;
        recU[b + 1] = materialOf(material);
        recU[b + 2] = NONE;
        recU[b + 3] = NONE;
        owners[k] = owner;
        const o = k * 6;
        bounds[o] = bounds[o + 1] = bounds[o + 2] = Infinity;
        bounds[o + 3] = bounds[o + 4] = bounds[o + 5] = -Infinity;
    };
    const grow = (x, y, z) => {
        const o = k * 6;
        if (x < bounds[o]) bounds[o] = x;
        if (y < bounds[o + 1]) bounds[o + 1] = y;
        if (z < bounds[o + 2]) bounds[o + 2] = z;
        if (x > bounds[o + 3]) bounds[o + 3] = x;
        if (y > bounds[o + 4]) bounds[o + 4] = y;
        if (z > bounds[o + 5]) bounds[o + 5] = z;
    };
    // Pads each axis by 1e-4 · max(1, |min|, |max|) so flat primitives keep
    // boxes rays can enter (src/Scene.mjs:381-387).
    const finish = () => {
        const o = k * 6;
        for (let a = 0; a < 3; ++a) {
            const pad = 1e-4 * Math.max(1.0, Math.abs(bounds[o + a]), Math.abs(bounds[o + 3 + a]));
            bounds[o + a] -= pad;
            bounds[o + 3 + a] += pad;
        }
        ++k;
    };
    const put = (slot, values) => {
        const b = k * PRIM_FLOATS + 4 + slot * 4;
        rec[b] = values[0]; rec[b + 1] = values[1]; rec[b + 2] = values[2]; rec[b + 3] = values[3];
    };

    for (const obj of objects) {
        if (obj instanceof Sphere) {
            const c = obj.center, r = obj.radius, ax = obj.axes();
            begin(obj, PRIMITIVE_KINDS.sphere, obj.material);
            put(0, [c.x, c.y, c.z, r]);
            put(1, [ax.x.x, ax.x.y, ax.x.z, 0.0]);
            put(2, [ax.north.x, ax.north.y, ax.north.z, 0.0]);
            grow(c.x - r, c.y - r, c.z - r);
            grow(c.x + r, c.y + r, c.z + r);
            finish();
        }
        else if (obj instanceof Quad) {
            const q = obj.corner, u = obj.u, v = obj.v;
            begin(obj, PRIMITIVE_KINDS.quad, obj.material);
            put(0, [q.x, q.y, q.z, 0.0]);
            put(1, [u.x, u.y, u.z, 0.0]);
            put(2, [v.x, v.y, v.z, 0.0]);
            for (const p of [q, q.add(u), q.add(v), q.add(u).add(v)])
                grow(p.x, p.y, p.z);
            finish();
        }
        else if (obj instanceof Box) {
            const m = obj.inverse;
            begin(obj, PRIMITIVE_KINDS.box, obj.material);
            for (let r = 0; r < 3; ++r)
                put(r, [m.value(r, 0), m.value(r, 1), m.value(r, 2), m.value(r, 3)]);
            for (const p of obj.corners())
                grow(p.x, p.y, p.z);
            finish();
        }
        else {
            const P = obj.positions, I = obj.indices, N = obj.normals, UV = obj.uvs;
            const flags = (N ? SMOOTH : 0) + (UV ? TEXTURED : 0);
            for (let i = 0; i < I.length; i += 3) {
                const a = P[I[i]], b = P[I[i + 1]], c = P[I[i + 2]];
                begin(obj, PRIMITIVE_KINDS.triangle, obj.material);
                put(0, [a.x, a.y, a.z, 0.0]);
                put(1, [b.x, b.y, b.z, 0.0]);
                put(2, [c.x, c.y, c.z, 0.0]);
                grow(a.x, a.y, a.z); grow(b.x, b.y, b.z); grow(c.x, c.y, c.z);
                if (flags) {
                    const blk = [];
                    for (let j = 0; j < 3; ++j) {
                        const n = N ? N[I[i + j]] : null;
                        blk.push(n ? n.x : 0.0, n ? n.y : 0.0, n ? n.z : 0.0, UV ? UV[I[i + j]].x : 0.0);
                    }
                    blk.push(UV ? UV[I[i]].y : 0.0, UV ? UV[I[i + 1]].y : 0.0, UV ? UV[I[i + 2]].y : 0.0, flags);
                    blockOf[k] = blockStore.length;
                    blockStore.push(blk);
                }
                finish();
            }
        }
    }

    // Volumes stay out of the BVH: rays pass through them.
    const isVolume = (i) => materialList[recU[i * PRIM_FLOATS + 1]].type == "volume";
    const surfaceIds = [], volumeIds = [];
    for (let i = 0; i < total; ++i)
        (isVolume(i) ? volumeIds : surfaceIds).push(i);

    const count = surfaceIds.length;
    const mins = new Float64Array(count * 3), maxs = new Float64Array(count * 3);
    for (let i = 0; i < count; ++i) {
        const id = surfaceIds[i];
        for (let a = 0; a < 3; ++a) {
            mins[i * 3 + a] = bounds[id * 6 + a];
            maxs[i * 3 + a] = bounds[id * 6 + 3 + a];
        }
    }
    const bvh = BVH.build(mins, maxs);

    const slots = Math.max(count, 1);
    const primitives = new ArrayBuffer((slots + volumeIds.length) * PRIM_FLOATS * 4);
    const pf = new Float32Array(primitives);
    const pu = new Uint32Array(primitives);
    if (count == 0) {
        // Nothing to bind would be an error: a root leaf over one primitive of no kind.
        pu[0] = NONE;
        new Uint32Array(bvh.nodes)[7] = 1;
    }

    const volumes = volumeIds.map((id, j) => {
        pf.set(rec.subarray(id * PRIM_FLOATS, (id + 1) * PRIM_FLOATS), (slots + j) * PRIM_FLOATS);
        return slots + j;
    });

    let nBlocks = 0;
    for (let i = 0; i < count; ++i)
        if (blockOf[surfaceIds[bvh.order[i]]] >= 0) ++nBlocks;
    const attributes = new Float32Array(nBlocks > 0 ? nBlocks * 16 : 4);
    const lightList = [];
    const packedOwners = new Array(count);
    let nb = 0;
    for (let i = 0; i < count; ++i) {
        const src = surfaceIds[bvh.order[i]];
        pf.set(rec.subarray(src * PRIM_FLOATS, (src + 1) * PRIM_FLOATS), i * PRIM_FLOATS);
        packedOwners[i] = owners[src];
        if (blockOf[src] >= 0) {
            pu[i * PRIM_FLOATS + 3] = nb * 4;
            attributes.set(blockStore[blockOf[src]], nb * 16);
            ++nb;
        }
        const kind = pu[i * PRIM_FLOATS];
        const m = materialList[pu[i * PRIM_FLOATS + 1]];
        if (m.emissive && (kind == PRIMITIVE_KINDS.sphere || kind == PRIMITIVE_KINDS.quad)) {
            pu[i * PRIM_FLOATS + 2] = lightList.length;
            lightList.push(i);
        }
    }
    blockStore = null;

    const materials = new ArrayBuffer(Math.max(materialList.length, 1) * MAT_FLOATS * 4);
    const mf = new Float32Array(materials);
    const mu = new Uint32Array(materials);
    materialList.forEach((m, i) => {
        const o = i * MAT_FLOATS;
        mf[o] = m.color.x; mf[o + 1] = m.color.y; mf[o + 2] = m.color.z;
        mu[o + 3] = MATERIAL_TYPES[m.type];
        mf[o + 4] = m.emission.x; mf[o + 5] = m.emission.y; mf[o + 6] = m.emission.z;
        mf[o + 7] = m.roughness;
        if (m.checker) {
            mf[o + 8] = m.checker.x; mf[o + 9] = m.checker.y; mf[o + 10] = m.checker.z;
            mf[o + 11] = m.checkerSize;
        }
        mf[o + 12] = m.ior;
        mu[o + 13] = textureOf(m.texture, textureIndex, textureList);
        mf[o + 14] = m.textureScale.x; mf[o + 15] = m.textureScale.y;
        mu[o + 16] = textureOf(m.normalMap, normalIndex, normalList);
        mf[o + 17] = m.normalScale;
        mu[o + 18] = textureOf(m.emissionTexture, textureIndex, textureList);
        mf[o + 19] = m.dispersion;
        mf[o + 20] = m.density;
        mf[o + 21] = m.anisotropy;
    });

    return {
        primitives,
        primitiveCount: count,
        nodes: bvh.nodes,
        nodeCount: bvh.nodeCount,
        materials,
        materialCount: materialList.length,
        attributes,
        textures: textureList,
        normalMaps: normalList,
        lights: new Uint32Array(lightList.length > 0 ? lightList : [NONE]),
        volumes,
        objects: packedOwners,
        materialList,
    };
}

// ---- FP64 raycast helpers (src/Scene.mjs:549-848) ----

function slabs64(f, base, o, inv) {
    let tNear = 0.0, tFar = Infinity;
    for (let a = 0; a < 3; ++a) {
        let t0 = (f[base + a] - o[a]) * inv[a];
        let t1 = (f[base + 4 + a] - o[a]) * inv[a];
        if (Number.isNaN(t0) || Number.isNaN(t1))
            continue;
        if (t0 > t1) { const t = t0; t0 = t1; t1 = t; }
        tNear = Math.max(tNear, t0);
        tFar = Math.min(tFar, t1);
    }
    return tNear <= tFar ? tNear : Infinity;
}

const sub = (a, b) => [a[0] - b[0], a[1] - b[1], a[2] - b[2]];
const dot = (a, b) => a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
const cross = (a, b) => [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]];

function intersect64(f, u32, p, o, d, tMax) {
    const b = p * PRIM_FLOATS, kind = u32[b], T_MIN = 1e-4;
    const vec = (i) => [f[b + i], f[b + i + 1], f[b + i + 2]];
    if (kind == PRIMITIVE_KINDS.sphere) {
        const oc = sub(o, vec(4)), r = f[b + 7];
        const a = dot(d, d), hb = dot(oc, d), c = dot(oc, oc) - r * r;
        const disc = hb * hb - a * c;
        if (disc < 0.0) return null;
        const sq = Math.sqrt(disc);
        let t = (-hb - sq) / a;
        if (t <= T_MIN || t >= tMax) {
            t = (-hb + sq) / a;
            if (t <= T_MIN || t >= tMax) return null;
        }
        return [t, 0.0, 0.0];
    }
    if (kind == PRIMITIVE_KINDS.quad) {
        const q = vec(4), qu = vec(8), qv = vec(12);
        const n = cross(qu, qv), denom = dot(n, d);
        if (Math.abs(denom) < 1e-12) return null;
        const t = dot(n, sub(q, o)) / denom;
        if (t <= T_MIN || t >= tMax) return null;
        const h = [o[0] + d[0] * t - q[0], o[1] + d[1] * t - q[1], o[2] + d[2] * t - q[2]];
        const nn = dot(n, n);
        const alpha = dot(n, cross(h, qv)) / nn, beta = dot(n, cross(qu, h)) / nn;
        if (alpha < 0.0 || alpha > 1.0 || beta < 0.0 || beta > 1.0) return null;
        return [t, alpha, beta];
    }
    if (kind == PRIMITIVE_KINDS.box) {
        let tNear = -Infinity, tFar = Infinity;
        for (let a = 0; a < 3; ++a) {
            const r = b + 4 + a * 4;
            const lo = f[r] * o[0] + f[r + 1] * o[1] + f[r + 2] * o[2] + f[r + 3];
            const ld = f[r] * d[0] + f[r + 1] * d[1] + f[r + 2] * d[2];
            const inv = 1.0 / ld;
            let t0 = (-1.0 - lo) * inv, t1 = (1.0 - lo) * inv;
            if (Number.isNaN(t0) || Number.isNaN(t1)) continue;
            if (t0 > t1) { const t = t0; t0 = t1; t1 = t; }
            tNear = Math.max(tNear, t0);
            tFar = Math.min(tFar, t1);
        }
        if (tNear > tFar) return null;
        const t = tNear > T_MIN ? tNear : tFar;
        if (t <= T_MIN || t >= tMax) return null;
        return [t, 0.0, 0.0];
    }
    if (kind != PRIMITIVE_KINDS.triangle) return null;
    const v0 = vec(4), e1 = sub(vec(8), v0), e2 = sub(vec(12), v0);
    const pv = cross(d, e2), det = dot(e1, pv);
    if (Math.abs(det) < 1e-12) return null;
    const id = 1.0 / det, tv = sub(o, v0);
    const bu = dot(tv, pv) * id;
    if (bu < 0.0 || bu > 1.0) return null;
    const qv = cross(tv, e1), bv = dot(d, qv) * id;
    if (bv < 0.0 || bu + bv > 1.0) return null;
    const t = dot(e2, qv) * id;
    if (t <= T_MIN || t >= tMax) return null;
    return [t, bu, bv];
}

function boxPoint(f, b, point) {
    const lp = [];
    for (let r = 0; r < 3; ++r) {
        const o = b + 4 + r * 4;
        lp.push(f[o] * point.x + f[o + 1] * point.y + f[o + 2] * point.z + f[o + 3]);
    }
    return lp;
}

function normals64(packed, f, u32, p, point, bu, bv) {
    const b = p * PRIM_FLOATS, kind = u32[b];
    const vec = (i) => new Vec3(f[b + i], f[b + i + 1], f[b + i + 2]);
    if (kind == PRIMITIVE_KINDS.sphere) {
        const n = point.subtract(vec(4)).scale(1.0 / f[b + 7]).normalized();
        return { geometric: n, shading: n };
    }
    if (kind == PRIMITIVE_KINDS.quad) {
        const n = vec(8).cross(vec(12)).normalized();
        return { geometric: n, shading: n };
    }
    if (kind == PRIMITIVE_KINDS.box) {
        const lp = boxPoint(f, b, point);
        let axis = 0;
        for (let a = 1; a < 3; ++a)
            if (Math.abs(lp[a]) > Math.abs(lp[axis])) axis = a;
        const sign = Math.sign(lp[axis]) || 1.0;
        const n = vec(4 + axis * 4).scale(sign).normalized();
        return { geometric: n, shading: n };
    }
    const v0 = vec(4);
    const geometric = vec(8).subtract(v0).cross(vec(12).subtract(v0)).normalized();
    const ai = u32[b + 3], af = packed.attributes;
    if (ai == NONE || !(af[(ai + 3) * 4 + 3] & SMOOTH))
        return { geometric, shading: geometric };
    const nrm = (j) => new Vec3(af[(ai + j) * 4], af[(ai + j) * 4 + 1], af[(ai + j) * 4 + 2]);
    let shading = nrm(0).scale(1.0 - bu - bv).add(nrm(1).scale(bu)).add(nrm(2).scale(bv)).normalized();
    if (shading.dot(geometric) < 0.0)
        shading = shading.negate();
    return { geometric, shading };
}

function tangents64(packed, f, u32, p, point, bu, bv) {
    const b = p * PRIM_FLOATS, kind = u32[b];
    const vec = (i) => new Vec3(f[b + i], f[b + i + 1], f[b + i + 2]);
    if (kind == PRIMITIVE_KINDS.sphere) {
        const d = point.subtract(vec(4)).normalized();
        let tangent = vec(12).cross(d);
        if (tangent.length() < 1e-6)
            tangent = vec(8);
        return { tangent, up: d.cross(tangent) };
    }
    if (kind == PRIMITIVE_KINDS.quad)
        return { tangent: vec(8), up: vec(12) };
    if (kind == PRIMITIVE_KINDS.box) {
        const lp = boxPoint(f, b, point);
        const ax = Math.abs(lp[0]), ay = Math.abs(lp[1]), az = Math.abs(lp[2]);
        if (ax >= ay && ax >= az) return { tangent: vec(12).scale(-Math.sign(lp[0])), up: vec(8) };
        if (ay >= az) return { tangent: vec(4), up: vec(12).scale(-Math.sign(lp[1])) };
        return { tangent: vec(4).scale(Math.sign(lp[2])), up: vec(8) };
    }
    const v0 = vec(4), e1 = vec(8).subtract(v0), e2 = vec(12).subtract(v0);
    const ai = u32[b + 3], af = packed.attributes;
    let du1 = [1, 0], du2 = [0, 1];
    if (ai != NONE && (af[(ai + 3) * 4 + 3] & TEXTURED)) {
        const uv = (j) => [af[(ai + j) * 4 + 3], af[(ai + 3) * 4 + j]];
        const a = uv(0), c = uv(1), d = uv(2);
        const t1 = [c[0] - a[0], c[1] - a[1]], t2 = [d[0] - a[0], d[1] - a[1]];
        if (Math.abs(t1[0] * t2[1] - t2[0] * t1[1]) > 1e-12) { du1 = t1; du2 = t2; }
    }
    const det = du1[0] * du2[1] - du2[0] * du1[1];
    const dpdu = e1.scale(du2[1]).subtract(e2.scale(du1[1])).scale(1.0 / det);
    const dpdv = e2.scale(du1[0]).subtract(e1.scale(du2[0])).scale(1.0 / det);
    return { tangent: dpdu, up: dpdv.negate() };
}

function uv64(packed, f, u32, p, point, bu, bv) {
    const b = p * PRIM_FLOATS, kind = u32[b];
    const vec = (i) => new Vec3(f[b + i], f[b + i + 1], f[b + i + 2]);
    if (kind == PRIMITIVE_KINDS.sphere)
        return sphereUv(point.subtract(vec(4)).normalized(), vec(8), vec(12));
    if (kind == PRIMITIVE_KINDS.quad)
        return new Vec2(bu, 1.0 - bv);
    if (kind == PRIMITIVE_KINDS.box) {
        const lp = boxPoint(f, b, point);
        const half = [0, 1, 2].map((r) => 1.0 / Math.hypot(f[b + 4 + r * 4], f[b + 5 + r * 4], f[b + 6 + r * 4]));
        return boxUv(lp, half);
    }
    const ai = u32[b + 3], af = packed.attributes;
    if (ai == NONE || !(af[(ai + 3) * 4 + 3] & TEXTURED))
        return new Vec2(bu, bv);
    const w = 1.0 - bu - bv;
    return new Vec2(af[ai * 4 + 3] * w + af[(ai + 1) * 4 + 3] * bu + af[(ai + 2) * 4 + 3] * bv,
        af[(ai + 3) * 4] * w + af[(ai + 3) * 4 + 1] * bu + af[(ai + 3) * 4 + 2] * bv);
}

/** The shading normal n tilted by a normal-map texel t, kept outside ng (src/Scene.mjs:863-879). */
export function bump(n, ng, tangent, up, t, strength) {
    let right = tangent.subtract(n.scale(n.dot(tangent)));
    if (right.length() < 1e-9)
        return n;
    right = right.normalized();
    let above = n.cross(right);
    if (above.dot(up) < 0.0)
        above = above.negate();
    const tilted = right.scale((t.x * 2.0 - 1.0) * strength)
        .add(above.scale((t.y * 2.0 - 1.0) * strength))
        .add(n.scale(t.z * 2.0 - 1.0))
        .normalized();
    return tilted.dot(ng) > 0.0 ? tilted : n;
}

/** Longitude and latitude on a globe with axes x and north (src/Scene.mjs:888-893). */
export function sphereUv(d, x, north) {
    const z = x.cross(north);
    const u = Math.atan2(-d.dot(z), d.dot(x)) / (2.0 * Math.PI);
    const v = Math.acos(Math.min(Math.max(d.dot(north), -1.0), 1.0)) / Math.PI;
    return new Vec2(u - Math.floor(u), v);
}

/** Texture coordinates on a box face in world units (src/Scene.mjs:904-914). */
export function boxUv(lp, half) {
    const x = lp[0], y = lp[1], z = lp[2];
    const ax = Math.abs(x), ay = Math.abs(y), az = Math.abs(z);
    if (ax >= ay && ax >= az)
        return new Vec2((1.0 - Math.sign(x) * z) * half[2], (1.0 - y) * half[1]);
    if (ay >= az)
        return new Vec2((1.0 + x) * half[0], (1.0 + Math.sign(y) * z) * half[2]);
    return new Vec2((1.0 + Math.sign(z) * x) * half[0], (1.0 - y) * half[1]);
}
