// Scene objects: Sphere (src/Sphere.mjs), Quad (src/Quad.mjs), Box
// (src/Box.mjs) and Mesh with its generators (src/Mesh.mjs:127-356). Same
// constructors, argument meanings and errors; written for Node 12.
import { Mat4, Vec2, Vec3 } from "./math.mjs";

export class Sphere {

    #center; #radius; #material; #rotation;

    constructor(center, radius, material, { rotation = Mat4.identity() } = {}) {
        if (!(radius > 0.0))
            throw new RangeError("A sphere's radius must be more than 0.");
        this.#center = center;
        this.#radius = radius;
        this.#material = material;
        this.#rotation = rotation;
    }

    get center() { return this.#center; }
    get radius() { return this.#radius; }
    get material() { return this.#material; }
    get rotation() { return this.#rotation; }

    /** The globe's +x (u = 0 on the equator) and north pole, unit length, in world space. */
    axes() {
        return {
            x: this.#rotation.multiplyDirection(new Vec3(1, 0, 0)).normalized(),
            north: this.#rotation.multiplyDirection(new Vec3(0, 1, 0)).normalized(),
        };
    }
}

/** The parallelogram corner, corner + u, corner + v, corner + u + v; front along u × v. */
export class Quad {

    #corner; #u; #v; #material;

    constructor(corner, u, v, material) {
        if (!(u.cross(v).length() > 0.0))
            throw new RangeError("A quad's edges must not be parallel.");
        this.#corner = corner;
        this.#u = u;
        this.#v = v;
        this.#material = material;
    }

    get corner() { return this.#corner; }
    get u() { return this.#u; }
    get v() { return this.#v; }
    get material() { return this.#material; }
    get normal() { return this.#u.cross(this.#v).normalized(); }
    get area() { return this.#u.cross(this.#v).length(); }
}

/** The cube [-1, 1]³ moved by an affine transform. */
export class Box {

    #transform; #inverse; #material;

    constructor(transform, material) {
        this.#transform = transform;
        this.#inverse = transform.inverse();
        this.#material = material;
    }

    static fromCorners(min, max, material, rotateY = 0.0) {
        const c = min.add(max).scale(0.5);
        const h = max.subtract(min).scale(0.5);
        return new Box(Mat4.translate(c.x, c.y, c.z)
            .multiplyMat4(Mat4.rotateY(rotateY))
            .multiplyMat4(Mat4.scale(h.x, h.y, h.z)), material);
    }

    /** A square-sectioned stick from a to b, thickness across. */
    static between(a, b, thickness, material) {
        const centre = a.add(b).scale(0.5);
        const along = b.subtract(a);
        const half = along.length() / 2.0;
        if (!(half > 0.0))
            throw new RangeError("A stick's ends must differ.");
        const z = along.normalized();
        const helper = Math.abs(z.y) < 0.9 ? new Vec3(0, 1, 0) : new Vec3(1, 0, 0);
        const x = helper.cross(z).normalized();
        const y = z.cross(x);
        const t = thickness / 2.0;
        return new Box(new Mat4(
            x.x * t, y.x * t, z.x * half, centre.x,
            x.y * t, y.y * t, z.y * half, centre.y,
            x.z * t, y.z * t, z.z * half, centre.z,
            0, 0, 0, 1), material);
    }

    get transform() { return this.#transform; }
    /** World to box space. */
    get inverse() { return this.#inverse; }
    get material() { return this.#material; }

    corners() {
        const out = [];
        for (const x of [-1.0, 1.0])
            for (const y of [-1.0, 1.0])
                for (const z of [-1.0, 1.0])
                    out.push(this.#transform.multiplyVector3(new Vec3(x, y, z)));
        return out;
    }
}

// Area-weighted vertex normals: each face's unnormalised cross product is
// summed into its three vertices (src/Mesh.mjs:358-383).
function vertexNormals(positions, indices) {
    const sums = new Float64Array(positions.length * 3);
    for (let i = 0; i < indices.length; i += 3) {
        const a = indices[i], b = indices[i + 1], c = indices[i + 2];
        const n = positions[b].subtract(positions[a]).cross(positions[c].subtract(positions[a]));
        for (const v of [a, b, c]) {
            sums[v * 3] += n.x;
            sums[v * 3 + 1] += n.y;
            sums[v * 3 + 2] += n.z;
        }
    }
    return positions.map((_, v) => {
        const n = new Vec3(sums[v * 3], sums[v * 3 + 1], sums[v * 3 + 2]);
        return n.length() > 0.0 ? n.normalized() : new Vec3(0.0, 1.0, 0.0);
    });
}

// Keeps triangles whose area exceeds 1e-12 (src/Mesh.mjs:385-400).
function withoutDegenerates(positions, indices) {
    const kept = [];
    for (let i = 0; i < indices.length; i += 3) {
        const a = indices[i], b = indices[i + 1], c = indices[i + 2];
        if (positions[b].subtract(positions[a]).cross(positions[c].subtract(positions[a])).length() > 1e-12)
            kept.push(a, b, c);
    }
    return kept;
}

/** A triangle mesh; fronts face along (b - a) × (c - a). */
export class Mesh {

    #positions; #indices; #normals; #uvs; #material;

    constructor(positions, indices, material, { shading = "smooth", normals = undefined, uvs = null } = {}) {
        if (indices.length % 3 != 0)
            throw new RangeError("A mesh needs three indices per triangle.");
        for (let i = 0; i < indices.length; ++i) {
            if (!(indices[i] >= 0 && indices[i] < positions.length && Number.isInteger(indices[i])))
                throw new RangeError(`Index ${indices[i]} has no vertex.`);
        }
        if (normals && normals.length != positions.length)
            throw new RangeError("A mesh needs one normal per vertex.");
        if (uvs && uvs.length != positions.length)
            throw new RangeError("A mesh needs one texture coordinate per vertex.");
        this.#positions = positions;
        this.#indices = Uint32Array.from(indices);
        this.#material = material;
        this.#normals = normals ? normals : (shading == "smooth" ? vertexNormals(positions, this.#indices) : null);
        this.#uvs = uvs;
    }

    get positions() { return this.#positions; }
    get indices() { return this.#indices; }
    get normals() { return this.#normals; }
    get uvs() { return this.#uvs; }
    get material() { return this.#material; }
    get triangleCount() { return this.#indices.length / 3; }

    /** Moved by a transform; a mirroring transform flips the winding. */
    transformed(transform) {
        const positions = this.#positions.map((p) => transform.multiplyVector3(p));
        let indices = this.#indices;
        if (transform.determinant() < 0.0) {
            indices = indices.slice();
            for (let i = 0; i < indices.length; i += 3) {
                const t = indices[i + 1];
                indices[i + 1] = indices[i + 2];
                indices[i + 2] = t;
            }
        }
        let normals;
        if (this.#normals) {
            const nm = transform.inverse().transpose();
            normals = this.#normals.map((n) => nm.multiplyDirection(n).normalized());
        }
        return new Mesh(positions, indices, this.#material, { shading: "flat", normals, uvs: this.#uvs });
    }

    /** A grid surface point(u, v) over [0,1]², open or closed (welded) each way. */
    static surface(point, uSegments, vSegments, material, { uWrap = "open", vWrap = "open", shading = "smooth" } = {}) {
        const uRows = uWrap == "open" ? uSegments + 1 : uSegments;
        const vRows = vWrap == "open" ? vSegments + 1 : vSegments;
        const welded = [];
        for (let i = 0; i < uRows; ++i)
            for (let j = 0; j < vRows; ++j)
                welded.push(point(i / uSegments, j / vSegments));
        const weld = (i, j) => (i % uRows) * vRows + (j % vRows);
        const cols = vSegments + 1;
        const at = (i, j) => i * cols + j;

        const positions = [];
        const uvs = [];
        for (let i = 0; i <= uSegments; ++i)
            for (let j = 0; j <= vSegments; ++j) {
                positions.push(welded[weld(i, j)]);
                uvs.push(new Vec2(i / uSegments, j / vSegments));
            }

        const indices = [];
        const weldedIndices = [];
        for (let i = 0; i < uSegments; ++i)
            for (let j = 0; j < vSegments; ++j) {
                indices.push(at(i, j), at(i + 1, j), at(i + 1, j + 1), at(i, j), at(i + 1, j + 1), at(i, j + 1));
                weldedIndices.push(weld(i, j), weld(i + 1, j), weld(i + 1, j + 1), weld(i, j), weld(i + 1, j + 1), weld(i, j + 1));
            }

        let normals;
        if (shading == "smooth") {
            const wn = vertexNormals(welded, Uint32Array.from(withoutDegenerates(welded, weldedIndices)));
            normals = [];
            for (let i = 0; i <= uSegments; ++i)
                for (let j = 0; j <= vSegments; ++j)
                    normals.push(wn[weld(i, j)]);
        }
        return new Mesh(positions, withoutDegenerates(positions, indices), material, { shading, normals, uvs });
    }

    /** A unit sphere from a subdivided icosahedron. */
    static icosphere(material, subdivisions = 3, { shading = "smooth" } = {}) {
        const g = (1.0 + Math.sqrt(5.0)) / 2.0;
        const positions = [
            [-1, g, 0], [1, g, 0], [-1, -g, 0], [1, -g, 0],
            [0, -1, g], [0, 1, g], [0, -1, -g], [0, 1, -g],
            [g, 0, -1], [g, 0, 1], [-g, 0, -1], [-g, 0, 1],
        ].map((p) => new Vec3(p[0], p[1], p[2]).normalized());
        let faces = [
            0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11,
            1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
            3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9,
            4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1,
        ];
        for (let s = 0; s < subdivisions; ++s) {
            const mids = new Map();
            const mid = (a, b) => {
                const key = a < b ? a * 1e7 + b : b * 1e7 + a;
                let k = mids.get(key);
                if (k === undefined) {
                    k = positions.length;
                    positions.push(positions[a].add(positions[b]).normalized());
                    mids.set(key, k);
                }
                return k;
            };
            const next = [];
            for (let i = 0; i < faces.length; i += 3) {
                const a = faces[i], b = faces[i + 1], c = faces[i + 2];
                const ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
                next.push(a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca);
            }
            faces = next;
        }
        return new Mesh(positions, faces, material, shading == "smooth" ? { normals: positions } : { shading });
    }

    /** A solid of revolution of profile (x: radius, y: height) about the y axis. */
    static lathe(profile, material, { segments = 64, shading = "smooth" } = {}) {
        if (profile.length < 2)
            throw new RangeError("A lathe's profile needs at least 2 points.");
        const steps = profile.length - 1;
        const turned = Mesh.surface((u, v) => {
            const p = profile[Math.round(v * steps)];
            const a = u * 2.0 * Math.PI;
            return new Vec3(p.x * Math.cos(a), p.y, -p.x * Math.sin(a));
        }, segments, steps, material, { uWrap: "closed", shading });
        return new Mesh(turned.positions, turned.indices, material, {
            shading: "flat",
            normals: turned.normals ? turned.normals : undefined,
            uvs: (turned.uvs || []).map((uv) => new Vec2(uv.x, 1.0 - uv.y)),
        });
    }

    /** A torus about the y axis. */
    static torus(material, { radius = 1.0, tube = 0.35, segments = 64, sides = 32 } = {}) {
        return Mesh.surface((u, v) => {
            const a = u * 2.0 * Math.PI;
            const b = v * 2.0 * Math.PI;
            const r = radius + tube * Math.cos(b);
            return new Vec3(r * Math.cos(a), tube * Math.sin(b), -r * Math.sin(a));
        }, segments, sides, material, { uWrap: "closed", vWrap: "closed" });
    }

    /** A tube round the closed curve curve(t), t in [0, 1), framed by parallel transport. */
    static tube(curve, material, { radius = 0.3, segments = 256, sides = 24 } = {}) {
        const centres = [];
        for (let s = 0; s < segments; ++s)
            centres.push(curve(s / segments));
        const tangent = (s) => centres[(s + 1) % segments].subtract(centres[(s + segments - 1) % segments]).normalized();
        const t0 = tangent(0);
        const helper = Math.abs(t0.y) < 0.9 ? new Vec3(0, 1, 0) : new Vec3(1, 0, 0);
        let normal = t0.cross(helper).normalized();
        const frames = [normal];
        for (let s = 1; s <= segments; ++s) {
            const t = tangent(s % segments);
            normal = normal.subtract(t.scale(normal.dot(t))).normalized();
            frames.push(normal);
        }
        const b0 = t0.cross(frames[0]);
        const twist = Math.atan2(frames[segments].dot(b0), frames[segments].dot(frames[0]));
        return Mesh.surface((u, v) => {
            const s = Math.round(u * segments) % segments;
            const t = tangent(s);
            const n = frames[s];
            const b = t.cross(n);
            const angle = 2.0 * Math.PI * v + twist * s / segments;
            return centres[s].add(n.scale(Math.cos(angle) * radius)).subtract(b.scale(Math.sin(angle) * radius));
        }, segments, sides, material, { uWrap: "closed", vWrap: "closed" });
    }
}
