// Scene objects: Sphere (src/Sphere.mjs), Quad (src/Quad.mjs), Box
// (src/Box.mjs) and Mesh with its generators (src/Mesh.mjs:127-356). The
// constructors, argument meanings and errors are the reference's API; the
// code is this module's own, for Node 12.
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

// ------------------------------------------------------------------ meshes
//
// The generators below build every parametric shape on one lattice: a grid
// of (uSegments + 1) × (vSegments + 1) vertices over [0, 1]², in which a
// closed direction's last row lies on its first (the seam). The shape is
// sampled once per distinct lattice site, so both sides of a seam share a
// position and a normal, while each grid vertex keeps its own texture
// coordinate (0 and 1 at a seam). Each grid square becomes two triangles,
// (00, 10, 11) and (00, 11, 01), squares row by row, and triangles of no
// area are dropped. The reference's generators (src/Mesh.mjs:127-356) give
// the same vertices in the same order; tests/golden/js_generators.json pins
// the bytes.

const AREA_EPSILON = 1e-12;

class Lattice {
    constructor(uSegments, vSegments, uClosed, vClosed) {
        this.u = uSegments;
        this.v = vSegments;
        this.uSites = uClosed ? uSegments : uSegments + 1;
        this.vSites = vClosed ? vSegments : vSegments + 1;
    }

    /** The distinct site that grid vertex (i, j) lies on. */
    site(i, j) { return (i % this.uSites) * this.vSites + (j % this.vSites); }

    /** Grid vertex (i, j)'s index. */
    vertex(i, j) { return i * (this.v + 1) + j; }

    /** Visits the grid's vertices (i, j) row by row. */
    eachVertex(f) {
        for (let i = 0; i <= this.u; ++i)
            for (let j = 0; j <= this.v; ++j)
                f(i, j);
    }

    /** The triangles of every square, as indices given by at(i, j). */
    triangles(at) {
        const out = [];
        for (let i = 0; i < this.u; ++i)
            for (let j = 0; j < this.v; ++j) {
                const a = at(i, j), b = at(i + 1, j), c = at(i + 1, j + 1), d = at(i, j + 1);
                out.push(a, b, c, a, c, d);
            }
        return out;
    }
}

// The unnormalised normal (b − a) × (c − a) of triangle t of an index list.
function faceNormal(positions, indices, t) {
    const a = positions[indices[t]];
    return positions[indices[t + 1]].subtract(a).cross(positions[indices[t + 2]].subtract(a));
}

// The index list without its triangles of (near) zero area.
function dropDegenerate(positions, indices) {
    const kept = [];
    for (let t = 0; t < indices.length; t += 3)
        if (faceNormal(positions, indices, t).length() > AREA_EPSILON)
            kept.push(indices[t], indices[t + 1], indices[t + 2]);
    return kept;
}

// Smooth normals: each vertex's faces' cross products summed (so larger
// faces weigh more), in face order, then normalised; a vertex on no face
// (or whose faces cancel) points up.
function smoothNormals(positions, indices) {
    const acc = new Float64Array(positions.length * 3);
    for (let t = 0; t < indices.length; t += 3) {
        const n = faceNormal(positions, indices, t);
        for (let k = 0; k < 3; ++k) {
            const o = indices[t + k] * 3;
            acc[o] += n.x;
            acc[o + 1] += n.y;
            acc[o + 2] += n.z;
        }
    }
    const up = new Vec3(0.0, 1.0, 0.0);
    return positions.map((_, v) => {
        const n = new Vec3(acc[v * 3], acc[v * 3 + 1], acc[v * 3 + 2]);
        return n.length() > 0.0 ? n.normalized() : up;
    });
}

// The icosahedron's vertices: for each coordinate plane p, the four points
// with ±1 on axis p and ±φ on the next axis (φ first positive, then
// negative; ±1 first negative), 0 on the third; and its 20 faces over them,
// wound outwards.
const PHI = (1.0 + Math.sqrt(5.0)) / 2.0;
function icosahedronVertices() {
    const out = [];
    for (let p = 0; p < 3; ++p)
        for (const along of [1, -1])
            for (const unit of [-1, 1]) {
                const c = [0, 0, 0];
                c[p] = unit;
                c[(p + 1) % 3] = along * PHI;
                out.push(new Vec3(c[0], c[1], c[2]).normalized());
            }
    return out;
}
const ICOSAHEDRON_FACES = [
    0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11,
    1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
    3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9,
    4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1,
];

/** A triangle mesh; fronts face along (b - a) × (c - a). */
export class Mesh {

    #positions; #indices; #normals; #uvs; #material;

    constructor(positions, indices, material, { shading = "smooth", normals = undefined, uvs = null } = {}) {
        if (indices.length % 3 != 0)
            throw new RangeError("A mesh needs three indices per triangle.");
        const bad = Array.prototype.find.call(indices, (i) => !(Number.isInteger(i) && i >= 0 && i < positions.length));
        if (bad !== undefined)
            throw new RangeError(`Index ${bad} has no vertex.`);
        if (normals && normals.length != positions.length)
            throw new RangeError("A mesh needs one normal per vertex.");
        if (uvs && uvs.length != positions.length)
            throw new RangeError("A mesh needs one texture coordinate per vertex.");
        this.#positions = positions;
        this.#indices = Uint32Array.from(indices);
        this.#material = material;
        this.#uvs = uvs;
        if (normals)
            this.#normals = normals;
        else
            this.#normals = shading == "smooth" ? smoothNormals(positions, this.#indices) : null;
    }

    get positions() { return this.#positions; }
    get indices() { return this.#indices; }
    get normals() { return this.#normals; }
    get uvs() { return this.#uvs; }
    get material() { return this.#material; }
    get triangleCount() { return this.#indices.length / 3; }

    /**
     * The mesh moved by an affine transform: positions by it, normals by its
     * inverse transpose (renormalised); a transform that mirrors (negative
     * determinant) swaps each triangle's last two corners to keep fronts out.
     */
    transformed(transform) {
        const mirrored = transform.determinant() < 0.0;
        const indices = Uint32Array.from(this.#indices);
        if (mirrored)
            for (let t = 0; t < indices.length; t += 3)
                [indices[t + 1], indices[t + 2]] = [indices[t + 2], indices[t + 1]];
        let normals;
        if (this.#normals) {
            const toNormals = transform.inverse().transpose();
            normals = this.#normals.map((n) => toNormals.multiplyDirection(n).normalized());
        }
        return new Mesh(this.#positions.map((p) => transform.multiplyVector3(p)), indices, this.#material,
            { shading: "flat", normals, uvs: this.#uvs });
    }

    /**
     * The surface point(u, v) over [0, 1]² on a uSegments × vSegments grid,
     * open or closed (welded at the seam) each way; smooth normals come from
     * the welded shape, so they are continuous across seams.
     */
    static surface(point, uSegments, vSegments, material, { uWrap = "open", vWrap = "open", shading = "smooth" } = {}) {
        const grid = new Lattice(uSegments, vSegments, uWrap != "open", vWrap != "open");
        const sites = [];
        for (let i = 0; i < grid.uSites; ++i)
            for (let j = 0; j < grid.vSites; ++j)
                sites.push(point(i / uSegments, j / vSegments));
        const positions = [], uvs = [];
        grid.eachVertex((i, j) => {
            positions.push(sites[grid.site(i, j)]);
            uvs.push(new Vec2(i / uSegments, j / vSegments));
        });
        let normals;
        if (shading == "smooth") {
            const siteTriangles = dropDegenerate(sites, grid.triangles((i, j) => grid.site(i, j)));
            const siteNormals = smoothNormals(sites, Uint32Array.from(siteTriangles));
            normals = [];
            grid.eachVertex((i, j) => normals.push(siteNormals[grid.site(i, j)]));
        }
        const triangles = dropDegenerate(positions, grid.triangles((i, j) => grid.vertex(i, j)));
        return new Mesh(positions, triangles, material, { shading, normals, uvs });
    }

    /**
     * The unit sphere as an icosahedron whose faces are split in four
     * `subdivisions` times, each new vertex the normalised midpoint of an
     * edge (shared by the edge's two faces).
     */
    static icosphere(material, subdivisions = 3, { shading = "smooth" } = {}) {
        const positions = icosahedronVertices();
        let faces = ICOSAHEDRON_FACES;
        for (let s = 0; s < subdivisions; ++s) {
            const midpoints = new Map();   // "lo,hi" → vertex
            const midpoint = (a, b) => {
                const key = a < b ? `${a},${b}` : `${b},${a}`;
                if (!midpoints.has(key)) {
                    midpoints.set(key, positions.length);
                    positions.push(positions[a].add(positions[b]).normalized());
                }
                return midpoints.get(key);
            };
            const split = [];
            for (let t = 0; t < faces.length; t += 3) {
                const [a, b, c] = [faces[t], faces[t + 1], faces[t + 2]];
                const ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
                split.push(a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca);
            }
            faces = split;
        }
        // On the unit sphere a vertex's position is its normal.
        return new Mesh(positions, faces, material, shading == "smooth" ? { normals: positions } : { shading });
    }

    /**
     * The solid swept by turning profile (x: radius, y: height) a full turn
     * about the y axis; texture v runs from the profile's last point down.
     */
    static lathe(profile, material, { segments = 64, shading = "smooth" } = {}) {
        if (profile.length < 2)
            throw new RangeError("A lathe's profile needs at least 2 points.");
        const steps = profile.length - 1;
        const swept = Mesh.surface((u, v) => {
            const p = profile[Math.round(v * steps)];
            const angle = u * 2.0 * Math.PI;
            return new Vec3(p.x * Math.cos(angle), p.y, -p.x * Math.sin(angle));
        }, segments, steps, material, { uWrap: "closed", shading });
        return new Mesh(swept.positions, swept.indices, material, {
            shading: "flat",
            normals: swept.normals || undefined,
            uvs: (swept.uvs || []).map((t) => new Vec2(t.x, 1.0 - t.y)),
        });
    }

    /** A torus about the y axis: ring radius `radius`, tube radius `tube`. */
    static torus(material, { radius = 1.0, tube = 0.35, segments = 64, sides = 32 } = {}) {
        const ring = (u, v) => {
            const around = u * 2.0 * Math.PI, across = v * 2.0 * Math.PI;
            const r = radius + tube * Math.cos(across);
            return new Vec3(r * Math.cos(around), tube * Math.sin(across), -r * Math.sin(around));
        };
        return Mesh.surface(ring, segments, sides, material, { uWrap: "closed", vWrap: "closed" });
    }

    /**
     * A tube of `radius` round the closed curve curve(t), t in [0, 1),
     * sampled at `segments` points. Its cross-sections are framed by parallel
     * transport of a normal along the curve's central-difference tangents;
     * the angle the transported normal has turned by when it comes round is
     * spread evenly along the tube, so the seam closes without a kink.
     */
    static tube(curve, material, { radius = 0.3, segments = 256, sides = 24 } = {}) {
        const centres = Array.from({ length: segments }, (_, k) => curve(k / segments));
        const tangent = (k) => centres[(k + 1) % segments].subtract(centres[(k + segments - 1) % segments]).normalized();
        const start = tangent(0);
        const helper = Math.abs(start.y) < 0.9 ? new Vec3(0, 1, 0) : new Vec3(1, 0, 0);
        const normals = [start.cross(helper).normalized()];
        for (let k = 1; k <= segments; ++k) {
            const t = tangent(k % segments), prev = normals[k - 1];
            normals.push(prev.subtract(t.scale(prev.dot(t))).normalized());
        }
        const binormal0 = start.cross(normals[0]);
        const twist = Math.atan2(normals[segments].dot(binormal0), normals[segments].dot(normals[0]));
        return Mesh.surface((u, v) => {
            const k = Math.round(u * segments) % segments;
            const n = normals[k];
            const b = tangent(k).cross(n);
            const angle = 2.0 * Math.PI * v + twist * k / segments;
            return centres[k].add(n.scale(Math.cos(angle) * radius)).subtract(b.scale(Math.sin(angle) * radius));
        }, segments, sides, material, { uWrap: "closed", vWrap: "closed" });
    }
}
