// Host-side frame state that the kernels read, computed in plain JS with no
// device: the lights table, light-tracing emitters, the 208-byte parameter
// block and the half-float environment map. Each follows the reference's
// Renderer (src/Renderer.mjs), which the kernels in this package replace:
//   lights table        src/Renderer.mjs:585-607, lightChances :1167-1179
//   emitters            #writeEmitters :670-717, deltaObjects :1187-1254
//   parameter block     #writeParams :925-991, layout src/shaders.mjs:9-24
//   environment upload  #uploadEnvironment :723-760, toHalf :1143-1160
import { Box, Mesh, Quad } from "./geometry.mjs";
import { Vec3 } from "./math.mjs";
import { PRIMITIVE_KINDS } from "./scene.mjs";

export const PARAMS_BYTES = 208;
/** The first bytes of the parameter block decide the image. */
export const IMAGE_KEY_BYTES = 168;
export const PHOTON_TABLE = 1048576;
export const PHOTON_CAPACITY = 524288;
export const MAX_TARGETS = 16;
export const ATROUS_STEPS = Object.freeze([1, 2, 4, 8, 16]);
export const VIEWS = Object.freeze({ color: 0, albedo: 1, normal: 2, depth: 3 });

const UNIFORM_SHARE = 0.25;

/** NEE's chance per light: 25% spread evenly, 75% by power. */
export function lightChances(powers) {
    let total = 0;
    for (const p of powers) total += p;
    const n = powers.length;
    if (!(total > 0) || !isFinite(total))
        return powers.map(() => 1 / n);
    return powers.map((p) => UNIFORM_SHARE / n + (1 - UNIFORM_SHARE) * p / total);
}

/**
 * The scene-level tables derived from a packed scene: each light's power
 * (luminance · π · area), and the u32 table the kernels read — light
 * primitive indices, then volume primitive indices, then the lights'
 * running chance totals as f32 bits (last forced to exactly 1).
 */
export function sceneTables(packed) {
    const lights = packed.lights[0] == 0xffffffff ? [] : Array.from(packed.lights);
    const pf = new Float32Array(packed.primitives);
    const pu = new Uint32Array(packed.primitives);
    const mf = new Float32Array(packed.materials);
    const powers = lights.map((prim) => {
        const b = prim * 16, m = pu[b + 1] * 24;
        const lum = 0.2126 * mf[m + 4] + 0.7152 * mf[m + 5] + 0.0722 * mf[m + 6];
        const v = (i) => new Vec3(pf[b + i], pf[b + i + 1], pf[b + i + 2]);
        const area = pu[b] == PRIMITIVE_KINDS.quad ? v(8).cross(v(12)).length() : 4 * Math.PI * Math.pow(pf[b + 7], 2);
        return { prim, power: lum * Math.PI * area };
    });
    const chances = lightChances(powers.map((l) => l.power));
    let running = 0;
    const cumulative = new Float32Array(chances.map((c) => (running += c)));
    if (cumulative.length > 0)
        cumulative[cumulative.length - 1] = 1;
    const list = lights.concat(packed.volumes, Array.from(new Uint32Array(cumulative.buffer)));
    const mats = packed.materialList;
    return {
        table: new Uint32Array(list.length > 0 ? list : [0xffffffff]),
        lightCount: lights.length,
        volumeCount: packed.volumes.length,
        powers,
        textured: packed.textures.length + packed.normalMaps.length > 0,
        dispersive: mats.some((m) => m.dispersion > 0),
    };
}

function boxAround(points) {
    const min = [Infinity, Infinity, Infinity], max = [-Infinity, -Infinity, -Infinity];
    for (const p of points)
        for (let a = 0; a < 3; ++a) {
            min[a] = Math.min(min[a], p[a]);
            max[a] = Math.max(max[a], p[a]);
        }
    return { min, max };
}

const halfDiagonal = (b) => Math.hypot(b.max[0] - b.min[0], b.max[1] - b.min[1], b.max[2] - b.min[2]) / 2;

/**
 * The mirrors and glass light tracing aims at: up to MAX_TARGETS bounding
 * spheres of groups (the largest group split across its longest centroid
 * extent until there are enough), the sphere round them all, and how far
 * out the sun's light starts. null when there are none.
 */
export function deltaObjects(scene) {
    const boxes = [];
    for (const object of scene.objects) {
        const m = object.material;
        if (!(m.type == "glass" || (m.type == "metal" && m.roughness < 0.01)))
            continue;
        let points;
        if (object instanceof Mesh) points = object.positions;
        else if (object instanceof Box) points = object.corners();
        else if (object instanceof Quad) {
            const c = object.corner, u = object.u, v = object.v;
            points = [c, c.add(u), c.add(v), c.add(u).add(v)];
        }
        else {
            const r = new Vec3(object.radius, object.radius, object.radius);
            points = [object.center.subtract(r), object.center.add(r)];
        }
        const box = boxAround(points.map((p) => [p.x, p.y, p.z]));
        box.centre = box.min.map((c, a) => (c + box.max[a]) / 2);
        boxes.push(box);
    }
    if (boxes.length == 0)
        return null;

    const groups = [boxes];
    while (groups.length < MAX_TARGETS) {
        let biggest = -1, size = 0;
        groups.forEach((g, i) => {
            const s = g.length > 1 ? halfDiagonal(boxAround([].concat(...g.map((b) => [b.min, b.max])))) : 0;
            if (s > size) { biggest = i; size = s; }
        });
        if (biggest < 0)
            break;
        const group = groups[biggest];
        const spread = boxAround(group.map((b) => b.centre));
        const extent = spread.max.map((c, a) => c - spread.min[a]);
        const axis = extent.indexOf(Math.max(...extent));
        group.sort((p, q) => p.centre[axis] - q.centre[axis]);
        groups.splice(biggest, 1, group.slice(0, group.length >> 1), group.slice(group.length >> 1));
    }

    const sphereAround = (box) => ({
        centre: new Vec3((box.min[0] + box.max[0]) / 2, (box.min[1] + box.max[1]) / 2, (box.min[2] + box.max[2]) / 2),
        radius: Math.max(halfDiagonal(box), 1e-3) * 1.001,
    });
    const targets = groups.map((g) => sphereAround(boxAround([].concat(...g.map((b) => [b.min, b.max])))));
    const all = sphereAround(boxAround([].concat(...boxes.map((b) => [b.min, b.max]))));
    const bounds = scene.bounds();
    const middle = bounds ? bounds.min.add(bounds.max).scale(0.5) : new Vec3(0, 0, 0);
    const reach = bounds ?
        Math.max(...targets.map((t) => t.centre.subtract(middle).length())) + bounds.max.subtract(bounds.min).length() / 2 + 1 : 1;
    return { centre: all.centre, radius: all.radius, reach, targets };
}

/**
 * The emitters buffer light tracing reads: a header vec4 (u32 sources, u32
 * targets, f32 reach, f32 circles' total area), a vec4 per source
 * (cumulative chance, u32 primitive or NONE for the sun, chance, 0), then a
 * vec4 per target (centre, radius). null when light tracing has nothing to
 * do. The buffer holds 2 + lights + MAX_TARGETS vec4s.
 */
export function emitters(sky, powers, delta) {
    if (!delta)
        return null;
    const circles = delta.targets.reduce((s, t) => s + Math.PI * t.radius * t.radius, 0);
    const sources = powers.slice();
    if (sky.sunStrength > 0) {
        const omc = 2.0 * Math.pow(Math.sin(sky.sunSize / 2.0), 2);
        const r = sky.sunRadiance;
        const lum = 0.2126 * r.x + 0.7152 * r.y + 0.0722 * r.z;
        sources.push({ prim: 0xffffffff, power: lum * 2 * Math.PI * omc * circles });
    }
    const total = sources.reduce((s, e) => s + e.power, 0);
    if (!(total > 0))
        return null;
    const data = new Float32Array((2 + powers.length + MAX_TARGETS) * 4);
    const u = new Uint32Array(data.buffer);
    u[0] = sources.length;
    u[1] = delta.targets.length;
    data[2] = delta.reach;
    data[3] = circles;
    let running = 0;
    sources.forEach((s, k) => {
        running += s.power / total;
        data[4 + k * 4] = k == sources.length - 1 ? 1.0 : running;
        u[4 + k * 4 + 1] = s.prim;
        data[4 + k * 4 + 2] = s.power / total;
    });
    delta.targets.forEach((t, j) => {
        data.set([t.centre.x, t.centre.y, t.centre.z, t.radius], 4 + (sources.length + j) * 4);
    });
    return data;
}

/**
 * Writes the 208-byte parameter block. state: { lightCount, volumeCount,
 * maxBounces, samplesPerFrame, exposure, fireflyClamp, view, denoise,
 * caustics, samples, delta }. Returns the block; bytes [0, 168) are the
 * image key.
 */
export function writeParams(buffer, scene, camera, width, height, state) {
    const f = new Float32Array(buffer);
    const u = new Uint32Array(buffer);
    const { right, up, forward } = camera.basis();
    const sky = scene.sky;
    const row = (r, v) => { f[r * 4] = v.x; f[r * 4 + 1] = v.y; f[r * 4 + 2] = v.z; };
    row(0, camera.position);
    f[3] = camera.aperture;
    row(1, right);
    f[7] = Math.tan(camera.fov / 2.0);
    row(2, up);
    f[11] = camera.effectiveFocusDistance;
    row(3, forward);
    f[15] = width / height;
    row(4, sky.zenith);
    f[19] = 2.0 * Math.pow(Math.sin(sky.sunSize / 2.0), 2);
    row(5, sky.horizon);
    u[23] = sky.sunStrength > 0.0 ? 1 : 0;
    row(6, sky.ground);
    f[27] = state.fireflyClamp;
    row(7, sky.sunDirection);
    u[31] = state.lightCount;
    row(8, sky.sunRadiance);
    u[35] = state.maxBounces;
    u[36] = VIEWS[state.view];
    u[37] = state.volumeCount;
    f[38] = sky.environment ? sky.environmentStrength : 0.0;
    f[39] = sky.environmentRotation;
    u[40] = width;
    u[41] = height;
    return buffer;
}

/** The per-frame part of the block: sample counters, exposure, flags and photon radius. */
export function writeFrameParams(buffer, camera, height, state) {
    const f = new Float32Array(buffer);
    const u = new Uint32Array(buffer);
    u[42] = state.samples;
    u[43] = state.samplesPerFrame;
    u[44] = state.samples + state.samplesPerFrame;
    f[45] = state.exposure;
    u[46] = state.denoise ? 1 : 0;
    u[47] = state.caustics ? 1 : 0;
    // About 3 pixels across at the mirrors and glass, shrinking as
    // frames^(-1/6); cells twice the radius.
    const sphere = state.delta;
    const distance = sphere ? Math.max(camera.position.subtract(sphere.centre).length(), sphere.radius) : 1;
    const start = Math.max(3 * 2 * Math.tan(camera.fov / 2) * distance / height, 1e-5);
    const frame = Math.floor(state.samples / Math.max(state.samplesPerFrame, 1)) + 1;
    f[48] = start * Math.pow(frame, -1 / 6);
    f[49] = 2 * f[48];
    return buffer;
}

/** Nearest half float as 16 bits; ≥ 65504 clamps to the largest finite half, ≤ 0 and NaN give 0. */
export function toHalf(x) {
    if (!(x > 0.0)) return 0;
    if (x >= 65504.0) return 0x7bff;
    if (x < 6.1035e-5) return Math.round(x / 5.9605e-8);
    const e = Math.floor(Math.log2(x));
    let mant = Math.round((x / Math.pow(2, e) - 1.0) * 1024);
    let biased = e + 15;
    if (mant == 1024) { mant = 0; biased++; }
    return (biased << 10) | mant;
}

/** The environment as the kernels read it: RGBA half floats and the importance table. */
export function environmentData(environment) {
    const map = environment ? environment : { width: 1, height: 1, data: new Float32Array(3) };
    const table = environment ? environment.importance() : { width: 1, height: 0, data: new Float32Array(1) };
    const rgba = new Uint16Array(map.width * map.height * 4);
    for (let i = 0; i < map.width * map.height; ++i) {
        for (let c = 0; c < 3; ++c)
            rgba[i * 4 + c] = toHalf(map.data[i * 3 + c]);
        rgba[i * 4 + 3] = toHalf(1.0);
    }
    return { rgba, width: map.width, height: map.height, table: table.data, tableWidth: table.width, tableHeight: table.height };
}

/** Largest texture layer size (src/Renderer.mjs:33). */
export const MAX_TEXTURE_SIZE = 2048;

/**
 * Textures as one array of RGBA8 layers, each resized to the largest one's
 * size (capped at MAX_TEXTURE_SIZE), as #uploadTextures packs them
 * (src/Renderer.mjs:637-661): sRGB colour for textures, plain numbers for
 * normal maps. null when there are none.
 */
export function textureLayers(textures, srgb) {
    if (textures.length == 0)
        return null;
    const width = Math.min(MAX_TEXTURE_SIZE, Math.max(1, ...textures.map((t) => t.width)));
    const height = Math.min(MAX_TEXTURE_SIZE, Math.max(1, ...textures.map((t) => t.height)));
    const data = new Uint8Array(width * height * 4 * textures.length);
    textures.forEach((t, i) => data.set(t.resized(width, height, srgb).data, i * width * height * 4));
    return { data, width, height, layers: textures.length };
}
