// Seeded workload scenes: the Cornell box with a glass ball (256×256
// raycast), 200 spheres on a checker floor (1080p path tracing), the
// triangle-heavy terrain/tori/lathes/tubes scene with fog and an
// environment map, and the caustics scene (water sheet, dispersive glass,
// spherical lamp, sun). Randomness is mulberry32, as the reference's demo
// scenes use it (scenes/helpers.mjs:11-20).
import { Camera } from "./camera.mjs";
import { Box, Mesh, Quad, Sphere } from "./geometry.mjs";
import { Material } from "./material.mjs";
import { Mat4, Vec2, Vec3 } from "./math.mjs";
import { Scene } from "./scene.mjs";
import { EnvironmentMap, Sky } from "./sky.mjs";

/** mulberry32: a repeatable sequence of numbers in [0, 1). */
export function seededRandom(seed) {
    let a = seed >>> 0;
    return () => {
        a = (a + 0x6d2b79f5) >>> 0;
        let t = a;
        t = Math.imul(t ^ (t >>> 15), t | 1);
        t ^= t + Math.imul(t ^ (t >>> 7), t | 61);
        return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
    };
}

const uniform = (rnd, lo, hi) => lo + (hi - lo) * rnd();

/**
 * Cornell box, 2 m room open at +z: white floor, ceiling and back (0.73),
 * red left, green right; a 0.5 × 0.5 m lamp of radiance 15 under the
 * ceiling; boxes 1.2 m and 0.6 m tall turned by seeded random yaws; a glass
 * ball of radius 0.3, ior 1.5. Nine primitives.
 */
export function cornell({ seed = 1 } = {}) {
    const rnd = seededRandom(seed);
    const white = Material.diffuse(new Vec3(0.73, 0.73, 0.73));
    const red = Material.diffuse(new Vec3(0.65, 0.05, 0.05));
    const green = Material.diffuse(new Vec3(0.12, 0.45, 0.15));
    const lamp = Material.light(new Vec3(1, 1, 1), 15.0);
    const scene = new Scene({ sky: Sky.black() });
    scene.add(
        new Quad(new Vec3(-1, 0, 1), new Vec3(2, 0, 0), new Vec3(0, 0, -2), white),
        new Quad(new Vec3(-1, 2, -1), new Vec3(2, 0, 0), new Vec3(0, 0, 2), white),
        new Quad(new Vec3(-1, 0, -1), new Vec3(2, 0, 0), new Vec3(0, 2, 0), white),
        new Quad(new Vec3(-1, 0, 1), new Vec3(0, 0, -2), new Vec3(0, 2, 0), red),
        new Quad(new Vec3(1, 0, -1), new Vec3(0, 0, 2), new Vec3(0, 2, 0), green),
        new Quad(new Vec3(-0.25, 1.985, -0.25), new Vec3(0.5, 0, 0), new Vec3(0, 0, 0.5), lamp),
        Box.fromCorners(new Vec3(-0.65, 0.0, -0.6), new Vec3(-0.05, 1.2, 0.0), white, rnd() * 2 * Math.PI),
        Box.fromCorners(new Vec3(0.05, 0.0, -0.05), new Vec3(0.65, 0.6, 0.55), white, rnd() * 2 * Math.PI),
        new Sphere(new Vec3(0.35, 0.9, 0.25), 0.3, Material.glass(1.5)));
    const camera = new Camera({ position: new Vec3(0, 1, 3.85), target: new Vec3(0, 1, 0), fov: 37 * Math.PI / 180 });
    return { scene, camera, width: 256, height: 256, maxBounces: 8 };
}

/**
 * 200 spheres (radius U[0.1, 0.4], resting on the floor) on a 20 × 20 m
 * checker quad with 1 m cells: 50% diffuse (albedo U[0.1, 0.9]³), 25% GGX
 * metal (roughness U[0, 0.5]), 15% glass (ior U[1.3, 1.7]), 10% lights
 * (radiance U[2, 10]); the default sky gradient and a sun of 0.5° across.
 */
export function spheres({ seed = 2, count = 200, width = 1920, height = 1080 } = {}) {
    const rnd = seededRandom(seed);
    const sky = new Sky({ sunStrength: 3.0, sunSize: 0.25 * Math.PI / 180, sunDirection: new Vec3(0.5, 0.8, 0.3) });
    const scene = new Scene({ sky });
    scene.add(new Quad(new Vec3(-10, 0, 10), new Vec3(20, 0, 0), new Vec3(0, 0, -20),
        Material.checkered(new Vec3(0.8, 0.8, 0.8), new Vec3(0.15, 0.15, 0.15), 1.0)));
    for (let i = 0; i < count; ++i) {
        const r = uniform(rnd, 0.1, 0.4);
        const x = uniform(rnd, -9.5, 9.5), z = uniform(rnd, -9.5, 9.5);
        const k = rnd();
        let m;
        if (k < 0.5)
            m = Material.diffuse(new Vec3(uniform(rnd, 0.1, 0.9), uniform(rnd, 0.1, 0.9), uniform(rnd, 0.1, 0.9)));
        else if (k < 0.75)
            m = Material.metal(new Vec3(uniform(rnd, 0.5, 1.0), uniform(rnd, 0.5, 1.0), uniform(rnd, 0.5, 1.0)), uniform(rnd, 0.0, 0.5));
        else if (k < 0.9)
            m = Material.glass(uniform(rnd, 1.3, 1.7));
        else
            m = Material.light(new Vec3(1, 1, 1), uniform(rnd, 2.0, 10.0));
        scene.add(new Sphere(new Vec3(x, r, z), r, m));
    }
    const camera = new Camera({ position: new Vec3(0, 3.2, 12.5), target: new Vec3(0, 0.3, 0), fov: 40 * Math.PI / 180 });
    return { scene, camera, width, height, maxBounces: 8 };
}

/** Value noise in [-1, 1], lattice values from mulberry32, quintic fade. */
export function valueNoise(seed) {
    const rnd = seededRandom(seed);
    const lattice = new Float64Array(256 * 256);
    for (let i = 0; i < lattice.length; ++i)
        lattice[i] = rnd() * 2.0 - 1.0;
    const at = (i, j) => lattice[(i & 255) * 256 + (j & 255)];
    const fade = (t) => t * t * t * (t * (t * 6.0 - 15.0) + 10.0);
    return (x, z) => {
        const i = Math.floor(x), j = Math.floor(z);
        const u = fade(x - i), v = fade(z - j);
        const a = at(i, j) + (at(i + 1, j) - at(i, j)) * u;
        const b = at(i, j + 1) + (at(i + 1, j + 1) - at(i, j + 1)) * u;
        return a + (b - a) * v;
    };
}

function fbm(noise, x, z, octaves) {
    let h = 0.0, amp = 1.0, freq = 0.15;
    for (let o = 0; o < octaves; ++o) {
        h += amp * noise(x * freq + o * 13.1, z * freq - o * 27.7);
        amp *= 0.5;
        freq *= 2.0;
    }
    return h;
}

/** A closed cubic Bézier loop through n random control points around centre. */
function bezierLoop(rnd, centre, size) {
    const pts = [];
    const n = 4;
    for (let i = 0; i < n * 3; ++i)
        pts.push(centre.add(new Vec3(uniform(rnd, -size, size), uniform(rnd, 0, size), uniform(rnd, -size, size))));
    return (t) => {
        const s = (t * n) % n, seg = Math.floor(s), u = s - seg;
        const p0 = pts[seg * 3], p1 = pts[seg * 3 + 1], p2 = pts[seg * 3 + 2], p3 = pts[(seg * 3 + 3) % pts.length];
        const a = (1 - u) * (1 - u) * (1 - u), b = 3 * u * (1 - u) * (1 - u), c = 3 * u * u * (1 - u), d = u * u * u;
        return p0.scale(a).add(p1.scale(b)).add(p2.scale(c)).add(p3.scale(d));
    };
}

/**
 * About 2.1M smooth triangles at the default size: a grid × grid heightfield
 * of 6-octave fBm (2·(grid-1)² triangles) plus seeded tori, lathe solids
 * and Bézier tubes, all GGX plastic (roughness U[0.05, 0.8]); a fog box
 * (σs 0.05/m, σa 0.005/m, g 0.7); a 2048 × 1024 environment of a gradient
 * and a Gaussian sun lobe peaking at 1e4.
 */
export function triangles({ seed = 3, grid = 1024, extras = 24, envWidth = 2048, envHeight = 1024, width = 1920, height = 1080 } = {}) {
    const rnd = seededRandom(seed);
    const plastic = () => Material.plastic(new Vec3(uniform(rnd, 0.2, 0.9), uniform(rnd, 0.2, 0.9), uniform(rnd, 0.2, 0.9)), uniform(rnd, 0.05, 0.8));
    const noise = valueNoise(seed + 100);
    const size = 40.0;
    const terrain = Mesh.surface((u, v) => {
        const x = (u - 0.5) * size, z = (v - 0.5) * size;
        return new Vec3(x, 1.5 * fbm(noise, x, z, 6), -z);
    }, grid - 1, grid - 1, plastic());

    const sunDir = new Vec3(0.3, 0.6, -0.5).normalized();
    const env = EnvironmentMap.procedural(envWidth, envHeight, (d) => {
        const t = Math.max(d.y, 0.0);
        const base = new Vec3(0.6, 0.7, 0.9).scale(1 - t).add(new Vec3(0.15, 0.3, 0.8).scale(t));
        const c = d.dot(sunDir);
        const lobe = 1e4 * Math.exp(-(1 - c) / 2e-5);
        return (d.y < 0 ? base.scale(0.3) : base).add(new Vec3(lobe, lobe * 0.95, lobe * 0.85));
    });
    const scene = new Scene({ sky: Sky.environment(env, 1.0, 0.0) });
    scene.add(terrain);
    for (let i = 0; i < extras; ++i) {
        const c = new Vec3(uniform(rnd, -15, 15), uniform(rnd, 1.5, 4), uniform(rnd, -15, 15));
        const kind = i % 3;
        if (kind == 0) {
            const R = uniform(rnd, 0.5, 2), r = uniform(rnd, 0.05, 0.4);
            scene.add(Mesh.torus(plastic(), { radius: R, tube: r, segments: 96, sides: 32 })
                .transformed(Mat4.translate(c.x, c.y, c.z).multiplyMat4(Mat4.rotateX(rnd() * 3))));
        }
        else if (kind == 1) {
            const prof = [new Vec2(0, 0)];
            for (let k = 1; k < 7; ++k) prof.push(new Vec2(uniform(rnd, 0.2, 1.2), k * 0.3));
            prof.push(new Vec2(0, 7 * 0.3));
            scene.add(Mesh.lathe(prof, plastic(), { segments: 96 }).transformed(Mat4.translate(c.x, c.y - 1, c.z)));
        }
        else {
            scene.add(Mesh.tube(bezierLoop(rnd, c, 2.0), plastic(), { radius: uniform(rnd, 0.05, 0.2), segments: 256, sides: 16 }));
        }
    }
    const sigmaS = 0.05, sigmaA = 0.005;
    scene.add(Box.fromCorners(new Vec3(-20, -2, -20), new Vec3(20, 6, 20),
        Material.volume(new Vec3(1, 1, 1).scale(sigmaS / (sigmaS + sigmaA)), sigmaS + sigmaA, 0.7)));
    const camera = new Camera({ position: new Vec3(0, 9, 22), target: new Vec3(0, 1, 0), fov: 45 * Math.PI / 180 });
    return { scene, camera, width, height, maxBounces: 12 };
}

/**
 * Caustics: a 10 × 10 m diffuse floor under a res × res-vertex water sheet
 * displaced by 16 seeded sines (amplitude ≤ 2 cm, wavelength U[0.1, 1] m),
 * a dispersive glass ball (Cauchy A 1.5, B 0.0042 µm²), a spherical lamp of
 * radius 5 cm and 100 W, and a sun.
 */
export function caustics({ seed = 4, res = 512, width = 1920, height = 1080 } = {}) {
    const rnd = seededRandom(seed);
    const waves = [];
    for (let i = 0; i < 16; ++i) {
        const lambda = uniform(rnd, 0.1, 1.0), angle = rnd() * 2 * Math.PI;
        waves.push({ k: 2 * Math.PI / lambda, cx: Math.cos(angle), cz: Math.sin(angle), a: uniform(rnd, 0.002, 0.02) / 4, phase: rnd() * 2 * Math.PI });
    }
    const water = Mesh.surface((u, v) => {
        const x = (u - 0.5) * 10, z = (v - 0.5) * 10;
        let h = 0.6;
        for (const w of waves) h += w.a * Math.sin(w.k * (w.cx * x + w.cz * z) + w.phase);
        return new Vec3(x, h, -z);
    }, res - 1, res - 1, Material.glass(1.33));
    const sky = new Sky({ sunStrength: 3.0, sunSize: 0.01, sunDirection: new Vec3(0.2, 0.9, 0.3) });
    const scene = new Scene({ sky });
    const B = 0.0042, A = 1.5;
    const ior = A + B / (0.5876 * 0.5876);
    const abbe = (ior - 1) / (B * (1 / (0.4861 * 0.4861) - 1 / (0.6563 * 0.6563)));
    const lampRadius = 0.05, power = 100.0;
    const radiance = power / (Math.PI * 4 * Math.PI * lampRadius * lampRadius);
    scene.add(
        new Quad(new Vec3(-5, 0, 5), new Vec3(10, 0, 0), new Vec3(0, 0, -10), Material.diffuse(new Vec3(0.8, 0.8, 0.8))),
        water,
        new Sphere(new Vec3(1.2, 1.4, 0.5), 0.5, Material.glass(ior, new Vec3(1, 1, 1), abbe)),
        new Sphere(new Vec3(-1.0, 3.0, 0.0), lampRadius, Material.light(new Vec3(1, 1, 1), radiance)));
    const camera = new Camera({ position: new Vec3(0, 4, 7), target: new Vec3(0, 0.5, 0), fov: 45 * Math.PI / 180 });
    return { scene, camera, width, height, maxBounces: 8 };
}

/** Random mixed objects (spheres, quads, turned boxes, tori), as a traversal stress scene. */
export function mixed({ seed = 11, count = 60 } = {}) {
    const rnd = seededRandom(seed);
    const white = Material.diffuse(new Vec3(0.8, 0.8, 0.8));
    const scene = new Scene();
    for (let i = 0; i < count; ++i) {
        const c = new Vec3(rnd() * 10 - 5, rnd() * 10 - 5, rnd() * 10 - 5);
        const kind = i % 4;
        if (kind == 0)
            scene.add(new Sphere(c, 0.2 + rnd(), white));
        else if (kind == 1)
            scene.add(new Quad(c, new Vec3(rnd() * 2, rnd(), 0), new Vec3(0, rnd(), rnd() * 2), white));
        else if (kind == 2)
            scene.add(Box.fromCorners(c, c.add(new Vec3(rnd() + 0.1, rnd() + 0.1, rnd() + 0.1)), white, rnd() * 3));
        else
            scene.add(Mesh.torus(white, { radius: 0.5, tube: 0.2, segments: 12, sides: 6 })
                .transformed(Mat4.translate(c.x, c.y, c.z).multiplyMat4(Mat4.rotateX(rnd() * 3))));
    }
    const camera = new Camera({ position: new Vec3(0, 0, 16), target: new Vec3(0, 0, 0), fov: 50 * Math.PI / 180 });
    return { scene, camera, width: 128, height: 128, maxBounces: 8 };
}

import { fixture } from "./fixtures.mjs";



const vec = (a) => new Vec3(a[0], a[1], a[2]);

function materialFrom(m = {}) {
    const o = Object.assign({}, m);
    for (const k of ["color", "emission", "checker"])
        if (Array.isArray(o[k])) o[k] = vec(o[k]);
    return new Material(o);
}

const mat4From = (t) => (t ? new Mat4(...t) : Mat4.identity());

/**
 * A scene from a JSON description, for tests: { objects: [...], sky, camera,
 * width, height, maxBounces }. Objects: {type: "sphere", center, radius},
 * {type: "quad", corner, u, v}, {type: "box", transform (16, row-major)},
 * {type: "boxCorners", min, max, rotateY}, {type: "stick", a, b, thickness},
 * {type: "icosphere", subdivisions, transform}, {type: "torus", radius, tube,
 * segments, sides, transform}; each with a material of Material's options.
 */
export function custom({ objects = [], sky = null, camera = {}, width = 64, height = 64, maxBounces = 8 } = {}) {
    const skyOpts = {};
    if (sky) for (const [k, v] of Object.entries(sky)) skyOpts[k] = Array.isArray(v) ? vec(v) : v;
    const scene = new Scene({ sky: sky === "black" ? Sky.black() : new Sky(skyOpts) });
    for (const o of objects) {
        const m = materialFrom(o.material);
        if (o.type == "sphere") scene.add(new Sphere(vec(o.center), o.radius, m, { rotation: mat4From(o.rotation) }));
        else if (o.type == "quad") scene.add(new Quad(vec(o.corner), vec(o.u), vec(o.v), m));
        else if (o.type == "box") scene.add(new Box(mat4From(o.transform), m));
        else if (o.type == "boxCorners") scene.add(Box.fromCorners(vec(o.min), vec(o.max), m, o.rotateY || 0));
        else if (o.type == "stick") scene.add(Box.between(vec(o.a), vec(o.b), o.thickness, m));
        else if (o.type == "icosphere") scene.add(Mesh.icosphere(m, o.subdivisions || 3).transformed(mat4From(o.transform)));
        else if (o.type == "torus") scene.add(Mesh.torus(m, o).transformed(mat4From(o.transform)));
        else throw new RangeError(`Unknown object type "${o.type}".`);
    }
    const cam = new Camera({
        position: vec(camera.position || [0, 0, 5]), target: vec(camera.target || [0, 0, 0]),
        fov: camera.fov || 40 * Math.PI / 180, aperture: camera.aperture || 0, focusDistance: camera.focusDistance || 0,
    });
    return { scene, camera: cam, width, height, maxBounces };
}

export const CONFIGS = Object.freeze({ cornell, spheres, triangles, caustics, mixed, fixture, custom });
