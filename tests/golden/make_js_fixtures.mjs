// Writes tests/golden/js_generators.json: SHA-256 digests of the exact bytes
// that the host's Texture and Mesh generators produce for a fixed set of
// inputs (pixels, sample() values, vertex positions, indices, normals, uvs),
// and of the packed scene buffers of the workloads built from them. The
// digests pin these producers: tests/test_js_host.py runs this with --check,
// which recomputes them and must match, so a rewrite keeps every byte.
//
//   node tests/golden/make_js_fixtures.mjs            (writes the fixture)
//   node tests/golden/make_js_fixtures.mjs --check    (compares, exit 1 on a difference)
import { createHash } from "crypto";
import { readFileSync, writeFileSync } from "fs";
import { dirname, join } from "path";
import { fileURLToPath } from "url";
import { Mesh } from "../../raymond_amd/js/geometry.mjs";
import { Texture } from "../../raymond_amd/js/texture.mjs";
import { Material } from "../../raymond_amd/js/material.mjs";
import { Mat4, Vec2, Vec3 } from "../../raymond_amd/js/math.mjs";
import { CONFIGS } from "../../raymond_amd/js/configs.mjs";
import { packForDevice } from "../../raymond_amd/js/renderer.mjs";

const here = dirname(fileURLToPath(import.meta.url));
const OUT = join(here, "js_generators.json");

function digest(...arrays) {
    const h = createHash("sha256");
    for (const a of arrays)
        h.update(Buffer.from(a.buffer, a.byteOffset, a.byteLength));
    return h.digest("hex");
}

const f64 = (vs, k) => Float64Array.from(vs.flatMap((v) => (k == 3 ? [v.x, v.y, v.z] : [v.x, v.y])));

function meshDigest(m) {
    return {
        triangles: m.triangleCount,
        vertices: m.positions.length,
        sha256: digest(f64(m.positions, 3), Uint32Array.from(m.indices), m.normals ? f64(m.normals, 3) : new Float64Array(0),
            m.uvs ? f64(m.uvs, 2) : new Float64Array(0)),
    };
}

function cases() {
    const out = {};
    const white = Material.diffuse(new Vec3(0.8, 0.8, 0.8));
    // Textures.
    const gradient = Texture.procedural(64, 48, (u, v) => new Vec3(u, v * v, 0.5 + 0.5 * Math.sin(u * 6.28)));
    const wide = Texture.procedural(37, 5, (u, v) => new Vec3(2 * u - 0.3, Math.exp(-4 * v), 0.002 * u));
    const bumps = Texture.normalMap(64, 64, (u, v) => 0.03 * Math.sin(u * Math.PI * 4) * Math.cos(v * Math.PI * 6));
    const ridge = Texture.normalMap(20, 12, (u, v) => 0.2 * Math.abs(u - 0.4) + 0.05 * v * v);
    out["texture.procedural 64x48"] = digest(gradient.data);
    out["texture.procedural 37x5"] = digest(wide.data);
    out["texture.normalMap 64x64"] = digest(bumps.data);
    out["texture.normalMap 20x12"] = digest(ridge.data);
    const points = [];
    for (let k = 0; k < 500; ++k)
        points.push([Math.sin(k * 12.9898) * 3.7, Math.cos(k * 78.233) * 2.3 - 0.5]);
    for (const [name, t] of [["gradient", gradient], ["wide", wide], ["bumps", bumps]])
        for (const srgb of [true, false]) {
            const vals = new Float64Array(points.length * 3);
            points.forEach(([u, v], i) => {
                const c = t.sample(u, v, srgb);
                vals.set([c.x, c.y, c.z], i * 3);
            });
            out[`texture.sample ${name} srgb=${srgb}`] = digest(vals);
        }
    out["texture.resized gradient 40x24 srgb"] = digest(gradient.resized(40, 24).data);
    out["texture.resized gradient 100x70 srgb"] = digest(gradient.resized(100, 70).data);
    out["texture.resized bumps 32x80 plain"] = digest(bumps.resized(32, 80, false).data);
    out["texture.resized same size"] = digest(gradient.resized(64, 48).data);
    // Meshes.
    const wave = (u, v) => new Vec3(u * 4 - 2, 0.1 * Math.sin(9 * u) * Math.cos(7 * v), -(v * 3 - 1.5));
    out["mesh.surface open smooth"] = meshDigest(Mesh.surface(wave, 17, 11, white));
    out["mesh.surface open flat"] = meshDigest(Mesh.surface(wave, 9, 6, white, { shading: "flat" }));
    const cyl = (u, v) => new Vec3(Math.cos(u * 2 * Math.PI), v * 2, -Math.sin(u * 2 * Math.PI));
    out["mesh.surface closed u"] = meshDigest(Mesh.surface(cyl, 24, 5, white, { uWrap: "closed" }));
    out["mesh.surface closed v"] = meshDigest(Mesh.surface((u, v) => cyl(v, u), 5, 24, white, { vWrap: "closed" }));
    out["mesh.surface degenerate rows"] = meshDigest(Mesh.surface((u, v) => new Vec3(u * v, v, u * (1 - v)), 8, 8, white));
    for (const s of [0, 1, 2, 3])
        out[`mesh.icosphere ${s}`] = meshDigest(Mesh.icosphere(white, s));
    out["mesh.icosphere 2 flat"] = meshDigest(Mesh.icosphere(white, 2, { shading: "flat" }));
    const profile = [new Vec2(0, 0), new Vec2(0.5, 0.2), new Vec2(0.8, 0.6), new Vec2(0.4, 1.1), new Vec2(0, 1.5)];
    out["mesh.lathe smooth"] = meshDigest(Mesh.lathe(profile, white, { segments: 40 }));
    out["mesh.lathe flat"] = meshDigest(Mesh.lathe(profile, white, { segments: 12, shading: "flat" }));
    out["mesh.torus"] = meshDigest(Mesh.torus(white, { radius: 1.3, tube: 0.27, segments: 48, sides: 20 }));
    out["mesh.torus default"] = meshDigest(Mesh.torus(white));
    const knot = (t) => {
        const a = t * 2 * Math.PI;
        return new Vec3(Math.sin(a) + 2 * Math.sin(2 * a), Math.cos(a) - 2 * Math.cos(2 * a), -Math.sin(3 * a));
    };
    out["mesh.tube knot"] = meshDigest(Mesh.tube(knot, white, { radius: 0.2, segments: 96, sides: 12 }));
    out["mesh.tube circle"] = meshDigest(Mesh.tube((t) => new Vec3(Math.cos(t * 2 * Math.PI), 0, Math.sin(t * 2 * Math.PI)), white));
    const turned = Mat4.translate(0.6, 0.35, 1.0).multiplyMat4(Mat4.rotateX(1.1));
    out["mesh.transformed torus"] = meshDigest(Mesh.torus(white, { segments: 20, sides: 8 }).transformed(turned));
    out["mesh.transformed mirrored"] = meshDigest(Mesh.icosphere(white, 1).transformed(Mat4.scale(-1, 2, 1)));
    // Packed scenes built from them (buffers exactly as uploaded).
    const scenes = {
        "triangles grid 64": ["triangles", { grid: 64, extras: 6, envWidth: 64, envHeight: 32, width: 96, height: 54 }],
        "caustics res 32": ["caustics", { res: 32, width: 96, height: 54 }],
        "mixed": ["mixed", {}],
        "fixture textured": ["fixture", { name: "textured", size: 32 }],
        "fixture bumpy": ["fixture", { name: "bumpy", size: 32 }],
        "fixture patternedLamps": ["fixture", { name: "patternedLamps", size: 32 }],
        "fixture lampInFog mesh": ["fixture", { name: "lampInFog", arg: "mesh", size: 32 }],
    };
    for (const [name, [config, opts]] of Object.entries(scenes)) {
        const made = CONFIGS[config](opts);
        const state = { maxBounces: made.maxBounces, samplesPerFrame: 1, exposure: 1, fireflyClamp: 0, view: "color",
            denoise: true, caustics: true, samples: 0 };
        const p = packForDevice(made.scene, made.camera, made.width, made.height, state);
        const parts = [new Uint32Array(p.primitives), new Uint32Array(p.nodes), new Uint32Array(p.materials), p.attributes,
            p.tables.table, new Uint8Array(p.params)];
        if (p.textures) parts.push(p.textures.data);
        if (p.normalMaps) parts.push(p.normalMaps.data);
        out[`packed ${name}`] = digest(...parts);
    }
    return out;
}

const now = cases();
if (process.argv.includes("--check")) {
    const want = JSON.parse(readFileSync(OUT, "utf8"));
    const bad = Object.keys(want).filter((k) => JSON.stringify(want[k]) != JSON.stringify(now[k]));
    const extra = Object.keys(now).filter((k) => !(k in want));
    if (bad.length || extra.length) {
        console.log(JSON.stringify({ differ: bad, unpinned: extra }));
        process.exit(1);
    }
    console.log(`all ${Object.keys(want).length} generator outputs match`);
}
else {
    writeFileSync(OUT, JSON.stringify(now, null, 1) + "\n");
    console.log(`wrote ${OUT}: ${Object.keys(now).length} cases`);
}
