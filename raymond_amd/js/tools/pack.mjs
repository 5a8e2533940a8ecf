// Packs a workload scene exactly as Renderer would upload it and writes the
// buffers to one file, for harnesses outside Node (the Python tests and
// bench read it with raymond_amd.scenefile).
//
//   node pack.mjs <config> <out.rmpk> [json options]
//
// options: { "scene": {…generator options}, "maxBounces": n, "samplesPerFrame": n,
//            "fireflyClamp": x, "view": "color", "denoise": true, "caustics": true,
//            "width": w, "height": h }
import { writeFileSync } from "fs";
import { CONFIGS } from "../configs.mjs";
import { packForDevice } from "../renderer.mjs";

const [, , name, out, json] = process.argv;
if (!name || !out || !CONFIGS[name]) {
    process.stderr.write(`usage: node pack.mjs <${Object.keys(CONFIGS).join("|")}> <out.rmpk> [json options]\n`);
    process.exit(2);
}
const opts = json ? JSON.parse(json) : {};
const t0 = Date.now();
const made = CONFIGS[name](opts.scene || {});
if (opts.camera) Object.assign(made.camera, opts.camera);
const width = opts.width || made.width, height = opts.height || made.height;
const state = {
    maxBounces: opts.maxBounces || made.maxBounces,
    samplesPerFrame: opts.samplesPerFrame || 1,
    exposure: 1.0,
    fireflyClamp: opts.fireflyClamp === undefined ? 0.0 : opts.fireflyClamp,
    view: opts.view || "color",
    denoise: opts.denoise === undefined ? true : opts.denoise,
    caustics: opts.caustics === undefined ? true : opts.caustics,
    samples: 0,
};
const packed = packForDevice(made.scene, made.camera, width, height, state);
const tPack = (Date.now() - t0) / 1000;

// Container: "RMPACK01", u32 JSON length, JSON, then 16-byte-aligned sections.
const sections = [];
const add = (sname, dtype, arr) => { if (arr) sections.push({ name: sname, dtype, bytes: new Uint8Array(arr.buffer, arr.byteOffset, arr.byteLength) }); };
add("primitives", "u32", new Uint32Array(packed.primitives));
add("nodes", "u32", new Uint32Array(packed.nodes));
add("materials", "u32", new Uint32Array(packed.materials));
add("attributes", "f32", packed.attributes);
add("lights", "u32", packed.tables.table);
add("emitters", "f32", packed.emitters);
add("params", "u8", new Uint8Array(packed.params));
add("env_rgba", "u16", packed.environment.rgba);
add("env_table", "f32", packed.environment.table);
if (packed.textures) add("textures", "u8", packed.textures.data);
// Which scene object each BVH primitive came from (object order of scene.add).
{
    const objects = made.scene.objects;
    const index = new Map(objects.map((o, i) => [o, i]));
    const owners = made.scene.pack().objects;
    add("object_of", "u32", Uint32Array.from(owners, (o) => index.get(o)));
}
if (packed.normalMaps) add("normal_maps", "u8", packed.normalMaps.data);

const meta = {
    config: name, options: opts, width, height,
    primitiveCount: packed.primitiveCount, primitiveSlots: packed.primitives.byteLength / 64,
    nodeCount: packed.nodeCount, materialCount: packed.materialCount,
    lightCount: packed.tables.lightCount, volumeCount: packed.tables.volumeCount,
    textured: packed.tables.textured, dispersive: packed.tables.dispersive,
    envWidth: packed.environment.width, envHeight: packed.environment.height,
    envTableWidth: packed.environment.tableWidth, envTableHeight: packed.environment.tableHeight,
    causticsActive: packed.causticsActive, packSeconds: tPack,
    textures: packed.textures ? [packed.textures.width, packed.textures.height, packed.textures.layers] : null,
    normalMaps: packed.normalMaps ? [packed.normalMaps.width, packed.normalMaps.height, packed.normalMaps.layers] : null,
    sections: [],
};
let offset = 0;
for (const s of sections) {
    meta.sections.push({ name: s.name, dtype: s.dtype, offset, bytes: s.bytes.byteLength });
    offset += Math.ceil(s.bytes.byteLength / 16) * 16;
}
const header = Buffer.from(JSON.stringify(meta), "utf8");
const pre = 8 + 4 + header.length;
const base = Math.ceil(pre / 16) * 16;
const file = Buffer.alloc(base + offset);
file.write("RMPACK01", 0, "latin1");
file.writeUInt32LE(header.length, 8);
header.copy(file, 12);
for (let i = 0; i < sections.length; ++i)
    Buffer.from(sections[i].bytes.buffer, sections[i].bytes.byteOffset, sections[i].bytes.byteLength).copy(file, base + meta.sections[i].offset);
writeFileSync(out, file);
process.stdout.write(JSON.stringify({ out, primitives: meta.primitiveCount, nodes: meta.nodeCount, seconds: tPack }) + "\n");
