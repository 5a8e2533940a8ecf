// Renderer: the reference's progressive path tracer API (src/Renderer.mjs:84-547)
// on one MI355X. Scene packing, the lights table, emitters and the parameter
// block are computed here in JS exactly as the reference's Renderer does;
// the frame itself (trace, light tracing, photon gather, denoise) runs as HIP
// kernels in libraymond_hip.so, reached through the N-API addon raymond.node.
// There is no other backend: without the addon or a GPU, create() rejects.
import { OPTION_LIGHT_PATHS, OPTION_PRECISION, OPTION_SCHEDULE, OPTION_WAVES, SCHEDULES, loadAddon, precisionFlags }
    from "./addon.mjs";
import {
    IMAGE_KEY_BYTES, PARAMS_BYTES, VIEWS, deltaObjects, emitters, environmentData,
    sceneTables, textureLayers, writeFrameParams, writeParams,
} from "./frame.mjs";

/** Buffers the reader reads back: accumulation, guides, denoised result. */
export const BUFFERS = Object.freeze({ accum: 0, albedo: 1, normal: 2, denoised: 3 });

export { loadAddon };

/**
 * Everything a frame needs from the host, computed without a device: the
 * packed scene, lights table, emitters, parameter block and environment.
 */
export function packForDevice(scene, camera, width, height, state) {
    const packed = scene.pack();
    const tables = sceneTables(packed);
    const delta = deltaObjects(scene);
    const emit = emitters(scene.sky, tables.powers, delta);
    const causticsActive = !!(state.caustics && camera.aperture == 0 && state.view == "color" &&
        tables.volumeCount == 0 && emit);
    const params = new ArrayBuffer(PARAMS_BYTES);
    const full = Object.assign({}, state, {
        lightCount: tables.lightCount, volumeCount: tables.volumeCount,
        caustics: causticsActive, denoise: state.denoise && state.view == "color", delta,
    });
    writeParams(params, scene, camera, width, height, full);
    writeFrameParams(params, camera, height, full);
    return {
        primitives: packed.primitives, primitiveCount: packed.primitiveCount,
        nodes: packed.nodes, nodeCount: packed.nodeCount,
        materials: packed.materials, materialCount: packed.materialCount,
        attributes: packed.attributes, tables, delta, emitters: emit, params,
        environment: environmentData(scene.sky.environment), causticsActive,
        textures: textureLayers(packed.textures, true), normalMaps: textureLayers(packed.normalMaps, false),
    };
}

export class Renderer {

    #addon; #ctx; #canvas;
    #scene = null; #camera = null;
    #maxBounces = 8; #samplesPerFrame = 1; #exposure = 1.0; #fireflyClamp = 20.0;
    #view = "color"; #denoise = true; #caustics = true; #precision = {}; #schedule = "split"; #waves = 4;
    #causticsActive = false;
    #samples = 0;
    #params = new ArrayBuffer(PARAMS_BYTES);
    #imageKey = new Uint8Array(IMAGE_KEY_BYTES);
    #uploadedScene = null; #uploadedVersion = -1; #sceneHandle = null;
    #tables = null; #delta = null;
    #uploadedEnvironment = undefined;
    #width = 0; #height = 0;
    #lightPaths = 0;
    #queue;
    #disposed = false;

    constructor(addon, ctx, canvas) {
        this.#addon = addon;
        this.#ctx = ctx;
        this.#canvas = canvas;
        const self = this;
        this.#queue = { onSubmittedWorkDone: () => self.#addon.synchronize(self.#ctx) };
    }

    /**
     * Opens the GPU for a canvas-like { width, height } object. Rejects with
     * an Error saying why if the addon or the device can't be opened.
     */
    static async create(canvas, { powerPreference = "high-performance", device = 0 } = {}) {
        void powerPreference;
        if (!canvas || !(canvas.width >= 0) || !(canvas.height >= 0))
            throw new Error("A renderer needs a canvas with a width and height.");
        const addon = loadAddon();
        const ctx = addon.open(device);
        return new Renderer(addon, ctx, canvas);
    }

    /** The device stand-in: device.queue.onSubmittedWorkDone() resolves when the GPU is idle. */
    get device() { return { queue: this.#queue }; }
    get canvas() { return this.#canvas; }
    get scene() { return this.#scene; }
    set scene(v) { this.#scene = v; }
    get camera() { return this.#camera; }
    set camera(v) { this.#camera = v; }
    get maxBounces() { return this.#maxBounces; }
    set maxBounces(v) { this.#maxBounces = Math.max(1, Math.min(64, Math.floor(v))); }
    get samplesPerFrame() { return this.#samplesPerFrame; }
    set samplesPerFrame(v) { this.#samplesPerFrame = Math.max(1, Math.min(64, Math.floor(v))); }
    get exposure() { return this.#exposure; }
    set exposure(v) { this.#exposure = Math.max(v, 0.0); }
    get fireflyClamp() { return this.#fireflyClamp; }
    set fireflyClamp(v) { this.#fireflyClamp = Math.max(v, 0.0); }
    get view() { return this.#view; }
    set view(v) {
        if (!Object.prototype.hasOwnProperty.call(VIEWS, v))
            throw new RangeError(`Unknown view "${v}".`);
        this.#view = v;
    }
    get denoise() { return this.#denoise; }
    set denoise(v) { this.#denoise = v; }
    get caustics() { return this.#caustics; }
    set caustics(v) { this.#caustics = v; }
    /**
     * Precision variants for the frames from now on, none by default:
     * { quantizedBvh, fp16Guides, compensated } (RM_OPTION_PRECISION).
     */
    /**
     * How the path-tracing pass maps paths to lanes: "per-pixel" (default),
     * "compact", "persistent" or "wavefront" — the same image, bit for bit.
     */
    get schedule() { return this.#schedule; }
    set schedule(v) {
        if (!Object.prototype.hasOwnProperty.call(SCHEDULES, v))
            throw new RangeError(`schedule is one of ${Object.keys(SCHEDULES).join(", ")}.`);
        this.#addon.setOption(this.#ctx, OPTION_SCHEDULE, SCHEDULES[v]);
        this.#schedule = v;
    }
    /**
     * Waves per SIMD the "per-pixel" path pass is compiled for: 4 (default)
     * or 5, the same image bit for bit. Five is faster on scenes whose frame
     * the path pass dominates (large meshes, glass).
     */
    get waves() { return this.#waves; }
    set waves(v) {
        if (v !== 4 && v !== 5)
            throw new RangeError("waves is 4 or 5.");
        this.#addon.setOption(this.#ctx, OPTION_WAVES, v);
        this.#waves = v;
    }
    /**
     * Light paths (and photon slots) a frame when caustics are on: 0 (the
     * default) for the reference's one a sample, at most 524,288; more
     * traces more photons a frame, and each carries less power.
     */
    get lightPaths() { return this.#lightPaths; }
    set lightPaths(v) {
        if (!(Number.isInteger(v) && v >= 0 && v <= 2 ** 28))
            throw new RangeError("lightPaths is a whole number from 0 to 2^28.");
        this.#addon.setOption(this.#ctx, OPTION_LIGHT_PATHS, v);
        this.#lightPaths = v;
        this.#samples = 0;
    }
    get precision() { return Object.assign({}, this.#precision); }
    set precision(v) {
        this.#addon.setOption(this.#ctx, OPTION_PRECISION, precisionFlags(v));
        this.#precision = Object.assign({}, v);
    }
    get samples() { return this.#samples; }

    reset() { this.#samples = 0; }

    /** Adds samplesPerFrame paths per pixel; does nothing without a scene and camera. */
    render() {
        const scene = this.#scene, camera = this.#camera;
        if (!scene || !camera)
            return;
        const width = Math.max(1, this.#canvas.width);
        const height = Math.max(1, this.#canvas.height);
        const addon = this.#addon, ctx = this.#ctx;

        if (scene != this.#uploadedScene || scene.version != this.#uploadedVersion) {
            const packed = scene.pack();
            this.#tables = sceneTables(packed);
            this.#delta = deltaObjects(scene);
            if (this.#sceneHandle)
                addon.releaseScene(ctx, this.#sceneHandle);
            this.#sceneHandle = addon.uploadScene(ctx, packed.primitives, packed.primitiveCount,
                packed.nodes, packed.nodeCount, packed.materials, packed.materialCount,
                packed.attributes, this.#tables.table);
            for (const [layers, normal] of [[textureLayers(packed.textures, true), false], [textureLayers(packed.normalMaps, false), true]])
                if (layers)
                    addon.uploadTextures(ctx, this.#sceneHandle, normal, layers.data, layers.width, layers.height, layers.layers);
            this.#uploadedScene = scene;
            this.#uploadedVersion = scene.version;
            this.#samples = 0;
        }

        if (scene.sky.environment !== this.#uploadedEnvironment) {
            const env = environmentData(scene.sky.environment);
            addon.uploadEnvironment(ctx, env.rgba, env.width, env.height, env.table, env.tableWidth, env.tableHeight);
            this.#uploadedEnvironment = scene.sky.environment;
            this.#samples = 0;
        }

        if (width != this.#width || height != this.#height) {
            addon.resize(ctx, width, height);
            this.#width = width;
            this.#height = height;
            this.#samples = 0;
        }

        const emit = emitters(scene.sky, this.#tables.powers, this.#delta);
        const caustics = !!(this.#caustics && camera.aperture == 0 && this.#view == "color" &&
            this.#tables.volumeCount == 0 && emit);
        if (caustics != this.#causticsActive) {
            this.#causticsActive = caustics;
            this.#samples = 0;
        }

        const state = {
            lightCount: this.#tables.lightCount, volumeCount: this.#tables.volumeCount,
            maxBounces: this.#maxBounces, samplesPerFrame: this.#samplesPerFrame, exposure: this.#exposure,
            fireflyClamp: this.#fireflyClamp, view: this.#view, denoise: denoising(this),
            caustics, delta: this.#delta, samples: 0,
        };
        writeParams(this.#params, scene, camera, width, height, state);
        const key = new Uint8Array(this.#params, 0, IMAGE_KEY_BYTES);
        let same = true;
        for (let i = 0; i < IMAGE_KEY_BYTES; ++i)
            if (key[i] != this.#imageKey[i]) { same = false; break; }
        if (!same) {
            this.#imageKey.set(key);
            this.#samples = 0;
        }
        state.samples = this.#samples;
        writeFrameParams(this.#params, camera, height, state);

        addon.frame(ctx, this.#sceneHandle, this.#params, caustics ? emit : null);
        this.#samples += this.#samplesPerFrame;
    }

    /**
     * Renders a few whole frames under each candidate schedule and occupancy
     * (rm_tune) and keeps the fastest for the frames from now on; the image
     * starts again. Blocks until the frames are done. Returns { schedule,
     * waves, persistentBlocks, stackDepth, blockShape, splitDepth, ms,
     * candidates: [the same for each] }. Not in the reference.
     */
    tune({ frames = 2 } = {}) {
        if (!this.#scene || !this.#camera)
            throw new Error("tune() needs a scene and a camera.");
        this.render();
        const emit = emitters(this.#scene.sky, this.#tables.powers, this.#delta);
        const result = this.#addon.tune(this.#ctx, this.#sceneHandle, this.#params,
            this.#causticsActive ? emit : null, frames);
        const names = Object.keys(SCHEDULES);
        const candidates = result.candidates.map((c) => ({
            schedule: names.find((k) => SCHEDULES[k] == c.schedule), waves: c.waves,
            persistentBlocks: c.persistentBlocks, stackDepth: c.stackDepth,
            blockShape: c.blockShape == 1 ? "rows" : "squares", splitDepth: c.splitDepth, ms: c.ms,
        }));
        const best = candidates[result.best];
        this.#schedule = best.schedule;
        this.#waves = best.waves;
        this.#samples = 0;
        return Object.assign({ candidates }, best);
    }

    /**
     * The average so far, 4 floats per pixel from the top left, in linear
     * light; with denoised, the denoised colour (needs denoise on and the
     * color view).
     */
    async readPixels({ denoised = false } = {}) {
        if (this.#width == 0 || this.#samples == 0)
            throw new Error("Nothing has been rendered yet.");
        if (denoised && !denoising(this))
            throw new Error("Only the color view is denoised, with denoise on.");
        const pixels = new Float32Array(this.#width * this.#height * 4);
        const samples = this.#samples;
        await this.#addon.read(this.#ctx, denoised ? BUFFERS.denoised : BUFFERS.accum, pixels);
        if (!denoised)
            for (let i = 0; i < pixels.length; ++i)
                pixels[i] /= samples;
        return pixels;
    }

    /** Frees the device memory; the renderer can't be used afterwards. */
    /**
     * What render() draws to the canvas for the last frame: the display
     * pass's tone-mapped sRGB image as W·H RGBA8 bytes (src/shaders.mjs:2100-2152).
     */
    async readDisplay() {
        if (this.#width == 0 || this.#samples == 0)
            throw new Error("Nothing has been rendered yet.");
        const pixels = new Uint8Array(this.#width * this.#height * 4);
        await this.#addon.display(this.#ctx, pixels);
        return pixels;
    }

    dispose() {
        if (this.#disposed)
            return;
        this.#disposed = true;
        if (this.#sceneHandle)
            this.#addon.releaseScene(this.#ctx, this.#sceneHandle);
        this.#sceneHandle = null;
        this.#addon.close(this.#ctx);
    }
}

function denoising(renderer) {
    return renderer.denoise && renderer.view == "color";
}
