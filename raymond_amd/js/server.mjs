// RenderServer: many render jobs on one MI355X through the library's job
// server (rm_server_* in include/raymond.h). A job is what a Renderer would
// render for a scene and camera, run until a target number of samples per
// pixel; jobs submitted while others run are admitted as slots and memory
// free up, and their frames overlap on the device. Scenes are packed and
// uploaded once and stay resident while the server is open.
import { OPTION_PRECISION, precisionFlags } from "./addon.mjs";
import { loadAddon, packForDevice } from "./renderer.mjs";

export class RenderServer {

    #addon; #server; #ctx;
    #scenes = new Map();   // Scene → { version, handle }
    #sizes = new Map();    // job id → { width, height, samples, denoise }
    #closed = false;

    /**
     * Opens a server on `device` with `slots` concurrent jobs within
     * `memoryBudget` bytes; `precision` ({ quantizedBvh, fp16Guides,
     * compensated }) applies to every job.
     */
    static async create({ device = 0, slots = 8, memoryBudget = 200 * 2 ** 30, precision = {} } = {}) {
        const addon = loadAddon();
        const server = addon.serverOpen(device, slots, memoryBudget);
        try {
            addon.serverSetOption(server, OPTION_PRECISION, precisionFlags(precision));
        }
        catch (error) {
            addon.serverClose(server);
            throw error;
        }
        return new RenderServer(addon, server);
    }

    constructor(addon, server) {
        this.#addon = addon;
        this.#server = server;
        this.#ctx = addon.serverContext(server);
    }

    /**
     * Queues a job and returns its id: `scene` seen through `camera` at
     * width × height, until `targetSamples` samples per pixel, with the
     * Renderer's settings (maxBounces, samplesPerFrame, exposure,
     * fireflyClamp, denoise, caustics). With `deadlineMs`, a job that has
     * not queued its last frame that many ms after submission ends, and
     * wait() rejects, as if cancelled then.
     */
    submit({ scene, camera, width, height, targetSamples, maxBounces = 8, samplesPerFrame = 1, exposure = 1.0,
        fireflyClamp = 20.0, denoise = true, caustics = true, deadlineMs = 0 }) {
        if (this.#closed)
            throw new Error("The server is closed.");
        if (!(targetSamples >= 1))
            throw new RangeError("A job needs a target of at least one sample per pixel.");
        if (!(deadlineMs >= 0))
            throw new RangeError("A deadline is a number of milliseconds, 0 for none.");
        const state = { maxBounces, samplesPerFrame, exposure, fireflyClamp, view: "color", denoise, caustics, samples: 0 };
        const packed = packForDevice(scene, camera, width, height, state);
        const handle = upload(this.#addon, this.#ctx, this.#scenes, scene, packed);
        const env = packed.environment;
        const hasMap = env.tableHeight > 0;
        const id = this.#addon.serverSubmit(this.#server, handle, packed.params, targetSamples,
            packed.causticsActive ? packed.emitters : null, hasMap ? env.rgba : null, env.width, env.height,
            hasMap ? env.table : null, env.tableWidth, env.tableHeight, deadlineMs);
        const frames = Math.ceil(targetSamples / samplesPerFrame);
        this.#sizes.set(id, { width, height, samples: frames * samplesPerFrame, denoise });
        return id;
    }

    /**
     * Resolves once job `id` is done with { pixels, stats }: pixels as
     * readPixels gives them (the denoised image when the job denoises, else
     * the average), stats its queued, first-frame and done times in ms.
     */
    async wait(id) {
        const job = this.#sizes.get(id);
        if (!job)
            throw new Error(`No job ${id} is waiting.`);
        this.#sizes.delete(id);
        const pixels = new Float32Array(job.width * job.height * 4);
        const stats = await this.#addon.serverWait(this.#server, id, pixels);
        if (!job.denoise)
            for (let i = 0; i < pixels.length; ++i)
                pixels[i] /= job.samples;
        return { pixels, stats };
    }

    /**
     * Cancels job `id`: a queued job leaves the queue, a running one stops
     * after the frames it has in flight; wait(id) then rejects with an
     * Error saying it was cancelled. A job whose last frame is queued
     * finishes with its image.
     */
    cancel(id) {
        if (!this.#sizes.has(id))
            throw new Error(`No job ${id} is waiting.`);
        this.#addon.serverCancel(this.#server, id);
    }

    /** Waits for every job, then frees the server and its scenes. */
    close() {
        if (this.#closed)
            return;
        this.#closed = true;
        this.#addon.serverClose(this.#server);
    }

}

/** The scene's device handle on the server's context, uploading it when new or changed. */
function upload(addon, ctx, scenes, scene, packed) {
    const known = scenes.get(scene);
    if (known && known.version == scene.version)
        return known.handle;
    if (known)
        addon.releaseScene(ctx, known.handle);
    const handle = addon.uploadScene(ctx, packed.primitives, packed.primitiveCount, packed.nodes, packed.nodeCount,
        packed.materials, packed.materialCount, packed.attributes, packed.tables.table);
    if (packed.textures)
        addon.uploadTextures(ctx, handle, false, packed.textures.data, packed.textures.width, packed.textures.height,
            packed.textures.layers);
    if (packed.normalMaps)
        addon.uploadTextures(ctx, handle, true, packed.normalMaps.data, packed.normalMaps.width,
            packed.normalMaps.height, packed.normalMaps.layers);
    scenes.set(scene, { version: scene.version, handle });
    return handle;
}
