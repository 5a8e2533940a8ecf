// Errors from the library reach JavaScript as Errors carrying its sentence:
// running out of device memory (through the addon and through
// Renderer.render), an argument the library refuses, and a cancelled served
// job, whose wait() rejects. Run on the GPU by tests/test_gpu_failures.py.
import { loadAddon } from "../../raymond_amd/js/addon.mjs";
import { Renderer, RenderServer } from "../../raymond_amd/js/index.mjs";
import { cornell, spheres } from "../../raymond_amd/js/configs.mjs";

function expectError(what, fn, pattern) {
    try {
        fn();
    }
    catch (e) {
        if (!(e instanceof Error) || !pattern.test(e.message))
            throw new Error(`${what}: threw "${e.message}", which doesn't match ${pattern}`);
        console.log(`${what}: ${e.message}`);
        return;
    }
    throw new Error(`${what}: didn't throw`);
}

async function main() {
    const addon = loadAddon();
    // Two 16383² images hold more than the GPU's memory: the second fails.
    const big = addon.open(0);
    const other = addon.open(0);
    addon.resize(big, 16383, 16383);
    expectError("second 16383² image", () => addon.resize(other, 16383, 16383), /memory.*\(code -3\)/);
    addon.close(other);
    // So does a Renderer asked for one while the first holds its memory.
    const { scene, camera } = cornell();
    const renderer = await Renderer.create({ width: 16383, height: 16383 });
    renderer.scene = scene;
    renderer.camera = camera;
    expectError("Renderer.render at 16383²", () => renderer.render(), /memory.*\(code -3\)/);
    renderer.dispose();
    addon.close(big);
    // An argument the library refuses.
    const ctx = addon.open(0);
    expectError("a 0×0 image", () => addon.resize(ctx, 0, 0), /width and height.*\(code -1\)/);
    addon.close(ctx);
    // A served job cancelled while it runs: its wait() rejects.
    const server = await RenderServer.create({ slots: 1 });
    const made = spheres({ width: 96, height: 54 });
    const id = server.submit({ scene: made.scene, camera: made.camera, width: 96, height: 54, targetSamples: 100000 });
    const done = server.submit({ scene: made.scene, camera: made.camera, width: 96, height: 54, targetSamples: 2 });
    server.cancel(id);
    let rejected = null;
    try {
        await server.wait(id);
    }
    catch (e) {
        rejected = e;
    }
    if (!(rejected instanceof Error) || !/cancelled/.test(rejected.message))
        throw new Error(`a cancelled job's wait should reject, got ${rejected && rejected.message}`);
    console.log(`cancelled job: ${rejected.message}`);
    const { pixels, stats } = await server.wait(done);
    if (stats.samples != 2 || !pixels.every(Number.isFinite))
        throw new Error("the job after the cancelled one should finish");
    server.close();
}

main().catch((e) => { console.error(e.stack); process.exit(1); });
