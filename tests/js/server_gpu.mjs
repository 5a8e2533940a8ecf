// Drives the JS RenderServer on the GPU: the same Cornell box job through the
// server (three copies in flight with a different job between them) and
// through a Renderer, frame for frame; prints whether the images agree.
import { Renderer, RenderServer } from "../../raymond_amd/js/index.mjs";
import { cornell, spheres } from "../../raymond_amd/js/configs.mjs";

async function main() {
    const frames = 5;
    const { scene, camera } = cornell();
    const other = spheres({ width: 48, height: 27 });
    const settings = { fireflyClamp: 0, caustics: false };

    const renderer = await Renderer.create({ width: 64, height: 64 });
    renderer.scene = scene;
    renderer.camera = camera;
    renderer.fireflyClamp = 0;
    renderer.caustics = false;
    for (let k = 0; k < frames; ++k)
        renderer.render();
    const alone = await renderer.readPixels({ denoised: true });
    const average = await renderer.readPixels();
    renderer.dispose();

    const server = await RenderServer.create({ slots: 2 });
    const job = { scene, camera, width: 64, height: 64, targetSamples: frames, ...settings };
    const ids = [server.submit(job),
        server.submit({ scene: other.scene, camera: other.camera, width: 48, height: 27, targetSamples: 3, ...settings }),
        server.submit(job), server.submit({ ...job, denoise: false })];
    const results = [];
    for (const id of ids)
        results.push(await server.wait(id));
    server.close();

    const same = (a, b) => a.length == b.length && a.every((v, i) => Object.is(v, b[i]));
    console.log(JSON.stringify({
        denoisedFirst: same(results[0].pixels, alone),
        denoisedSecond: same(results[2].pixels, alone),
        average: same(results[3].pixels, average),
        stats: results.map((r) => r.stats),
    }));
}

main().catch((e) => { console.error(e.stack); process.exit(1); });
