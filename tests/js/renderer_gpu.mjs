// Drives the JS Renderer through the N-API addon on the GPU: renders the
// Cornell box for a few frames and writes the averaged and denoised pixels to
// the files given, and the display pass's image as a PNG, for the Python side
// to compare with the C ABI's output.
import { writeFileSync } from "fs";
import { Renderer, encodePng } from "../../raymond_amd/js/index.mjs";
import { cornell } from "../../raymond_amd/js/configs.mjs";

async function main() {
    const [, , outAvg, outDen, frames, outPng] = process.argv;
    const { scene, camera } = cornell();
    const renderer = await Renderer.create({ width: 64, height: 64 });
    renderer.scene = scene;
    renderer.camera = camera;
    renderer.fireflyClamp = 0;
    renderer.caustics = false;
    let threw = false;
    try { await renderer.readPixels(); } catch (e) { threw = /Nothing has been rendered yet/.test(e.message); }
    if (!threw) throw new Error("readPixels before render should throw");
    for (let k = 0; k < Number(frames); ++k)
        renderer.render();
    await renderer.device.queue.onSubmittedWorkDone();
    const avg = await renderer.readPixels();
    const den = await renderer.readPixels({ denoised: true });
    const shown = await renderer.readDisplay();
    renderer.view = "albedo";
    let bad = false;
    try { await renderer.readPixels({ denoised: true }); } catch (e) { bad = /Only the color view/.test(e.message); }
    if (!bad) throw new Error("denoised readback outside the colour view should throw");
    writeFileSync(outAvg, Buffer.from(avg.buffer));
    writeFileSync(outDen, Buffer.from(den.buffer));
    writeFileSync(outPng, encodePng(64, 64, shown));
    console.log(JSON.stringify({ samples: renderer.samples }));
    // The other schedules and the precision variants run through the same calls.
    let rejected = false;
    try { renderer.schedule = "diagonal"; } catch (e) { rejected = e instanceof RangeError; }
    if (!rejected) throw new Error("an unknown schedule should be a RangeError");
    rejected = false;
    try { renderer.waves = 6; } catch (e) { rejected = e instanceof RangeError; }
    if (!rejected) throw new Error("waves other than 4 or 5 should be a RangeError");
    renderer.waves = 5;
    renderer.view = "color";
    renderer.render();
    renderer.schedule = "wavefront";
    renderer.precision = { quantizedBvh: true, fp16Guides: true, compensated: true };
    renderer.render();
    renderer.render();
    const varied = await renderer.readPixels({ denoised: true });
    if (!varied.every(Number.isFinite)) throw new Error("the varied frames should be finite");
    // The tuner renders whole frames under each candidate and keeps the fastest.
    const tuned = renderer.tune({ frames: 1 });
    if (tuned.candidates.length != 12) throw new Error("the tuner should try twelve candidates");
    if (tuned.ms != Math.min(...tuned.candidates.map((c) => c.ms))) throw new Error("the tuner should keep the fastest");
    if (renderer.schedule != tuned.schedule || renderer.waves != tuned.waves || renderer.samples != 0)
        throw new Error("the renderer should take the tuned schedule and start the image again");
    renderer.render();
    if (!(await renderer.readPixels()).every(Number.isFinite)) throw new Error("frames after tuning should be finite");
    // A light-path budget for the caustics, past the reference's one a pixel.
    rejected = false;
    try { renderer.lightPaths = -1; } catch (e) { rejected = e instanceof RangeError; }
    if (!rejected) throw new Error("a negative light-path budget should be a RangeError");
    renderer.caustics = true;
    renderer.lightPaths = 50000;
    renderer.render();
    renderer.render();
    if (renderer.lightPaths != 50000 || !(await renderer.readPixels()).every(Number.isFinite))
        throw new Error("frames with a light-path budget should be finite");
    renderer.dispose();
}

main().catch((e) => { console.error(e.stack); process.exit(1); });
