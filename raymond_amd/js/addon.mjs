// Loads raymond.node, the N-API addon over libraymond_hip.so (both built in
// raymond_amd/lib). Loading needs no GPU: the BVH builder runs on the host,
// and the device calls fail with an Error when no GPU is visible.
import { createRequire } from "module";
import { dirname, join } from "path";
import { fileURLToPath } from "url";

const here = dirname(fileURLToPath(import.meta.url));
let addonCache = null;

/** The addon, loaded once; throws an Error naming the path when it can't load. */
export function loadAddon() {
    if (addonCache)
        return addonCache;
    const require = createRequire(import.meta.url);
    const path = process.env.RAYMOND_ADDON || join(here, "..", "lib", "raymond.node");
    try {
        addonCache = require(path);
    }
    catch (error) {
        throw new Error(`The Raymond HIP addon isn't built or can't load (${path}): ${error.message}`);
    }
    return addonCache;
}

/** RM_OPTION_PRECISION. */
export const OPTION_PRECISION = 2;

/**
 * RM_PRECISION_* flags from { quantizedBvh, fp16Guides, compensated }: the
 * 8-bit child-quantized BVH, FP16 denoiser guides, compensated FP32 sums.
 */
export function precisionFlags({ quantizedBvh = false, fp16Guides = false, compensated = false } = {}) {
    return (quantizedBvh ? 1 : 0) | (fp16Guides ? 2 : 0) | (compensated ? 4 : 0);
}

/** RM_OPTION_WAVES: waves per SIMD of the one-thread-per-pixel path pass, 4 or 5. */
export const OPTION_WAVES = 5;

/** RM_OPTION_LIGHT_PATHS: light paths a frame with caustics on (0: the reference's count). */
export const OPTION_LIGHT_PATHS = 6;

/** RM_OPTION_SCHEDULE and its values. */
export const OPTION_SCHEDULE = 0;
export const SCHEDULES = Object.freeze({ "per-pixel": 0, compact: 1, persistent: 2, wavefront: 3, sorted: 4, split: 5 });
