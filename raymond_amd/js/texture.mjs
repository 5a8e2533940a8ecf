// Texture: an RGBA8 image for the kernels' texture layers. Colour textures
// hold sRGB-encoded bytes and normal maps plain bytes; both are filtered
// bilinearly with repeat wrapping, the sampler the kernels and the oracle
// implement (the reference's is src/Texture.mjs:164-187). The API
// (procedural, normalMap, sample, resized, width/height/data) is the
// reference's; the byte values it produces are pinned by
// tests/golden/js_generators.json.
import { Vec3 } from "./math.mjs";

// Byte decoders, tabulated once: the sRGB transfer function (IEC 61966-2-1)
// to linear light, and the plain byte / 255 of normal maps.
const DECODE_SRGB = new Float64Array(256);
const DECODE_PLAIN = new Float64Array(256);
for (let b = 0; b < 256; ++b) {
    const v = b / 255;
    DECODE_PLAIN[b] = v;
    DECODE_SRGB[b] = v <= 0.04045 ? v / 12.92 : Math.pow((v + 0.055) / 1.055, 2.4);
}

// Byte encoders: linear light to an sRGB byte, and a plain [0, 1] value.
function encodeSrgb(linear) {
    const v = Math.min(Math.max(linear, 0), 1);
    return Math.round((v <= 0.0031308 ? v * 12.92 : 1.055 * Math.pow(v, 1 / 2.4) - 0.055) * 255);
}
const encodePlain = (v) => Math.round(v * 255);

// x modulo n in [0, n), for integers of either sign (repeat wrapping).
function repeat(x, n) {
    const r = x % n;
    return r < 0 ? r + n : r;
}

// A width × height RGBA8 buffer whose pixel (x, y) gets rgb(x, y) — a
// [r, g, b] of bytes — and opaque alpha, visiting pixels row by row.
function fillPixels(width, height, rgb) {
    const pixels = new Uint8Array(width * height * 4);
    let at = 0;
    for (let y = 0; y < height; ++y)
        for (let x = 0; x < width; ++x, at += 4) {
            const [r, g, b] = rgb(x, y);
            pixels[at] = r;
            pixels[at + 1] = g;
            pixels[at + 2] = b;
            pixels[at + 3] = 255;
        }
    return pixels;
}

export class Texture {

    #width; #height; #pixels;

    constructor(width, height, data) {
        if (!(Number.isInteger(width) && Number.isInteger(height) && width > 0 && height > 0))
            throw new RangeError("A texture needs a whole number of pixels each way.");
        if (data.length != width * height * 4)
            throw new RangeError(`A ${width} × ${height} texture needs ${width * height * 4} bytes.`);
        this.#width = width;
        this.#height = height;
        this.#pixels = Uint8Array.from(data);
    }

    get width() { return this.#width; }
    get height() { return this.#height; }
    get data() { return this.#pixels; }

    /** A texture whose pixels show color(u, v) (linear Vec3) at their centres. */
    static procedural(width, height, color) {
        return new Texture(width, height, fillPixels(width, height, (x, y) => {
            const c = color((x + 0.5) / width, (y + 0.5) / height);
            return [encodeSrgb(c.x), encodeSrgb(c.y), encodeSrgb(c.z)];
        }));
    }

    /**
     * A normal map of the height field heightAt(u, v): the slope at each
     * pixel by central differences (wrapping round the edges), scaled by
     * the width so that heights are in units of the texture's width, and the
     * unit normal (−∂h/∂x, ∂h/∂y, 1) stored as bytes of n·0.5 + 0.5.
     */
    static normalMap(width, height, heightAt) {
        const field = new Float64Array(width * height);
        for (let y = 0, i = 0; y < height; ++y)
            for (let x = 0; x < width; ++x, ++i)
                field[i] = heightAt((x + 0.5) / width, (y + 0.5) / height);
        const h = (x, y) => field[repeat(y, height) * width + repeat(x, width)];
        const slope = (ahead, behind) => (ahead - behind) * width / 2.0;
        return new Texture(width, height, fillPixels(width, height, (x, y) => {
            const sx = slope(h(x + 1, y), h(x - 1, y));
            const sy = slope(h(x, y + 1), h(x, y - 1));
            const norm = Math.hypot(sx, sy, 1.0);
            return [encodePlain(-sx / norm * 0.5 + 0.5), encodePlain(sy / norm * 0.5 + 0.5),
                encodePlain(1.0 / norm * 0.5 + 0.5)];
        }));
    }

    /**
     * The bilinearly filtered colour at (u, v), repeating: in linear light
     * for colour textures, or the plain byte values / 255 with srgb false.
     */
    sample(u, v, srgb = true) {
        const decode = srgb ? DECODE_SRGB : DECODE_PLAIN;
        const w = this.#width, h = this.#height, px = this.#pixels;
        const fx0 = u * w - 0.5, fy0 = v * h - 0.5;
        const x0 = Math.floor(fx0), y0 = Math.floor(fy0);
        const fx = fx0 - x0, fy = fy0 - y0;
        const left = repeat(x0, w), right = repeat(x0 + 1, w);
        const top = repeat(y0, h) * w, bottom = repeat(y0 + 1, h) * w;
        // The four texels in the order (x0, y0), (x0+1, y0), (x0, y0+1), (x0+1, y0+1).
        const texels = [(top + left) * 4, (top + right) * 4, (bottom + left) * 4, (bottom + right) * 4];
        const weights = [(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy];
        const rgb = [0, 0, 0];
        for (let k = 0; k < 4; ++k)
            for (let c = 0; c < 3; ++c)
                rgb[c] += weights[k] * decode[px[texels[k] + c]];
        return new Vec3(rgb[0], rgb[1], rgb[2]);
    }

    /**
     * This texture resampled to width × height by sample() at the new pixel
     * centres: re-encoded as sRGB for colour, as plain bytes with srgb false.
     */
    resized(width, height, srgb = true) {
        if (width == this.#width && height == this.#height)
            return this;
        const encode = srgb ? encodeSrgb : encodePlain;
        return new Texture(width, height, fillPixels(width, height, (x, y) => {
            const c = this.sample((x + 0.5) / width, (y + 0.5) / height, srgb);
            return [encode(c.x), encode(c.y), encode(c.z)];
        }));
    }
}
