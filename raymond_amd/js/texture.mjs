// Texture (src/Texture.mjs): sRGB8 RGBA images, the bilinear repeating
// sample() the kernels mirror (src/Texture.mjs:164-187), and resized().
import { Vec3 } from "./math.mjs";

const SRGB_TO_LINEAR = Float64Array.from({ length: 256 }, (_, i) => {
    const c = i / 255;
    return c <= 0.04045 ? c / 12.92 : Math.pow((c + 0.055) / 1.055, 2.4);
});

function toSrgbByte(c) {
    const x = Math.min(Math.max(c, 0), 1);
    const s = x <= 0.0031308 ? x * 12.92 : 1.055 * Math.pow(x, 1 / 2.4) - 0.055;
    return Math.round(s * 255);
}

const wrap = (a, n) => ((a % n) + n) % n;

export class Texture {

    #width; #height; #data;

    constructor(width, height, data) {
        if (!(Number.isInteger(width) && Number.isInteger(height) && width > 0 && height > 0))
            throw new RangeError("A texture needs a whole number of pixels each way.");
        if (data.length != width * height * 4)
            throw new RangeError(`A ${width} × ${height} texture needs ${width * height * 4} bytes.`);
        this.#width = width;
        this.#height = height;
        this.#data = Uint8Array.from(data);
    }

    get width() { return this.#width; }
    get height() { return this.#height; }
    get data() { return this.#data; }

    /** A texture from color(u, v) → linear Vec3 at each pixel centre. */
    static procedural(width, height, color) {
        const data = new Uint8Array(width * height * 4);
        for (let y = 0; y < height; ++y)
            for (let x = 0; x < width; ++x) {
                const c = color((x + 0.5) / width, (y + 0.5) / height);
                const i = (y * width + x) * 4;
                data[i] = toSrgbByte(c.x);
                data[i + 1] = toSrgbByte(c.y);
                data[i + 2] = toSrgbByte(c.z);
                data[i + 3] = 255;
            }
        return new Texture(width, height, data);
    }

    /** A normal map from a height field heightAt(u, v), by central differences. */
    static normalMap(width, height, heightAt) {
        const h = new Float64Array(width * height);
        for (let y = 0; y < height; ++y)
            for (let x = 0; x < width; ++x)
                h[y * width + x] = heightAt((x + 0.5) / width, (y + 0.5) / height);
        const at = (x, y) => h[wrap(y, height) * width + wrap(x, width)];
        const data = new Uint8Array(width * height * 4);
        for (let y = 0; y < height; ++y)
            for (let x = 0; x < width; ++x) {
                const right = (at(x + 1, y) - at(x - 1, y)) * width / 2.0;
                const down = (at(x, y + 1) - at(x, y - 1)) * width / 2.0;
                const len = Math.hypot(right, down, 1.0);
                const i = (y * width + x) * 4;
                data[i] = Math.round((-right / len * 0.5 + 0.5) * 255);
                data[i + 1] = Math.round((down / len * 0.5 + 0.5) * 255);
                data[i + 2] = Math.round((1.0 / len * 0.5 + 0.5) * 255);
                data[i + 3] = 255;
            }
        return new Texture(width, height, data);
    }

    /** Bilinear, repeating, in linear light (or plain bytes / 255 with srgb false). */
    sample(u, v, srgb = true) {
        const decode = srgb ? (b) => SRGB_TO_LINEAR[b] : (b) => b / 255;
        const x = u * this.#width - 0.5, y = v * this.#height - 0.5;
        const x0 = Math.floor(x), y0 = Math.floor(y);
        const fx = x - x0, fy = y - y0;
        let r = 0, g = 0, b = 0;
        const taps = [[0, 0, (1 - fx) * (1 - fy)], [1, 0, fx * (1 - fy)], [0, 1, (1 - fx) * fy], [1, 1, fx * fy]];
        for (const [dx, dy, w] of taps) {
            const i = (wrap(y0 + dy, this.#height) * this.#width + wrap(x0 + dx, this.#width)) * 4;
            r += w * decode(this.#data[i]);
            g += w * decode(this.#data[i + 1]);
            b += w * decode(this.#data[i + 2]);
        }
        return new Vec3(r, g, b);
    }

    resized(width, height, srgb = true) {
        if (width == this.#width && height == this.#height)
            return this;
        if (srgb)
            return Texture.procedural(width, height, (u, v) => this.sample(u, v));
        const data = new Uint8Array(width * height * 4);
        for (let y = 0; y < height; ++y)
            for (let x = 0; x < width; ++x) {
                const c = this.sample((x + 0.5) / width, (y + 0.5) / height, false);
                const i = (y * width + x) * 4;
                data[i] = Math.round(c.x * 255);
                data[i + 1] = Math.round(c.y * 255);
                data[i + 2] = Math.round(c.z * 255);
                data[i + 3] = 255;
            }
        return new Texture(width, height, data);
    }
}
