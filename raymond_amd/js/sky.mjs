// Sky (src/Sky.mjs) and EnvironmentMap (src/EnvironmentMap.mjs): the
// gradient and sun, the equirectangular map, its importance table
// (src/EnvironmentMap.mjs:202-247) and the CPU mirrors of what the kernels
// sample (sampleDirection/pdf/sample, :256-308).
import { Vec3 } from "./math.mjs";

export class EnvironmentMap {

    #width; #height; #data; #importance = null;

    constructor(width, height, data) {
        if (!(Number.isInteger(width) && Number.isInteger(height) && width > 0 && height > 0))
            throw new RangeError("An environment map needs a whole number of pixels each way.");
        if (data.length != width * height * 3)
            throw new RangeError(`A ${width} × ${height} environment map needs ${width * height * 3} floats.`);
        this.#width = width;
        this.#height = height;
        // NaN, infinities and negative values become 0.
        this.#data = Float32Array.from(data, (c) => Number.isFinite(c) ? Math.max(c, 0.0) : 0.0);
    }

    get width() { return this.#width; }
    get height() { return this.#height; }
    get data() { return this.#data; }

    /** The unit direction at (u, v): u round the horizon from -z, v from +y down. */
    static direction(u, v) {
        const phi = (u - 0.5) * 2.0 * Math.PI;
        const theta = v * Math.PI;
        return new Vec3(Math.sin(theta) * Math.sin(phi), Math.cos(theta), -Math.sin(theta) * Math.cos(phi));
    }

    static uv(d) {
        const u = 0.5 + Math.atan2(d.x, -d.z) / (2.0 * Math.PI);
        return [u - Math.floor(u), Math.acos(Math.min(Math.max(d.y, -1.0), 1.0)) / Math.PI];
    }

    /** A map filled by radiance(direction) at each pixel centre. */
    static procedural(width, height, radiance) {
        const data = new Float32Array(width * height * 3);
        for (let y = 0; y < height; ++y)
            for (let x = 0; x < width; ++x) {
                const c = radiance(EnvironmentMap.direction((x + 0.5) / width, (y + 0.5) / height));
                data[(y * width + x) * 3] = c.x;
                data[(y * width + x) * 3 + 1] = c.y;
                data[(y * width + x) * 3 + 2] = c.z;
            }
        return new EnvironmentMap(width, height, data);
    }

    /**
     * The table for picking directions in proportion to luminance × sin θ:
     * cw × (ch + 1) floats, cw = min(width, 512), ch = min(height, 256, cw).
     * Row y < ch is row y's CDF across; row ch holds the CDF down the rows.
     */
    importance() {
        if (this.#importance)
            return this.#importance;
        const W = this.#width, H = this.#height, src = this.#data;
        const cw = Math.min(W, 512), ch = Math.min(H, 256, cw);
        const table = new Float32Array(cw * (ch + 1));
        const rows = new Float64Array(ch);
        for (let y = 0; y < ch; ++y) {
            const y0 = Math.floor(y * H / ch), y1 = Math.max(y0 + 1, Math.floor((y + 1) * H / ch));
            const s = Math.sin((y + 0.5) / ch * Math.PI);
            let sum = 0.0;
            for (let x = 0; x < cw; ++x) {
                const x0 = Math.floor(x * W / cw), x1 = Math.max(x0 + 1, Math.floor((x + 1) * W / cw));
                let light = 0.0;
                for (let py = y0; py < y1; ++py)
                    for (let px = x0; px < x1; ++px) {
                        const i = (py * W + px) * 3;
                        light += 0.2126 * src[i] + 0.7152 * src[i + 1] + 0.0722 * src[i + 2];
                    }
                sum += light / ((y1 - y0) * (x1 - x0)) * s;
                table[y * cw + x] = sum;
            }
            for (let x = 0; x < cw; ++x)
                table[y * cw + x] = sum > 0 ? table[y * cw + x] / sum : (x + 1) / cw;
            rows[y] = sum;
        }
        let total = 0.0;
        for (let y = 0; y < ch; ++y) total += rows[y];
        let running = 0.0;
        for (let y = 0; y < ch; ++y) {
            running += rows[y];
            table[ch * cw + y] = total > 0 ? running / total : (y + 1) / ch;
        }
        table[ch * cw + ch - 1] = 1.0;
        for (let y = 0; y < ch; ++y)
            table[y * cw + cw - 1] = 1.0;
        this.#importance = { width: cw, height: ch, data: table };
        return this.#importance;
    }

    sampleDirection(xi) {
        const { width: cw, height: ch, data } = this.importance();
        const y = search(data, ch * cw, ch, xi[0]);
        const x = search(data, y * cw, cw, xi[1]);
        const direction = EnvironmentMap.direction((x + xi[2]) / cw, (y + xi[3]) / ch);
        return { direction, pdf: this.pdf(direction) };
    }

    pdf(d) {
        const { width: cw, height: ch, data } = this.importance();
        const [u, v] = EnvironmentMap.uv(d);
        const x = Math.min(Math.floor(u * cw), cw - 1), y = Math.min(Math.floor(v * ch), ch - 1);
        const row = data[ch * cw + y] - (y > 0 ? data[ch * cw + y - 1] : 0.0);
        const cell = data[y * cw + x] - (x > 0 ? data[y * cw + x - 1] : 0.0);
        const s = Math.sin(v * Math.PI);
        return s > 0 ? row * cell * cw * ch / (2.0 * Math.PI * Math.PI * s) : 0.0;
    }

    /** Bilinear radiance towards unit d, wrapping round the horizon, clamped at the poles. */
    sample(d) {
        const W = this.#width, H = this.#height;
        const [u, v] = EnvironmentMap.uv(d);
        const x = u * W - 0.5;
        const y = Math.min(Math.max(v * H - 0.5, 0.0), H - 1);
        const x0 = Math.floor(x), y0 = Math.floor(y);
        const fx = x - x0, fy = y - y0;
        let r = 0, g = 0, b = 0;
        const taps = [[0, 0, (1 - fx) * (1 - fy)], [1, 0, fx * (1 - fy)], [0, 1, (1 - fx) * fy], [1, 1, fx * fy]];
        for (const [dx, dy, w] of taps) {
            const px = ((x0 + dx) % W + W) % W;
            const py = Math.min(y0 + dy, H - 1);
            const i = (py * W + px) * 3;
            r += w * this.#data[i];
            g += w * this.#data[i + 1];
            b += w * this.#data[i + 2];
        }
        return new Vec3(r, g, b);
    }
}

function search(data, offset, n, xi) {
    let lo = 0, hi = n - 1;
    while (lo < hi) {
        const mid = (lo + hi) >> 1;
        if (data[offset + mid] > xi) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

/** The light from everything a ray escapes to. Immutable. */
export class Sky {

    #o;

    constructor(options = {}) {
        const o = Object.assign({
            zenith: new Vec3(0.25, 0.45, 0.9),
            horizon: new Vec3(0.85, 0.9, 1.0),
            ground: new Vec3(0.3, 0.28, 0.25),
            sunDirection: new Vec3(0.5, 0.8, 0.3),
            sunColor: new Vec3(1.0, 0.95, 0.85),
            sunStrength: 0.0,
            sunSize: 0.05,
            environment: null,
            environmentStrength: 1.0,
            environmentRotation: 0.0,
        }, options);
        o.sunDirection = o.sunDirection.normalized();
        o.sunStrength = Math.max(o.sunStrength, 0.0);
        o.sunSize = Math.min(Math.max(o.sunSize, 0.001), Math.PI / 4.0);
        o.environmentStrength = Math.max(o.environmentStrength, 0.0);
        this.#o = Object.freeze(o);
    }

    static environment(environment, strength = 1.0, rotation = 0.0) {
        return new Sky({ environment, environmentStrength: strength, environmentRotation: rotation });
    }

    static black() {
        const k = new Vec3(0.0, 0.0, 0.0);
        return new Sky({ zenith: k, horizon: k, ground: k });
    }

    get zenith() { return this.#o.zenith; }
    get horizon() { return this.#o.horizon; }
    get ground() { return this.#o.ground; }
    get sunDirection() { return this.#o.sunDirection; }
    get sunColor() { return this.#o.sunColor; }
    get sunStrength() { return this.#o.sunStrength; }
    /** The sun's angular radius in radians. */
    get sunSize() { return this.#o.sunSize; }
    get environment() { return this.#o.environment; }
    get environmentStrength() { return this.#o.environmentStrength; }
    get environmentRotation() { return this.#o.environmentRotation; }

    /** Strength spread over the disc's solid angle (src/Sky.mjs:132-136). */
    get sunRadiance() {
        const omega = 2.0 * Math.PI * (1.0 - Math.cos(this.#o.sunSize));
        return this.#o.sunColor.scale(this.#o.sunStrength / omega);
    }

    /** The sky's radiance towards unit d, without the sun (src/Sky.mjs:144-154). */
    radiance(d) {
        const o = this.#o;
        if (o.environment) {
            const [u, v] = EnvironmentMap.uv(d);
            const turned = EnvironmentMap.direction(u - o.environmentRotation / (2.0 * Math.PI), v);
            return o.environment.sample(turned).scale(o.environmentStrength);
        }
        const t = Math.sqrt(Math.min(Math.abs(d.y), 1.0));
        return d.y >= 0.0 ? o.horizon.lerp(o.zenith, t) : o.horizon.lerp(o.ground, t);
    }

    with(options) {
        return new Sky(Object.assign({}, this.#o, options));
    }
}
