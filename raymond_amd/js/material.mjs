// Materials: type codes, clamps and the Cauchy dispersion term follow
// src/Material.mjs:53-60 (codes), :111-128 (clamps), :196-209 (dispersion).
import { Vec2, Vec3 } from "./math.mjs";

/** The numeric code for each type, as the kernels see it. */
export const MATERIAL_TYPES = Object.freeze({
    diffuse: 0,
    metal: 1,
    plastic: 2,
    glass: 3,
    light: 4,
    volume: 5,
});

// Fraunhofer d, F and C lines in micrometres, where an Abbe number is defined.
const LINE_D = 0.5876;
const LINE_F = 0.4861;
const LINE_C = 0.6563;

const clamp = (x, lo, hi) => Math.min(Math.max(x, lo), hi);

/** What a surface is made of. Immutable. */
export class Material {

    #o;

    constructor(options = {}) {
        const o = Object.assign({
            type: "diffuse",
            color: new Vec3(0.8, 0.8, 0.8),
            roughness: 0.0,
            ior: 1.5,
            emission: new Vec3(0.0, 0.0, 0.0),
            checker: null,
            checkerSize: 1.0,
            texture: null,
            textureScale: new Vec2(1.0, 1.0),
            normalMap: null,
            normalScale: 1.0,
            emissionTexture: null,
            abbe: 0.0,
            density: 1.0,
            anisotropy: 0.0,
        }, options);

        if (!Object.prototype.hasOwnProperty.call(MATERIAL_TYPES, o.type))
            throw new RangeError(`Unknown material type "${o.type}".`);

        o.roughness = clamp(o.roughness, 0.0, 1.0);
        o.ior = Math.max(o.ior, 1.0);
        o.abbe = Math.max(o.abbe, 0.0);
        o.density = Math.max(o.density, 0.0);
        o.anisotropy = clamp(o.anisotropy, -0.99, 0.99);
        this.#o = Object.freeze(o);
    }

    get type() { return this.#o.type; }
    get color() { return this.#o.color; }
    get roughness() { return this.#o.roughness; }
    get ior() { return this.#o.ior; }
    get emission() { return this.#o.emission; }
    get checker() { return this.#o.checker; }
    get checkerSize() { return this.#o.checkerSize; }
    get texture() { return this.#o.texture; }
    get textureScale() { return this.#o.textureScale; }
    get normalMap() { return this.#o.normalMap; }
    get normalScale() { return this.#o.normalScale; }
    get emissionTexture() { return this.#o.emissionTexture; }
    get abbe() { return this.#o.abbe; }
    get density() { return this.#o.density; }
    get anisotropy() { return this.#o.anisotropy; }

    /** Cauchy's B (µm²) from the ior at the d line and the Abbe number; 0 for none. */
    get dispersion() {
        if (this.#o.type != "glass" || this.#o.abbe <= 0.0)
            return 0.0;
        return (this.#o.ior - 1.0) / (this.#o.abbe * (Math.pow(LINE_F, -2) - Math.pow(LINE_C, -2)));
    }

    /** The index of refraction at a wavelength in nanometres. */
    iorAt(wavelength) {
        const l = wavelength / 1000.0;
        return this.#o.ior + this.dispersion * (Math.pow(l, -2) - Math.pow(LINE_D, -2));
    }

    withTexture(texture, scale = 1.0, color = new Vec3(1.0, 1.0, 1.0)) {
        return copyWith(this, {
            color, checker: null, texture,
            textureScale: typeof scale == "number" ? new Vec2(scale, scale) : scale,
        });
    }

    withNormalMap(normalMap, strength = 1.0) {
        return copyWith(this, { normalMap, normalScale: strength });
    }

    withEmission(color, strength = 1.0, texture = null) {
        return copyWith(this, { emission: color.scale(strength), emissionTexture: texture });
    }

    /** A light: gives off light and is sampled directly. */
    get emissive() { return this.#o.type == "light" && this.glows; }

    /** Gives off any light at all. */
    get glows() { return this.#o.emission.x + this.#o.emission.y + this.#o.emission.z > 0.0; }

    /** The options this material was made with, after clamping. */
    get options() { return this.#o; }

    static diffuse(color) { return new Material({ type: "diffuse", color }); }
    static metal(color, roughness = 0.0) { return new Material({ type: "metal", color, roughness }); }
    static plastic(color, roughness = 0.2) { return new Material({ type: "plastic", color, roughness }); }
    static glass(ior = 1.5, tint = new Vec3(1.0, 1.0, 1.0), abbe = 0.0) {
        return new Material({ type: "glass", color: tint, ior, abbe });
    }
    static volume(color, density, anisotropy = 0.0) {
        return new Material({ type: "volume", color, density, anisotropy });
    }
    static light(color, strength = 1.0) { return new Material({ type: "light", emission: color.scale(strength) }); }
    static checkered(color, other, size = 1.0) {
        return new Material({ type: "diffuse", color, checker: other, checkerSize: size });
    }
}

function copyWith(material, changes) {
    return new Material(Object.assign({}, material.options, changes));
}

/** CSS sRGB hex ("#ff8800" or "#f80") to linear RGB. */
export function hexColor(hex) {
    let digits = hex.replace(/^#/, "");
    if (digits.length == 3)
        digits = digits.split("").map((d) => d + d).join("");
    if (!/^[0-9a-f]{6}$/i.test(digits))
        throw new RangeError(`"${hex}" is not a hex colour.`);
    const channel = (i) => {
        const c = parseInt(digits.slice(i * 2, i * 2 + 2), 16) / 255.0;
        return c <= 0.04045 ? c / 12.92 : Math.pow((c + 0.055) / 1.055, 2.4);
    };
    return new Vec3(channel(0), channel(1), channel(2));
}
