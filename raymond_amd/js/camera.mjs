// Thin-lens camera (src/Camera.mjs): basis() from Mat4.lookAt, and ray /
// pixelRay computed as the kernels' camera ray does (src/Camera.mjs:115-161).
import { Mat4, Vec3 } from "./math.mjs";

export class Camera {

    #position; #target; #up; #fov; #aperture; #focusDistance;

    constructor({
        position = new Vec3(0.0, 0.0, 5.0),
        target = new Vec3(0.0, 0.0, 0.0),
        up = new Vec3(0.0, 1.0, 0.0),
        fov = 40.0 * Math.PI / 180.0,
        aperture = 0.0,
        focusDistance = 0.0,
    } = {}) {
        this.#position = position;
        this.#target = target;
        this.#up = up;
        this.#fov = fov;
        this.#aperture = Math.max(aperture, 0.0);
        this.#focusDistance = Math.max(focusDistance, 0.0);
    }

    get position() { return this.#position; }
    set position(v) { this.#position = v; }
    get target() { return this.#target; }
    set target(v) { this.#target = v; }
    get up() { return this.#up; }
    set up(v) { this.#up = v; }
    /** Vertical field of view in radians. */
    get fov() { return this.#fov; }
    set fov(v) { this.#fov = v; }
    /** The lens radius. */
    get aperture() { return this.#aperture; }
    set aperture(v) { this.#aperture = Math.max(v, 0.0); }
    get focusDistance() { return this.#focusDistance; }
    set focusDistance(v) { this.#focusDistance = Math.max(v, 0.0); }

    get effectiveFocusDistance() {
        return this.#focusDistance > 0.0 ? this.#focusDistance : this.#target.subtract(this.#position).length();
    }

    get matrix() { return Mat4.lookAt(this.#position, this.#target, this.#up); }

    /** Unit right, up and forward (towards the target) in world space. */
    basis() {
        const m = this.matrix;
        return {
            right: new Vec3(m.value(0, 0), m.value(1, 0), m.value(2, 0)),
            up: new Vec3(m.value(0, 1), m.value(1, 1), m.value(2, 1)),
            forward: new Vec3(-m.value(0, 2), -m.value(1, 2), -m.value(2, 2)),
        };
    }

    /** The pinhole ray through image point (x, y) in [-1, 1]², y up. */
    ray(x, y, aspectRatio) {
        const { right, up, forward } = this.basis();
        const t = Math.tan(this.#fov / 2.0);
        const direction = forward.add(right.scale(x * t * aspectRatio)).add(up.scale(y * t)).normalized();
        return { origin: this.#position, direction };
    }

    /** The ray through the centre of pixel (px, py), counted from the top left. */
    pixelRay(px, py, width, height) {
        return this.ray(2.0 * (px + 0.5) / width - 1.0, 1.0 - 2.0 * (py + 0.5) / height, width / height);
    }
}
