// Vectors and 4×4 matrices for scene description (FP64, immutable values).
// API follows src/Vec2.mjs, src/Vec3.mjs, src/Vec4.mjs and src/Mat4.mjs of the
// reference (names, argument order, row-major constructor).

export class Vec2 {
    constructor(x = 0.0, y = 0.0) {
        this.x = x;
        this.y = y;
        Object.freeze(this);
    }
}

export class Vec3 {
    constructor(x = 0.0, y = 0.0, z = 0.0) {
        this.x = x;
        this.y = y;
        this.z = z;
        Object.freeze(this);
    }

    add(o) { return new Vec3(this.x + o.x, this.y + o.y, this.z + o.z); }
    subtract(o) { return new Vec3(this.x - o.x, this.y - o.y, this.z - o.z); }
    scale(s) { return new Vec3(this.x * s, this.y * s, this.z * s); }
    multiply(o) { return new Vec3(this.x * o.x, this.y * o.y, this.z * o.z); }
    negate() { return new Vec3(-this.x, -this.y, -this.z); }
    lerp(o, t) { return new Vec3(this.x + (o.x - this.x) * t, this.y + (o.y - this.y) * t, this.z + (o.z - this.z) * t); }

    cross(o) {
        return new Vec3(this.y * o.z - this.z * o.y, this.z * o.x - this.x * o.z, this.x * o.y - this.y * o.x);
    }

    dot(o) { return this.x * o.x + this.y * o.y + this.z * o.z; }
    length() { return Math.sqrt(this.dot(this)); }

    normalized() {
        const l = this.length();
        return l > 0.0 ? this.scale(1.0 / l) : new Vec3(0.0, 0.0, 0.0);
    }

    toString() { return `Vec3(${this.x}, ${this.y}, ${this.z})`; }
}

export class Vec4 {
    constructor(x = 0.0, y = 0.0, z = 0.0, w = 0.0) {
        this.x = x;
        this.y = y;
        this.z = z;
        this.w = w;
        Object.freeze(this);
    }

    dot(o) { return this.x * o.x + this.y * o.y + this.z * o.z + this.w * o.w; }
    lerp(o, t) {
        return new Vec4(this.x + (o.x - this.x) * t, this.y + (o.y - this.y) * t,
            this.z + (o.z - this.z) * t, this.w + (o.w - this.w) * t);
    }
}

// 3×3 minor of a row-major 4×4 array with row r and column c removed.
function minor(m, r, c) {
    const v = [];
    for (let i = 0; i < 4; ++i) {
        if (i == r) continue;
        for (let j = 0; j < 4; ++j) {
            if (j == c) continue;
            v.push(m[i * 4 + j]);
        }
    }
    return v[0] * (v[4] * v[8] - v[5] * v[7]) - v[1] * (v[3] * v[8] - v[5] * v[6]) + v[2] * (v[3] * v[7] - v[4] * v[6]);
}

function cofactor(m, r, c) {
    return ((r + c) % 2 == 0 ? 1 : -1) * minor(m, r, c);
}

/** A 4×4 matrix; the constructor takes the 16 values row by row. */
export class Mat4 {

    #m;

    constructor(...values) {
        if (values.length != 16)
            throw new RangeError("A Mat4 takes 16 values, row by row.");
        this.#m = values.slice();
    }

    value(row, column) {
        checkIndex(row, column);
        return this.#m[row * 4 + column];
    }

    setValue(row, column, value) {
        checkIndex(row, column);
        this.#m[row * 4 + column] = value;
    }

    row(i) { return new Vec4(this.#m[i * 4], this.#m[i * 4 + 1], this.#m[i * 4 + 2], this.#m[i * 4 + 3]); }
    column(i) { return new Vec4(this.#m[i], this.#m[4 + i], this.#m[8 + i], this.#m[12 + i]); }

    determinant() {
        let d = 0.0;
        for (let c = 0; c < 4; ++c)
            d += this.#m[c] * cofactor(this.#m, 0, c);
        return d;
    }

    inverse() {
        const det = this.determinant();
        if (det == 0.0)
            throw new RangeError("The matrix has no inverse.");
        const out = new Array(16);
        for (let r = 0; r < 4; ++r)
            for (let c = 0; c < 4; ++c)
                out[c * 4 + r] = cofactor(this.#m, r, c) / det;
        return new Mat4(...out);
    }

    multiplyMat4(o) {
        const a = this.#m;
        const out = new Array(16).fill(0.0);
        for (let r = 0; r < 4; ++r)
            for (let c = 0; c < 4; ++c) {
                let s = 0.0;
                for (let k = 0; k < 4; ++k)
                    s += a[r * 4 + k] * o.value(k, c);
                out[r * 4 + c] = s;
            }
        return new Mat4(...out);
    }

    multiplyVector3(v) {
        const m = this.#m;
        return new Vec3(m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3],
            m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7],
            m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11]);
    }

    multiplyDirection(v) {
        const m = this.#m;
        return new Vec3(m[0] * v.x + m[1] * v.y + m[2] * v.z,
            m[4] * v.x + m[5] * v.y + m[6] * v.z,
            m[8] * v.x + m[9] * v.y + m[10] * v.z);
    }

    transpose() {
        const m = this.#m;
        const out = new Array(16);
        for (let r = 0; r < 4; ++r)
            for (let c = 0; c < 4; ++c)
                out[c * 4 + r] = m[r * 4 + c];
        return new Mat4(...out);
    }

    static identity() {
        return new Mat4(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1);
    }

    static translate(x, y, z) {
        return new Mat4(1, 0, 0, x, 0, 1, 0, y, 0, 0, 1, z, 0, 0, 0, 1);
    }

    static scale(x, y, z) {
        return new Mat4(x, 0, 0, 0, 0, y, 0, 0, 0, 0, z, 0, 0, 0, 0, 1);
    }

    static rotateX(theta) {
        const s = Math.sin(theta), c = Math.cos(theta);
        return new Mat4(1, 0, 0, 0, 0, c, -s, 0, 0, s, c, 0, 0, 0, 0, 1);
    }

    static rotateY(theta) {
        const s = Math.sin(theta), c = Math.cos(theta);
        return new Mat4(c, 0, s, 0, 0, 1, 0, 0, -s, 0, c, 0, 0, 0, 0, 1);
    }

    static rotateZ(theta) {
        const s = Math.sin(theta), c = Math.cos(theta);
        return new Mat4(c, -s, 0, 0, s, c, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1);
    }

    // Camera to world for a camera at eye looking at target down its -z
    // (src/Mat4.mjs:307-318): columns right, true up, -look, eye.
    static lookAt(eye, target, up) {
        const look = target.subtract(eye).normalized();
        const right = look.cross(up).normalized();
        const trueUp = right.cross(look);
        return new Mat4(
            right.x, trueUp.x, -look.x, eye.x,
            right.y, trueUp.y, -look.y, eye.y,
            right.z, trueUp.z, -look.z, eye.z,
            0, 0, 0, 1);
    }
}

function checkIndex(row, column) {
    if (!(row >= 0 && row < 4 && column >= 0 && column < 4 && Number.isInteger(row) && Number.isInteger(column)))
        throw new RangeError(`No element (${row}, ${column}) in a 4×4 matrix.`);
}
