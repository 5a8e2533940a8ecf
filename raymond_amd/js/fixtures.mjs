// Small scenes with exact or cross-checkable answers, for the closed-form and
// consistency tests. Geometry, materials and cameras are the reference's test
// scenes (tests/browser/fixture.mjs:35-292), which its closed forms are
// stated for; each returns { scene, camera, width, height, maxBounces }.
import { Camera } from "./camera.mjs";
import { Box, Mesh, Quad, Sphere } from "./geometry.mjs";
import { Material } from "./material.mjs";
import { Mat4, Vec2, Vec3 } from "./math.mjs";
import { Scene } from "./scene.mjs";
import { EnvironmentMap, Sky } from "./sky.mjs";
import { Texture } from "./texture.mjs";

const BLACK = new Vec3(0, 0, 0);
const WHITE = new Vec3(1, 1, 1);
const WHITE_SKY = new Sky({ zenith: WHITE, horizon: WHITE, ground: WHITE });
export const FLOOR_ALBEDO = 0.5;

const GRADIENT = Texture.procedural(64, 64, (u, v) => new Vec3(u, v, 0.5 + 0.5 * Math.sin(u * 6.28)));
const GRID = Texture.procedural(64, 64, (u, v) => {
    const i = Math.floor(u * 4), j = Math.floor(v * 4);
    return new Vec3((i + 1) / 4, (j + 1) / 4, ((i + j) % 2) * 0.8 + 0.1);
});
const BUMPS = Texture.normalMap(64, 64, (u, v) => 0.03 * Math.sin(u * Math.PI * 4) * Math.cos(v * Math.PI * 6));
const PATTERN = Texture.procedural(64, 64, (u, v) => new Vec3(1 - 0.9 * u, 0.3 + 0.4 * v, 0.1 + 0.9 * u * v));

const floor = (material = Material.diffuse(new Vec3(FLOOR_ALBEDO, FLOOR_ALBEDO, FLOOR_ALBEDO))) =>
    new Quad(new Vec3(-20, 0, 20), new Vec3(40, 0, 0), new Vec3(0, 0, -40), material);

const LOOK_DOWN = { position: new Vec3(0, 4, 4), target: new Vec3(0, 0, 0), fov: Math.PI / 3 };
const FRONT = { position: new Vec3(0, 0, 3.2), target: new Vec3(0, 0, 0), fov: Math.PI / 3 };
const TEXTURED_VIEW = { position: new Vec3(0.8, 2.6, 3.8), target: new Vec3(0, 0.4, 0), fov: Math.PI / 3 };

const MATERIALS = {
    diffuse: Material.diffuse(new Vec3(0.7, 0.7, 0.7)),
    plastic: Material.plastic(new Vec3(0.6, 0.6, 0.6), 0.3),
    metal: Material.metal(new Vec3(0.8, 0.8, 0.8), 0.3),
};

// An equirectangular environment with a soft warm lobe and a cool band,
// smooth enough to check against by quadrature.
function environment() {
    const sun = new Vec3(0.4, 0.6, -0.7).normalized();
    return EnvironmentMap.procedural(64, 32, (d) => {
        const c = Math.max(d.dot(sun), 0);
        const lobe = 6.0 * Math.pow(c, 8);
        const band = 0.4 + 0.3 * Math.max(d.y, 0);
        return new Vec3(band + lobe, band + 0.8 * lobe, band + 0.2 + 0.5 * lobe);
    });
}

const SCENES = {
    furnaceDiffuse: () => [new Scene({ sky: WHITE_SKY }).add(new Sphere(new Vec3(0, 0, 0), 1, Material.diffuse(new Vec3(0.8, 0.8, 0.8)))), FRONT],
    furnaceGlass: () => [new Scene({ sky: WHITE_SKY }).add(new Sphere(new Vec3(0, 0, 0), 1, Material.glass(1.5))), FRONT],
    furnaceDispersive: () => [new Scene({ sky: WHITE_SKY }).add(new Sphere(new Vec3(0, 0, 0), 1, Material.glass(1.5, WHITE, 20))), FRONT],
    furnaceFog: () => [new Scene({ sky: WHITE_SKY }).add(new Sphere(new Vec3(0, 0, 0), 1, Material.volume(WHITE, 2.0, 0.5))), FRONT],
    blackSmoke: () => [new Scene({ sky: WHITE_SKY }).add(
        new Sphere(new Vec3(-0.7, 0, 0), 0.8, Material.volume(BLACK, 1.5)),
        new Box(Mat4.translate(0.6, 0, 0).multiplyMat4(Mat4.rotateY(0.6)).multiplyMat4(Mat4.rotateX(0.4)).multiplyMat4(Mat4.scale(0.6, 0.5, 0.4)),
            Material.volume(BLACK, 1.5))), FRONT],
    sphereLight: () => [new Scene({ sky: Sky.black() }).add(floor(), new Sphere(new Vec3(0.5, 1.5, -1), 0.5, Material.light(WHITE, 6.0))), LOOK_DOWN],
    quadLight: () => [new Scene({ sky: Sky.black() }).add(floor(),
        new Quad(new Vec3(-0.5, 1.5, -1.5), new Vec3(1.2, 0, 0), new Vec3(0, 0, 0.8), Material.light(WHITE, 6.0))), LOOK_DOWN],
    sun: () => [new Scene({ sky: new Sky({ zenith: BLACK, horizon: BLACK, ground: BLACK, sunStrength: 2.0, sunDirection: new Vec3(0, 1, 0), sunSize: 0.05 }) })
        .add(floor()), LOOK_DOWN],
    lightAsSphere: (m = "diffuse") => [new Scene({ sky: Sky.black() }).add(floor(MATERIALS[m]),
        new Sphere(new Vec3(0, 1.6, -0.5), 1.0, Material.light(WHITE, 2.0))), LOOK_DOWN],
    lightAsMesh: (m = "diffuse") => [new Scene({ sky: Sky.black() }).add(floor(MATERIALS[m]),
        Mesh.icosphere(Material.light(WHITE, 2.0), 5).transformed(Mat4.translate(0, 1.6, -0.5))), LOOK_DOWN],
    lampInFog: (as = "light") => [new Scene({ sky: Sky.black() }).add(floor(),
        Box.fromCorners(new Vec3(-3, 0, -3), new Vec3(3, 2.5, 2), Material.volume(new Vec3(0.9, 0.9, 0.9), 0.5, 0.4)),
        as == "light" ? new Sphere(new Vec3(0, 1.4, -0.5), 0.6, Material.light(WHITE, 3.0)) :
            Mesh.icosphere(Material.light(WHITE, 3.0), 5).transformed(Mat4.translate(0, 1.4, -0.5).multiplyMat4(Mat4.scale(0.6, 0.6, 0.6)))), LOOK_DOWN],
    environmentFloor: () => [new Scene({ sky: Sky.environment(environment(), 1.5, 0.7) }).add(floor()), LOOK_DOWN],
    environmentSky: () => [new Scene({ sky: Sky.environment(environment(), 1.5, 0.7) }),
        { position: new Vec3(0, 0, 0), target: new Vec3(0.3, 0.25, -1), fov: 2.2 }],
    causticLamp: () => [new Scene({ sky: Sky.black() }).add(floor(), new Sphere(new Vec3(0, 0.9, 0), 0.5, Material.glass(1.5)),
        new Sphere(new Vec3(0.3, 2.6, -0.2), 0.15, Material.light(WHITE, 60.0))), LOOK_DOWN],
    causticSun: () => [new Scene({ sky: new Sky({ zenith: BLACK, horizon: BLACK, ground: BLACK, sunStrength: 2.0, sunDirection: new Vec3(0.3, 1, 0.1), sunSize: 0.03 }) })
        .add(floor(), new Sphere(new Vec3(0, 0.9, 0), 0.5, Material.glass(1.5))), LOOK_DOWN],
    causticMirror: () => [new Scene({ sky: Sky.black() }).add(floor(),
        new Quad(new Vec3(-1, 0.2, -1.6), new Vec3(2, 0, 0), new Vec3(0, 1.2, 0.5), Material.metal(new Vec3(0.9, 0.9, 0.9), 0)),
        new Quad(new Vec3(0.25, 1.4, 1.4), new Vec3(-0.5, 0, 0), new Vec3(0, 0.5, 0.2), Material.light(WHITE, 30.0))), LOOK_DOWN],
    causticThroughGlass: () => [new Scene({ sky: Sky.black() }).add(floor(), new Sphere(new Vec3(0, 0.9, 0), 0.5, Material.glass(1.5)),
        new Sphere(new Vec3(0.3, 2.6, -0.2), 0.3, Material.light(WHITE, 20.0)),
        Box.fromCorners(new Vec3(-3, 0, 1.6), new Vec3(3, 3.5, 1.62), Material.glass(1.5))),
        { position: new Vec3(0, 2.4, 2.4), target: new Vec3(0, 0, 0), fov: Math.PI / 4 }],
    room: () => [new Scene({ sky: Sky.black() }).add(
        floor(Material.checkered(new Vec3(0.8, 0.8, 0.8), new Vec3(0.15, 0.15, 0.15), 0.5)),
        new Quad(new Vec3(-2, 0, -2), new Vec3(4, 0, 0), new Vec3(0, 3, 0), Material.diffuse(new Vec3(0.7, 0.3, 0.2))),
        new Sphere(new Vec3(-0.6, 0.5, -0.4), 0.5, Material.diffuse(new Vec3(0.3, 0.6, 0.9))),
        Box.fromCorners(new Vec3(0.2, 0, -0.8), new Vec3(1.0, 0.8, 0.0), Material.plastic(new Vec3(0.9, 0.9, 0.9), 0.3), 0.4),
        new Quad(new Vec3(-0.6, 2.5, -0.6), new Vec3(1.2, 0, 0), new Vec3(0, 0, 1.2), Material.light(WHITE, 4.0))),
        { position: new Vec3(0.5, 2.2, 3.2), target: new Vec3(0, 0.4, -0.5), fov: Math.PI / 3 }],
    textured: () => [new Scene({ sky: Sky.black() }).add(
        new Quad(new Vec3(-10, 0, 10), new Vec3(20, 0, 0), new Vec3(0, 0, -20), Material.diffuse(WHITE).withTexture(GRID, 6)),
        new Sphere(new Vec3(-1.0, 0.7, 0), 0.7, Material.diffuse(WHITE).withTexture(GRADIENT), { rotation: Mat4.rotateZ(0.4).multiplyMat4(Mat4.rotateY(1.0)) }),
        Box.fromCorners(new Vec3(0.3, 0, -0.9), new Vec3(1.3, 1.0, 0.1), Material.plastic(new Vec3(1, 0.8, 0.8)).withTexture(GRID, new Vec2(1, 2), new Vec3(1, 0.8, 0.8)), 0.5),
        Mesh.torus(Material.diffuse(WHITE).withTexture(GRADIENT, 3), { radius: 0.4, tube: 0.18, segments: 32, sides: 16 })
            .transformed(Mat4.translate(0.6, 0.35, 1.0).multiplyMat4(Mat4.rotateX(1.1))),
        new Quad(new Vec3(-0.6, 2.5, -0.6), new Vec3(1.2, 0, 0), new Vec3(0, 0, 1.2), Material.light(WHITE, 4.0))), TEXTURED_VIEW],
    bumpy: () => [new Scene({ sky: Sky.black() }).add(
        new Quad(new Vec3(-10, 0, 10), new Vec3(20, 0, 0), new Vec3(0, 0, -20), Material.diffuse(WHITE).withTexture(GRID, 6).withNormalMap(BUMPS, 1.5)),
        new Sphere(new Vec3(-1.0, 0.7, 0), 0.7, Material.diffuse(WHITE).withNormalMap(BUMPS, 2), { rotation: Mat4.rotateZ(0.4).multiplyMat4(Mat4.rotateY(1.0)) }),
        Box.fromCorners(new Vec3(0.3, 0, -0.9), new Vec3(1.3, 1.0, 0.1), Material.plastic(WHITE).withTexture(GRID, new Vec2(1, 2)).withNormalMap(BUMPS, 2), 0.5),
        Mesh.torus(Material.diffuse(WHITE).withTexture(GRADIENT, 3).withNormalMap(BUMPS, 2), { radius: 0.4, tube: 0.18, segments: 32, sides: 16 })
            .transformed(Mat4.translate(0.6, 0.35, 1.0).multiplyMat4(Mat4.rotateX(1.1))),
        new Quad(new Vec3(-0.6, 2.5, -0.6), new Vec3(1.2, 0, 0), new Vec3(0, 0, 1.2), Material.light(WHITE, 4.0))), TEXTURED_VIEW],
    // A lamp with a pattern over a floor: a light (sampled directly) or a
    // black diffuse quad glowing with the same pattern (found only by
    // bouncing); the floor under it must come out alike, quarter by quarter.
    patternedLamp: (as = "light") => [new Scene({ sky: Sky.black() }).add(floor(),
        new Quad(new Vec3(-1.0, 1.2, -1.5), new Vec3(2.0, 0, 0), new Vec3(0, 0, 1.4),
            (as == "light" ? Material.light(WHITE) : Material.diffuse(BLACK)).withEmission(WHITE, 4, PATTERN))), LOOK_DOWN],
    // A white strip lamp behind a glass ball, with Abbe number `abbe` (0:
    // no dispersion): dispersion fringes the strip's edges with colour.
    stripThroughGlass: (abbe = "0") => [new Scene({ sky: Sky.black() }).add(
        new Quad(new Vec3(-0.3, -2, -3), new Vec3(0.6, 0, 0), new Vec3(0, 4, 0), Material.light(WHITE, 1.0)),
        new Sphere(new Vec3(0, 0, 0), 1, Material.glass(1.5, WHITE, Number(abbe)))), FRONT],
    patternedLamps: () => [new Scene({ sky: Sky.black() }).add(
        new Quad(new Vec3(-1.5, -1, 0), new Vec3(1.4, 0, 0), new Vec3(0, 2, 0), Material.light(WHITE).withEmission(WHITE, 1, PATTERN)),
        new Sphere(new Vec3(0.8, 0, 0), 0.8, Material.light(WHITE).withEmission(WHITE, 1, PATTERN), { rotation: Mat4.rotateY(0.7) })),
        { position: new Vec3(0, 0, 2.6), target: new Vec3(0, 0, 0), fov: Math.PI / 3 }],
};

/** fixture({ name, arg, size, bounces }) */
export function fixture({ name, arg, size = 32, bounces = 8 } = {}) {
    if (!SCENES[name])
        throw new RangeError(`No fixture "${name}".`);
    const [scene, cam] = SCENES[name](arg);
    return { scene, camera: new Camera(cam), width: size, height: size, maxBounces: bounces };
}

export const FIXTURES = Object.freeze(Object.keys(SCENES));
