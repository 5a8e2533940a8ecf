// The public API, under the reference's names (src/index.mjs:3-20).
export { Vec2, Vec3, Vec4, Mat4 } from "./math.mjs";
export { Material, MATERIAL_TYPES, hexColor } from "./material.mjs";
export { Sphere, Quad, Box, Mesh } from "./geometry.mjs";
export { Texture } from "./texture.mjs";
export { Sky, EnvironmentMap } from "./sky.mjs";
export { Camera } from "./camera.mjs";
export { BVH } from "./bvh.mjs";
export { Scene, PRIMITIVE_KINDS, NONE, sphereUv, boxUv, bump } from "./scene.mjs";
export { Renderer } from "./renderer.mjs";
export { RenderServer } from "./server.mjs";
export { encodePng } from "./png.mjs";
