// Writes tests/golden/voxelrt_camera.json: the reference's own camera
// matrices (VoxelRT src/camera.js:47-57 Camera.matrices over src/math.js)
// for a set of poses and aspect ratios, as the float32 words main.js:407
// stores into the uniform block (f32.set).  Inputs (pose, aspect; fov 70 deg,
// near 0.05, far 400 as camera.js:10-12) and outputs only.  Run once where
// the reference tree is mounted:
//   node tests/golden/make_camera_golden.mjs
import { writeFileSync } from 'fs';
import { dirname, join } from 'path';
import { fileURLToPath } from 'url';

const REF = '/root/reference/src/camera.js';
const HERE = dirname(fileURLToPath(import.meta.url));

function f32words(m) {
  const f = Float32Array.from(m);
  return Array.from(new Uint32Array(f.buffer));
}

async function main() {
  const { Camera } = await import(REF);
  // seeded poses: spawn-like, orbit-like and steep ones
  let z = 0x2545F491;
  const rnd = () => {
    z ^= z << 13; z >>>= 0;
    z ^= z >>> 17;
    z ^= z << 5; z >>>= 0;
    return z / 4294967296;
  };
  const poses = [
    { pos: [8, 4.775, 2.75], yaw: 0, pitch: -0.05 },
    { pos: [4, 3.275, -0.875], yaw: 0, pitch: -0.05 },
    { pos: [12.5, 9.25, 3.0], yaw: 2.4, pitch: -0.35 },
    { pos: [0.03, 0.5, 15.9], yaw: -1.57, pitch: 1.55 },
  ];
  for (let i = 0; i < 12; i++)
    poses.push({ pos: [rnd() * 64, rnd() * 16, rnd() * 64], yaw: (rnd() - 0.5) * 12, pitch: (rnd() - 0.5) * 3.1 });
  const aspects = [1920 / 1080, 1, 640 / 360, 2560 / 1080, 33 / 17];
  const cases = [];
  for (const p of poses)
    for (const aspect of aspects) {
      // the reference's method on a plain state object (the constructor
      // wires DOM input handlers; matrices() reads only these fields)
      const cam = { pos: p.pos, yaw: p.yaw, pitch: p.pitch, fov: (70 * Math.PI) / 180, near: 0.05, far: 400 };
      const { viewProj, invViewProj } = Camera.prototype.matrices.call(cam, aspect);
      cases.push({ pos: p.pos, yaw: p.yaw, pitch: p.pitch, aspect,
                   view_proj: f32words(viewProj), inv_view_proj: f32words(invViewProj) });
    }
  writeFileSync(join(HERE, 'voxelrt_camera.json'),
                JSON.stringify({ source: 'VoxelRT src/camera.js:47-57, src/math.js', cases }) + '\n');
  console.log(`${cases.length} cases`);
}

main().catch((e) => {
  console.error(e);
  process.exit(1);
});
