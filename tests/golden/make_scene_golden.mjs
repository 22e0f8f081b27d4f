// Writes tests/golden/voxelrt_scene.json: digests of the reference's own
// generateScene(N, variant) (VoxelRT src/scene.js:232-476) for N = 64, 128,
// 256 and both variants -- the packed voxel grid, the brick occupancy words
// and the brick masks as SHA-256 of their little-endian u32 bytes, the count
// of solid voxels, and the spawn pose.  Outputs only (the inputs are N and the
// variant name).  Run once where the reference tree is mounted:
//   node tests/golden/make_scene_golden.mjs
// scene.js calls performance.now(), a browser global that Node 12 keeps in
// perf_hooks; it is provided before the module loads.
import { createHash } from 'crypto';
import { writeFileSync } from 'fs';
import { performance } from 'perf_hooks';
import { dirname, join } from 'path';
import { fileURLToPath } from 'url';

const REF = '/root/reference/src/scene.js';
const HERE = dirname(fileURLToPath(import.meta.url));

function digest(u32) {
  const h = createHash('sha256');
  h.update(Buffer.from(u32.buffer, u32.byteOffset, u32.byteLength));
  return h.digest('hex');
}

async function main() {
  globalThis.performance = performance;
  const { generateScene } = await import(REF);
  const out = { source: 'VoxelRT src/scene.js:232-476 generateScene(N, variant)', scenes: [] };
  for (const variant of ['default', 'lamps']) {
    for (const n of [64, 128, 256]) {
      const s = generateScene(n, variant);
      let solid = 0;
      for (let i = 0; i < s.vox.length; i++) if (s.vox[i] !== 0) solid++;
      out.scenes.push({
        n, variant,
        vox_sha256: digest(s.vox),
        bricks_sha256: digest(s.bricks),
        masks_sha256: digest(s.brickMasks),
        solid,
        spawn: { pos: s.spawn.pos, yaw: s.spawn.yaw, pitch: s.spawn.pitch },
      });
      console.log(`n=${n} ${variant}: ${solid} solid voxels`);
    }
  }
  writeFileSync(join(HERE, 'voxelrt_scene.json'), JSON.stringify(out, null, 1) + '\n');
}

main().catch((e) => {
  console.error(e);
  process.exit(1);
});
