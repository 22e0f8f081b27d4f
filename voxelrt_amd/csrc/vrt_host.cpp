// vrt_host.cpp -- host helpers of libvoxelrt_hip: camera matrices and
// uniforms (src/math.js, src/camera.js:47-56, src/main.js:399-426) and the
// seeded scene generators.  No device work here.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/voxelrt.h"

namespace {

// ---------------------------------------------------------------------------
// camera: math.js in double, each Float32Array store rounded to f32
// ---------------------------------------------------------------------------
struct Basis {
  double fw[3], rt[3], up[3];
};
Basis camera_basis(double yaw, double pitch) {  // math.js:96-107
  Basis b;
  const double cp = std::cos(pitch), sp = std::sin(pitch);
  const double cy = std::cos(yaw), sy = std::sin(yaw);
  b.fw[0] = sy * cp; b.fw[1] = sp; b.fw[2] = cy * cp;
  b.rt[0] = -cy; b.rt[1] = 0.0; b.rt[2] = sy;
  b.up[0] = b.rt[1] * b.fw[2] - b.rt[2] * b.fw[1];
  b.up[1] = b.rt[2] * b.fw[0] - b.rt[0] * b.fw[2];
  b.up[2] = b.rt[0] * b.fw[1] - b.rt[1] * b.fw[0];
  return b;
}

void mat4_mul(const float* a, const float* b, float* out) {  // math.js:60-70
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += (double)a[k * 4 + r] * (double)b[c * 4 + k];
      out[c * 4 + r] = (float)s;
    }
}

// ---------------------------------------------------------------------------
// deterministic hashing for the seeded generators
// ---------------------------------------------------------------------------
inline uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(mix64(seed ^ 0x5DEECE66Dull)) {}
  uint64_t next() { return s = mix64(s); }
  double uf() { return u01(next()); }
  int range(int lo, int hi) { return lo + (int)(uf() * (double)(hi - lo + 1)); }
};

inline uint32_t pack(int r, int g, int b, int e = 0) {  // scene.js:18 M()
  return ((uint32_t)r | ((uint32_t)g << 8) | ((uint32_t)b << 16) | ((uint32_t)e << 24));
}
// linear albedo -> authored byte (unpackMaterial squares byte/255)
inline int alb_byte(double a) {
  return (int)std::lround(255.0 * std::sqrt(std::max(0.0, std::min(1.0, a))));
}

struct Vol {
  uint32_t* v;
  int nx, ny, nz;
  bool inb(int x, int y, int z) const {
    return x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz;
  }
  uint64_t idx(int x, int y, int z) const {
    return (uint64_t)x + (uint64_t)y * nx + (uint64_t)z * (uint64_t)nx * ny;
  }
  void set(int x, int y, int z, uint32_t m) {
    if (inb(x, y, z)) v[idx(x, y, z)] = m;
  }
  uint32_t get(int x, int y, int z) const { return inb(x, y, z) ? v[idx(x, y, z)] : 0u; }
  template <class F>
  void box(int x0, int y0, int z0, int x1, int y1, int z1, F f) {
    for (int z = z0; z <= z1; ++z)
      for (int y = y0; y <= y1; ++y)
        for (int x = x0; x <= x1; ++x) set(x, y, z, f(x, y, z));
  }
  void boxm(int x0, int y0, int z0, int x1, int y1, int z1, uint32_t m) {
    box(x0, y0, z0, x1, y1, z1, [m](int, int, int) { return m; });
  }
};

// ---------------------------------------------------------------------------
// VRT_SCENE_VOXELRT: generateScene(N, variant), scene.js:202-409, in JS
// semantics (doubles, Math.floor/round, ToInt32 wrap, Math.imul)
// ---------------------------------------------------------------------------
struct JsNoise {  // scene.js:203-222
  int32_t seed;
  static int32_t to_i32(int64_t v) { return (int32_t)(uint32_t)(uint64_t)v; }
  double hash(double x, double y) const {
    // (x*374761393 + y*668265263 + seed*1442695041) | 0, exact in double here
    int64_t s = (int64_t)x * 374761393ll + (int64_t)y * 668265263ll + (int64_t)seed * 1442695041ll;
    int32_t h = to_i32(s);
    uint32_t hu = (uint32_t)h;
    int32_t t = (int32_t)(hu ^ (hu >> 13));
    h = (int32_t)((uint32_t)t * 1274126177u);  // Math.imul
    hu = (uint32_t)h;
    return (double)(hu ^ (hu >> 16)) / 4294967296.0;
  }
  static double smooth(double t) { return t * t * (3 - 2 * t); }
  double noise2(double x, double y) const {
    const double xi = std::floor(x), yi = std::floor(y);
    const double xf = smooth(x - xi), yf = smooth(y - yi);
    const double a = hash(xi, yi), b = hash(xi + 1, yi);
    const double c = hash(xi, yi + 1), d = hash(xi + 1, yi + 1);
    return a + (b - a) * xf + (c - a) * yf + (a - b - c + d) * xf * yf;
  }
  double operator()(double x, double y) const {
    double v = 0, amp = 0.5, f = 1;
    for (int o = 0; o < 4; ++o) {
      v += noise2(x * f, y * f) * amp;
      amp *= 0.5;
      f *= 2;
    }
    return v;
  }
};

inline int js_round(double x) { return (int)std::floor(x + 0.5); }  // Math.round
inline int V(double meters) { return js_round(meters / 0.0625); }

struct VoxelrtLayout {
  int N, cx, cz, x0, x1, z0, z1, floorY, y1, roomW, roomH;
  std::vector<int32_t> heights;
  int groundAt(int x, int z) const {
    x = std::max(0, std::min(N - 1, x));
    z = std::max(0, std::min(N - 1, z));
    return heights[(size_t)x + (size_t)z * N];
  }
};

namespace M {
inline uint32_t grass(int v) { return pack(88 + v, 130 + v, 60); }
inline uint32_t dirt(int v) { return pack(110 + v, 82 + v, 58); }
inline uint32_t rock(int v) { return pack(120 + v, 120 + v, 125 + v); }
inline uint32_t leaf(int v) { return pack(50 + v, 120 + v, 45); }
const uint32_t white = pack(225, 222, 215), red = pack(200, 45, 40), green = pack(55, 175, 60);
const uint32_t floorA = pack(180, 150, 110), floorB = pack(150, 118, 84);
const uint32_t boxTall = pack(200, 200, 210), boxShort = pack(210, 170, 90);
const uint32_t trunk = pack(95, 68, 42), lampPost = pack(60, 60, 65);
const uint32_t lampWarm = pack(255, 190, 110, 255), lampCyan = pack(90, 220, 255, 255);
const uint32_t lampMagenta = pack(255, 90, 200, 255), ceilLight = pack(255, 240, 220, 210);
const uint32_t lampHot = pack(255, 214, 150, 255), panelOrange = pack(255, 140, 40, 200);
const uint32_t panelTeal = pack(40, 220, 200, 165), panelViolet = pack(190, 80, 255, 245);
const uint32_t panelLime = pack(150, 255, 60, 125), dimStrip = pack(140, 180, 255, 45);
const uint32_t pillar = pack(118, 116, 126), shelf = pack(140, 105, 70);
const uint32_t lampWarmDim = pack(255, 190, 110, 64), lampCyanDim = pack(90, 220, 255, 64);
const uint32_t lampMagentaDim = pack(255, 90, 200, 64);
}  // namespace M

VoxelrtLayout voxelrt_generate(int N, int variant, uint32_t* vox) {
  VoxelrtLayout L;
  L.N = N;
  Vol vol{vox, N, N, N};
  if (vox) std::memset(vox, 0, (size_t)N * N * N * 4);
  JsNoise noise{1337};
  L.heights.assign((size_t)N * N, 0);
  const int baseH = (int)std::floor(N * 0.16);
  for (int z = 0; z < N; ++z)
    for (int x = 0; x < N; ++x) {
      const int h = baseH + (int)std::floor(noise(x * 0.02, z * 0.02) * N * 0.10);
      L.heights[(size_t)x + (size_t)z * N] = h;
      if (!vox) continue;
      for (int y = 0; y <= h; ++y) {
        const int v = (x * 7 + y * 13 + z * 5) % 9;
        uint32_t m;
        if (y == h) m = M::grass(v);
        else if (y > h - 4) m = M::dirt(v);
        else m = M::rock(v);
        vol.set(x, y, z, m);
      }
    }
  L.roomW = std::min(V(5), (int)std::floor(N * 0.45));
  L.roomH = std::min(V(3.2), (int)std::floor(N * 0.3));
  L.cx = (int)std::floor(N * 0.5);
  L.cz = (int)std::floor(N * 0.55);
  L.x0 = L.cx - (L.roomW >> 1); L.x1 = L.cx + (L.roomW >> 1);
  L.z0 = L.cz - (L.roomW >> 1); L.z1 = L.cz + (L.roomW >> 1);
  L.floorY = L.groundAt(L.cx, L.cz) + 1;
  L.y1 = L.floorY + L.roomH;
  if (!vox) return L;
  const int T = 2, cx = L.cx, cz = L.cz, x0 = L.x0, x1 = L.x1, z0 = L.z0, z1 = L.z1;
  const int floorY = L.floorY, y1 = L.y1, roomW = L.roomW, roomH = L.roomH;
  vol.box(x0 - T, floorY, z0 - T, x1 + T, floorY, z1 + T, [](int x, int, int z) { return M::dirt((x + z) % 9); });
  vol.boxm(x0 - T, floorY + 1, z0 - T, x1 + T, y1 + T + 4, z1 + T, 0u);
  vol.box(x0 - T, floorY, z0 - T, x1 + T, floorY, z1 + T,
          [](int x, int, int z) { return ((x >> 3) + (z >> 3)) % 2 ? M::floorA : M::floorB; });
  vol.boxm(x0 - T, floorY + 1, z1 + 1, x1 + T, y1, z1 + T, M::white);
  vol.boxm(x0 - T, floorY + 1, z0 - T, x1 + T, y1, z0 - 1, M::white);
  vol.boxm(x0 - T, floorY + 1, z0 - T, x0 - 1, y1, z1 + T, M::red);
  vol.boxm(x1 + 1, floorY + 1, z0 - T, x1 + T, y1, z1 + T, M::green);
  vol.boxm(x0 - T, y1 + 1, z0 - T, x1 + T, y1 + T, z1 + T, M::white);
  const int doorHalf = (int)std::floor(roomW * 0.28);
  vol.boxm(cx - doorHalf, floorY + 1, z0 - T, cx + doorHalf, y1 - 2, z0 - 1, 0u);
  const int winY0 = floorY + (int)std::floor(roomH * 0.45);
  const int winY1 = floorY + (int)std::floor(roomH * 0.8);
  vol.boxm(x1 + 1, winY0, cz - (int)std::floor(roomW * 0.3), x1 + T, winY1, cz + (int)std::floor(roomW * 0.3), 0u);
  if (variant == 1) {  // scene.js:303-361
    const int stripLen = (int)std::floor(roomW * 0.5);
    const int sx0 = cx - (stripLen >> 1), sx1 = sx0 + stripLen - 1;
    const int stripDz = (int)std::floor(roomW * 0.22);
    for (int dz : {-stripDz, 0, stripDz}) vol.boxm(sx0, y1, cz + dz - 1, sx1, y1, cz + dz + 1, M::dimStrip);
    vol.boxm(sx0, y1 - 4, z1, sx1, y1 - 3, z1, M::dimStrip);
    const int pW = 5, pH = 3;
    const int pMidY = floorY + (int)std::floor(roomH * 0.35);
    vol.boxm(x0, pMidY, cz - 9, x0, pMidY + pH, cz - 9 + pW, M::panelOrange);
    vol.boxm(x1, floorY + (int)std::floor(roomH * 0.2), cz - 22, x1,
             floorY + (int)std::floor(roomH * 0.2) + pH, cz - 22 + pW, M::panelTeal);
    vol.boxm(cx + 10, floorY + (int)std::floor(roomH * 0.5), z1, cx + 10 + pW,
             floorY + (int)std::floor(roomH * 0.5) + pH, z1, M::panelViolet);
    vol.boxm(x0 + 6, floorY + (int)std::floor(roomH * 0.4), z0, x0 + 6 + pW,
             floorY + (int)std::floor(roomH * 0.4) + pH, z0, M::panelLime);
    const int pAx = cx - (int)std::floor(roomW * 0.12), pAz = cz + (int)std::floor(roomW * 0.15);
    vol.boxm(pAx - 1, floorY + 1, pAz - 1, pAx + 2, y1, pAz + 2, M::pillar);
    const int pBx = cx + (int)std::floor(roomW * 0.22), pBz = cz - (int)std::floor(roomW * 0.12);
    vol.boxm(pBx - 1, floorY + 1, pBz - 1, pBx + 2, floorY + (int)std::floor(roomH * 0.75), pBz + 2, M::pillar);
    const int shelfY = floorY + (int)std::floor(roomH * 0.3);
    vol.boxm(x1 - 10, shelfY, cz + 10, x1 - 1, shelfY + 1, cz + 26, M::shelf);
    auto pendant = [&](int lx, int lz) {
      for (int y = y1 - 2; y <= y1; ++y) vol.set(lx, y, lz, M::lampPost);
      vol.boxm(lx, y1 - 4, lz, lx, y1 - 3, lz, M::lampHot);
    };
    pendant(cx + (int)std::floor(roomW * 0.19), cz - (int)std::floor(roomW * 0.16));
    pendant(cx - (int)std::floor(roomW * 0.3), cz + (int)std::floor(roomW * 0.3));
    auto standing = [&](int lx, int lz) {
      vol.set(lx, floorY + 1, lz, M::lampPost);
      vol.boxm(lx, floorY + 2, lz, lx, floorY + 3, lz, M::lampHot);
    };
    standing(x1 - 5, cz + 18);
    standing(pAx - 4, z1 - 4);
    standing(pBx + 4, pBz - 4);
    standing(x0 + 3, z1 - 3);
    vol.boxm(cx - 24, y1 - 6, z0, cx - 24, y1 - 5, z0, M::lampHot);
    vol.boxm(cx - 6, floorY + 8, z1, cx - 6, floorY + 9, z1, M::lampHot);
  } else {
    const int lHalf = (int)std::floor(roomW * 0.18);
    vol.boxm(cx - lHalf, y1, cz - lHalf, cx + lHalf, y1, cz + lHalf, M::ceilLight);
  }
  const int tallW = (int)std::floor(roomW * 0.16), tallH = (int)std::floor(roomH * 0.55);
  vol.boxm(cx - (int)std::floor(roomW * 0.28) - tallW, floorY + 1, cz + (int)std::floor(roomW * 0.05),
           cx - (int)std::floor(roomW * 0.28), floorY + tallH, cz + (int)std::floor(roomW * 0.05) + tallW,
           M::boxTall);
  const int shortW = (int)std::floor(roomW * 0.18), shortH = (int)std::floor(roomH * 0.28);
  vol.boxm(cx + (int)std::floor(roomW * 0.10), floorY + 1, cz - (int)std::floor(roomW * 0.25),
           cx + (int)std::floor(roomW * 0.10) + shortW, floorY + shortH,
           cz - (int)std::floor(roomW * 0.25) + shortW, M::boxShort);
  auto tree = [&](int tx, int tz) {  // scene.js:377-391
    const int g = L.groundAt(tx, tz);
    const int trunkH = V(1.4) + ((tx * 31 + tz * 17) % V(0.8));
    for (int y = g + 1; y <= g + trunkH; ++y) {
      vol.set(tx, y, tz, M::trunk);
      vol.set(tx + 1, y, tz, M::trunk);
    }
    const int cy = g + trunkH + V(0.35);
    const int r = V(0.7) + ((tx + tz) % 3);
    for (int dz = -r; dz <= r; ++dz)
      for (int dy = -r; dy <= r; ++dy)
        for (int dx = -r; dx <= r; ++dx) {
          const double d2 = dx * dx + dy * dy * 1.6 + dz * dz;
          if (d2 <= r * r && ((dx * 13 + dy * 7 + dz * 11) % 5 != 0 || d2 < r * r * 0.5))
            vol.set(tx + dx, cy + dy, tz + dz, M::leaf((dx * 3 + dy * 5 + dz * 7) % 12 + ((dx + dy + dz) % 2) * 6));
        }
  };
  const double spots[6][2] = {{0.15, 0.2}, {0.82, 0.25}, {0.2, 0.8}, {0.86, 0.78}, {0.12, 0.55}, {0.65, 0.12}};
  for (auto& s : spots) tree((int)std::floor(N * s[0]), (int)std::floor(N * s[1]));
  auto lantern = [&](int lx, int lz, uint32_t mat) {  // scene.js:396-401
    const int g = L.groundAt(lx, lz);
    const int h = V(1.1);
    for (int y = g + 1; y <= g + h; ++y) vol.set(lx, y, lz, M::lampPost);
    vol.boxm(lx - 1, g + h + 1, lz - 1, lx + 1, g + h + 3, lz + 1, mat);
  };
  const bool lamps = variant == 1;
  lantern((int)std::floor(N * 0.35), (int)std::floor(N * 0.25), lamps ? M::lampWarmDim : M::lampWarm);
  lantern((int)std::floor(N * 0.68), (int)std::floor(N * 0.30), lamps ? M::lampCyanDim : M::lampCyan);
  lantern((int)std::floor(N * 0.25), (int)std::floor(N * 0.65), lamps ? M::lampMagentaDim : M::lampMagenta);
  if (!lamps) vol.boxm(x0 + 2, floorY + 1, z1 - 4, x0 + 4, floorY + 3, z1 - 2, M::lampWarm);
  return L;
}

// ---------------------------------------------------------------------------
// VRT_SCENE_WORLD: seeded terrain + Cornell room + boxes + trees + lanterns
// ---------------------------------------------------------------------------
struct WorldLayout {
  int rx0, rx1, rz0, rz1, floorY, y1, R, RH, cx, cz;
  bool room;
};

struct ValueNoise {  // seeded 2D value noise, 4 octaves
  uint64_t seed;
  double lattice(int64_t x, int64_t y) const {
    return u01(mix64(seed ^ mix64((uint64_t)x * 0x9E3779B1ull + (uint64_t)y * 0x85EBCA77ull)));
  }
  double noise2(double x, double y) const {
    const double xi = std::floor(x), yi = std::floor(y);
    double xf = x - xi, yf = y - yi;
    xf = xf * xf * (3 - 2 * xf);
    yf = yf * yf * (3 - 2 * yf);
    const double a = lattice((int64_t)xi, (int64_t)yi), b = lattice((int64_t)xi + 1, (int64_t)yi);
    const double c = lattice((int64_t)xi, (int64_t)yi + 1), d = lattice((int64_t)xi + 1, (int64_t)yi + 1);
    return a + (b - a) * xf + (c - a) * yf + (a - b - c + d) * xf * yf;
  }
  double operator()(double x, double y) const {  // in [0, 1)
    double v = 0, amp = 0.5, f = 1, norm = 0;
    for (int o = 0; o < 4; ++o) {
      v += noise2(x * f, y * f) * amp;
      norm += amp;
      amp *= 0.5;
      f *= 2;
    }
    return v / norm;
  }
};

WorldLayout world_generate(const vrt_scene_desc& d, uint32_t* vox) {
  const int nx = (int)d.nx, ny = (int)d.ny, nz = (int)d.nz;
  Vol vol{vox, nx, ny, nz};
  if (vox) std::memset(vox, 0, (size_t)nx * ny * nz * 4);
  ValueNoise noise{mix64(d.seed + 17)};
  const double freq = 6.0 / std::max(nx, nz);  // ~6 hills across the map
  const float hmin = d.h_min > 0 ? d.h_min : 0.0625f * ny;
  const float hmax = d.h_max > d.h_min ? d.h_max : 0.3125f * ny;
  std::vector<int32_t> heights((size_t)nx * nz);
  for (int z = 0; z < nz; ++z)
    for (int x = 0; x < nx; ++x) {
      const double n = noise(x * freq, z * freq);
      int h = (int)std::floor(hmin + n * (hmax - hmin + 1.0));
      h = std::max(0, std::min(ny - 1, std::min(h, (int)hmax)));
      heights[(size_t)x + (size_t)z * nx] = h;
      if (!vox) continue;
      // green albedo 0.2 .. 0.5 varying with a second noise field
      const double a = 0.2 + 0.3 * noise(x * freq * 7.0 + 101.0, z * freq * 7.0 - 57.0);
      const uint32_t m = pack(alb_byte(0.45 * a), alb_byte(a), alb_byte(0.25 * a));
      for (int y = 0; y <= h; ++y) vol.set(x, y, z, m);
    }
  auto groundAt = [&](int x, int z) {
    x = std::max(0, std::min(nx - 1, x));
    z = std::max(0, std::min(nz - 1, z));
    return heights[(size_t)x + (size_t)z * nx];
  };
  WorldLayout L{};
  L.R = (int)d.room;
  L.RH = d.room_h ? (int)d.room_h : (int)d.room;
  L.room = d.room > 0;
  L.cx = nx / 2;
  L.cz = (int)(nz * 0.55);
  const uint32_t white = pack(alb_byte(0.75), alb_byte(0.75), alb_byte(0.75));
  const uint32_t red = pack(alb_byte(0.63), alb_byte(0.065), alb_byte(0.05));
  const uint32_t greenw = pack(alb_byte(0.14), alb_byte(0.45), alb_byte(0.091));
  const uint32_t panel = pack(255, 240, 220, 200);
  if (L.room) {
    const int R = L.R, T = 2;
    L.rx0 = L.cx - R / 2; L.rx1 = L.rx0 + R - 1;
    L.rz0 = L.cz - R / 2; L.rz1 = L.rz0 + R - 1;
    L.floorY = groundAt(L.cx, L.cz) + 1;
    L.y1 = L.floorY + L.RH;  // ceiling layer
    if (vox) {
      const int x0 = L.rx0, x1 = L.rx1, z0 = L.rz0, z1 = L.rz1, fy = L.floorY, y1 = L.y1;
      vol.boxm(x0 - T, fy - 1, z0 - T, x1 + T, fy - 1, z1 + T, white);                // floor slab
      vol.boxm(x0 - T, fy, z0 - T, x1 + T, std::min(ny - 1, y1 + T + 6), z1 + T, 0u); // carve
      vol.boxm(x0 - T, fy, z1 + 1, x1 + T, y1, z1 + T, white);                        // back
      vol.boxm(x0 - T, fy, z0 - T, x1 + T, y1, z0 - 1, white);                        // front
      vol.boxm(x0 - T, fy, z0 - T, x0 - 1, y1, z1 + T, red);                          // left
      vol.boxm(x1 + 1, fy, z0 - T, x1 + T, y1, z1 + T, greenw);                       // right
      vol.boxm(x0 - T, y1, z0 - T, x1 + T, y1 + T - 1, z1 + T, white);                // ceiling
      const int door = std::max(2, R * 7 / 25);                                       // doorway
      vol.boxm(L.cx - door, fy, z0 - T, L.cx + door, std::max(fy, y1 - 1 - R / 8), z0 - 1, 0u);
      const int ph = std::max(1, R / 6);                                              // light panel
      vol.boxm(L.cx - ph, y1, L.cz - ph, L.cx + ph, y1, L.cz + ph, panel);
      // boxes: tall then short (scene.js:369-374 proportions), more alternate
      for (uint32_t b = 0; b < d.boxes; ++b) {
        const bool tall = (b % 2) == 0;
        const int bw = std::max(2, tall ? R * 4 / 25 : R * 9 / 50);
        const int bh = std::max(2, tall ? (L.RH * 11) / 20 : (L.RH * 7) / 25);
        const int off = (int)(b / 2) * (bw + 2);
        const int bx = tall ? L.cx - R * 7 / 25 - bw + off : L.cx + R / 10 + off;
        const int bz = tall ? L.cz + R / 20 : L.cz - R / 4;
        vol.boxm(bx, fy, bz, bx + bw, fy + bh - 1, bz + bw,
                 tall ? pack(alb_byte(0.75), alb_byte(0.75), alb_byte(0.8)) : pack(210, 170, 90));
      }
    }
  }
  Rng rng(d.seed * 0x2545F4914F6CDD1Dull + 7);
  auto in_room = [&](int x, int z, int pad) {
    return L.room && x >= L.rx0 - 2 - pad && x <= L.rx1 + 2 + pad && z >= L.rz0 - 2 - pad && z <= L.rz1 + 2 + pad;
  };
  const double s = std::max(nx, nz) / 256.0;  // object scale
  const uint32_t trunk = pack(95, 68, 42), post = pack(60, 60, 65);
  for (uint32_t t = 0; t < d.trees; ++t) {
    int tx = 0, tz = 0;
    for (int tries = 0; tries < 64; ++tries) {
      tx = rng.range(4, nx - 5);
      tz = rng.range(4, nz - 5);
      if (!in_room(tx, tz, (int)(14 * s) + 2)) break;
    }
    if (!vox) continue;
    const int g = groundAt(tx, tz);
    const int trunkH = std::max(3, (int)((22 + rng.range(0, 12)) * s));
    for (int y = g + 1; y <= std::min(ny - 1, g + trunkH); ++y) {
      vol.set(tx, y, tz, trunk);
      vol.set(tx + 1, y, tz, trunk);
    }
    const int cy = g + trunkH + std::max(1, (int)(6 * s));
    const int r = std::max(2, (int)((11 + rng.range(0, 2)) * s));
    for (int dz = -r; dz <= r; ++dz)
      for (int dy = -r; dy <= r; ++dy)
        for (int dx = -r; dx <= r; ++dx) {
          const double d2 = dx * dx + dy * dy * 1.6 + dz * dz;
          if (d2 <= r * r && ((dx * 13 + dy * 7 + dz * 11) % 5 != 0 || d2 < r * r * 0.5)) {
            const int v = (dx * 3 + dy * 5 + dz * 7) % 12 + ((dx + dy + dz) % 2) * 6;
            vol.set(tx + dx, cy + dy, tz + dz, pack(50 + v, 120 + v, 45));
          }
        }
  }
  const uint32_t lamp_mats[3] = {pack(255, 190, 110, 255), pack(90, 220, 255, 255), pack(255, 90, 200, 255)};
  for (uint32_t l = 0; l < d.lanterns; ++l) {
    int lx = 0, lz = 0;
    for (int tries = 0; tries < 64; ++tries) {
      lx = rng.range(3, nx - 4);
      lz = rng.range(3, nz - 4);
      if (!in_room(lx, lz, 3)) break;
    }
    if (!vox) continue;
    const int g = groundAt(lx, lz);
    const int h = std::max(3, (int)(18 * s));
    for (int y = g + 1; y <= std::min(ny - 1, g + h); ++y) vol.set(lx, y, lz, post);
    vol.boxm(lx - 1, g + h + 1, lz - 1, lx + 1, g + h + 3, lz + 1, lamp_mats[l % 3]);
  }
  return L;
}

void random_bricks(const vrt_scene_desc& d, uint32_t* vox) {
  const int nx = (int)d.nx, ny = (int)d.ny, nz = (int)d.nz;
  const int bgx = nx / 8, bgy = ny / 8, bgz = nz / 8;
  std::memset(vox, 0, (size_t)nx * ny * nz * 4);
  for (int bz = 0; bz < bgz; ++bz)
    for (int by = 0; by < bgy; ++by)
      for (int bx = 0; bx < bgx; ++bx) {
        const uint64_t bi = (uint64_t)bx + (uint64_t)by * bgx + (uint64_t)bz * bgx * bgy;
        if (!(u01(mix64(d.seed * 0x9E3779B97F4A7C15ull ^ (bi * 2 + 1))) < d.p_brick)) continue;
        for (int l = 0; l < 512; ++l) {
          const uint64_t h = mix64((d.seed + 0x632BE59BD9B4E019ull) ^ (bi * 512 + (uint64_t)l) * 0xD1B54A32D192ED03ull);
          if (!(u01(h) < d.p_voxel)) continue;
          const int x = bx * 8 + (l & 7), y = by * 8 + ((l >> 3) & 7), z = bz * 8 + (l >> 6);
          vox[(uint64_t)x + (uint64_t)y * nx + (uint64_t)z * nx * ny] = ((uint32_t)(h >> 32) & 0x00FFFFFFu) | 1u;
        }
      }
}

bool desc_ok(const vrt_scene_desc* d) {
  if (!d) return false;
  if (d->kind == VRT_SCENE_VOXELRT) return d->nx >= 8 && d->nx % 8 == 0 && d->nx <= 1024;
  return d->nx >= 8 && d->ny >= 8 && d->nz >= 8 && d->nx % 8 == 0 && d->ny % 8 == 0 &&
         d->nz % 8 == 0 && (uint64_t)d->nx * d->ny * d->nz <= (1ull << 31);
}

}  // namespace

extern "C" {

void vrt_camera_matrices(const double pos[3], double yaw, double pitch, double fov_y, double aspect,
                         double near_z, double far_z, float view_proj[16], float inv_view_proj[16]) {
  const Basis b = camera_basis(yaw, pitch);
  const double z[3] = {-b.fw[0], -b.fw[1], -b.fw[2]};
  auto dotd = [](const double* a, const double* c) { return a[0] * c[0] + a[1] * c[1] + a[2] * c[2]; };
  float proj[16] = {0}, iproj[16] = {0}, c2w[16], w2c[16];
  const double f = 1.0 / std::tan(fov_y / 2.0);  // math.js:73-93
  proj[0] = (float)(f / aspect);
  proj[5] = (float)f;
  proj[10] = (float)(far_z / (near_z - far_z));
  proj[11] = -1.0f;
  proj[14] = (float)((near_z * far_z) / (near_z - far_z));
  iproj[0] = (float)(aspect / f);
  iproj[5] = (float)(1.0 / f);
  iproj[11] = (float)((near_z - far_z) / (near_z * far_z));
  iproj[14] = -1.0f;
  iproj[15] = (float)(1.0 / near_z);
  const double c2wd[16] = {b.rt[0], b.rt[1], b.rt[2], 0, b.up[0], b.up[1], b.up[2], 0,
                           -b.fw[0], -b.fw[1], -b.fw[2], 0, pos[0], pos[1], pos[2], 1};
  const double w2cd[16] = {b.rt[0], b.up[0], z[0], 0, b.rt[1], b.up[1], z[1], 0, b.rt[2], b.up[2], z[2], 0,
                           -dotd(b.rt, pos), -dotd(b.up, pos), -dotd(z, pos), 1};
  for (int i = 0; i < 16; ++i) {
    c2w[i] = (float)c2wd[i];
    w2c[i] = (float)w2cd[i];
  }
  mat4_mul(proj, w2c, view_proj);
  mat4_mul(c2w, iproj, inv_view_proj);
}

void vrt_uniforms_init(vrt_uniforms* u, const double pos[3], double yaw, double pitch,
                       const float prev_view_proj[16], const double prev_cam_pos[3], uint32_t width,
                       uint32_t height, uint32_t frame, uint32_t bounces, uint32_t flags,
                       double sun_azimuth, double sun_elevation, double sun_cos_radius,
                       float max_history) {
  std::memset(u, 0, sizeof(*u));
  float vp[16], ivp[16];
  vrt_camera_matrices(pos, yaw, pitch, 70.0 * M_PI / 180.0, (double)width / (double)height, 0.05, 400.0, vp, ivp);
  std::memcpy(u->inv_view_proj, ivp, sizeof ivp);
  std::memcpy(u->prev_view_proj, prev_view_proj ? prev_view_proj : vp, sizeof vp);
  const double* pp = prev_cam_pos ? prev_cam_pos : pos;
  for (int i = 0; i < 3; ++i) {
    u->cam_pos[i] = (float)pos[i];
    u->prev_cam_pos[i] = (float)pp[i];
  }
  double sd[3] = {std::cos(sun_elevation) * std::cos(sun_azimuth), std::sin(sun_elevation),
                  std::cos(sun_elevation) * std::sin(sun_azimuth)};
  const double l = std::sqrt(sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2]);
  for (int i = 0; i < 3; ++i) u->sun_dir[i] = (float)(sd[i] / l);
  u->sun_dir[3] = (float)sun_cos_radius;
  u->sun_radiance[0] = 5.0f;  // main.js:412
  u->sun_radiance[1] = 4.5f;
  u->sun_radiance[2] = 3.8f;
  u->sun_radiance[3] = 1.0f;
  u->params0[0] = frame;
  u->params0[1] = bounces;
  u->params0[2] = flags;
  u->params1[0] = (float)width;
  u->params1[1] = (float)height;
  u->params1[2] = 1.0f;  // exposure
  u->params4[0] = 30.0f;
  u->params4[1] = max_history;
  u->params4[2] = 24.0f;
}

int vrt_generate_scene(const vrt_scene_desc* desc, uint32_t* vox) {
  if (!desc_ok(desc) || !vox) return VRT_ERR_INVALID;
  try {
    switch (desc->kind) {
      case VRT_SCENE_RANDOM_BRICKS: random_bricks(*desc, vox); return VRT_OK;
      case VRT_SCENE_WORLD: world_generate(*desc, vox); return VRT_OK;
      case VRT_SCENE_VOXELRT: voxelrt_generate((int)desc->nx, (int)desc->variant, vox); return VRT_OK;
      default: return VRT_ERR_INVALID;
    }
  } catch (...) {
    return VRT_ERR_NO_MEMORY;
  }
}

int vrt_scene_spawn(const vrt_scene_desc* desc, double pos[3], double* yaw, double* pitch) {
  if (!desc_ok(desc) || !pos || !yaw || !pitch) return VRT_ERR_INVALID;
  try {
    if (desc->kind == VRT_SCENE_VOXELRT) {  // scene.js:463-472
      VoxelrtLayout L = voxelrt_generate((int)desc->nx, (int)desc->variant, nullptr);
      pos[0] = L.cx * 0.0625;
      pos[1] = (L.groundAt(L.cx, L.z0 - V(3)) + 1) * 0.0625 + 1.65;
      pos[2] = (L.z0 - V(3.5)) * 0.0625;
      *yaw = 0.0;
      *pitch = -0.05;
      return VRT_OK;
    }
    if (desc->kind == VRT_SCENE_WORLD) {
      WorldLayout L = world_generate(*desc, nullptr);
      if (L.room) {
        pos[0] = (L.cx + 0.5) * 0.0625;
        pos[1] = (L.floorY + L.RH * 0.45) * 0.0625;
        pos[2] = (L.rz0 - L.R * 0.9) * 0.0625;
      } else {
        pos[0] = desc->nx * 0.5 * 0.0625;
        pos[1] = desc->ny * 0.6 * 0.0625;
        pos[2] = desc->nz * 0.1 * 0.0625;
      }
      *yaw = 0.0;
      *pitch = -0.05;
      return VRT_OK;
    }
    pos[0] = desc->nx * 0.5 * 0.0625;
    pos[1] = desc->ny * 0.5 * 0.0625;
    pos[2] = -desc->nz * 0.25 * 0.0625;
    *yaw = 0.0;
    *pitch = 0.0;
    return VRT_OK;
  } catch (...) {
    return VRT_ERR_NO_MEMORY;
  }
}

}  // extern "C"
