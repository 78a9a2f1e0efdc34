// Test program of the SOP cook driver (mpmsnow_amd/cpp/SnowCook.hpp, after the reference's
// houdiniPlugin/SOP_MPMSim.cpp:138-372).
//   test_cook seed <out>   seeding only (no GPU): writes the seeded positions
//   test_cook cook <out>   on the GPU: cooks the start frame and one frame of a snowball falling
//                          onto a ground level set, checks the SOP's control flow, writes the
//                          positions after the frame
// <out>: int32 count, then count x 3 float32 positions.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "SnowCook.hpp"

using MpmSim::LevelSetCollisionObject;
using MpmSim::Vector3f;

namespace {

int failures = 0;
#define CHECK(c)                                                       \
  do {                                                                 \
    if (!(c)) {                                                        \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);         \
      ++failures;                                                      \
    }                                                                  \
  } while (0)

// the scene (the Python side restates it): a sphere of radius 0.1 at (0.5, 0.3, 0.5) sampled
// on a lattice of voxel 1/64 over [0.35, 0.65] x [0.15, 0.45] x [0.35, 0.65], with the velocity
// field (0.2 (z - 0.5), -1, 0); and a ground at y = 0.18 on a lattice of voxel 1/16 over
// [0, 1] x [0, 0.5] x [0, 1]
std::shared_ptr<LevelSetCollisionObject> snowball() {
  const int d[3] = {20, 20, 20};
  const float vox = 1.0f / 64.0f;
  const Vector3f o(0.35f, 0.15f, 0.35f);
  std::vector<float> phi((size_t)d[0] * d[1] * d[2]), vel(3 * phi.size());
  for (int k = 0; k < d[2]; ++k)
    for (int j = 0; j < d[1]; ++j)
      for (int i = 0; i < d[0]; ++i) {
        const size_t q = (size_t)i + d[0] * ((size_t)j + d[1] * k);
        const float x = o[0] + vox * i, y = o[1] + vox * j, z = o[2] + vox * k;
        const float dx = x - 0.5f, dy = y - 0.3f, dz = z - 0.5f;
        phi[q] = std::sqrt(dx * dx + dy * dy + dz * dz) - 0.1f;
        vel[3 * q] = 0.2f * (z - 0.5f);
        vel[3 * q + 1] = -1.0f;
        vel[3 * q + 2] = 0.0f;
      }
  return std::make_shared<LevelSetCollisionObject>(d, o, vox, phi, vel);
}
std::shared_ptr<LevelSetCollisionObject> ground() {
  const int d[3] = {17, 9, 17};
  const float vox = 1.0f / 16.0f;
  const Vector3f o(0.0f, 0.0f, 0.0f);
  std::vector<float> phi((size_t)d[0] * d[1] * d[2]);
  for (int k = 0; k < d[2]; ++k)
    for (int j = 0; j < d[1]; ++j)
      for (int i = 0; i < d[0]; ++i) phi[(size_t)i + d[0] * ((size_t)j + d[1] * k)] = vox * j - 0.18f;
  return std::make_shared<LevelSetCollisionObject>(d, o, vox, phi);
}

MpmSimCook::CookParameters params() {
  MpmSimCook::CookParameters p;
  p.gridSize = 1.0f / 32.0f;
  p.subSteps = 40;
  p.startTime = 0.0;
  p.tolerance = 1e-5f;
  p.maxIterations = 200;
  return p;
}

void write(const char* path, const std::vector<Vector3f>& x) {
  FILE* f = std::fopen(path, "wb");
  const int n = (int)x.size();
  std::fwrite(&n, 4, 1, f);
  for (const Vector3f& p : x) std::fwrite(p.v, 4, 3, f);
  std::fclose(f);
}

struct CountingInterrupt : MpmSimCook::Interrupt {
  int calls = 0, fireAt;
  explicit CountingInterrupt(int n) : fireAt(n) {}
  bool opInterrupt() override { return ++calls >= fireAt; }
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::printf("usage: test_cook seed|cook <out>\n");
    return 2;
  }
  const std::string mode = argv[1];
  if (mode == "seed") {
    std::vector<Vector3f> x;
    std::vector<float> m;
    MpmSimCook::seedParticles(*snowball(), params().gridSize / 2.0f, x, m);
    CHECK(!x.empty() && m.size() == x.size());
    const float h = params().gridSize / 2.0f;
    for (float mi : m) CHECK(mi == 400.0f * h * h * h);
    write(argv[2], x);
    std::printf("seeded %d particles\n", (int)x.size());
  } else if (mode == "cook") {
    const double frame = 1.0 / 24.0;
    MpmSimCook::SnowCook cook(params(), snowball());
    cook.setColliders({ground()});
    // the start frame: a fresh simulation, no step
    std::vector<Vector3f> x0 = cook.cook(0.0);
    CHECK(cook.hasSim() && cook.stepsTaken() == 0);
    std::vector<Vector3f> seeds;
    std::vector<float> m;
    MpmSimCook::seedParticles(*snowball(), params().gridSize / 2.0f, seeds, m);
    CHECK(x0.size() == seeds.size());
    for (size_t i = 0; i < x0.size(); ++i)
      for (int a = 0; a < 3; ++a) CHECK(x0[i][a] == seeds[i][a]);
    // one frame: 40 substeps of frame / 40
    std::vector<Vector3f> x1 = cook.cook(frame);
    CHECK(cook.stepsTaken() == 40 && cook.errors().empty());
    float ymin = 1e9f;
    for (const Vector3f& p : x1) ymin = std::fmin(ymin, p[1]);
    CHECK(ymin < 0.2f);  // the ball reached the ground during the frame
    write(argv[2], x1);
    // a cook at or before the previous cook time is the SOP's error
    bool threw = false;
    try {
      cook.cook(frame);
    } catch (const std::exception&) {
      threw = true;
    }
    CHECK(threw);
    // the start time again restarts the simulation at the seeds
    std::vector<Vector3f> xr = cook.cook(0.0);
    CHECK(xr.size() == seeds.size() && xr[0][1] == seeds[0][1]);
    // an interrupt stops the substeps early; the positions stay finite
    MpmSimCook::SnowCook cook2(params(), snowball());
    cook2.setColliders({ground()});
    cook2.cook(0.0);
    CountingInterrupt boss(12);
    std::vector<Vector3f> xi = cook2.cook(frame, &boss);
    CHECK(cook2.stepsTaken() > 0 && cook2.stepsTaken() < 40);
    for (const Vector3f& p : xi) CHECK(std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]));
    std::printf("cooked frame: %d particles, lowest y %.4f; interrupted cook took %d substeps\n", (int)x1.size(),
                ymin, cook2.stepsTaken());
  } else {
    std::printf("unknown mode %s\n", mode.c_str());
    return 2;
  }
  if (failures) {
    std::printf("%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("ok %s\n", mode.c_str());
  return 0;
}
