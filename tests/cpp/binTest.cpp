// Mirror of the reference's test/binTest_Broadphase.cu: two identical
// spheres at every site of an n^3 lattice give exactly n^3 pairs.
#include <cmath>
#include <cstdlib>
#include "collision_detection.h"

struct Sphere {
  double r;
  real3 pos;
  uint index;
};

int main(int argc, char** argv) {
  uint numSpheresPerSide = 10, bpa = 10;
  if (argc >= 2) numSpheresPerSide = atoi(argv[1]);
  if (argc >= 3) bpa = atoi(argv[2]);
  double sphereRad = 0.25, sphereSpacing = 1;
  std::vector<Sphere> spheres;
  Sphere sphere;
  int index = 0;
  for (uint k = 0; k < numSpheresPerSide; k++)
    for (uint j = 0; j < numSpheresPerSide; j++)
      for (uint i = 0; i < numSpheresPerSide; i++) {
        sphere.pos = make_real3(i * sphereSpacing, j * sphereSpacing, k * sphereSpacing);
        sphere.r = sphereRad;
        sphere.index = index++;
        spheres.push_back(sphere);
        sphere.index = index++;
        spheres.push_back(sphere);
      }
  custom_vector<real3> aabb_data;
  for (auto& s : spheres) aabb_data.push_back(make_real3(s.pos.x - s.r, s.pos.y - s.r, s.pos.z - s.r));
  for (auto& s : spheres) aabb_data.push_back(make_real3(s.pos.x + s.r, s.pos.y + s.r, s.pos.z + s.r));
  custom_vector<long long> potentialCollisions;
  Broadphase broadphaseManager;
  broadphaseManager.setBinsPerAxis(I3(bpa, bpa, bpa));
  broadphaseManager.detectPossibleCollisions(aabb_data, potentialCollisions);
  std::cout << broadphaseManager.getNumPossibleContacts() << " possible collisions" << std::endl;
  bool ok = broadphaseManager.getNumPossibleContacts() == pow(numSpheresPerSide, 3);
  for (size_t i = 0; ok && i < potentialCollisions.size(); ++i)
    ok = potentialCollisions[i] == (((long long)(2 * i) << 32) | (long long)(2 * i + 1));
  std::cout << (ok ? "TEST PASSED" : "TEST FAILED") << std::endl;
  return ok ? 0 : 1;
}
