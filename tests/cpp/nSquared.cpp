// Mirror of the reference's test/nSquared_Broadphase.cu: every true sphere
// contact (dist^2 < (rA+rB)^2) of 300,000 glibc-rand() spheres in
// [-30,30]^3 (r in [0.1,1], never seeded) is a candidate pair on a 30^3
// grid.  The true contacts come from a uniform-grid search instead of the
// reference's O(N^2) loop; the membership check is a binary search.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <unordered_map>
#include "collision_detection.h"

static double getRandomNumber(double min, double max) {
  double x = rand() / static_cast<double>(RAND_MAX);
  return min + (x * (max - min));
}

int main(int argc, char** argv) {
  uint numSpheres = 300000;
  if (argc >= 2) numSpheres = atoi(argv[1]);
  std::vector<real3> pos(numSpheres);
  std::vector<double> rad(numSpheres);
  for (uint i = 0; i < numSpheres; i++) {
    double x = getRandomNumber(-30, 30), y = getRandomNumber(-30, 30), z = getRandomNumber(-30, 30);
    pos[i] = make_real3(x, y, z);
    rad[i] = getRandomNumber(.1, 1);
  }
  custom_vector<real3> aabb_data;
  for (uint i = 0; i < numSpheres; i++) aabb_data.push_back(R3(pos[i].x - rad[i], pos[i].y - rad[i], pos[i].z - rad[i]));
  for (uint i = 0; i < numSpheres; i++) aabb_data.push_back(R3(pos[i].x + rad[i], pos[i].y + rad[i], pos[i].z + rad[i]));
  custom_vector<long long> potentialCollisions;
  Broadphase broadphaseManager;
  broadphaseManager.setBinsPerAxis(I3(30, 30, 30));
  broadphaseManager.detectPossibleCollisions(aabb_data, potentialCollisions);
  std::cout << broadphaseManager.getNumPossibleContacts() << " possible collisions" << std::endl;
  if (!std::is_sorted(potentialCollisions.begin(), potentialCollisions.end())) {
    std::cout << "TEST FAILED (pair list not sorted)" << std::endl;
    return 1;
  }
  // true contacts by cell hashing (cell size 2 >= max diameter)
  std::unordered_map<long long, std::vector<uint>> cells;
  auto key = [](int a, int b, int c) { return ((long long)(a + 64) << 32) | ((long long)(b + 64) << 16) | (c + 64); };
  for (uint i = 0; i < numSpheres; i++)
    cells[key((int)floor(pos[i].x / 2), (int)floor(pos[i].y / 2), (int)floor(pos[i].z / 2))].push_back(i);
  size_t truth = 0, missed = 0;
  for (uint i = 0; i < numSpheres; i++) {
    int cx = (int)floor(pos[i].x / 2), cy = (int)floor(pos[i].y / 2), cz = (int)floor(pos[i].z / 2);
    for (int dx = -1; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dz = -1; dz <= 1; dz++) {
          auto it = cells.find(key(cx + dx, cy + dy, cz + dz));
          if (it == cells.end()) continue;
          for (uint j : it->second) {
            if (j <= i) continue;
            double ddx = pos[i].x - pos[j].x, ddy = pos[i].y - pos[j].y, ddz = pos[i].z - pos[j].z;
            double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
            if (d2 < (rad[i] + rad[j]) * (rad[i] + rad[j])) {
              ++truth;
              long long p = ((long long)i << 32) | j;
              if (!std::binary_search(potentialCollisions.begin(), potentialCollisions.end(), p)) ++missed;
            }
          }
        }
  }
  std::cout << truth << " actual collisions" << std::endl;
  if (missed) {
    std::cout << "TEST FAILED (" << missed << " missed collisions)" << std::endl;
    return 1;
  }
  std::cout << "TEST PASSED" << std::endl;
  return 0;
}
