// Mirror of the reference's test/test1.cu: N identical AABBs [0,1]^3 on a
// 1x1x1 grid must give N(N-1)/2 candidate pairs.
#include <cassert>
#include <cstdlib>
#include "collision_detection.h"

int main(int argc, char** argv) {
  int numObjects = 10;
  if (argc == 2) numObjects = atoi(argv[1]);
  custom_vector<real3> aabb_data;
  custom_vector<long long> potentialCollisions;
  int numObjectsFactorial = 0;
  for (int i = 0; i < numObjects; i++) {
    aabb_data.push_back(make_real3(0, 0, 0));
    numObjectsFactorial += i;
  }
  for (int i = 0; i < numObjects; i++) aabb_data.push_back(make_real3(1, 1, 1));
  Broadphase broadphaseManager;
  broadphaseManager.setBinsPerAxis(I3(1, 1, 1));
  broadphaseManager.detectPossibleCollisions(aabb_data, potentialCollisions);
  if ((int)potentialCollisions.size() != numObjectsFactorial) {
    std::cout << "Test Failed: " << potentialCollisions.size() << " != " << numObjectsFactorial << std::endl;
    return 1;
  }
  std::cout << "Test Passed!" << std::endl;
  return 0;
}
