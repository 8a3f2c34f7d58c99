// The full reference call chain through the drop-in classes:
// GenerateAABB -> detectPossibleCollisions -> DoNarrowphase, on a seeded
// scene of mixed shapes, then host_MPR_Store on the same pairs.  Inputs and
// outputs are written to argv[1] so the test driver can compare them with
// the CPU oracle.
#include <cmath>
#include <cstdio>
#include "collision_detection.h"

static uint64_t st = 0x9E3779B97F4A7C15ull;
static double urand() {  // splitmix64 -> [0,1)
  uint64_t z = (st += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (z >> 11) * (1.0 / 9007199254740992.0);
}

template <class T> static void put(FILE* f, const std::vector<T>& v) {
  uint64_t n = v.size();
  fwrite(&n, 8, 1, f);
  if (n) fwrite(v.data(), sizeof(T), n, f);
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const int n = argc >= 3 ? atoi(argv[2]) : 6000;
  const shape_type kinds[5] = {SPHERE, BOX, ELLIPSOID, CAPSULE, CYLINDER};
  custom_vector<shape_type> T(n);
  custom_vector<real3> A(n), B(n), C(n);
  custom_vector<real4> R(n);
  custom_vector<uint> ID(n);
  custom_vector<real3> pos(n);
  custom_vector<real4> rot(n);
  custom_vector<bool> active(n);
  double L = 0.22 * std::cbrt((double)n / 1000.0);
  for (int i = 0; i < n; ++i) {
    T[i] = kinds[(int)(urand() * 5) % 5];
    double s = 0.01 * (0.6 + 0.8 * urand());
    B[i] = R3(s, s * (0.5 + urand()), s * (0.5 + urand()));
    if (T[i] == SPHERE) B[i] = R3(s, 0, 0);
    A[i] = R3(0.002 * (urand() - 0.5), 0.002 * (urand() - 0.5), 0);
    double a = urand() * 6.283185307179586, c = std::cos(a / 2), sn = std::sin(a / 2);
    R[i] = R4(c, 0, sn, 0);
    ID[i] = i;
    pos[i] = R3(L * urand(), L * urand(), L * urand());
    double q[4] = {urand() - 0.5, urand() - 0.5, urand() - 0.5, urand() - 0.5};
    double qn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    rot[i] = R4(q[0] / qn, q[1] / qn, q[2] / qn, q[3] / qn);
    active[i] = (i % 7) != 3;
  }
  AABBGenerator gen;
  custom_vector<real3> aabb;
  gen.GenerateAABB(T, A, B, C, R, ID, pos, rot, aabb);
  custom_vector<real3> aabb_unshifted = aabb;
  Broadphase bp;
  bp.setBinsPerAxis(I3(12, 12, 12));
  custom_vector<long long> pairs;
  bp.detectPossibleCollisions(aabb, pairs);
  custom_vector<long long> pairs_sorted = pairs;
  Narrowphase np;
  np.SetCollisionEnvelope(0.0005);
  custom_vector<real3> norm, cpta, cptb;
  custom_vector<real> dpth;
  custom_vector<int2> bids;
  uint nc = 0;
  np.DoNarrowphase(T, A, B, C, R, ID, active, pos, rot, pairs, norm, cpta, cptb, dpth, bids, nc);
  // Narrowphase::host_MPR_Store over the unpermuted pair list: per-pair
  // outputs at the pair index, contact_active cleared for contacts only.
  const size_t P = pairs_sorted.size();
  custom_vector<uint> cact(P, 1);
  custom_vector<real3> pn(P, R3(NAN)), ppa(P, R3(NAN)), ppb(P, R3(NAN));
  custom_vector<real> pd(P, NAN);
  custom_vector<int2> pid(P, I2(-1, -1));
  np.host_MPR_Store(T.data(), A.data(), B.data(), C.data(), R.data(), ID.data(), active.data(), pos.data(), rot.data(),
                    pairs_sorted.data(), cact.data(), pn.data(), ppa.data(), ppb.data(), pd.data(), pid.data());
  FILE* f = fopen(argv[1], "wb");
  put(f, T); put(f, A); put(f, B); put(f, R); put(f, ID); put(f, pos); put(f, rot); put(f, active);
  put(f, aabb_unshifted); put(f, aabb); put(f, pairs_sorted); put(f, pairs);
  put(f, norm); put(f, cpta); put(f, cptb); put(f, dpth); put(f, bids);
  put(f, cact); put(f, pn); put(f, ppa); put(f, ppb); put(f, pd); put(f, pid);
  fclose(f);
  printf("%d shapes, %zu pairs, %u contacts\n", n, pairs.size(), nc);
  return 0;
}
