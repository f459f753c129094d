// Seeded synthetic scenes. The reference has no random scene generator; its
// benchmark scenes are built through the construction graph
// (crates/mechanic-bench/src/main.rs:442-524). These rows follow the shape its
// compile step gives a single-part compound (crates/mechanic-core/src/compile.rs:
// root at the centre of mass with identity rotation 333-334, collider carrying
// the part rotation 889-895, cuboid mass and inertia 841-853).
#include <cmath>
#include <cstring>

#include "../../include/mechanic_hip.h"

namespace {

struct SplitMix64 {
  uint64_t state;
  uint64_t next() {
    state += 0x9e3779b97f4a7c15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  // Uniform in [0, 1) with 24 random bits: (x >> 40) * 2^-24.
  double uniform() { return (double)(next() >> 40) * (1.0 / 16777216.0); }
};

}  // namespace

extern "C" int mh_scene_cuboid_pile(const mh_pile_params* prm, mh_transform* poses, float* inverse_masses,
                                    mh_mass* masses, mh_spatial_inertia* inertias, mh_collider* colliders,
                                    size_t capacity) {
  if (!prm) return MH_ERR_INVALID_ARGUMENT;
  const uint64_t count = (uint64_t)prm->nx * prm->ny * prm->nz;
  if (count > capacity || !poses || !inverse_masses || !masses || !colliders) return MH_ERR_INVALID_ARGUMENT;
  if (!(prm->edge > 0.0f) || !(prm->density > 0.0f)) return MH_ERR_INVALID_ARGUMENT;
  SplitMix64 rng{prm->seed};
  const float e = prm->edge;
  // Cuboid mass and principal inertia, f32 as compile.rs:843-851 evaluates them.
  const float mass = prm->density * e * e * e;
  const float inertia = mass * (e * e + e * e) / 12.0f;
  const float inv_mass = 1.0f / mass;
  const float inv_inertia = 1.0f / inertia;
  const double half_x = 0.5 * (double)(prm->nx - 1), half_z = 0.5 * (double)(prm->nz - 1);
  const double two_pi = 6.283185307179586476925286766559;
  uint64_t body = 0;
  for (uint32_t iy = 0; iy < prm->ny; ++iy) {
    for (uint32_t iz = 0; iz < prm->nz; ++iz) {
      for (uint32_t ix = 0; ix < prm->nx; ++ix, ++body) {
        const double jx = (2.0 * rng.uniform() - 1.0) * prm->jitter;
        const double jy = (2.0 * rng.uniform() - 1.0) * prm->jitter;
        const double jz = (2.0 * rng.uniform() - 1.0) * prm->jitter;
        const double yaw = prm->yaw ? rng.uniform() * two_pi : 0.0;
        mh_transform& t = poses[body];
        t.position[0] = (float)(((double)ix - half_x) * prm->spacing + jx);
        t.position[1] = (float)((double)prm->drop_height + 0.5 * e + (double)iy * prm->spacing + jy);
        t.position[2] = (float)(((double)iz - half_z) * prm->spacing + jz);
        t.position[3] = 0.0f;
        t.rotation[0] = t.rotation[1] = t.rotation[2] = 0.0f;
        t.rotation[3] = 1.0f;
        inverse_masses[body] = inv_mass;
        mh_mass& m = masses[body];
        std::memset(&m, 0, sizeof(m));
        m.inverse_mass[0] = inv_mass;
        m.inverse_inertia_x[0] = inv_inertia;
        m.inverse_inertia_y[1] = inv_inertia;
        m.inverse_inertia_z[2] = inv_inertia;
        if (inertias) {
          mh_spatial_inertia& s = inertias[body];
          std::memset(&s, 0, sizeof(s));
          s.mass[0] = mass;
          s.inertia_x[0] = inertia;
          s.inertia_y[1] = inertia;
          s.inertia_z[2] = inertia;
        }
        mh_collider& c = colliders[body];
        std::memset(&c, 0, sizeof(c));
        c.local_rotation[1] = (float)std::sin(0.5 * yaw);
        c.local_rotation[3] = (float)std::cos(0.5 * yaw);
        c.half_extents[0] = c.half_extents[1] = c.half_extents[2] = 0.5f * e;
        c.metadata[0] = (uint32_t)body;
        c.metadata[1] = (uint32_t)body;
      }
    }
  }
  return MH_OK;
}
