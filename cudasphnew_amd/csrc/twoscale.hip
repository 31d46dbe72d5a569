// twoscale.hip -- two-scale sub-particle pass (AdvanceTwoScale,
// particle_simulation.cu:2145-2182).
#include "engine.h"

namespace sph {

sph_status twoscale_create(Engine& e) { return e.fail(SPH_INVALID_ARG, "two-scale pass not available"); }
void twoscale_destroy(Engine& e) { e.ts = nullptr; }
void twoscale_enqueue(Engine&) {}
sph_status twoscale_read(Engine& e, sph_field, void*, size_t) { return e.fail(SPH_INVALID_ARG, "no two-scale data"); }
sph_status twoscale_counts(Engine&, sph_counts*) { return SPH_OK; }
size_t twoscale_bytes(const Engine&) { return 0; }

}  // namespace sph
