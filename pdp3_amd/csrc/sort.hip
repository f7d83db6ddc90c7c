// Stable counting sort of particles by cell (placeholder).
#include "context.hpp"

namespace pdp3 {
int sort_species(Ctx &, Species &) { return 0; }
}  // namespace pdp3
