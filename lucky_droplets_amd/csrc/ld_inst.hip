// One Engine<T, N> instantiation per object file (built with -DLD_T, -DLD_TN, -DLD_N).
#include "ld_engine.hpp"

#if !defined(LD_T) || !defined(LD_TN) || !defined(LD_N)
#error "build with -DLD_T=<float|double> -DLD_TN=<f32|f64> -DLD_N=<grid size>"
#endif

#define LD_FACT2(a, b) make_engine_##a##_##b
#define LD_FACT(a, b) LD_FACT2(a, b)

namespace ld {
EngineBase* LD_FACT(LD_TN, LD_N)(const ld_params& p) { return new Engine<LD_T, LD_N>(p); }
}  // namespace ld
