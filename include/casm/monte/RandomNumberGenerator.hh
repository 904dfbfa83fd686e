// Drop-in counterpart of libcasm-monte's RandomNumberGenerator.hh:14-42: a
// shared engine (std::mt19937_64 by default, definitions.hh:17) and the two
// draws the Monte Carlo methods make, through the libstdc++ distributions the
// reference uses.
#ifndef CASM_MI355X_monte_RandomNumberGenerator
#define CASM_MI355X_monte_RandomNumberGenerator

#include "casm/monte/definitions.hh"

namespace CASM {
namespace monte {

template <typename EngineType = default_engine_type>
struct RandomNumberGenerator {
  std::shared_ptr<EngineType> engine;
  RandomNumberGenerator(std::shared_ptr<EngineType> _engine = std::shared_ptr<EngineType>())
      : engine(_engine) {
    if (engine == nullptr) {
      engine = std::make_shared<EngineType>();
      std::random_device device;
      engine->seed(device());
    }
  }
  /// uniformly distributed integer in [0, maximum_value]
  template <typename IntType>
  IntType random_int(IntType maximum_value) {
    return std::uniform_int_distribution<IntType>(0, maximum_value)(*engine);
  }
  /// uniformly distributed real in [0, maximum_value)
  template <typename RealType>
  RealType random_real(RealType maximum_value) {
    return std::uniform_real_distribution<RealType>(0, maximum_value)(*engine);
  }
};

}  // namespace monte
}  // namespace CASM

#endif
