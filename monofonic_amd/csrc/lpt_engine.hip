// lpt_engine.hip -- fused nLPT source engine (placeholder until implemented)
#include "ops.hpp"

namespace mfx {

size_t lpt_workspace_bytes(int n, int prec, int order) { return 16; }

template <typename T>
void lpt_sources(Ctx& c, const cx<T>* phi, cx<T>* phi2, cx<T>* phi3, cx<T>* a3x, cx<T>* a3y, cx<T>* a3z, int n,
                 double boxlen, int order, void* ws) {
  throw Error(1, "fused LPT engine not available");
}

template void lpt_sources<float>(Ctx&, const cx<float>*, cx<float>*, cx<float>*, cx<float>*, cx<float>*, cx<float>*,
                                 int, double, int, void*);
template void lpt_sources<double>(Ctx&, const cx<double>*, cx<double>*, cx<double>*, cx<double>*, cx<double>*,
                                  cx<double>*, int, double, int, void*);
}  // namespace mfx
