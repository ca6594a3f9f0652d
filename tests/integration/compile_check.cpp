// Compiles tests/integration/hip_binding.hpp against include/afq.h and links it with
// libafq_hip.so.  The few SAFIRE names the binding reads are declared here with the meaning the
// reference gives them (config.h WALKER_TYPES / SpinTypes, WalkerConfig.hpp walker_data, a
// device array with data() / extent() / strides(), a walker set with SlaterMatrices /
// getProperty); main() builds a walker buffer in the reference's row layout on the host and
// checks that walkers_view points at its Slater blocks and property columns with the row as the
// stride.  Nothing here touches a GPU: the library's entry points are only linked.
#include <array>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace sfqmc::afqmc {
using RealType = double;
using ComplexType = std::complex<double>;
enum WALKER_TYPES { UNDEFINED_WALKER_TYPE, CLOSED, COLLINEAR, NONCOLLINEAR };
enum SpinTypes { Alpha, Beta };
enum MEMORY_SPACE { HOST_MEMORY, DEVICE_MEMORY };
enum walker_data { SM, UR, DR, VR, WEIGHT, PHASE, PSEUDO_ELOC_, E1_, EXX_, EJ_, OVLP, LOGSCL_UP,
                   LOGSCL_DN, IS_UNITARY, SMN, FIELDS, WEIGHT_FAC, WEIGHT_HISTORY, THETA };
namespace utils {
inline void check(bool ok, const char* msg) {
  if (!ok) {
    std::fprintf(stderr, "check failed: %s\n", msg);
    std::abort();
  }
}
}  // namespace utils

// strided view of rank R over T (the layout-erased views the reference passes around)
template <class T, int R>
struct view {
  T* p;
  std::array<long, R> ext, str;
  T* data() const { return p; }
  long extent(int i) const { return ext[i]; }
  std::array<long, R> shape() const { return ext; }
  std::array<long, R> strides() const { return str; }
};

namespace memory {
template <MEMORY_SPACE M, class T, int R>
struct buffered_array {
  std::vector<T> v;
  std::array<long, R> ext;
  explicit buffered_array(std::array<long, R> e) : ext(e) {
    long n = 1;
    for (long x : e) n *= x;
    v.resize(n);
  }
  T* data() { return v.data(); }
  long extent(int i) const { return ext[i]; }
};
}  // namespace memory

// a walker set over one complex row per walker (WalkerSetBase.cpp:75-121)
struct RowWalkerSet {
  int nw, nmo, na, nb, row;
  std::array<int, 19> displ{};
  std::vector<ComplexType> buf;
  RowWalkerSet(int nw_, int nmo_, int na_, int nb_) : nw(nw_), nmo(nmo_), na(na_), nb(nb_) {
    int c = 0;
    displ[SM] = c;
    c += nmo * (na + nb);
    for (walker_data d : {WEIGHT, PHASE, PSEUDO_ELOC_, E1_, EXX_, EJ_, OVLP, LOGSCL_UP, LOGSCL_DN,
                          IS_UNITARY, THETA})
      displ[d] = c++;
    row = c;
    buf.resize((size_t)nw * row);
  }
  int size() const { return nw; }
  WALKER_TYPES walkerType() const { return nb > 0 ? COLLINEAR : CLOSED; }
  view<ComplexType, 3> SlaterMatrices(SpinTypes s) {
    int i0 = s == Alpha ? displ[SM] : displ[SM] + nmo * na;
    long nc = s == Alpha ? na : nb;
    return {buf.data() + i0, {nw, nmo, nc}, {row, nc, 1}};
  }
  view<ComplexType, 1> getProperty(walker_data id) {
    return {buf.data() + displ[id], {nw}, {row}};
  }
};
}  // namespace sfqmc::afqmc

#include "hip_binding.hpp"

using namespace sfqmc::afqmc;

// every binding template instantiated (linked, never called here)
void instantiate(afq_ctx* ctx, RowWalkerSet& w) {
  view<double, 3> L{nullptr, {8, 8, 20}, {160, 20, 1}};
  view<double, 2> h{nullptr, {8, 8}, {8, 1}};
  view<ComplexType, 2> P{nullptr, {3, 8}, {8, 1}}, G{nullptr, {4, 24}, {24, 1}},
      v{nullptr, {4, 20}, {20, 1}}, E{nullptr, {4, 3}, {3, 1}}, K{nullptr, {4, 20}, {20, 1}},
      X{nullptr, {4, 20}, {20, 1}}, P1{nullptr, {8, 8}, {8, 1}};
  view<ComplexType, 1> vmf{nullptr, {20}, {1}}, ov{nullptr, {4}, {1}};
  view<ComplexType, 3> V{nullptr, {4, 8, 8}, {64, 8, 1}}, C{nullptr, {4, 3, 8}, {24, 8, 1}};
  HipCholeskyOps ops(ctx, COLLINEAR, 8, 3, 2, L, h, P, 0.5, true);
  ops.vbias(G, v, 0.005);
  auto Vh = ops.vHS(X, 0.005);
  (void)Vh;
  ops.energy(E, G, 0);
  ops.energy(Beta, E, G, 0, K);
  ops.one_body_propagator(0.005, vmf, P1);
  ops.runtime_optimization(G);
  ops.update_potentials(1, 2.0);
  auto A = w.SlaterMatrices(Alpha);
  hip_det_ops::MixedDensityMatrix(ctx, P, A, C, ov, true);
  hip_det_ops::Log_Overlap(ctx, P, A, ov);
  hip_det_ops::Propagate(ctx, 'N', P1, V, A, 6);
  hip_det_ops::orthogonalize(ctx, A, ov);
  afq_walkers aw = walkers_view(w, true);
  afq_prop* prop = nullptr;
  afq_prop_step(prop, &aw, nullptr, nullptr);
}

int main() {
  const int nw = 5, nmo = 8, na = 3, nb = 2;
  RowWalkerSet w(nw, nmo, na, nb);
  afq_walkers aw = walkers_view(w, true);
  const double* base = reinterpret_cast<const double*>(w.buf.data());
  auto off = [&](const double* p) { return (long)(p - base) / 2; };
  bool ok = aw.W == nw && aw.phi_stride == w.row && aw.prop_stride == w.row &&
            off(aw.Phi) == 0 && off(aw.Phi_b) == nmo * na &&
            off(aw.weight) == w.displ[WEIGHT] && off(aw.phase) == w.displ[PHASE] &&
            off(aw.eloc) == w.displ[PSEUDO_ELOC_] && off(aw.ovlp) == w.displ[OVLP] &&
            off(aw.theta) == w.displ[THETA] && off(aw.E1) == w.displ[E1_] &&
            off(aw.EXX) == w.displ[EXX_] && off(aw.EJ) == w.displ[EJ_] &&
            w.row == nmo * (na + nb) + 11;
  std::printf("walkers_view: row %d, phi_stride %lld, prop_stride %lld: %s\n", w.row,
              (long long)aw.phi_stride, (long long)aw.prop_stride, ok ? "ok" : "MISMATCH");
  if (std::getenv("AFQ_BINDING_NEVER_SET_THIS")) instantiate(nullptr, w);
  return ok ? 0 : 1;
}
