// C ABI (include/afq.h) over the device classes.  Exceptions never cross the boundary: each
// entry point converts afq::Error into its status code and keeps the message per thread.
#include <mutex>
#include <string>

#include "../../include/afq.h"
#include "ops.hpp"
#include "tuner.hpp"

using afq::z64;

struct afq_ctx {
  afq::Context c;
  afq_ctx(int dev, hipStream_t s) : c(dev, s) {}
};
struct afq_ham {
  afq_ctx* ctx;
  afq::CholeskyOps* ops;
};
struct afq_phmsd {
  afq_ham* ham;
  afq::PhmsdTrial* t;
};
struct afq_nomsd {
  afq_ham* ham;
  afq::NomsdTrial* t;
};
struct afq_prop {
  afq_ham* ham;
  afq::Propagator* p;
};

// Workspace of the allocation-free entry points.  Each GPU holds its own copy until the process
// ends; host code never uses this name as an address (it is a host-side stand-in), only the
// device address looked up per device when a context is created.
static __device__ __align__(256) unsigned char scratch[AFQ_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Looks up this device's copy of `scratch` once; the device must be current.
void register_scratch(int device) {
  AFQ_CHECK(device >= 0 && device < kMaxDevices, afq::kErrArg, "device index out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  AFQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  AFQ_CHECK(p != nullptr, afq::kErrHip, "no device address for the allocation-free workspace");
  g_scratch_ptr[device] = p;
}

void* device_scratch(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  void* p = (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
  AFQ_CHECK(p != nullptr, afq::kErrState, "allocation-free workspace not set up on this device");
  return p;
}

// Sets the context's allocation-free mode for one call.
struct NoAllocScope {
  afq::Context& c;
  explicit NoAllocScope(afq::Context& ctx) : c(ctx) { c.set_no_alloc(true); }
  ~NoAllocScope() { c.set_no_alloc(false); }
};

thread_local std::string g_err;

template <class F>
int guard(F&& f) {
  try {
    f();
    g_err.clear();
    return AFQ_OK;
  } catch (const afq::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return AFQ_ERR_STATE;
  } catch (...) {
    g_err = "unknown error";
    return AFQ_ERR_STATE;
  }
}

#define REQUIRE(cond, msg) AFQ_CHECK((cond), afq::kErrArg, (msg))

inline const z64* Z(const double* p) { return reinterpret_cast<const z64*>(p); }
inline z64* Z(double* p) { return reinterpret_cast<z64*>(p); }

afq::WalkerState to_state(const afq_walkers* w) {
  REQUIRE(w->phi_stride >= 0 && w->prop_stride >= 0, "negative walker stride");
  afq::WalkerState st{Z(w->Phi), Z(w->Phi_b), Z(w->weight), Z(w->phase), Z(w->eloc),
                      Z(w->ovlp), Z(w->theta), w->W};
  st.phi_bs = st.phi_b_bs = (long)w->phi_stride;
  st.prop_bs = w->prop_stride > 0 ? (long)w->prop_stride : 1;
  st.e1 = Z(w->E1);
  st.exx = Z(w->EXX);
  st.ej = Z(w->EJ);
  st.e_bs = st.prop_bs;
  return st;
}

// batch stride of a Slater block: the caller's, or N n when 0
long block_stride(int64_t stride, int N, int n) {
  REQUIRE(stride == 0 || stride >= (int64_t)N * n, "walker stride below N * n");
  return stride ? (long)stride : (long)N * n;
}

void check_walkers(const afq_prop* p, const afq_walkers* w) {
  REQUIRE(w != nullptr, "walkers is NULL");
  REQUIRE(w->Phi && w->weight && w->phase && w->eloc && w->ovlp && w->theta, "walker array is NULL");
  REQUIRE(w->W > 0, "W must be positive");
  if (p->ham->ops->walker_type() == afq::kCollinear && p->ham->ops->ndown() > 0)
    REQUIRE(w->Phi_b != nullptr, "COLLINEAR walkers need Phi_b");
}

void check_precision(int precision) {
  REQUIRE(precision == AFQ_FP64 || precision == AFQ_MIXED, "precision must be AFQ_FP64 or AFQ_MIXED");
}

// the two spin blocks' batch strides, and Phi_b present when the walkers are COLLINEAR
struct NomsdBlocks {
  long bs_a, bs_b;
};
NomsdBlocks nomsd_blocks(const afq_nomsd* trial, const double* Phi_a, const double* Phi_b,
                         int64_t phi_stride, int W) {
  REQUIRE(trial && Phi_a, "NULL argument");
  REQUIRE(W > 0 && phi_stride >= 0, "W must be positive, the stride non-negative");
  const afq::CholeskyOps& h = *trial->ham->ops;
  const bool col = h.walker_type() == afq::kCollinear;
  if (col && h.ndown() > 0) REQUIRE(Phi_b != nullptr, "COLLINEAR walkers need Phi_b");
  return {block_stride(phi_stride, h.N(), h.nup()),
          col ? block_stride(phi_stride, h.N(), h.ndown()) : 0};
}


}  // namespace

extern "C" {

int afq_abi_version(void) { return 302; }

const char* afq_last_error(void) { return g_err.c_str(); }

int afq_ctx_create(int device, void* stream, afq_ctx** out) {
  return guard([&] {
    REQUIRE(out != nullptr, "out is NULL");
    int n = 0;
    AFQ_HIP(hipGetDeviceCount(&n));
    AFQ_CHECK(device >= 0 && device < n, afq::kErrArg, "no such device");
    auto* c = new afq_ctx(device, reinterpret_cast<hipStream_t>(stream));  // makes device current
    try {
      register_scratch(device);
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

int afq_ctx_destroy(afq_ctx* ctx) {
  return guard([&] { delete ctx; });
}

int afq_ctx_synchronize(afq_ctx* ctx) {
  return guard([&] {
    REQUIRE(ctx, "ctx is NULL");
    AFQ_HIP(hipStreamSynchronize(ctx->c.stream()));
  });
}

int afq_tuning_report(char* buf, int len) {
  return guard([&] {
    REQUIRE(buf && len > 0, "NULL argument");
    std::string r = afq::Tuner::get().report();
    size_t n = std::min<size_t>(r.size(), (size_t)len - 1);
    r.copy(buf, n);
    buf[n] = 0;
  });
}

int afq_tune_force(const char* site, int candidate) {
  return guard([&] { afq::Tuner::get().force(site ? std::string(site) : std::string(), candidate); });
}

int afq_ctx_scratch_bytes(afq_ctx* ctx, uint64_t* bytes) {
  return guard([&] {
    REQUIRE(ctx && bytes, "NULL argument");
    *bytes = ctx->c.scratch_bytes();
  });
}

int afq_ham_create(afq_ctx* ctx, int walker_type, int N, int nup, int ndown, int M,
                   const double* Likn, const double* h1, const double* PsiT, double E0,
                   int precision, afq_ham** out) {
  return guard([&] {
    REQUIRE(ctx && out && Likn && h1 && PsiT, "NULL argument");
    auto* h = new afq_ham{ctx, nullptr};
    try {
      h->ops = new afq::CholeskyOps(ctx->c, walker_type, N, nup, ndown, M, Likn, h1, Z(PsiT), E0,
                                    precision);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int afq_ham_create_spin(afq_ctx* ctx, int walker_type, int N, int nup, int ndown, int M,
                        const double* Likn, int nspin_L, const double* h1, int nspin_h1,
                        const double* PsiT, double E0, int precision, afq_ham** out) {
  return guard([&] {
    REQUIRE(ctx && out && Likn && h1 && PsiT, "NULL argument");
    auto* h = new afq_ham{ctx, nullptr};
    try {
      h->ops = new afq::CholeskyOps(ctx->c, walker_type, N, nup, ndown, M, Likn, h1, Z(PsiT), E0,
                                    precision, nspin_L, nspin_h1);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int afq_ham_destroy(afq_ham* ham) {
  return guard([&] {
    if (!ham) return;
    delete ham->ops;
    delete ham;
  });
}

int afq_ham_tensors(afq_ham* ham, const double** Lnak, const double** haj, const double** vexx,
                    int* real_trial) {
  return guard([&] {
    REQUIRE(ham, "ham is NULL");
    if (Lnak) *Lnak = reinterpret_cast<const double*>(ham->ops->Lnak());
    if (haj) *haj = reinterpret_cast<const double*>(ham->ops->haj());
    if (vexx) *vexx = ham->ops->vexx();
    if (real_trial) *real_trial = ham->ops->real_trial() ? 1 : 0;
  });
}

int afq_vbias(afq_ham* ham, const double* G, int W, double dt, int precision, double* v) {
  return guard([&] {
    REQUIRE(ham && G && v, "NULL argument");
    REQUIRE(W >= 0, "W");
    ham->ops->vbias(Z(G), W, dt, Z(v), (afq::Precision)precision);
  });
}

int afq_vhs(afq_ham* ham, const void* X, int W, double dt, int precision, void* V) {
  return guard([&] {
    REQUIRE(ham && X && V, "NULL argument");
    REQUIRE(W >= 0, "W");
    ham->ops->vHS(X, W, dt, V, (afq::Precision)precision);
  });
}

int afq_energy(afq_ham* ham, const double* G, int W, int addH1, int addEJ, int addEXX,
               int precision, double* E) {
  return guard([&] {
    REQUIRE(ham && G && E, "NULL argument");
    ham->ops->energy(Z(G), W, Z(E), addH1 != 0, addEJ != 0, addEXX != 0, (afq::Precision)precision);
  });
}

int afq_energy_spin(afq_ham* ham, int spin, const double* G, int W, int addH1, int addEJ,
                    int addEXX, int precision, double* E, double* EJn) {
  return guard([&] {
    REQUIRE(ham && G && E, "NULL argument");
    REQUIRE(W >= 0, "W");
    ham->ops->energy_spin(spin, Z(G), W, Z(E), Z(EJn), addH1 != 0, addEJ != 0, addEXX != 0,
                          (afq::Precision)precision);
  });
}

int afq_vmf(afq_ham* ham, double dt, double* vMF) {
  return guard([&] {
    REQUIRE(ham && vMF, "NULL argument");
    ham->ops->vMF(dt, Z(vMF));
  });
}

int afq_one_body_propagator(afq_ham* ham, const double* vMF, double dt, double* P1) {
  return guard([&] {
    REQUIRE(ham && vMF && P1, "NULL argument");
    ham->ops->one_body_propagator(Z(vMF), dt, Z(P1));
  });
}

int afq_mixed_density_matrix(afq_ctx* ctx, const double* PsiT, int n, int N, const double* Phi,
                             int64_t phi_stride, int W, double* G, double* ovlp, int compact) {
  return guard([&] {
    REQUIRE(ctx && PsiT && Phi && G && ovlp, "NULL argument");
    REQUIRE(n >= 0 && n <= 64 && N >= n && W >= 0, "invalid sizes");
    long gbs = compact ? (long)n * N : (long)N * N;
    afq::det_ops::mixed_density_matrix(ctx->c, Z(PsiT), n, N, Z(Phi), block_stride(phi_stride, N, n),
                                       W, Z(G), gbs, Z(ovlp), compact != 0);
  });
}

int afq_mixed_density_matrix_noalloc(afq_ctx* ctx, const double* PsiT, int n, int N,
                                     const double* Phi, int64_t phi_stride, int W, double* G,
                                     double* ovlp, int compact) {
  return guard([&] {
    REQUIRE(ctx && PsiT && Phi && G && ovlp, "NULL argument");
    REQUIRE(n >= 0 && n <= 64 && N >= n && W >= 0, "invalid sizes");
    const size_t nT = (size_t)W * n * n, nGc = compact ? 0 : (size_t)W * n * N;
    const size_t need = (nT + nGc) * sizeof(z64);
    AFQ_CHECK(need <= AFQ_NOALLOC_SCRATCH_BYTES, afq::kErrShape,
              "allocation-free density matrix needs " + std::to_string(need) +
                  " bytes of workspace, above its limit of " +
                  std::to_string((size_t)AFQ_NOALLOC_SCRATCH_BYTES));
    const long bs = block_stride(phi_stride, N, n);
    z64* ws = static_cast<z64*>(device_scratch(ctx->c.device()));
    long gbs = compact ? (long)n * N : (long)N * N;
    NoAllocScope na(ctx->c);
    afq::det_ops::mixed_density_matrix(ctx->c, Z(PsiT), n, N, Z(Phi), bs, W, Z(G), gbs, Z(ovlp),
                                       compact != 0, 1.0, nullptr, ws, compact ? nullptr : ws + nT);
  });
}

int afq_log_overlap(afq_ctx* ctx, const double* PsiT, int n, int N, const double* Phi,
                    int64_t phi_stride, int W, double* ovlp) {
  return guard([&] {
    REQUIRE(ctx && PsiT && Phi && ovlp, "NULL argument");
    REQUIRE(n >= 0 && n <= 64 && N >= n && W >= 0, "invalid sizes");
    afq::det_ops::log_overlap(ctx->c, Z(PsiT), n, N, Z(Phi), block_stride(phi_stride, N, n), W,
                              Z(ovlp));
  });
}

int afq_inverse_logdet(afq_ctx* ctx, const double* A, int nb, int n, double* ovlp, double* Ainv) {
  return guard([&] {
    REQUIRE(ctx && A && ovlp && Ainv, "NULL argument");
    REQUIRE(n >= 1 && n <= 64 && nb >= 0, "invalid sizes");
    afq::det_ops::inverse_logdet(ctx->c, Z(A), nb, n, Z(ovlp), Z(Ainv));
  });
}

int afq_propagate(afq_ctx* ctx, const double* P1, const void* V, double* Phi, int64_t phi_stride,
                  int W, int N, int n, int order, char TA, int precision) {
  return guard([&] {
    REQUIRE(ctx && P1 && V && Phi, "NULL argument");
    REQUIRE(W >= 0 && N > 0 && n >= 0 && order >= 0, "invalid sizes");
    REQUIRE(TA == 'N' || TA == 'H', "TA must be 'N' or 'H'");
    REQUIRE(precision == AFQ_FP64 || precision == AFQ_MIXED, "precision");
    afq::det_ops::PropagateArgs a{W, N, n, order, TA, Z(P1), nullptr, nullptr, nullptr, V,
                                  (long)N * N, Z(Phi), block_stride(phi_stride, N, n),
                                  (afq::Precision)precision};
    if (precision == AFQ_MIXED) {
      afq::c32* p32 = afq::scratch_as<afq::c32>(ctx->c, "capi_P1c32", (size_t)N * N);
      afq::cast_z64_c32(ctx->c.stream(), (long)N * N, Z(P1), p32);
      a.P1c32 = p32;
    }
    afq::det_ops::propagate(ctx->c, a);
  });
}

int afq_orthogonalize(afq_ctx* ctx, double* Phi, int64_t phi_stride, int W, int N, int n,
                      double* ldet) {
  return guard([&] {
    REQUIRE(ctx && Phi && ldet, "NULL argument");
    REQUIRE(n >= 0 && n <= 64 && N >= n && W >= 0, "invalid sizes");
    afq::det_ops::orthogonalize(ctx->c, Z(Phi), block_stride(phi_stride, N, n), W, N, n, Z(ldet));
  });
}

int afq_field_uniforms(afq_ctx* ctx, uint64_t seed, uint32_t stream, uint32_t step, int W, int M,
                       double* u) {
  return guard([&] {
    REQUIRE(ctx && u, "NULL argument");
    afq::field_uniforms(ctx->c.stream(), seed, stream, step, W, M, u);
  });
}

int afq_construct_x(afq_ctx* ctx, const double* u, const double* vMF, int W, int M, double bound,
                    int project_vbias, int free_projection, double* X, double* HW, double* mf) {
  return guard([&] {
    REQUIRE(ctx && u && vMF && X && HW && mf, "NULL argument");
    afq::FieldParams fp{bound, project_vbias, free_projection, 0, 0, 0, nullptr};
    afq::construct_x(ctx->c.stream(), u, Z(vMF), W, M, fp, Z(X), Z(HW), Z(mf), nullptr);
  });
}

int afq_walker_update(afq_ctx* ctx, int W, const afq_update_params* p, const double* Eshift,
                      const double* new_ovlp, const double* mf, const double* HW, double* weight,
                      double* phase, double* eloc, double* ovlp, double* theta) {
  return guard([&] {
    REQUIRE(ctx && p && Eshift && new_ovlp && mf && HW && weight && phase && eloc && ovlp && theta,
            "NULL argument");
    afq::UpdateParams up{p->dt, p->E0, p->lower_cut, p->upper_cut, p->weight_floor, p->weight_frac,
                         p->target_pop, p->free_projection, p->use_cp, 0};
    afq::walker_update(ctx->c.stream(), W, up, Eshift, Z(new_ovlp), Z(mf), Z(HW), Z(weight),
                       Z(phase), Z(eloc), Z(ovlp), Z(theta));
  });
}

int afq_average_pseudo_energy(afq_ctx* ctx, int W, const double* weight, const double* eloc,
                              double dt, double relax, double* out, double* Eshift) {
  return guard([&] {
    REQUIRE(ctx && weight && eloc, "NULL argument");
    afq::average_pseudo_energy(ctx->c.stream(), W, Z(weight), Z(eloc), dt, relax, out, Eshift);
  });
}

static afq::PropagatorParams to_params(const afq_prop_params* p) {
  afq::PropagatorParams pp;
  pp.dt = p->dt;
  pp.order = p->order;
  pp.vbias_bound = p->vbias_bound;
  pp.lower_cut = p->lower_cut;
  pp.upper_cut = p->upper_cut;
  pp.weight_floor = p->weight_floor;
  pp.weight_frac = p->weight_frac;
  pp.target_pop = p->target_pop;
  pp.free_projection = p->free_projection;
  pp.use_cp = p->use_cp;
  pp.project_force_bias = p->project_force_bias;
  pp.local_energy = p->local_energy;
  pp.seed = p->seed;
  pp.stream = p->stream;
  return pp;
}

int afq_prop_create(afq_ham* ham, int W, const afq_prop_params* p, afq_prop** out) {
  return guard([&] {
    REQUIRE(ham && p && out, "NULL argument");
    auto* pr = new afq_prop{ham, nullptr};
    try {
      pr->p = new afq::Propagator(*ham->ops, W, to_params(p));
    } catch (...) {
      delete pr;
      throw;
    }
    *out = pr;
  });
}

int afq_prop_create_phmsd(afq_phmsd* trial, int W, const afq_prop_params* p, afq_prop** out) {
  return guard([&] {
    REQUIRE(trial && p && out, "NULL argument");
    auto* pr = new afq_prop{trial->ham, nullptr};
    try {
      pr->p = new afq::Propagator(*trial->ham->ops, W, to_params(p), trial->t);
    } catch (...) {
      delete pr;
      throw;
    }
    *out = pr;
  });
}

int afq_prop_create_nomsd(afq_nomsd* trial, int W, const afq_prop_params* p, afq_prop** out) {
  return guard([&] {
    REQUIRE(trial && p && out, "NULL argument");
    auto* pr = new afq_prop{trial->ham, nullptr};
    try {
      pr->p = new afq::Propagator(*trial->ham->ops, W, to_params(p), trial->t);
    } catch (...) {
      delete pr;
      throw;
    }
    *out = pr;
  });
}

int afq_prop_destroy(afq_prop* prop) {
  return guard([&] {
    if (!prop) return;
    delete prop->p;
    delete prop;
  });
}

int afq_prop_views(afq_prop* prop, const double** vMF, const double** P1) {
  return guard([&] {
    REQUIRE(prop, "prop is NULL");
    if (vMF) *vMF = reinterpret_cast<const double*>(prop->p->vMF());
    if (P1) *P1 = reinterpret_cast<const double*>(prop->p->P1());
  });
}

int afq_prop_step(afq_prop* prop, const afq_walkers* w, const double* u, const double* Eshift) {
  return guard([&] {
    REQUIRE(prop && Eshift, "NULL argument");
    check_walkers(prop, w);
    prop->p->step(to_state(w), u, Eshift);
  });
}

int afq_prop_orthogonalize(afq_prop* prop, const afq_walkers* w) {
  return guard([&] {
    REQUIRE(prop, "prop is NULL");
    check_walkers(prop, w);
    prop->p->orthogonalize(to_state(w));
  });
}

int afq_prop_set_step(afq_prop* prop, uint32_t step) {
  return guard([&] {
    REQUIRE(prop, "prop is NULL");
    prop->p->set_step(step);
  });
}

int afq_prop_energy(afq_prop* prop, const afq_walkers* w, int precision, double* E, double* ovlp) {
  return guard([&] {
    REQUIRE(prop && E && ovlp, "NULL argument");
    check_walkers(prop, w);
    prop->p->energy(to_state(w), Z(E), Z(ovlp), (afq::Precision)precision);
  });
}

int afq_prop_step_graph(afq_prop* prop, const afq_walkers* w, double* Eshift, int ortho,
                        int popctl, double relax, double* energy_out) {
  return guard([&] {
    REQUIRE(prop && Eshift, "NULL argument");
    check_walkers(prop, w);
    prop->p->step_graph(to_state(w), Eshift, ortho != 0, popctl != 0, relax, energy_out);
  });
}

int afq_prop_population_control(afq_prop* prop, const afq_walkers* w, double r) {
  return guard([&] {
    REQUIRE(prop, "prop is NULL");
    check_walkers(prop, w);
    REQUIRE(r > 0.0 && r < 1.0, "r must be in (0, 1)");
    double* rd = afq::scratch_as<double>(prop->ham->ctx->c, "capi_comb_r", 1);
    AFQ_HIP(hipMemcpyAsync(rd, &r, sizeof(double), hipMemcpyHostToDevice, prop->ham->ctx->c.stream()));
    AFQ_HIP(hipStreamSynchronize(prop->ham->ctx->c.stream()));
    prop->p->population_control(to_state(w), rd);
  });
}

int afq_prop_bound_stats(afq_prop* prop, int64_t* counts, int reset) {
  return guard([&] {
    REQUIRE(prop && counts, "NULL argument");
    long long c[9];
    prop->p->bound_stats(c, reset != 0);
    for (int i = 0; i < 9; ++i) counts[i] = (int64_t)c[i];
  });
}

int afq_prop_status(afq_prop* prop, int* collapsed) {
  return guard([&] {
    REQUIRE(prop && collapsed, "NULL argument");
    *collapsed = prop->p->status();
  });
}

int afq_prop_enable_backprop(afq_prop* prop, int nback) {
  return guard([&] {
    REQUIRE(prop, "prop is NULL");
    prop->p->enable_backprop(nback);
  });
}

int afq_prop_bp_anchor(afq_prop* prop, const afq_walkers* w) {
  return guard([&] {
    REQUIRE(prop, "prop is NULL");
    check_walkers(prop, w);
    prop->p->bp_anchor(to_state(w));
  });
}

int afq_prop_backpropagate(afq_prop* prop, const afq_walkers* w, int nbp, int ortho, int restore,
                           double* G, double* weights) {
  return guard([&] {
    REQUIRE(prop && G && weights, "NULL argument");
    check_walkers(prop, w);
    REQUIRE(restore >= 0 && restore <= 2, "restore must be 0, 1 or 2");
    prop->p->backpropagate(to_state(w), nbp, ortho, restore, Z(G), Z(weights));
  });
}

int afq_prop_bp_views(afq_prop* prop, const double** fields, const double** factors, int* nback) {
  return guard([&] {
    REQUIRE(prop && fields && factors && nback, "NULL argument");
    *fields = reinterpret_cast<const double*>(prop->p->bp_fields());
    *factors = reinterpret_cast<const double*>(prop->p->bp_factors());
    *nback = prop->p->backprop_length();
  });
}

int afq_prop_bp_measure(afq_prop* prop, const afq_walkers* w, int nbp, int ortho, int restore,
                        double* onerdm, double* diag2rdm) {
  return guard([&] {
    REQUIRE(prop, "NULL argument");
    check_walkers(prop, w);
    prop->p->bp_measure(to_state(w), nbp, ortho, restore, onerdm ? Z(onerdm) : nullptr,
                        diag2rdm ? Z(diag2rdm) : nullptr);
  });
}

int afq_prop_mixed_measure(afq_prop* prop, const afq_walkers* w, double* onerdm, double* diag2rdm) {
  return guard([&] {
    REQUIRE(prop, "NULL argument");
    check_walkers(prop, w);
    prop->p->mixed_measure(to_state(w), onerdm ? Z(onerdm) : nullptr,
                           diag2rdm ? Z(diag2rdm) : nullptr);
  });
}

int afq_prop_number_of_references(afq_prop* prop, int* nref) {
  return guard([&] {
    REQUIRE(prop && nref, "NULL argument");
    *nref = prop->p->number_of_references();
  });
}

int afq_prop_vhs_timing(afq_prop* prop, int on) {
  return guard([&] {
    REQUIRE(prop, "NULL argument");
    prop->p->set_vhs_timing(on != 0);
  });
}

int afq_prop_vhs_ms(afq_prop* prop, float* ms) {
  return guard([&] {
    REQUIRE(prop && ms, "NULL argument");
    *ms = prop->p->vhs_ms();
  });
}

int afq_prop_te_anchor(afq_prop* prop, int follow_walkers) {
  return guard([&] {
    REQUIRE(prop, "NULL argument");
    prop->p->te_anchor(follow_walkers != 0);
  });
}

int afq_prop_te_propagate(afq_prop* prop, const afq_walkers* w, int nsteps) {
  return guard([&] {
    REQUIRE(prop, "NULL argument");
    check_walkers(prop, w);
    prop->p->te_propagate(to_state(w), nsteps);
  });
}

int afq_prop_te_measure(afq_prop* prop, const afq_walkers* w, int restore_steps, int transpose_m,
                        double* onerdm, double* diag2rdm) {
  return guard([&] {
    REQUIRE(prop, "NULL argument");
    check_walkers(prop, w);
    prop->p->te_measure(to_state(w), restore_steps, transpose_m != 0,
                        onerdm ? Z(onerdm) : nullptr, diag2rdm ? Z(diag2rdm) : nullptr);
  });
}

int afq_prop_te_views(afq_prop* prop, const double** X, const double** Y, const double** M,
                      const double** P1inv) {
  return guard([&] {
    REQUIRE(prop && X && Y && M && P1inv, "NULL argument");
    *X = reinterpret_cast<const double*>(prop->p->te_operator(0));
    *Y = reinterpret_cast<const double*>(prop->p->te_operator(1));
    *M = reinterpret_cast<const double*>(prop->p->te_operator(2));
    *P1inv = reinterpret_cast<const double*>(prop->p->P1_inverse());
  });
}

int afq_rdm_measure(afq_ctx* ctx, int walker_type, int N, int W, const double* G,
                    const double* weights, const double* coeff, int accumulate, double* onerdm,
                    double* diag2rdm) {
  return guard([&] {
    REQUIRE(ctx && G && weights, "NULL argument");
    REQUIRE(walker_type == AFQ_CLOSED || walker_type == AFQ_COLLINEAR, "walker type");
    REQUIRE(N > 0 && W > 0, "invalid sizes");
    const int nsp = walker_type == AFQ_COLLINEAR ? 2 : 1;
    afq::rdm_measure(ctx->c, W, nsp, N, nsp == 2, Z(G), (long)nsp * N * N, Z(weights),
                     coeff ? Z(coeff) : nullptr, accumulate != 0, onerdm ? Z(onerdm) : nullptr,
                     diag2rdm ? Z(diag2rdm) : nullptr);
  });
}

/* ---- PHMSD ---- */

int afq_phmsd_create(afq_ham* ham, const afq_ph_excitations* ex, afq_phmsd** out) {
  return guard([&] {
    REQUIRE(ham && ex && out, "NULL argument");
    afq::PhExcitations e;
    for (int s = 0; s < 2; ++s) {
      e.nel[s] = ex->nel[s];
      e.nconf[s] = ex->nconf[s];
      e.rank[s] = ex->rank[s];
      e.exc_start[s] = ex->exc_start[s];
      e.exc[s] = ex->exc[s];
      e.refc[s] = ex->refc[s];
    }
    e.ndet = ex->ndet;
    e.det_a = ex->det_a;
    e.det_b = ex->det_b;
    e.ci = Z(ex->ci);
    auto* t = new afq_phmsd{ham, nullptr};
    try {
      t->t = new afq::PhmsdTrial(*ham->ops, e);
    } catch (...) {
      delete t;
      throw;
    }
    *out = t;
  });
}

int afq_phmsd_destroy(afq_phmsd* trial) {
  return guard([&] {
    if (!trial) return;
    delete trial->t;
    delete trial;
  });
}

int afq_phmsd_mixed_density_matrix(afq_phmsd* trial, const double* Phi_a, const double* Phi_b,
                                   int64_t phi_stride, int W, double* G, double* ovlp) {
  return guard([&] {
    REQUIRE(trial && Phi_a && Phi_b && G && ovlp, "NULL argument");
    REQUIRE(W > 0 && phi_stride >= 0, "W must be positive, the stride non-negative");
    trial->t->mixed_density_matrix(Z(Phi_a), Z(Phi_b), W, Z(G), Z(ovlp), (long)phi_stride,
                                   (long)phi_stride);
  });
}

int afq_phmsd_log_overlap(afq_phmsd* trial, const double* Phi_a, const double* Phi_b,
                          int64_t phi_stride, int W, double* ovlp) {
  return guard([&] {
    REQUIRE(trial && Phi_a && Phi_b && ovlp, "NULL argument");
    REQUIRE(W > 0 && phi_stride >= 0, "W must be positive, the stride non-negative");
    trial->t->log_overlap(Z(Phi_a), Z(Phi_b), W, Z(ovlp), (long)phi_stride, (long)phi_stride);
  });
}

int afq_phmsd_energy(afq_phmsd* trial, const double* Phi_a, const double* Phi_b,
                     int64_t phi_stride, int W, double* E, double* ovlp) {
  return guard([&] {
    REQUIRE(trial && Phi_a && Phi_b && E && ovlp, "NULL argument");
    REQUIRE(W > 0 && phi_stride >= 0, "W must be positive, the stride non-negative");
    trial->t->energy(Z(Phi_a), Z(Phi_b), W, Z(E), Z(ovlp), (long)phi_stride, (long)phi_stride);
  });
}

int afq_phmsd_vmf(afq_phmsd* trial, double dt, double* vMF) {
  return guard([&] {
    REQUIRE(trial && vMF, "NULL argument");
    trial->t->vMF(dt, Z(vMF));
  });
}

/* ---- multi-determinant NOMSD ---- */

int afq_nomsd_create(afq_ham* ham, int ndet, const double* PsiT, const double* ci,
                     afq_nomsd** out) {
  return guard([&] {
    REQUIRE(ham && PsiT && ci && out, "NULL argument");
    REQUIRE(ndet >= 1, "ndet must be positive");
    auto* t = new afq_nomsd{ham, nullptr};
    try {
      t->t = new afq::NomsdTrial(*ham->ops, ndet, Z(PsiT), Z(ci));
    } catch (...) {
      delete t;
      throw;
    }
    *out = t;
  });
}

int afq_nomsd_destroy(afq_nomsd* trial) {
  return guard([&] {
    if (!trial) return;
    delete trial->t;
    delete trial;
  });
}

int afq_nomsd_log_overlap(afq_nomsd* trial, const double* Phi_a, const double* Phi_b,
                          int64_t phi_stride, int W, double* ovlp) {
  return guard([&] {
    const NomsdBlocks b = nomsd_blocks(trial, Phi_a, Phi_b, phi_stride, W);
    REQUIRE(ovlp, "NULL argument");
    trial->t->log_overlap(Z(Phi_a), Z(Phi_b), W, Z(ovlp), b.bs_a, b.bs_b);
  });
}

int afq_nomsd_mixed_density_matrix(afq_nomsd* trial, const double* Phi_a, const double* Phi_b,
                                   int64_t phi_stride, int W, const double* log_m, double* G,
                                   double* ovlp) {
  return guard([&] {
    const NomsdBlocks b = nomsd_blocks(trial, Phi_a, Phi_b, phi_stride, W);
    REQUIRE(log_m && G && ovlp, "NULL argument");
    trial->t->mixed_density_matrix(Z(Phi_a), Z(Phi_b), W, Z(log_m), Z(G), Z(ovlp), b.bs_a,
                                   b.bs_b);
  });
}

int afq_nomsd_vbias(afq_nomsd* trial, const double* Phi_a, const double* Phi_b,
                    int64_t phi_stride, int W, const double* log_m, double dt, int precision,
                    double* v) {
  return guard([&] {
    const NomsdBlocks b = nomsd_blocks(trial, Phi_a, Phi_b, phi_stride, W);
    REQUIRE(log_m && v, "NULL argument");
    check_precision(precision);
    trial->t->vbias(Z(Phi_a), Z(Phi_b), W, Z(log_m), dt, Z(v), (afq::Precision)precision,
                    b.bs_a, b.bs_b);
  });
}

int afq_nomsd_energy(afq_nomsd* trial, const double* Phi_a, const double* Phi_b,
                     int64_t phi_stride, int W, int precision, double* E, double* ovlp) {
  return guard([&] {
    const NomsdBlocks b = nomsd_blocks(trial, Phi_a, Phi_b, phi_stride, W);
    REQUIRE(E && ovlp, "NULL argument");
    check_precision(precision);
    trial->t->energy(Z(Phi_a), Z(Phi_b), W, Z(E), Z(ovlp), (afq::Precision)precision, b.bs_a,
                     b.bs_b);
  });
}

int afq_nomsd_vmf(afq_nomsd* trial, double dt, double* vMF) {
  return guard([&] {
    REQUIRE(trial && vMF, "NULL argument");
    trial->t->vMF(dt, Z(vMF));
  });
}

int afq_vbias_full(afq_ham* ham, const double* G, int W, double dt, int precision, double* v) {
  return guard([&] {
    REQUIRE(ham && G && v, "NULL argument");
    REQUIRE(W >= 0, "W");
    check_precision(precision);
    ham->ops->vbias_full(Z(G), W, dt, Z(v), (afq::Precision)precision);
  });
}

}  // extern "C"
