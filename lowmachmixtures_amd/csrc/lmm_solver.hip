// Context, multigrid drivers, GMRES / flexible GMRES and the C ABI of the
// MI355X Stokes library.  Host code only orchestrates: every field operation
// is a gfx950 kernel on the context stream; the host keeps the GMRES
// Hessenberg scalars (gmres.f90:55-59) and the stopping logic.
#include <hip/hip_runtime.h>

#include <chrono>
#include <atomic>
#include <mutex>
#include <cstdlib>
#include <thread>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <algorithm>
#include <map>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/lmm.h"
#include "lmm_cnn.h"
#include "lmm_species.h"
#include "lmm_kernels.h"

using namespace lmm;

namespace lmm {
void launch_fold_cells(hipStream_t s, const double *src, int nsrc, double *partial);
}

// ---------------------------------------------------------------------------
// context and fields
// ---------------------------------------------------------------------------

template <typename T>
struct SLevel {
  Grid g;
  T *alpha[3] = {nullptr, nullptr, nullptr};  // theta*alpha (owned)
  T *beta = nullptr, *gamma = nullptr;        // owned on coarse levels (and level 0 in FP32)
  T *bed[3] = {nullptr, nullptr, nullptr};
  T *phi[3] = {nullptr, nullptr, nullptr};
  T *rhs[3] = {nullptr, nullptr, nullptr};
  T *res[3] = {nullptr, nullptr, nullptr};
  StagCoef<T> K() const {
    StagCoef<T> k;
    for (int d = 0; d < 3; d++) {
      k.alpha[d] = alpha[d];
      k.bed[d] = bed[d];
    }
    k.beta = beta;
    k.gamma = gamma;
    return k;
  }
};

template <typename T>
struct CLevel {
  Grid g;
  T *a[3] = {nullptr, nullptr, nullptr};
  T *phi = nullptr, *rhs = nullptr, *res = nullptr;
};

struct ProfFamily {
  int64_t launches = 0;
  double ms = 0.0;
  double bytes = 0.0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

struct lmm_field {
  lmm_ctx *ctx;
  int kind;
  int ncomp;
  double *d;
};

struct lmm_ctx {
  Grid g;
  lmm_params prm;
  hipStream_t stream = nullptr;
  int nl_s = 0, nl_c = 0;
  std::vector<SLevel<double>> s64;
  std::vector<CLevel<double>> c64;
  std::vector<SLevel<float>> s32;
  std::vector<CLevel<float>> c32;
  bool lvl32 = false;
  double *partial = nullptr;   // RED_BLOCKS*4
  double *partial2 = nullptr;  // per-block cell sums
  double *scal = nullptr;      // device scalars
  double *hscal = nullptr;     // pinned host mirror
  double *ainv[3] = {nullptr, nullptr, nullptr};
  double *alpha_theta[3] = {nullptr, nullptr, nullptr};
  double *scaled_beta = nullptr, *scaled_gamma = nullptr, *scaled_bed[3] = {nullptr, nullptr, nullptr};
  // Krylov storage (Stokes vectors [u0|u1|(u2)|p])
  int kry_m = -1;
  double *V = nullptr, *r = nullptr, *w = nullptr, *tmp = nullptr, *b = nullptr, *x = nullptr;
  float *Z32 = nullptr, *b32 = nullptr;
  double *mac_rhs = nullptr, *phi_cc = nullptr;
  float *mac_rhs32 = nullptr;
  // stochastic noise held by the context (init_m_stochastic)
  int n_rngs = 0;
  double *Wcc = nullptr, *Wed = nullptr;
  double *dweights = nullptr;
  double *stoch_work = nullptr;  // filtered stochastic fluxes (filtering_width > 0)
  // Smoother kernels: fused red-black sweeps on levels with at least
  // fused_min_tiles workgroups in a launch (32^3 and up), colour-by-colour kernels
  // elsewhere.  Both give the same bits.  LMM_SMOOTHER=colour in the
  // environment selects the colour kernels everywhere, LMM_SMOOTHER=fused
  // the fused sweep on every level whose extents allow it;
  // LMM_FUSED_MIN_TILES and LMM_FUSED_TARGET (workgroups aimed for when the
  // z-slab thickness is chosen) tune the choice.
  bool fused = true;
  int fused_min_tiles = 64;
  // coarse levels (at most COARSE_MAX_CELLS cells) of both multigrids run
  // in one single-workgroup launch; LMM_COARSE=0 in the environment keeps
  // one launch per pass on every level (same bits)
  bool coarse_one_launch = true;
  long coarse_max_cells = COARSE_MAX_CELLS;  // largest level inside the coarse launch
  // the flexible GMRES applies A to its FP32 preconditioned vectors and adds
  // them to x without widening them into FP64 buffers first (same bits);
  // LMM_Z32=0 in the environment widens them first
  bool z32_direct = true;
  // the staggered down-leg computes the residual and restricts it in one
  // pass where the level allows (same bits); LMM_RESID_RESTRICT=0 in the
  // environment keeps the two passes
  bool fused_resid_restrict = true;
  // the fused sweep runs the three face components in one launch
  // (LMM_FUSED_X3=0 in the environment: one launch per component; same bits)
  bool fused3 = true;
  // GMRES fuses each MGS axpy with the next reduction, the basis scaling
  // with the FP32 conversion of the preconditioner input and the restart
  // residual b - Ax into the operator (same bits); LMM_KRYLOV_FUSED=0 in the
  // environment keeps separate passes
  bool fused_krylov = true;
  // the flexible GMRES orthogonalises by classical Gram-Schmidt (one
  // multi-dot and one update pass per step); LMM_FLEX_GS=mgs in the
  // environment keeps modified Gram-Schmidt (different rounding)
  bool flex_cgs = true;
  double *partial_cgs = nullptr;  // RED_BLOCKS*CGS_STRIDE
  // fused-sweep form, z-slab target and operator brick depth of this
  // context's launches (from the environment at creation;
  // lmm_ctx_set_kernel_options changes them)
  Tuning tune;
  lmm_kernel_options kopt;
  // profiling: 0 off, 1 every family, 2 only the kernel-exact fused sweeps
  int prof = 0;
  std::map<std::string, ProfFamily> fam;
  std::vector<hipEvent_t> evpool;  // events reused across scopes
  std::string err;
  int vcycles = 0;
  // learned coarse-grid correction (precision LMM_LEARNED)
  lmm_cnn *cnn = nullptr;
  bool learned_on = false;    // the FP32 preconditioner applies the learned stage
  int learned_applied = 0;
  float *gen32 = nullptr;     // training-pair work: phi* and e (6 face arrays)
  // species knobs (lmm_ctx_set_species; reference defaults until set) and
  // the mass noise of init_mass_stochastic ([n_rngs][dim][ns][N])
  bool species_set = false;
  lmm_species_params species{};
  int n_rngs_mass = 0;
  double *Wmass = nullptr;
  double *dweights_mass = nullptr;
  double *sp_work = nullptr;  // species scratch, grown on demand
  long sp_work_n = 0;
  double *iw = nullptr;       // inertial-step temporaries, grown on demand
  long iw_n = 0;
  // set while a composite entry point (the inertial step) calls other entry
  // points: their completion checks skip the stream synchronisation, the
  // composite one synchronises once at its end
  bool defer_sync = false;
  double hweights_mass[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // weights last copied to dweights_mass
  bool weights_mass_valid = false;
  // HIP graphs of the FP32 preconditioner application, one per output slot
  // (kernel option precon_graph); dropped whenever what they captured may
  // have changed (parameters, kernel options, storage)
  bool precon_graph = false;
  struct PGraph {
    hipGraphExec_t exec = nullptr;
    double theta = 0.0;
    int vcycles = 0;    // V-cycles one application counts (added on each replay)
    bool warm = false;  // one eager application done: storage allocated
  };
  std::map<const float *, PGraph> pgraphs;
};

// A network bound to a context: the correction acts on staggered level
// `level` of the context's hierarchy.
struct lmm_cnn {
  lmm_ctx *ctx = nullptr;
  CnnNet *net = nullptr;
  lmm_cnn_config cfg;
  uint16_t *xin = nullptr;  // BF16 input of one application
  float *yout = nullptr;    // FP32 output of one application
  double *ss = nullptr;     // sum of squares of the coarse residual
};

static int set_err(lmm_ctx *c, int code, const char *fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

#define HIPCHK(ctx, call)                                                                    \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return set_err((ctx), LMM_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_));  \
  } while (0)

// Device allocation, zeroed in order on the owning context's stream (a
// blocking stream, so it is also ordered with the legacy null stream).
template <typename T>
static T *dalloc(hipStream_t stream, long n) {
  void *p = nullptr;
  size_t bytes = sizeof(T) * (size_t)(n > 0 ? n : 1);
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  if (hipMemsetAsync(p, 0, bytes, stream) != hipSuccess) {
    hipFree(p);
    return nullptr;
  }
  return (T *)p;
}

static void dfree(void *p) {
  if (p) hipFree(p);
}

// ------------------------- profiling ------------------------------------

// Times the launches made inside its lifetime with an event pair on the
// context's stream.  A kernel-exact scope (exact = true) instead hands the
// pair to the next fused-sweep launch, whose dispatch timestamps it; a pair
// no launch took is dropped, not counted.
struct ProfScope {
  lmm_ctx *c;
  ProfFamily *f = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  bool exact = false;
  double bytes = 0.0;
  ProfScope(lmm_ctx *ctx, const char *name, double nbytes, bool kernel_exact = false) : c(ctx) {
    if (!c->prof || (c->prof == 2 && !kernel_exact)) return;
    f = &c->fam[name];
    exact = kernel_exact;
    bytes = nbytes;
    a = take();
    b = take();
    if (exact) *launch_events() = {a, b};
    else hipEventRecord(a, c->stream);
  }
  ~ProfScope() {
    if (!f) return;
    if (exact) {
      LaunchEvents *ev = launch_events();
      if (ev->start == a) {  // not taken by a launch
        *ev = {nullptr, nullptr};
        c->evpool.push_back(a);
        c->evpool.push_back(b);
        return;
      }
    } else {
      hipEventRecord(b, c->stream);
    }
    f->launches++;
    f->bytes += bytes;
    f->pending.push_back({a, b});
  }
  hipEvent_t take() {
    if (c->evpool.empty()) {
      hipEvent_t e;
      hipEventCreate(&e);
      return e;
    }
    hipEvent_t e = c->evpool.back();
    c->evpool.pop_back();
    return e;
  }
};

static void prof_flush(lmm_ctx *c) {
  for (auto &kv : c->fam) {
    for (auto &ev : kv.second.pending) {
      hipEventSynchronize(ev.second);
      float ms = 0.f;
      hipEventElapsedTime(&ms, ev.first, ev.second);
      kv.second.ms += ms;
      c->evpool.push_back(ev.first);
      c->evpool.push_back(ev.second);
    }
    kv.second.pending.clear();
  }
}

// ------------------------- helpers --------------------------------------

static int compute_nlevs(const Grid &g, int minwidth) {
  int nl = -1;
  for (int d = 0; d < g.dim; d++) {
    int temp = g.n[d], rdir = 1;
    while (temp % 2 == 0 && temp / minwidth != 1) {
      temp /= 2;
      rdir++;
    }
    nl = (nl == -1) ? rdir : std::min(nl, rdir);
  }
  return nl;
}

static long kind_count(const lmm_ctx *c) { return c->g.N; }

static bool is_face(int kind) { return kind >= LMM_FACE_X && kind <= LMM_FACE_Z; }

template <typename T>
static int alloc_levels(lmm_ctx *c, std::vector<SLevel<T>> &S, std::vector<CLevel<T>> &CC) {
  int dim = c->g.dim;
  S.assign(c->nl_s, SLevel<T>());
  Grid g = c->g;
  for (int n = 0; n < c->nl_s; n++) {
    S[n].g = g;
    long N = g.N;
    for (int d = 0; d < dim; d++) {
      S[n].alpha[d] = dalloc<T>(c->stream, N);
      S[n].res[d] = dalloc<T>(c->stream, N);
      if (n > 0 || sizeof(T) == 4) {
        S[n].phi[d] = dalloc<T>(c->stream, N);
        S[n].rhs[d] = dalloc<T>(c->stream, N);
      }
      if (!S[n].alpha[d] || !S[n].res[d]) return LMM_ERR_NOMEM;
      if ((n > 0 || sizeof(T) == 4) && (!S[n].phi[d] || !S[n].rhs[d])) return LMM_ERR_NOMEM;
    }
    if (n > 0 || sizeof(T) == 4) {
      S[n].beta = dalloc<T>(c->stream, N);
      S[n].gamma = dalloc<T>(c->stream, N);
      if (!S[n].beta || !S[n].gamma) return LMM_ERR_NOMEM;
      for (int e = 0; e < (dim == 2 ? 1 : 3); e++) {
        S[n].bed[e] = dalloc<T>(c->stream, N);
        if (!S[n].bed[e]) return LMM_ERR_NOMEM;
      }
    }
    g = coarsen(g);
  }
  CC.assign(c->nl_c, CLevel<T>());
  g = c->g;
  for (int n = 0; n < c->nl_c; n++) {
    CC[n].g = g;
    long N = g.N;
    for (int d = 0; d < dim; d++) {
      CC[n].a[d] = dalloc<T>(c->stream, N);
      if (!CC[n].a[d]) return LMM_ERR_NOMEM;
    }
    if (n > 0 || sizeof(T) == 4) {
      CC[n].phi = dalloc<T>(c->stream, N);
      CC[n].rhs = dalloc<T>(c->stream, N);
      if (!CC[n].phi || !CC[n].rhs) return LMM_ERR_NOMEM;
    }
    CC[n].res = dalloc<T>(c->stream, N);
    if (!CC[n].res) return LMM_ERR_NOMEM;
    g = coarsen(g);
  }
  return LMM_OK;
}

template <typename T>
static void free_levels(std::vector<SLevel<T>> &S, std::vector<CLevel<T>> &CC, bool own0) {
  for (size_t n = 0; n < S.size(); n++) {
    for (int d = 0; d < 3; d++) {
      dfree(S[n].alpha[d]);
      dfree(S[n].res[d]);
      if (n > 0 || own0) {
        dfree(S[n].phi[d]);
        dfree(S[n].rhs[d]);
        dfree(S[n].bed[d]);
      }
    }
    if (n > 0 || own0) {
      dfree(S[n].beta);
      dfree(S[n].gamma);
    }
  }
  for (size_t n = 0; n < CC.size(); n++) {
    for (int d = 0; d < 3; d++) dfree(CC[n].a[d]);
    if (n > 0 || own0) {
      dfree(CC[n].phi);
      dfree(CC[n].rhs);
    }
    dfree(CC[n].res);
  }
  S.clear();
  CC.clear();
}

static void drop_graphs(lmm_ctx *c) {
  for (auto &kv : c->pgraphs)
    if (kv.second.exec) hipGraphExecDestroy(kv.second.exec);
  c->pgraphs.clear();
}

static int ensure_krylov(lmm_ctx *c, int m) {
  if (c->kry_m >= m) return LMM_OK;
  drop_graphs(c);  // the slots move
  long L = (long)(c->g.dim + 1) * c->g.N;
  dfree(c->V);
  dfree(c->Z32);
  c->V = dalloc<double>(c->stream, L * (m + 1));
  if (!c->V) return set_err(c, LMM_ERR_NOMEM, "cannot allocate %d Krylov vectors", m + 1);
  if (c->prm.precision != LMM_FP64) {
    c->Z32 = dalloc<float>(c->stream, L * m);
    if (!c->Z32) return set_err(c, LMM_ERR_NOMEM, "cannot allocate FP32 preconditioned vectors");
  }
  c->kry_m = m;
  return LMM_OK;
}

// ------------------------- coefficient staging ---------------------------

struct CoefIn {
  const double *alpha[3];
  const double *beta;
  const double *bed[3];
  const double *gamma;
};

static int ne_of(int dim) { return dim == 2 ? 1 : 3; }

// Level-0 FP64 coefficients (alpha*theta, optional scaled copies), coarse
// levels by the reference's restrictions (stag_mg_solver.f90:200-230), the
// face coefficients 1/alpha of the preconditioner (apply_precon.f90:74-76)
// and the cell-centred hierarchy; FP32 copies when requested.
static void prepare_coefs(lmm_ctx *c, const CoefIn &in, double theta_alpha, double scale, bool need_cc,
                          bool need32) {
  hipStream_t s = c->stream;
  const Grid &g = c->g;
  int dim = g.dim;
  auto &S = c->s64;
  for (int d = 0; d < dim; d++) {
    // alpha_fc_mg(1) = alpha_fc * theta_alpha
    launch_scale_copy(s, S[0].alpha[d], in.alpha[d], g.N, theta_alpha);
  }
  if (scale != 1.0) {
    launch_scale_copy(s, c->scaled_beta, in.beta, g.N, scale);
    launch_scale_copy(s, c->scaled_gamma, in.gamma, g.N, scale);
    for (int e = 0; e < ne_of(dim); e++) launch_scale_copy(s, c->scaled_bed[e], in.bed[e], g.N, scale);
    S[0].beta = c->scaled_beta;
    S[0].gamma = c->scaled_gamma;
    for (int e = 0; e < ne_of(dim); e++) S[0].bed[e] = c->scaled_bed[e];
  } else {
    S[0].beta = const_cast<double *>(in.beta);
    S[0].gamma = const_cast<double *>(in.gamma);
    for (int e = 0; e < ne_of(dim); e++) S[0].bed[e] = const_cast<double *>(in.bed[e]);
  }
  for (int n = 1; n < c->nl_s; n++) {
    const Grid &gc = S[n].g;
    launch_cc_restrict<double>(s, gc, S[n - 1].beta, S[n].beta);
    launch_cc_restrict<double>(s, gc, S[n - 1].gamma, S[n].gamma);
    launch_stag_restrict<double>(s, gc, S[n - 1].alpha, S[n].alpha, 1);
    if (dim == 2)
      launch_nodal_restrict<double>(s, gc, S[n - 1].bed[0], S[n].bed[0]);
    else
      launch_edge_restrict<double>(s, gc, S[n - 1].bed, S[n].bed);
  }
  if (need_cc) {
    for (int d = 0; d < dim; d++) launch_reciprocal(s, c->ainv[d], in.alpha[d], g.N);
    auto &C = c->c64;
    for (int d = 0; d < dim; d++) hipMemcpyAsync(C[0].a[d], c->ainv[d], sizeof(double) * g.N, hipMemcpyDeviceToDevice, s);
    for (int n = 1; n < c->nl_c; n++) launch_stag_restrict<double>(s, C[n].g, C[n - 1].a, C[n].a, 1);
  }
  if (need32) {
    auto &S32 = c->s32;
    for (int n = 0; n < c->nl_s; n++) {
      long N = S[n].g.N;
      for (int d = 0; d < dim; d++) launch_convert<float, double>(s, S32[n].alpha[d], S[n].alpha[d], N);
      launch_convert<float, double>(s, S32[n].beta, S[n].beta, N);
      launch_convert<float, double>(s, S32[n].gamma, S[n].gamma, N);
      for (int e = 0; e < ne_of(dim); e++) launch_convert<float, double>(s, S32[n].bed[e], S[n].bed[e], N);
    }
    auto &C = c->c64;
    auto &C32 = c->c32;
    for (int n = 0; n < c->nl_c; n++)
      for (int d = 0; d < dim; d++) launch_convert<float, double>(s, C32[n].a[d], C[n].a[d], C[n].g.N);
  }
}

// ------------------------- staggered multigrid ---------------------------

static double comp_sweep_bytes(const Grid &g, int visc, size_t es) {
  // compulsory traffic of one sweep (both colours) of one face component:
  // read phi, rhs and alpha of the component, beta_cc, the edge (3D) or node
  // (2D) arrays that touch the component (dim-1 of them, visc_type < 0),
  // gamma and the other components' phi where |visc_type| = 3 couples them;
  // write phi of the component.
  int dim = g.dim;
  int arrays = 5 + (visc < 0 ? dim - 1 : 0) + ((visc == 3 || visc == -3) ? dim : 0);
  return (double)arrays * (double)g.N * (double)es;
}

template <typename T>
static void smooth(lmm_ctx *c, SLevel<T> &L, T *const phi[3], T *const rhs[3], int nsweeps, bool zero_in = false) {
  const lmm_params &p = c->prm;
  int dim = L.g.dim;
  StagCoef<T> K = L.K();
  // the finest level is reported on its own: it carries the bulk of the bytes
  const char *fam = (L.g.N == c->g.N) ? "stag_smooth_l0" : "stag_smooth";
  for (int m = 0; m < nsweeps; m++) {
    if (p.stag_mg_smoother == 0) {
      ProfScope ps(c, fam, dim * comp_sweep_bytes(L.g, p.visc_type, sizeof(T)));
      launch_stag_jacobi<T>(c->stream, L.g, p.visc_type, (T)p.stag_mg_omega, phi, rhs, L.res, K);
      for (int d = 0; d < dim; d++)
        hipMemcpyAsync(phi[d], L.res[d], sizeof(T) * L.g.N, hipMemcpyDeviceToDevice, c->stream);
      continue;
    }
    const bool x3 = c->fused && c->fused3 && fused_rb3_ok(L.g, p.visc_type, c->fused_min_tiles);
    const bool x2 = c->fused && fused_rb2d_ok(L.g, p.visc_type, c->fused_min_tiles);
    const bool x2c = c->fused && fused_rb2d_comp_ok(L.g, p.visc_type, c->fused_min_tiles);
    if (x2 || x2c || x3 || (c->fused && fused_rb_ok(L.g, p.visc_type, c->fused_min_tiles))) {
      // sweep m reads one buffer and writes the other (phi <-> L.res); an
      // odd sweep count ends with a copy back
      T *const *src = (m % 2 == 0) ? phi : L.res;
      T *const *dst = (m % 2 == 0) ? L.res : phi;
      if (x2) {
        // both components in one launch: phi, rhs, alpha of each component,
        // beta_cc and the nodes (visc_type -1); phi written per component
        const bool zin = zero_in && m == 0;
        const double bytes = (double)(4 * dim + 1 + (p.visc_type < 0 ? 1 : 0) - (zin ? dim : 0)) * L.g.N * sizeof(T);
        ProfScope pf(c, "stag_rb_fused", bytes);
        ProfScope ps(c, fam, bytes, true);
        launch_stag_rb2d<T>(c->stream, L.g, p.visc_type, (T)p.stag_mg_omega, src, dst, rhs, K, zin);
      } else if (x3) {
        // all components in one launch: each array is compulsory once per
        // sweep (beta_cc and the edges shared by the components)
        // (a sweep from a zero guess does not read phi)
        const bool zin = zero_in && m == 0;
        const double bytes = (double)(4 * dim + 1 + (p.visc_type < 0 ? 3 : 0) - (zin ? dim : 0)) * L.g.N * sizeof(T);
        ProfScope pf(c, "stag_rb_fused", bytes);
        ProfScope ps(c, fam, bytes, true);
        launch_stag_rb_fused3<T>(c->stream, L.g, p.visc_type, (T)p.stag_mg_omega, src, dst, rhs, K,
                                 c->fused_min_tiles, zin);
      } else {
        // one component at a time; visc_type 2 and 3 read the other
        // components in their state at this component's turn of the colour
        // order: the ones before it already swept (dst), the later ones old
        for (int a = 0; a < dim; a++) {
          const T *cur[3] = {nullptr, nullptr, nullptr};
          for (int b = 0; b < dim; b++) cur[b] = (b < a) ? dst[b] : src[b];
          ProfScope pf(c, "stag_rb_fused", comp_sweep_bytes(L.g, p.visc_type, sizeof(T)));
          ProfScope ps(c, fam, comp_sweep_bytes(L.g, p.visc_type, sizeof(T)), true);
          if (x2c)
            launch_stag_rb2d_comp<T>(c->stream, L.g, p.visc_type, a, (T)p.stag_mg_omega, src[a], dst[a], rhs[a], K,
                                     cur[1 - a]);
          else
            launch_stag_rb_fused<T>(c->stream, L.g, p.visc_type, a, (T)p.stag_mg_omega, src[a], dst[a], rhs[a], K,
                                    c->fused_min_tiles, cur);
        }
      }
      if (m == nsweeps - 1 && m % 2 == 0)
        for (int a = 0; a < dim; a++)
          hipMemcpyAsync(phi[a], L.res[a], sizeof(T) * L.g.N, hipMemcpyDeviceToDevice, c->stream);
      continue;
    }
    for (int color = 1; color <= 2 * dim; color++) {
      // a colour pass touches every cache line of the component's arrays
      ProfScope ps(c, fam, comp_sweep_bytes(L.g, p.visc_type, sizeof(T)));
      launch_stag_gs_colour<T>(c->stream, L.g, p.visc_type, (color - 1) / 2, (color + 1) % 2,
                               (T)p.stag_mg_omega, phi, K, rhs[(color - 1) / 2]);
    }
  }
}

// max |rhs - L phi| per component on level 0 (FP64), synchronises
static void resid_inf(lmm_ctx *c, SLevel<double> &L, double *const phi[3], double *const rhs[3], double out[3]) {
  launch_stag_residual<double>(c->stream, L.g, c->prm.visc_type, phi, rhs, L.res, L.K());
  launch_face_maxabs(c->stream, L.g, L.res, c->partial);
  launch_finalize_max(c->stream, c->partial, c->scal + 4000);
  hipMemcpyAsync(c->hscal, c->scal + 4000, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  for (int d = 0; d < 3; d++) out[d] = c->hscal[d];
}

// max |rhs| per component on level 0 (FP64), synchronises
static void rhs_inf(lmm_ctx *c, SLevel<double> &L, double *const rhs[3], double out[3]) {
  launch_face_maxabs(c->stream, L.g, rhs, c->partial);
  launch_finalize_max(c->stream, c->partial, c->scal + 4000);
  hipMemcpyAsync(c->hscal, c->scal + 4000, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  for (int d = 0; d < 3; d++) out[d] = c->hscal[d];
}

// One call of stag_mg_solver (stag_mg_solver.f90:37-868) on prepared levels.
// phi0/rhs0 are the level-0 arrays; phi_zero tells that phi0 is zero on
// entry, in which case the initial-residual exit of :276-306 cannot trigger
// unless rhs is zero, and a V-cycle of zero data returns zero exactly.
// A level whose sweeps run as all-component fused sweeps can start its first
// smoothing of a cycle from a zero field without reading it: then its phi
// need not be zeroed first (every point of the sweep's output is written).
template <typename T>
static bool sweep_takes_zeros(lmm_ctx *c, const SLevel<T> &L) {
  const lmm_params &p = c->prm;
  return c->fused && p.stag_mg_smoother == 1 &&
         ((c->fused3 && fused_rb3_ok(L.g, p.visc_type, c->fused_min_tiles)) ||
          fused_rb2d_ok(L.g, p.visc_type, c->fused_min_tiles));
}

template <typename T>
static void stag_vcycles(lmm_ctx *c, std::vector<SLevel<T>> &S, T *const phi0[3], T *const rhs0[3],
                         bool phi_zero) {
  const lmm_params &p = c->prm;
  int dim = c->g.dim;
  int nl = c->nl_s;
  double resid0[3] = {1, 1, 1};
  bool need_check = (p.stag_mg_max_vcycles > 1) || !phi_zero;
  if constexpr (std::is_same<T, double>::value) {
   if (need_check) {
    // With a zero guess (apply_precon zeroes x_u first, apply_precon.f90:88-90)
    // the initial residual L 0 - rhs is -rhs bit for bit, so max |rhs| is
    // taken without reading phi0, whose contents are not the guess.
    if (phi_zero) rhs_inf(c, S[0], rhs0, resid0);
    else resid_inf(c, S[0], phi0, rhs0, resid0);
    bool allzero = true, anyzero = false;
    double mx = 0;
    for (int d = 0; d < dim; d++) {
      if (resid0[d] != 0.0) allzero = false;
      else anyzero = true;
      mx = std::max(mx, resid0[d]);
    }
    if (allzero) {
      // stag_mg_solver.f90:276-306 returns the guess: zero here
      if (phi_zero)
        for (int d = 0; d < dim; d++) launch_fill<T>(c->stream, phi0[d], S[0].g.N, (T)0);
      return;
    }
    if (anyzero)
      for (int d = 0; d < dim; d++) resid0[d] = mx;
   }
  }
  // levels nc..nl-1 run in the single-workgroup coarse launch (Gauss-Seidel
  // smoother only); nc = nl when there is none
  int nc = nl;
  if (c->coarse_one_launch && p.stag_mg_smoother == 1)
    for (int n = 1; n < nl; n++)
      if (S[n].g.N <= c->coarse_max_cells && nl - n <= COARSE_MAX_LEVELS) {
        nc = n;
        break;
      }
  std::vector<CoarseStagLevel<T>> clv;
  for (int n = nc; n < nl; n++) {
    CoarseStagLevel<T> L;
    L.g = S[n].g;
    for (int d = 0; d < 3; d++) {
      L.phi[d] = S[n].phi[d];
      L.rhs[d] = S[n].rhs[d];
      L.res[d] = S[n].res[d];
    }
    L.K = S[n].K();
    clv.push_back(L);
  }
  // Levels 0..nc-1 are smoothed here; a level's first smoothing in a cycle
  // is the down-leg one, or the bottom one for the last level (nc = nl).
  // When that smoothing has at least one sweep and takes zeros, the level's
  // phi is not zeroed: phi0 (zero guess, first cycle, no initial residual
  // check) and levels 1..nc-1 (zeroed every cycle).
  auto takes_zeros = [&](int n) {
    const int first = (n < nl - 1) ? p.stag_mg_nsmooths_down : p.stag_mg_nsmooths_bottom;
    return first >= 1 && sweep_takes_zeros(c, S[n]);
  };
  const bool zero0 = phi_zero && !need_check && takes_zeros(0);
  if (phi_zero && !zero0)
    for (int d = 0; d < dim; d++) launch_fill<T>(c->stream, phi0[d], S[0].g.N, (T)0);
  auto zin = [&](int n, int vc) { return (n == 0) ? (zero0 && vc == 1) : takes_zeros(n); };
  for (int vc = 1; vc <= p.stag_mg_max_vcycles; vc++) {
    for (int n = 1; n < nc; n++)
      if (!takes_zeros(n))
        for (int d = 0; d < dim; d++) launch_fill<T>(c->stream, S[n].phi[d], S[n].g.N, (T)0);
    for (int n = 0; n < std::min(nc, nl - 1); n++) {
      T *const *ph = (n == 0) ? phi0 : S[n].phi;
      T *const *rh = (n == 0) ? rhs0 : S[n].rhs;
      smooth<T>(c, S[n], ph, rh, p.stag_mg_nsmooths_down, zin(n, vc));
      {
        // read phi, rhs and the coefficients of every component, write the
        // coarse right-hand side
        const double fine = dim * comp_sweep_bytes(S[n].g, p.visc_type, sizeof(T)) - (double)dim * S[n].g.N * sizeof(T);
        const double bytes = fine + (double)dim * S[n + 1].g.N * sizeof(T);
        ProfScope ps(c, (n == 0) ? "stag_restrict_l0" : "stag_restrict", bytes);
        if (c->fused_resid_restrict && p.stag_mg_smoother == 1 &&
            stag_resid_restrict_ok(S[n].g, S[n + 1].g, p.visc_type)) {
          ProfScope pf(c, "stag_rr_fused", bytes);
          launch_stag_resid_restrict<T>(c->stream, S[n].g, S[n + 1].g, p.visc_type, ph, rh, S[n + 1].rhs,
                                        S[n].K());
        } else {
          launch_stag_residual<T>(c->stream, S[n].g, p.visc_type, ph, rh, S[n].res, S[n].K());
          launch_stag_restrict<T>(c->stream, S[n + 1].g, S[n].res, S[n + 1].rhs, 0);
        }
      }
    }
    if (nc < nl) {
      ProfScope ps(c, "stag_coarse", 0.0);
      launch_stag_coarse_vcycle<T>(c->stream, nl - nc, clv.data(), p.visc_type, (T)p.stag_mg_omega,
                                   p.stag_mg_nsmooths_down, p.stag_mg_nsmooths_up, p.stag_mg_nsmooths_bottom);
    } else {
      int n = nl - 1;
      T *const *ph = (n == 0) ? phi0 : S[n].phi;
      T *const *rh = (n == 0) ? rhs0 : S[n].rhs;
      smooth<T>(c, S[n], ph, rh, p.stag_mg_nsmooths_bottom, zin(n, vc));
    }
    for (int n = std::min(nc, nl - 1) - 1; n >= 0; n--) {
      T *const *ph = (n == 0) ? phi0 : S[n].phi;
      T *const *rh = (n == 0) ? rhs0 : S[n].rhs;
      {
        ProfScope ps(c, "stag_prolong", 0.0);
        launch_stag_prolong<T>(c->stream, S[n].g, ph, S[n + 1].phi);
      }
      smooth<T>(c, S[n], ph, rh, p.stag_mg_nsmooths_up);
    }
    c->vcycles += dim;
    if constexpr (std::is_same<T, double>::value) {
     if (vc < p.stag_mg_max_vcycles) {
      double r[3];
      resid_inf(c, S[0], phi0, rhs0, r);
      bool done = true;
      for (int d = 0; d < dim; d++)
        if (!(r[d] / resid0[d] < p.stag_mg_rel_tol)) done = false;
      if (done) break;
     }
    }
  }
}

// ------------------------- cell-centred multigrid ------------------------

// abs_eps of the cell-centred multigrid's stopping test (macproject.f90:423)
static constexpr double CC_ABS_EPS = 1e-16;

// max |a| over a cell array of grid g (FP64), synchronises
static double cell_inf(lmm_ctx *c, const Grid &g, const double *a) {
  Grid g1 = g;
  g1.dim = 1;  // the face reduction over one array
  const double *r[3] = {a, nullptr, nullptr};
  launch_face_maxabs(c->stream, g1, r, c->partial);
  launch_finalize_max(c->stream, c->partial, c->scal + 4030);
  hipMemcpyAsync(c->hscal, c->scal + 4030, sizeof(double), hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  return c->hscal[0];
}

template <typename T>
static void cc_vcycles(lmm_ctx *c, std::vector<CLevel<T>> &C, T *phi0, const T *rhs0) {
  const lmm_params &p = c->prm;
  int nl = c->nl_c;
  auto fused2d_at = [&](int n) { return c->fused && cc_rb2d_ok(C[n].g, c->fused_min_tiles); };
  auto fused_at = [&](int n) {
    return fused2d_at(n) || (c->fused && fused_rb_tz(C[n].g, c->fused_min_tiles, 4, 1) > 0);
  };
  // a level's first down-leg sweep from its zero guess can take the zeros
  // without reading them (every point of its output is written): then phi is
  // not zeroed first
  auto takes_zeros = [&](int n) {
    const int first = (n < nl - 1) ? p.mg_nsmooths_down : p.mg_nsmooths_bottom;
    return first >= 1 && fused_at(n) && (fused2d_at(n) || fused_rb_takes_zeros());
  };
  // The stopping test of the cell-centred multigrid: max |rhs - L phi| <=
  // mg_rel_tol * max |rhs| or <= 1e-16, the eps and abs_eps that
  // mgt_macproj_precon_build gives the tower (macproject.f90:402-425),
  // checked before the first cycle and after each cycle but the last.  It is
  // evaluated (FP64 only) when more than one cycle may run; with one cycle
  // the cycle always runs.
  const bool check = std::is_same<T, double>::value && p.mg_max_vcycles > 1;
  double bnorm = 0.0;
  if constexpr (std::is_same<T, double>::value) {
    if (check) {
      bnorm = cell_inf(c, c->g, rhs0);
      if (bnorm <= p.mg_rel_tol * bnorm || bnorm <= CC_ABS_EPS) {
        launch_fill<T>(c->stream, phi0, c->g.N, (T)0);
        return;
      }
    }
  }
  if (!takes_zeros(0)) launch_fill<T>(c->stream, phi0, c->g.N, (T)0);
  auto rb = [&](int n, int sweeps, bool zero_in) {
    T *ph = (n == 0) ? phi0 : C[n].phi;
    const T *rh = (n == 0) ? rhs0 : C[n].rhs;
    if (fused_at(n)) {
      // fused red-black sweeps, ping-pong between phi and C[n].res
      for (int m = 0; m < sweeps; m++) {
        // read phi, rhs and the dim face coefficients, write phi
        const bool zin = (m == 0 && zero_in);  // from a zero guess: phi not read
        ProfScope ps(c, (n == 0) ? "cc_smooth_l0" : "cc_smooth",
                     (double)(3 + C[n].g.dim - (zin ? 1 : 0)) * C[n].g.N * sizeof(T), true);
        const T *src = zin ? nullptr : (m % 2 == 0) ? ph : C[n].res;
        T *dst = (m % 2 == 0) ? C[n].res : ph;
        if (fused2d_at(n)) launch_cc_rb2d<T>(c->stream, C[n].g, src, dst, rh, C[n].a);
        else launch_cc_rb_fused<T>(c->stream, C[n].g, src, dst, rh, C[n].a, c->fused_min_tiles);
      }
      if (sweeps % 2 == 1) hipMemcpyAsync(ph, C[n].res, sizeof(T) * C[n].g.N, hipMemcpyDeviceToDevice, c->stream);
      return;
    }
    for (int m = 0; m < sweeps; m++)
      for (int colour = 0; colour < 2; colour++) {
        ProfScope ps(c, "cc_smooth", 0.0);
        launch_cc_rb_colour<T>(c->stream, C[n].g, colour, ph, rh, C[n].a);
      }
  };
  int nc = nl;
  if (c->coarse_one_launch)
    for (int n = 1; n < nl; n++)
      if (C[n].g.N <= c->coarse_max_cells && nl - n <= COARSE_MAX_LEVELS) {
        nc = n;
        break;
      }
  std::vector<CoarseCCLevel<T>> clv;
  for (int n = nc; n < nl; n++) {
    CoarseCCLevel<T> L;
    L.g = C[n].g;
    L.phi = C[n].phi;
    L.rhs = C[n].rhs;
    L.res = C[n].res;
    for (int d = 0; d < 3; d++) L.a[d] = C[n].a[d];
    clv.push_back(L);
  }
  for (int cyc = 0; cyc < p.mg_max_vcycles; cyc++) {
    for (int n = 1; n < nc; n++)
      if (!takes_zeros(n)) launch_fill<T>(c->stream, C[n].phi, C[n].g.N, (T)0);
    for (int n = 0; n < std::min(nc, nl - 1); n++) {
      rb(n, p.mg_nsmooths_down, takes_zeros(n) && (n > 0 || cyc == 0));  // first smoothing of the level
      T *ph = (n == 0) ? phi0 : C[n].phi;
      const T *rh = (n == 0) ? rhs0 : C[n].rhs;
      if (c->fused_resid_restrict) {
        // read phi, rhs and the dim face coefficients, write the coarse rhs
        ProfScope ps(c, (n == 0) ? "cc_restrict_l0" : "cc_restrict",
                     ((double)(2 + C[n].g.dim) * C[n].g.N + (double)C[n + 1].g.N) * sizeof(T));
        launch_cc_resid_restrict<T>(c->stream, C[n].g, ph, rh, C[n].a, C[n + 1].g, C[n + 1].rhs);
      } else {
        launch_cc_residual<T>(c->stream, C[n].g, ph, rh, C[n].res, C[n].a);
        launch_cc_restrict<T>(c->stream, C[n + 1].g, C[n].res, C[n + 1].rhs);
      }
    }
    if (nc < nl) {
      ProfScope ps(c, "cc_coarse", 0.0);
      launch_cc_coarse_vcycle<T>(c->stream, nl - nc, clv.data(), p.mg_nsmooths_down, p.mg_nsmooths_up,
                                 p.mg_nsmooths_bottom);
    } else {
      rb(nl - 1, p.mg_nsmooths_bottom, takes_zeros(nl - 1) && (nl - 1 > 0 || cyc == 0));
    }
    for (int n = std::min(nc, nl - 1) - 1; n >= 0; n--) {
      T *ph = (n == 0) ? phi0 : C[n].phi;
      launch_cc_prolong_add<T>(c->stream, C[n].g, ph, C[n + 1].phi);
      rb(n, p.mg_nsmooths_up, false);
    }
    c->vcycles += 1;
    if constexpr (std::is_same<T, double>::value) {
      if (check && cyc + 1 < p.mg_max_vcycles) {
        launch_cc_residual<T>(c->stream, C[0].g, phi0, rhs0, C[0].res, C[0].a);
        const double r = cell_inf(c, c->g, C[0].res);
        if (r <= p.mg_rel_tol * bnorm || r <= CC_ABS_EPS) break;
      }
    }
  }
}

// ------------------------- Stokes pieces ---------------------------------

struct SV {
  double *u[3];
  double *p;
};

static SV sv_of(lmm_ctx *c, double *base) {
  SV v;
  long N = c->g.N;
  for (int d = 0; d < 3; d++) v.u[d] = (d < c->g.dim) ? base + d * N : nullptr;
  v.p = base + c->g.dim * N;
  return v;
}

template <typename T>
struct SVT {
  T *u[3];
  T *p;
};

template <typename T>
static SVT<T> svt_of(lmm_ctx *c, T *base) {
  SVT<T> v;
  long N = c->g.N;
  for (int d = 0; d < 3; d++) v.u[d] = (d < c->g.dim) ? base + d * N : nullptr;
  v.p = base + c->g.dim * N;
  return v;
}

static void apply_matrix_dev(lmm_ctx *c, const SV &x, const SV &b) {
  ProfScope ps(c, "apply_matrix", (double)c->g.N * 8.0 * (c->g.dim == 3 ? 15 : 10));
  launch_apply_matrix<double>(c->stream, c->g, c->prm.visc_type, x.u, x.p, b.u, b.p, c->s64[0].K());
}

// r = b - A x with b the context's right-hand side (k_bmax_neg's
// arithmetic on A x; one pass where the fused kernel applies)
static void residual_dev(lmm_ctx *c, const SV &x, double *r) {
  SV b = sv_of(c, c->b), rv = sv_of(c, r);
  const int visc = c->prm.visc_type;
  if (c->fused_krylov && c->g.dim == 3 && (visc == -1 || visc == 1)) {
    // read x, b and the coefficients, write r
    ProfScope ps(c, "apply_matrix", (double)c->g.N * 8.0 * 19);
    launch_apply_matrix_resid(c->stream, c->g, visc, x.u, x.p, b.u, b.p, rv.u, rv.p, c->s64[0].K());
    return;
  }
  apply_matrix_dev(c, x, rv);
  launch_bmax_neg(c->stream, r, c->b, (long)(c->g.dim + 1) * c->g.N);
}

// FP64 A z of an FP32 Stokes vector z (same bits as widening z first)
static void apply_matrix_z32(lmm_ctx *c, float *z, const SV &b) {
  SVT<float> zf = svt_of<float>(c, z);
  if (c->z32_direct) {
    ProfScope ps(c, "apply_matrix", (double)c->g.N * (c->g.dim == 3 ? 11 * 8.0 + 4 * 4.0 : 10 * 8.0));
    if (launch_apply_matrix_x32(c->stream, c->g, c->prm.visc_type, zf.u, zf.p, b.u, b.p, c->s64[0].K())) return;
  }
  long L = (long)(c->g.dim + 1) * c->g.N;
  launch_convert<double, float>(c->stream, c->r, z, L);
  apply_matrix_dev(c, sv_of(c, c->r), b);
}

// The learned coarse-grid correction (no counterpart in the reference) of a
// preconditioned Stokes vector z = M(v): the residual v - A z (FP32),
// restricted to level lc (velocity with the 12-point weights, pressure with
// the 8-cell mean), goes through the network with the level's
// coefficients; its prediction of the error, prolongated back to the finest
// level (staggered and cell-centred prolongation), is added to z.
static void learned_stage(lmm_ctx *c, float *const vu[3], const float *vp, float *const zu[3], float *zp) {
  hipStream_t s = c->stream;
  auto &S = c->s32;
  auto &C = c->c32;
  lmm_cnn *nn = c->cnn;
  const int lc = nn->cfg.level, visc = c->prm.visc_type;
  const long N = c->g.N;
  ProfScope ps(c, "learned", 0.0);
  float *ru[3] = {S[0].res[0], S[0].res[1], S[0].res[2]};
  float *rp = C[0].res;
  launch_apply_matrix<float>(s, S[0].g, visc, zu, zp, ru, rp, S[0].K());
  for (int d = 0; d < 3; d++) launch_sub_f32(s, ru[d], vu[d], ru[d], N);
  launch_sub_f32(s, rp, vp, rp, N);
  launch_stag_restrict<float>(s, S[1].g, ru, S[1].rhs, 0);
  launch_cc_restrict<float>(s, C[1].g, rp, C[1].rhs);
  for (int n = 1; n < lc; n++) {
    launch_stag_restrict<float>(s, S[n + 1].g, S[n].rhs, S[n + 1].rhs, 0);
    launch_cc_restrict<float>(s, C[n + 1].g, C[n].rhs, C[n + 1].rhs);
  }
  const Grid &gc = S[lc].g;
  launch_stokes_ss(s, gc, S[lc].rhs, C[lc].rhs, c->partial, nn->ss);
  launch_cnn_pack(s, gc, S[lc].rhs, C[lc].rhs, S[lc].beta, S[lc].alpha, nn->ss, nn->xin);
  const int gn[3] = {gc.n[0], gc.n[1], gc.n[2]};
  cnn_forward(nn->net, 1, gn, nn->xin, nn->yout);
  launch_cnn_unpack(s, gc, nn->yout, nn->ss, c->partial, nn->ss + 2, S[lc].phi, C[lc].phi);
  for (int n = lc - 1; n >= 1; n--) {
    for (int d = 0; d < 3; d++) launch_fill<float>(s, S[n].phi[d], S[n].g.N, 0.f);
    launch_fill<float>(s, C[n].phi, C[n].g.N, 0.f);
    launch_stag_prolong<float>(s, S[n].g, S[n].phi, S[n + 1].phi);
    launch_cc_prolong_add<float>(s, C[n].g, C[n].phi, C[n + 1].phi);
  }
  launch_stag_prolong<float>(s, S[0].g, zu, S[1].phi);
  launch_cc_prolong_add<float>(s, C[0].g, zp, C[1].phi);
  c->learned_applied++;
}

// apply_precon.f90:29-195,434-460 with precon_type 1 (T = FP64 or FP32).
template <typename T>
static void apply_precon_t(lmm_ctx *c, std::vector<SLevel<T>> &S, std::vector<CLevel<T>> &C,
                           T *const bu[3], const T *bp, T *const xu[3], T *xp, T *mac_rhs, T *phi,
                           double theta_alpha) {
  hipStream_t s = c->stream;
  const Grid &g = c->g;
  int dim = g.dim;
  stag_vcycles<T>(c, S, xu, bu, true);  // x_u from a zero guess
  launch_div_plus<T>(s, g, xu, bp, mac_rhs, 1);
  cc_vcycles<T>(c, C, phi, mac_rhs);
  T *ainv[3] = {C[0].a[0], C[0].a[1], C[0].a[2]};
  const T *beta = S[0].beta, *gamma = S[0].gamma;
  // beta/gamma of the preconditioner are the unscaled-by-theta level-0 cell
  // coefficients (scale_factor already folded in by prepare_coefs)
  int nb = (int)((g.N + 255) / 256);
  launch_precon_final<T>(s, g, c->prm.visc_type, xu, xp, phi, mac_rhs, ainv, beta, gamma, theta_alpha,
                         c->partial2);
  launch_fold_cells(s, c->partial2, nb, c->partial);
  launch_finalize(s, g, c->partial, c->scal + 4010, 3, 1.0);
  // the pressure mean (and velocity means when theta_alpha = 0) are
  // accumulated and subtracted in FP64 also for FP32 fields
  launch_sub_scalar<T>(s, xp, g.N, c->scal + 4010);
  if (theta_alpha == 0.0) {
    launch_sum_faces_weighted<T>(s, g, xu, c->partial);
    launch_finalize(s, g, c->partial, c->scal + 4020, 4, 1.0);
    launch_sub_scalar3<T>(s, g, xu, c->scal + 4020);
  }
}

static void precon64(lmm_ctx *c, const SV &b, const SV &x, double theta_alpha) {
  apply_precon_t<double>(c, c->s64, c->c64, b.u, b.p, x.u, x.p, c->mac_rhs, c->phi_cc, theta_alpha);
}

// ------------------------- reductions to host ----------------------------

// dot / norm of Stokes vectors; result in device slot, optional host copy
static void stokes_dot_dev(lmm_ctx *c, const double *a, const double *b, double *slot, int op) {
  ProfScope ps(c, "krylov", (double)c->g.N * 8.0 * 2 * (c->g.dim + 1));
  launch_stokes_dot(c->stream, c->g, a, b, c->partial);
  launch_finalize(c->stream, c->g, c->partial, slot, op, c->prm.p_norm_weight);
}

// Modified Gram-Schmidt of w against V_0..V_{i-1} (gmres.f90:338-358): the
// entries H(1..i, i) into H[0..i-1] and |w| (with its velocity and pressure
// parts) into nrm[0..2] (device slots).  Fused form: the first dot, then each
// axmy fused with the next dot or, last, with |w|^2; the same arithmetic and
// summation order as the separate passes, so the same bits.
template <typename VF>
static void mgs(lmm_ctx *c, const VF &Vk, int i, double *H, double *nrm) {
  hipStream_t s = c->stream;
  const long L = (long)(c->g.dim + 1) * c->g.N;
  if (!c->fused_krylov) {
    for (int k = 1; k <= i; k++) {
      stokes_dot_dev(c, c->w, Vk(k - 1), H + k - 1, 0);
      ProfScope ps(c, "krylov", (double)L * 8.0 * 3);
      launch_axmy(s, c->w, Vk(k - 1), L, H + k - 1);
    }
    stokes_dot_dev(c, c->w, c->w, nrm, 1);
    return;
  }
  stokes_dot_dev(c, c->w, Vk(0), H, 0);
  for (int k = 1; k <= i; k++) {
    const bool last = (k == i);
    ProfScope ps(c, "krylov", (double)L * 8.0 * (last ? 3 : 4));
    launch_axmy_dot(s, c->g, c->w, Vk(k - 1), H + k - 1, last ? nullptr : Vk(k), c->partial);
    launch_finalize(s, c->g, c->partial, last ? nrm : H + k, last ? 1 : 0, c->prm.p_norm_weight);
  }
}

static void fetch(lmm_ctx *c, const double *dev, int n, double *host) {
  hipMemcpyAsync(c->hscal, dev, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  memcpy(host, c->hscal, sizeof(double) * n);
}

// ------------------------- GMRES host scalars (gmres.f90:517-605) --------

static void rotmat(double a, double b, double &cs, double &sn) {
  if (b == 0.0) {
    cs = 1.0;
    sn = 0.0;
  } else if (std::fabs(b) > std::fabs(a)) {
    double temp = a / b;
    sn = 1.0 / std::sqrt(1.0 + temp * temp);
    cs = temp * sn;
  } else {
    double temp = b / a;
    cs = 1.0 / std::sqrt(1.0 + temp * temp);
    sn = temp * cs;
  }
}

struct Hess {
  int m;
  std::vector<double> H, cs, sn, s, y;
  explicit Hess(int mm) : m(mm), H((size_t)(mm + 1) * mm, 0.0), cs(mm), sn(mm), s(mm + 1), y(mm) {}
  double &h(int r, int col) { return H[(size_t)(r - 1) + (size_t)(m + 1) * (col - 1)]; }
  void least_squares(int i) {
    for (int k = 1; k <= i - 1; k++) {
      double temp = cs[k - 1] * h(k, i) + sn[k - 1] * h(k + 1, i);
      h(k + 1, i) = -sn[k - 1] * h(k, i) + cs[k - 1] * h(k + 1, i);
      h(k, i) = temp;
    }
    rotmat(h(i, i), h(i + 1, i), cs[i - 1], sn[i - 1]);
    double temp = cs[i - 1] * s[i - 1];
    s[i] = -sn[i - 1] * s[i - 1];
    s[i - 1] = temp;
    h(i, i) = cs[i - 1] * h(i, i) + sn[i - 1] * h(i + 1, i);
    h(i + 1, i) = 0.0;
  }
  void solve(int k) {
    y[k] = s[k] / h(k + 1, k + 1);
    for (int i = k; i >= 1; i--) {
      double dp = 0.0;
      for (int q = i + 1; q <= k + 1; q++) dp = dp + h(i, q) * y[q - 1];
      y[i - 1] = (s[i - 1] - dp) / h(i, i);
    }
  }
};

// ------------------------- field checks ----------------------------------

static bool field_ok(lmm_ctx *c, const lmm_field *f, int kind, int ncomp = 1) {
  return f && f->ctx == c && f->kind == kind && f->ncomp >= ncomp;
}

static int check_stokes(lmm_ctx *c, lmm_field *const u[3], const lmm_field *p, const char *what) {
  for (int d = 0; d < c->g.dim; d++)
    if (!field_ok(c, u[d], LMM_FACE_X + d)) return set_err(c, LMM_ERR_ARG, "%s: face component %d missing or wrong kind", what, d);
  if (p && !field_ok(c, p, LMM_CELL)) return set_err(c, LMM_ERR_ARG, "%s: cell field missing or wrong kind", what);
  return LMM_OK;
}

static int check_coefs(lmm_ctx *c, lmm_field *const alpha[3], const lmm_field *beta, lmm_field *const bed[3],
                       const lmm_field *gamma) {
  int dim = c->g.dim;
  for (int d = 0; d < dim; d++)
    if (!field_ok(c, alpha[d], LMM_FACE_X + d)) return set_err(c, LMM_ERR_ARG, "alpha_fc[%d] missing or wrong kind", d);
  if (!field_ok(c, beta, LMM_CELL)) return set_err(c, LMM_ERR_ARG, "beta missing or wrong kind");
  if (!field_ok(c, gamma, LMM_CELL)) return set_err(c, LMM_ERR_ARG, "gamma missing or wrong kind");
  if (dim == 2) {
    if (!field_ok(c, bed[0], LMM_NODE)) return set_err(c, LMM_ERR_ARG, "beta_ed[0] must be a node field in 2D");
  } else {
    for (int e = 0; e < 3; e++)
      if (!field_ok(c, bed[e], LMM_EDGE_XY + e)) return set_err(c, LMM_ERR_ARG, "beta_ed[%d] missing or wrong kind", e);
  }
  return LMM_OK;
}

static CoefIn coef_in(lmm_ctx *c, lmm_field *const alpha[3], lmm_field *beta, lmm_field *const bed[3],
                      lmm_field *gamma) {
  CoefIn ci;
  for (int d = 0; d < 3; d++) {
    ci.alpha[d] = (d < c->g.dim) ? alpha[d]->d : nullptr;
    ci.bed[d] = (d < ne_of(c->g.dim)) ? bed[d]->d : nullptr;
  }
  ci.beta = beta->d;
  ci.gamma = gamma->d;
  return ci;
}

static int ensure_lvl32(lmm_ctx *c) {
  if (c->lvl32) return LMM_OK;
  int rc = alloc_levels<float>(c, c->s32, c->c32);
  if (rc) return set_err(c, rc, "cannot allocate FP32 multigrid levels");
  long L = (long)(c->g.dim + 1) * c->g.N;
  c->b32 = dalloc<float>(c->stream, L);
  c->mac_rhs32 = dalloc<float>(c->stream, c->g.N);
  if (!c->b32 || !c->mac_rhs32) return set_err(c, LMM_ERR_NOMEM, "cannot allocate FP32 work vectors");
  c->lvl32 = true;
  return LMM_OK;
}

// ---------------------------------------------------------------------------
// GMRES, reference algorithm (gmres.f90:30-607), FP64
// ---------------------------------------------------------------------------

static int gmres_ref(lmm_ctx *c, double theta_alpha, lmm_gmres_info *info) {
  const lmm_params &p = c->prm;
  const Grid &g = c->g;
  int dm = g.dim, m = p.gmres_max_inner;
  long L = (long)(dm + 1) * g.N;
  hipStream_t s = c->stream;
  SV b = sv_of(c, c->b), x = sv_of(c, c->x), r = sv_of(c, c->r), w = sv_of(c, c->w), tmp = sv_of(c, c->tmp);
  auto Vk = [&](int k) { return c->V + (long)k * L; };
  Hess hs(m);
  double *dH = c->scal;  // (m+1)*m slots, column-major like hs.H
  // |M^-1 b|
  precon64(c, b, tmp, theta_alpha);
  stokes_dot_dev(c, c->tmp, c->tmp, c->scal + 3000, 1);
  stokes_dot_dev(c, c->b, c->b, c->scal + 3004, 1);
  double hb[8];
  fetch(c, c->scal + 3000, 8, hb);
  double norm_pre_b = hb[0], norm_b = hb[4];
  info->norm_pre_rhs = norm_pre_b;
  info->norm_b = norm_b;
  if (norm_b <= 0.0) {
    launch_fill<double>(s, c->x, L, 0.0);
    info->status = LMM_OK;
    return LMM_OK;
  }
  int total_iter = 0, iter = 0, i_copy = 0;
  double norm_resid = 0, norm_init_resid = 0, norm_resid_Stokes = 0;
  bool converged = false;
  for (;;) {
    residual_dev(c, x, c->tmp);
    stokes_dot_dev(c, c->tmp, c->tmp, c->scal + 3008, 1);
    precon64(c, tmp, r, theta_alpha);
    stokes_dot_dev(c, c->r, c->r, c->scal + 3012, 1);
    fetch(c, c->scal + 3008, 8, hb);
    norm_resid_Stokes = hb[0];
    norm_resid = hb[4];
    if (iter == 0) norm_init_resid = norm_resid;
    if (p.gmres_verbose >= 1)
      printf("lmm gmres outer %d total %d |Pr| %.10e |r| %.10e\n", iter, total_iter, norm_resid, norm_resid_Stokes);
    if (total_iter >= p.gmres_max_iter) break;
    if (total_iter >= p.gmres_min_iter) {
      double mn = std::min(norm_pre_b, norm_init_resid);
      if (norm_resid <= p.gmres_rel_tol * mn) {
        converged = true;
        break;
      }
      if (norm_resid <= p.gmres_abs_tol) {
        converged = true;
        break;
      }
    }
    if (iter >= p.gmres_max_outer) break;
    iter++;
    // V(1) = r / |r|
    launch_div_by(s, Vk(0), c->r, L, c->scal + 3012);
    std::fill(hs.s.begin(), hs.s.end(), 0.0);
    hs.s[0] = norm_resid;
    int i;
    bool brk = false;
    for (i = 1; i <= m; i++) {
      total_iter++;
      i_copy = i;
      SV vi = sv_of(c, Vk(i - 1));
      apply_matrix_dev(c, vi, tmp);
      precon64(c, tmp, w, theta_alpha);
      // H(1..i, i), then H(i+1,i) = |w|
      double *col = dH + (long)(m + 1) * (i - 1);
      double *sl = c->scal + 3100;
      mgs(c, Vk, i, col, sl);
      hipMemcpyAsync(col + i, sl, sizeof(double), hipMemcpyDeviceToDevice, s);
      launch_div_by(s, Vk(i), c->w, L, sl);
      fetch(c, dH + (long)(m + 1) * (i - 1), i + 1, &hs.h(1, i));
      if (hs.h(i + 1, i) == 0.0) {
        brk = true;
        break;
      }
      hs.least_squares(i);
      double est = std::fabs(hs.s[i]);
      if (p.gmres_verbose >= 2) printf("lmm gmres inner %d est %.10e\n", total_iter, est);
      if (total_iter >= p.gmres_max_iter) break;
      if (total_iter >= p.gmres_min_iter) {
        double mn = std::min(norm_pre_b, norm_init_resid);
        if (est <= p.gmres_rel_tol * mn || est <= p.gmres_abs_tol) break;
      }
    }
    if (brk) {
      info->total_iter = total_iter;
      return set_err(c, LMM_ERR_BREAKDOWN, "gmres: error in orthogonalization");
    }
    hs.solve(i_copy - 1);
    memcpy(c->hscal, hs.y.data(), sizeof(double) * i_copy);
    hipMemcpyAsync(c->scal + 3200, c->hscal, sizeof(double) * i_copy, hipMemcpyHostToDevice, s);
    std::vector<double *> vl(i_copy);
    for (int k = 0; k < i_copy; k++) vl[k] = Vk(k);
    launch_update_sol(s, c->x, vl.data(), c->scal + 3200, i_copy, L);
    hipStreamSynchronize(s);  // hscal reused
  }
  info->total_iter = total_iter;
  info->outer_iter = iter;
  info->norm_resid = norm_resid;
  info->norm_resid_true = norm_resid_Stokes;
  info->status = converged ? LMM_OK : LMM_NOT_CONVERGED;
  return LMM_OK;
}

// ---------------------------------------------------------------------------
// Flexible GMRES with FP32 V-cycles (right preconditioned, MGS, FP64 outer)
// ---------------------------------------------------------------------------

// The FP32 preconditioner application z32 = M(b32) is a fixed sequence of
// device work for a given output slot, theta_alpha and configuration (the FP32
// V-cycles have no host-side stopping tests).  With precon_graph it is
// captured into a HIP graph at its second use per slot (the first runs
// eagerly, so any lazy allocation happens outside capture) and replayed
// afterwards: the same kernels with the same arguments, hence the same bits.
// The V-cycles clear memory with fill kernels, never hipMemsetAsync.
static bool graphs_usable(const lmm_ctx *c) {
  return c->precon_graph && !c->prof && !(c->learned_on && c->cnn) && !debug_sync_on();
}

static void precon32(lmm_ctx *c, const double *v64, float *z32, double theta_alpha, bool have32) {
  hipStream_t s = c->stream;
  long L = (long)(c->g.dim + 1) * c->g.N;
  if (!have32) launch_convert<float, double>(s, c->b32, v64, L);
  SVT<float> b = svt_of<float>(c, c->b32), z = svt_of<float>(c, z32);
  if (graphs_usable(c)) {
    lmm_ctx::PGraph &e = c->pgraphs[z32];
    if (e.exec && e.theta != theta_alpha) {
      hipGraphExecDestroy(e.exec);
      e.exec = nullptr;
    }
    if (!e.exec && e.warm) {
      hipGraph_t gr = nullptr;
      if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        const int vc0 = c->vcycles;
        apply_precon_t<float>(c, c->s32, c->c32, b.u, b.p, z.u, z.p, c->mac_rhs32, c->c32[0].phi, theta_alpha);
        e.vcycles = c->vcycles - vc0;
        c->vcycles = vc0;  // counted when the graph runs
        if (hipStreamEndCapture(s, &gr) == hipSuccess && gr) {
          if (hipGraphInstantiate(&e.exec, gr, nullptr, nullptr, 0) != hipSuccess) e.exec = nullptr;
          hipGraphDestroy(gr);
        }
        if (e.exec) e.theta = theta_alpha;
      }
      if (!e.exec) {  // capture unavailable: stay eager for this slot
        hipGetLastError();
        c->precon_graph = false;
        apply_precon_t<float>(c, c->s32, c->c32, b.u, b.p, z.u, z.p, c->mac_rhs32, c->c32[0].phi, theta_alpha);
        return;
      }
    }
    if (e.exec) {
      if (hipGraphLaunch(e.exec, s) != hipSuccess) report_error("hipGraphLaunch of the FP32 preconditioner failed");
      c->vcycles += e.vcycles;
      return;
    }
    e.warm = true;
  }
  apply_precon_t<float>(c, c->s32, c->c32, b.u, b.p, z.u, z.p, c->mac_rhs32, c->c32[0].phi, theta_alpha);
  if (c->learned_on && c->cnn) learned_stage(c, b.u, b.p, z.u, z.p);
}

// One classical Gram-Schmidt step of the flexible GMRES.  The stored basis
// vectors V_k are not normalised: the unit vector is v_k = sig[k] V_k
// (sig[0] = 1, V_0 = r/|r|).  On entry w = A Z_{i-1}, where Z_{i-1} = M(V_{i-1})
// stands for z_{i-1} = sig[i-1] Z_{i-1}, so A z_{i-1} = sig[i-1] w.  One pass
// forms d_k = <w, V_k> for every k < i, giving H(k+1, i) = sig[i-1] sig[k] d_k;
// a second writes V_i = w - sum_k sig[k]^2 d_k V_k (= the orthogonalised
// vector / sig[i-1]) with its FP32 copy (the next preconditioner input) and
// |V_i|, so H(i+1, i) = sig[i-1] |V_i| and sig[i] = 1/|V_i|.  Two passes over
// the basis per step instead of 2i, and no separate normalisation pass.
template <typename VF>
static void cgs_step(lmm_ctx *c, const VF &Vk, int i, std::vector<double> &sig, Hess &hs) {
  hipStream_t s = c->stream;
  const Grid &g = c->g;
  const long L = (long)(g.dim + 1) * g.N;
  double *dslot = c->scal + 3300, *cslot = c->scal + 3340, *nslot = c->scal + 3380;
  std::vector<double> sig2(i);
  for (int k = 0; k < i; k++) sig2[k] = sig[k] * sig[k];
  for (int k0 = 0; k0 < i; k0 += CGS_MAXV) {
    const int nv = std::min(CGS_MAXV, i - k0);
    const double *vl[CGS_MAXV];
    for (int k = 0; k < nv; k++) vl[k] = Vk(k0 + k);
    ProfScope ps(c, "krylov", (double)L * 8.0 * (nv + 1));
    launch_cgs_dots(s, g, c->w, vl, nv, c->partial_cgs);
    launch_cgs_finalize(s, g, c->partial_cgs, nv, sig2.data() + k0, c->prm.p_norm_weight, dslot + k0, cslot + k0);
  }
  for (int k0 = 0; k0 < i; k0 += CGS_MAXV) {
    const int nv = std::min(CGS_MAXV, i - k0);
    const bool last = (k0 + nv == i);
    const double *vl[CGS_MAXV];
    for (int k = 0; k < nv; k++) vl[k] = Vk(k0 + k);
    ProfScope ps(c, "krylov", (double)L * (8.0 * (nv + 2) + (last ? 4.0 : 0.0)));
    launch_cgs_update(s, g, (k0 == 0) ? c->w : Vk(i), vl, nv, cslot + k0, Vk(i), last ? c->b32 : nullptr, c->partial);
  }
  launch_finalize(s, g, c->partial, nslot, 1, c->prm.p_norm_weight);
  double hb[3 * 32 + 3];
  fetch(c, dslot, 83, hb);
  for (int k = 0; k < i; k++) hs.h(k + 1, i) = sig[i - 1] * sig[k] * hb[k];
  const double nv = hb[80];
  hs.h(i + 1, i) = sig[i - 1] * nv;
  sig[i] = (nv > 0.0) ? 1.0 / nv : 0.0;
}

static int gmres_flex(lmm_ctx *c, double theta_alpha, lmm_gmres_info *info) {
  const lmm_params &p = c->prm;
  const Grid &g = c->g;
  int dm = g.dim, m = p.gmres_max_inner;
  long L = (long)(dm + 1) * g.N;
  hipStream_t s = c->stream;
  SV b = sv_of(c, c->b), x = sv_of(c, c->x), tmp = sv_of(c, c->tmp), w = sv_of(c, c->w);
  auto Vk = [&](int k) { return c->V + (long)k * L; };
  auto Zk = [&](int k) { return c->Z32 + (long)k * L; };
  Hess hs(m);
  std::vector<double> sig(m + 1, 1.0);  // scale of each stored basis vector
  double *dH = c->scal;
  double hb[8];
  stokes_dot_dev(c, c->b, c->b, c->scal + 3004, 1);
  fetch(c, c->scal + 3004, 4, hb);
  double norm_b = hb[0];
  info->norm_b = norm_b;
  info->norm_pre_rhs = 0.0;
  if (norm_b <= 0.0) {
    launch_fill<double>(s, c->x, L, 0.0);
    return LMM_OK;
  }
  int total_iter = 0, iter = 0;
  double rnorm = 0;
  bool converged = false;
  // the learned stage runs until it fails: a non-finite Hessenberg column
  // (the step is redone with the plain V-cycle) or an inner step that cuts
  // the residual estimate by less than 10 % (later steps use the plain
  // V-cycle); either way the flexible GMRES keeps its basis
  c->learned_on = (p.precision == LMM_LEARNED && c->cnn != nullptr);
  c->learned_applied = 0;
  info->learned_fallback = 0;
  for (;;) {
    residual_dev(c, x, c->tmp);
    stokes_dot_dev(c, c->tmp, c->tmp, c->scal + 3008, 1);
    fetch(c, c->scal + 3008, 4, hb);
    rnorm = hb[0];
    if (p.gmres_verbose >= 1) printf("lmm fgmres outer %d total %d |r|/|b| %.10e\n", iter, total_iter, rnorm / norm_b);
    if (rnorm <= p.gmres_rel_tol * norm_b || rnorm <= p.gmres_abs_tol) {
      converged = true;
      break;
    }
    if (total_iter >= p.gmres_max_iter || iter >= p.gmres_max_outer) break;
    iter++;
    // b32_fresh: b32 holds V(i-1) in FP32 (written by the pass that made V(i-1))
    bool b32_fresh = c->fused_krylov;
    if (c->fused_krylov) launch_div_by_f32(s, Vk(0), c->b32, c->tmp, L, c->scal + 3008);
    else launch_div_by(s, Vk(0), c->tmp, L, c->scal + 3008);
    std::fill(hs.s.begin(), hs.s.end(), 0.0);
    std::fill(sig.begin(), sig.end(), 1.0);
    hs.s[0] = rnorm;
    int i, used = 0;
    double prev_est = rnorm;
    for (i = 1; i <= m; i++) {
      total_iter++;
      used = i;
      for (int attempt = 0; attempt < 2; attempt++) {
        // a redone step converts V(i-1) again: its first try overwrote b32
        precon32(c, Vk(i - 1), Zk(i - 1), theta_alpha, attempt == 0 && b32_fresh);
        apply_matrix_z32(c, Zk(i - 1), w);
        if (c->flex_cgs) {
          cgs_step(c, Vk, i, sig, hs);  // writes V(i) and its FP32 copy
          b32_fresh = true;
        } else {
          double *col = dH + (long)(m + 1) * (i - 1);
          double *sl = c->scal + 3100;
          mgs(c, Vk, i, col, sl);
          hipMemcpyAsync(col + i, sl, sizeof(double), hipMemcpyDeviceToDevice, s);
          if (c->fused_krylov) launch_div_by_f32(s, Vk(i), c->b32, c->w, L, sl);
          else launch_div_by(s, Vk(i), c->w, L, sl);
          b32_fresh = c->fused_krylov;
          fetch(c, dH + (long)(m + 1) * (i - 1), i + 1, &hs.h(1, i));
        }
        bool finite = true;
        for (int k = 1; k <= i + 1; k++) finite = finite && std::isfinite(hs.h(k, i));
        if (finite || !c->learned_on) break;
        c->learned_on = false;  // redo the step with the plain V-cycle
        info->learned_fallback = 1;
      }
      bool happy = (hs.h(i + 1, i) == 0.0);  // exact solution in the space
      hs.least_squares(i);
      double est = std::fabs(hs.s[i]);
      if (c->learned_on && !(est < 0.9 * prev_est)) {
        c->learned_on = false;
        info->learned_fallback = 1;
      }
      prev_est = est;
      if (happy || est <= p.gmres_rel_tol * norm_b || total_iter >= p.gmres_max_iter) break;
    }
    hs.solve(used - 1);
    // x += sum_k y_k z_k with z_k = sig_k Z_k  (Z in FP32, accumulation in FP64)
    for (int k = 0; k < used; k++) hs.y[k] *= sig[k];
    memcpy(c->hscal, hs.y.data(), sizeof(double) * used);
    hipMemcpyAsync(c->scal + 3200, c->hscal, sizeof(double) * used, hipMemcpyHostToDevice, s);
    if (c->z32_direct) {
      std::vector<float *> zl(used);
      for (int k = 0; k < used; k++) zl[k] = Zk(k);
      ProfScope ps(c, "krylov", (double)L * (16.0 + 4.0 * used));
      launch_update_sol_z32(s, c->x, zl.data(), c->scal + 3200, used, L);
    } else {
      for (int k = 0; k < used; k++) {
        launch_convert<double, float>(s, c->r, Zk(k), L);
        double *one = c->r;
        launch_update_sol(s, c->x, &one, c->scal + 3200 + k, 1, L);
      }
    }
    hipStreamSynchronize(s);
  }
  info->total_iter = total_iter;
  info->outer_iter = iter;
  info->norm_resid = rnorm;
  info->norm_resid_true = rnorm;
  info->status = converged ? LMM_OK : LMM_NOT_CONVERGED;
  info->learned_applied = c->learned_applied;
  c->learned_on = false;
  (void)tmp;
  return LMM_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

// Kernel-choice switches of a new context, from the environment (read, not
// written: each context keeps its own copy, so a context created later, or
// on another thread, changes nothing in the others).  The debug switch is
// process-wide and read once.
static lmm_kernel_options env_options() {
  static std::once_flag once;
  std::call_once(once, [] {
    const char *dbg = getenv("LMM_DEBUG_SYNC");
    set_debug_sync(dbg && dbg[0] == '1');
  });
  lmm_kernel_options o;
  lmm_kernel_options_default(&o);
  const char *sm = getenv("LMM_SMOOTHER");
  if (sm && strcmp(sm, "colour") == 0) o.fused_sweeps = 0;
  if (sm && strcmp(sm, "fused") == 0) o.fused_min_tiles = 1;
  const char *mt = getenv("LMM_FUSED_MIN_TILES");
  if (mt && atoi(mt) > 0) o.fused_min_tiles = atoi(mt);
  const char *tg = getenv("LMM_FUSED_TARGET");
  if (tg && atoi(tg) > 0) o.fused_target = atoi(tg);
  const char *fk = getenv("LMM_FUSED_KERNEL");
  if (fk && strcmp(fk, "march") == 0) o.fused_kernel = 0;
  if (fk && strcmp(fk, "block") == 0) o.fused_kernel = 1;
  const char *x3 = getenv("LMM_FUSED_X3");
  if (x3 && x3[0] == '0') o.fused_x3 = 0;
  const char *rr = getenv("LMM_RESID_RESTRICT");
  if (rr && rr[0] == '0') o.resid_restrict = 0;
  const char *kf = getenv("LMM_KRYLOV_FUSED");
  if (kf && kf[0] == '0') o.krylov_fused = 0;
  const char *z3 = getenv("LMM_Z32");
  if (z3 && z3[0] == '0') o.z32_direct = 0;
  const char *ab = getenv("LMM_AM_BRICK");
  if (ab) o.am_brick = atoi(ab);
  const char *co = getenv("LMM_COARSE");
  if (co && co[0] == '0') o.coarse_one_launch = 0;
  const char *gs = getenv("LMM_FLEX_GS");
  if (gs && strcmp(gs, "mgs") == 0) o.flex_cgs = 0;
  const char *pf = getenv("LMM_SWEEP_PREFETCH");
  if (pf && pf[0] == '1') o.sweep_prefetch = 1;
  const char *pg = getenv("LMM_PRECON_GRAPH");
  if (pg && pg[0] == '0') o.precon_graph = 0;
  if (pg && pg[0] == '1') o.precon_graph = 1;
  return o;
}

static void apply_options(lmm_ctx *c, const lmm_kernel_options &o) {
  c->fused = o.fused_sweeps != 0;
  c->fused_min_tiles = o.fused_min_tiles;
  c->fused3 = o.fused_x3 != 0;
  c->fused_resid_restrict = o.resid_restrict != 0;
  c->fused_krylov = o.krylov_fused != 0;
  c->z32_direct = o.z32_direct != 0;
  c->coarse_one_launch = o.coarse_one_launch != 0;
  c->flex_cgs = o.flex_cgs != 0;
  c->tune.rb_variant = o.fused_kernel;
  c->tune.rb_target = o.fused_target;
  c->tune.am_dz = o.am_brick;
  c->tune.rb_prefetch = o.sweep_prefetch;
  const char *pairs = getenv("LMM_SWEEP_PAIRS");
  c->tune.rb_pairs = (pairs && pairs[0] == '1') ? 1 : 2;
  const char *pt = getenv("LMM_PROLONG_TILED");
  c->tune.prolong_tiled = (pt && pt[0] == '0') ? 0 : 1;
  const char *cs = getenv("LMM_COARSE_SPLIT");
  c->tune.coarse_split = (cs && cs[0] == '0') ? 0 : 1;
  const char *r3 = getenv("LMM_RR_THREE");
  c->tune.rr_three = (r3 && r3[0] == '0') ? 0 : 1;
  c->precon_graph = o.precon_graph != 0;
  c->kopt = o;
  drop_graphs(c);
}

extern "C" {

void lmm_params_default(lmm_params *p) {
  memset(p, 0, sizeof(*p));
  p->visc_type = 1;
  p->precon_type = 1;
  p->p_norm_weight = 1.0;
  p->scale_factor = 1.0;
  p->mg_max_vcycles = 1;
  p->mg_minwidth = 2;
  p->mg_nsmooths_down = 2;
  p->mg_nsmooths_up = 2;
  p->mg_nsmooths_bottom = 8;
  p->mg_rel_tol = 1e-9;
  p->stag_mg_max_vcycles = 1;
  p->stag_mg_minwidth = 2;
  p->stag_mg_nsmooths_down = 2;
  p->stag_mg_nsmooths_up = 2;
  p->stag_mg_nsmooths_bottom = 8;
  p->stag_mg_smoother = 1;
  p->stag_mg_omega = 1.0;
  p->stag_mg_rel_tol = 1e-9;
  p->gmres_rel_tol = 1e-9;
  p->gmres_abs_tol = 0.0;
  p->gmres_max_outer = 20;
  p->gmres_max_inner = 5;
  p->gmres_max_iter = 100;
  p->gmres_min_iter = 1;
  p->gmres_verbose = 0;
  p->precision = LMM_FP64;
  p->stoch_stress_form = 1;
  p->filtering_width = 0;
}

static int check_params(lmm_ctx *c, const lmm_params *p) {
  int v = p->visc_type;
  if (!(v == 1 || v == -1 || v == 2 || v == -2 || v == 3 || v == -3))
    return set_err(c, LMM_ERR_ARG, "visc_type %d not in +-1,+-2,+-3", v);
  if (p->precon_type != 1) return set_err(c, LMM_ERR_UNSUPPORTED, "precon_type %d: only 1 is implemented", p->precon_type);
  if (p->gmres_max_inner < 1 || p->gmres_max_inner > 30) return set_err(c, LMM_ERR_ARG, "gmres_max_inner out of range 1..30");
  if (p->stag_mg_smoother != 0 && p->stag_mg_smoother != 1) return set_err(c, LMM_ERR_UNSUPPORTED, "stag_mg_smoother %d", p->stag_mg_smoother);
  if (p->precision < LMM_FP64 || p->precision > LMM_LEARNED)
    return set_err(c, LMM_ERR_UNSUPPORTED, "precision %d not available", p->precision);
  if (p->stag_mg_minwidth < 1 || p->mg_minwidth < 1) return set_err(c, LMM_ERR_ARG, "minwidth must be >= 1");
  const int fw = p->filtering_width;
  if (!(fw == 0 || fw == 1 || fw == 2 || fw == 4))
    return set_err(c, LMM_ERR_ARG, "filtering_width %d: must be 0, 1, 2 or 4 (multifab_filter.f90)", fw);
  return LMM_OK;
}

int lmm_ctx_create(int dim, const int n_cells[3], double dx, const lmm_params *p, lmm_ctx **out) {
  if (!out || !n_cells || (dim != 2 && dim != 3) || !(dx > 0)) return LMM_ERR_ARG;
  *out = nullptr;
  lmm_params dp;
  lmm_params_default(&dp);
  if (!p) p = &dp;
  lmm_ctx *c = new lmm_ctx();
  apply_options(c, env_options());
  // LMM_COARSE_MAX_CELLS: a smaller coarse-launch threshold, for measurement
  if (const char *cm = getenv("LMM_COARSE_MAX_CELLS")) {
    const long v = atol(cm);
    if (v >= 8 && v <= COARSE_MAX_CELLS) c->coarse_max_cells = v;
  }
  c->g = make_grid(dim, n_cells, dx);
  int rc = check_params(c, p);
  if (rc) {
    delete c;
    return rc;
  }
  c->prm = *p;
  for (int d = 0; d < dim; d++)
    if (c->g.n[d] < 2) {
      delete c;
      return LMM_ERR_ARG;
    }
  // kernels index points (and dim*N face points) with 32-bit integers
  if ((double)c->g.N * dim >= 2147483647.0) {
    delete c;
    return LMM_ERR_UNSUPPORTED;
  }
  c->nl_s = compute_nlevs(c->g, p->stag_mg_minwidth);
  c->nl_c = compute_nlevs(c->g, p->mg_minwidth);
  // every staggered level needs even extents (coloured sweeps on periodic faces)
  {
    Grid g = c->g;
    for (int n = 0; n < std::max(c->nl_s, c->nl_c); n++) {
      for (int d = 0; d < dim; d++)
        if (g.n[d] % 2 != 0) {
          delete c;
          return LMM_ERR_UNSUPPORTED;
        }
      g = coarsen(g);
    }
  }
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return LMM_ERR_DEVICE;
  }
  rc = alloc_levels<double>(c, c->s64, c->c64);
  long N = c->g.N, L = (long)(dim + 1) * N;
  c->partial = dalloc<double>(c->stream, RED_BLOCKS * 4);
  c->partial_cgs = dalloc<double>(c->stream, (long)RED_BLOCKS * CGS_STRIDE);
  c->partial2 = dalloc<double>(c->stream, (N + 255) / 256 + 1);
  c->scal = dalloc<double>(c->stream, 8192);
  if (hipHostMalloc((void **)&c->hscal, 8192 * sizeof(double), hipHostMallocDefault) != hipSuccess) c->hscal = nullptr;
  for (int d = 0; d < dim; d++) {
    c->ainv[d] = dalloc<double>(c->stream, N);
  }
  c->scaled_beta = dalloc<double>(c->stream, N);
  c->scaled_gamma = dalloc<double>(c->stream, N);
  for (int e = 0; e < ne_of(dim); e++) c->scaled_bed[e] = dalloc<double>(c->stream, N);
  c->r = dalloc<double>(c->stream, L);
  c->w = dalloc<double>(c->stream, L);
  c->tmp = dalloc<double>(c->stream, L);
  c->b = dalloc<double>(c->stream, L);
  c->x = dalloc<double>(c->stream, L);
  c->mac_rhs = dalloc<double>(c->stream, N);
  c->phi_cc = dalloc<double>(c->stream, N);
  bool ok = !rc && c->partial && c->partial_cgs && c->partial2 && c->scal && c->hscal && c->r && c->w && c->tmp && c->b && c->x &&
            c->mac_rhs && c->phi_cc && c->scaled_beta && c->scaled_gamma;
  for (int d = 0; d < dim; d++) ok = ok && c->ainv[d];
  for (int e = 0; e < ne_of(dim); e++) ok = ok && c->scaled_bed[e];
  if (ok && hipStreamSynchronize(c->stream) != hipSuccess) ok = false;
  if (!ok) {
    lmm_ctx_destroy(c);
    return LMM_ERR_NOMEM;
  }
  *out = c;
  return LMM_OK;
}

int lmm_ctx_destroy(lmm_ctx *c) {
  if (!c) return LMM_ERR_ARG;
  if (c->stream) hipStreamSynchronize(c->stream);
  drop_graphs(c);
  prof_flush(c);
  for (hipEvent_t e : c->evpool) hipEventDestroy(e);
  free_levels<double>(c->s64, c->c64, false);
  free_levels<float>(c->s32, c->c32, true);
  dfree(c->gen32);
  for (void *p : {(void *)c->partial, (void *)c->partial_cgs, (void *)c->partial2, (void *)c->scal, (void *)c->scaled_beta,
                  (void *)c->scaled_gamma, (void *)c->V, (void *)c->r, (void *)c->w, (void *)c->tmp, (void *)c->b,
                  (void *)c->x, (void *)c->Z32, (void *)c->b32, (void *)c->mac_rhs, (void *)c->phi_cc,
                  (void *)c->mac_rhs32, (void *)c->Wcc, (void *)c->Wed, (void *)c->dweights, (void *)c->stoch_work,
                  (void *)c->Wmass, (void *)c->dweights_mass, (void *)c->sp_work, (void *)c->iw})
    dfree(p);
  for (int d = 0; d < 3; d++) {
    dfree(c->ainv[d]);
    dfree(c->scaled_bed[d]);
  }
  if (c->hscal) hipHostFree(c->hscal);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return LMM_OK;
}

int lmm_ctx_set_params(lmm_ctx *c, const lmm_params *p) {
  if (!c || !p) return LMM_ERR_ARG;
  int rc = check_params(c, p);
  if (rc) return rc;
  if (p->stag_mg_minwidth != c->prm.stag_mg_minwidth || p->mg_minwidth != c->prm.mg_minwidth)
    return set_err(c, LMM_ERR_UNSUPPORTED, "minwidth is fixed at context creation");
  hipStreamSynchronize(c->stream);
  drop_graphs(c);
  if (p->gmres_max_inner != c->prm.gmres_max_inner) c->kry_m = std::min(c->kry_m, p->gmres_max_inner - 1);
  if (p->precision != c->prm.precision) {
    dfree(c->V);
    dfree(c->Z32);
    c->V = nullptr;
    c->Z32 = nullptr;
    c->kry_m = -1;
  }
  c->prm = *p;
  return LMM_OK;
}

int lmm_ctx_get_params(const lmm_ctx *c, lmm_params *p) {
  if (!c || !p) return LMM_ERR_ARG;
  *p = c->prm;
  return LMM_OK;
}

void lmm_kernel_options_default(lmm_kernel_options *o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->fused_sweeps = 1;
  o->fused_min_tiles = 64;
  o->fused_kernel = 2;
  o->fused_target = 768;
  o->fused_x3 = 1;
  o->resid_restrict = 1;
  o->coarse_one_launch = 1;
  o->am_brick = 32;
  o->z32_direct = 1;
  o->krylov_fused = 1;
  o->flex_cgs = 1;
  o->sweep_prefetch = 0;
  o->precon_graph = 0;
}

int lmm_ctx_get_kernel_options(const lmm_ctx *c, lmm_kernel_options *o) {
  if (!c || !o) return LMM_ERR_ARG;
  *o = c->kopt;
  return LMM_OK;
}

int lmm_ctx_set_kernel_options(lmm_ctx *c, const lmm_kernel_options *o) {
  if (!c || !o) return LMM_ERR_ARG;
  if (o->fused_min_tiles < 1 || o->fused_target < 1 || o->fused_kernel < 0 || o->fused_kernel > 2 || o->am_brick < 0)
    return set_err(c, LMM_ERR_ARG, "kernel options out of range");
  hipStreamSynchronize(c->stream);
  apply_options(c, *o);
  return LMM_OK;
}

}  // extern "C"

// Device time of `reps` smoothing sweeps of the finest staggered level plus
// as many fused sweeps of the finest cell-centred level, on the context's own
// level storage (scratch arrays where level 0 borrows the caller's fields).
// The values are whatever the buffers hold: the kernels have no
// data-dependent control flow, so the time does not depend on them, and
// every buffer written is rewritten before its next use by a solve.
template <typename T>
static int time_fine_sweeps(lmm_ctx *c, std::vector<SLevel<T>> &S, std::vector<CLevel<T>> &C, int reps,
                            float *ms) {
  const long N = c->g.N;
  const int dim = c->g.dim, ne = ne_of(dim);
  std::vector<void *> tmp;
  auto need = [&](T *p) -> T * {
    if (p) return p;
    T *q = dalloc<T>(c->stream, N);
    tmp.push_back(q);
    return q;
  };
  SLevel<T> L = S[0];
  T *phi[3] = {nullptr, nullptr, nullptr}, *rhs[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < dim; d++) {
    phi[d] = need(L.phi[d]);
    rhs[d] = need(L.rhs[d]);
  }
  L.beta = need(L.beta);
  L.gamma = need(L.gamma);
  for (int e = 0; e < ne; e++) L.bed[e] = need(L.bed[e]);
  CLevel<T> &Z = C[0];
  T *cin = need(Z.phi), *crhs = need(Z.rhs);
  for (void *q : tmp)
    if (!q) {
      for (void *r : tmp) dfree(r);
      return LMM_ERR_NOMEM;
    }
  const bool cc_fused = c->fused && fused_rb_tz(Z.g, c->fused_min_tiles, 4, 1) > 0;
  auto run = [&]() {
    smooth<T>(c, L, phi, rhs, reps);
    if (cc_fused)
      for (int m = 0; m < reps; m++) launch_cc_rb_fused<T>(c->stream, Z.g, cin, Z.res, crhs, Z.a, c->fused_min_tiles);
  };
  run();  // warm-up
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  hipEventRecord(a, c->stream);
  run();
  hipEventRecord(b, c->stream);
  hipEventSynchronize(b);
  hipEventElapsedTime(ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  for (void *q : tmp) dfree(q);
  if (first_error()[0]) return set_err(c, LMM_ERR_DEVICE, "kernel %s", first_error());
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(c, LMM_ERR_DEVICE, "autotune: %s", hipGetErrorString(e));
  return LMM_OK;
}

extern "C" {

int lmm_ctx_autotune(lmm_ctx *c, int *fused_target) {
  if (!c) return LMM_ERR_ARG;
  TuningScope ts(&c->tune);
  hipStreamSynchronize(c->stream);
  const bool mixed = c->prm.precision != LMM_FP64;
  if (mixed)
    if (int rc = ensure_lvl32(c)) return rc;
  static const int cand[] = {256, 384, 512, 768, 1024, 1536, 2048};
  const int keep = c->tune.rb_target;
  int best = keep;
  float best_ms = 0.f;
  for (int t : cand) {
    c->tune.rb_target = t;
    float ms = 0.f;
    int rc = mixed ? time_fine_sweeps<float>(c, c->s32, c->c32, 4, &ms)
                   : time_fine_sweeps<double>(c, c->s64, c->c64, 4, &ms);
    if (rc) {
      c->tune.rb_target = keep;
      return rc;
    }
    if (best_ms == 0.f || ms < best_ms) {
      best_ms = ms;
      best = t;
    }
  }
  lmm_kernel_options o = c->kopt;
  o.fused_target = best;
  apply_options(c, o);  // drops any captured graph: their slab shapes change
  if (fused_target) *fused_target = best;
  return LMM_OK;
}

int lmm_ctx_synchronize(lmm_ctx *c) {
  if (!c) return LMM_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LMM_OK;
}

int lmm_ctx_nlevels(const lmm_ctx *c, int *ns, int *nc) {
  if (!c) return LMM_ERR_ARG;
  if (ns) *ns = c->nl_s;
  if (nc) *nc = c->nl_c;
  return LMM_OK;
}

const char *lmm_ctx_last_error(const lmm_ctx *c) { return c ? c->err.c_str() : "null context"; }

int lmm_ctx_set_profiling(lmm_ctx *c, int on) {
  if (!c || on < 0 || on > 2) return LMM_ERR_ARG;
  c->prof = on;
  return LMM_OK;
}

int lmm_ctx_kernel_stats(lmm_ctx *c, const char *family, int64_t *launches, double *total_ms, double *bytes) {
  if (!c || !family) return LMM_ERR_ARG;
  prof_flush(c);
  auto it = c->fam.find(family);
  if (it == c->fam.end()) {
    if (launches) *launches = 0;
    if (total_ms) *total_ms = 0;
    if (bytes) *bytes = 0;
    return LMM_OK;
  }
  if (launches) *launches = it->second.launches;
  if (total_ms) *total_ms = it->second.ms;
  if (bytes) *bytes = it->second.bytes;
  return LMM_OK;
}

int lmm_ctx_reset_stats(lmm_ctx *c) {
  if (!c) return LMM_ERR_ARG;
  prof_flush(c);
  c->fam.clear();
  return LMM_OK;
}

// ---------------- fields ----------------

int lmm_field_create(lmm_ctx *c, int kind, int ncomp, lmm_field **out) {
  if (!c || !out || ncomp < 1) return LMM_ERR_ARG;
  int dim = c->g.dim;
  bool ok = kind == LMM_CELL || (kind >= LMM_FACE_X && kind < LMM_FACE_X + dim) || (kind == LMM_NODE && dim == 2) ||
            (kind >= LMM_EDGE_XY && kind <= LMM_EDGE_YZ && dim == 3);
  if (!ok) return set_err(c, LMM_ERR_ARG, "field kind %d invalid for dim %d", kind, dim);
  double *d = dalloc<double>(c->stream, kind_count(c) * ncomp);
  if (!d) return set_err(c, LMM_ERR_NOMEM, "field allocation failed");
  lmm_field *f = new lmm_field{c, kind, ncomp, d};
  *out = f;
  return LMM_OK;
}

int lmm_field_destroy(lmm_field *f) {
  if (!f) return LMM_ERR_ARG;
  hipStreamSynchronize(f->ctx->stream);
  dfree(f->d);
  delete f;
  return LMM_OK;
}

int lmm_field_info(const lmm_field *f, int *kind, int *ncomp, int64_t *count) {
  if (!f) return LMM_ERR_ARG;
  if (kind) *kind = f->kind;
  if (ncomp) *ncomp = f->ncomp;
  if (count) *count = kind_count(f->ctx);
  return LMM_OK;
}

int lmm_field_upload(lmm_field *f, const double *host) {
  if (!f || !host) return LMM_ERR_ARG;
  lmm_ctx *c = f->ctx;
  HIPCHK(c, hipMemcpyAsync(f->d, host, sizeof(double) * kind_count(c) * f->ncomp, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LMM_OK;
}

int lmm_field_download(const lmm_field *f, double *host) {
  if (!f || !host) return LMM_ERR_ARG;
  lmm_ctx *c = f->ctx;
  HIPCHK(c, hipMemcpyAsync(host, f->d, sizeof(double) * kind_count(c) * f->ncomp, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LMM_OK;
}

static void nodal_flags(int kind, int nod[3]) {
  nod[0] = nod[1] = nod[2] = 0;
  if (kind == LMM_FACE_X) nod[0] = 1;
  if (kind == LMM_FACE_Y) nod[1] = 1;
  if (kind == LMM_FACE_Z) nod[2] = 1;
  if (kind == LMM_NODE) nod[0] = nod[1] = 1;
  if (kind == LMM_EDGE_XY) nod[0] = nod[1] = 1;
  if (kind == LMM_EDGE_XZ) nod[0] = nod[2] = 1;
  if (kind == LMM_EDGE_YZ) nod[1] = nod[2] = 1;
}

static int fab_check(const lmm_field *f, int comp, const int lo[3], const int hi[3], int ng, long ext[3],
                     int nod[3]) {
  lmm_ctx *c = f->ctx;
  if (comp < 0 || comp >= f->ncomp || ng < 0) return set_err(c, LMM_ERR_ARG, "fab: bad comp or ng");
  nodal_flags(f->kind, nod);
  int dim = c->g.dim;
  for (int d = 0; d < 3; d++) {
    if (d < dim) {
      if (lo[d] != 0 || hi[d] - lo[d] + 1 != c->g.n[d] + nod[d])
        return set_err(c, LMM_ERR_ARG, "fab: box [%d,%d] in dir %d does not match the context box", lo[d], hi[d], d);
      ext[d] = c->g.n[d] + nod[d] + 2 * ng;
    } else {
      ext[d] = 1;
    }
  }
  return LMM_OK;
}

int lmm_field_upload_fab(lmm_field *f, int comp, const double *host, const int lo[3], const int hi[3], int ng) {
  if (!f || !host || !lo || !hi) return LMM_ERR_ARG;
  lmm_ctx *c = f->ctx;
  long ext[3];
  int nod[3];
  int rc = fab_check(f, comp, lo, hi, ng, ext, nod);
  if (rc) return rc;
  const Grid &g = c->g;
  int ngz = g.dim == 3 ? ng : 0;
  std::vector<double> buf(g.N);
  for (int k = 0; k < g.n[2]; k++)
    for (int j = 0; j < g.n[1]; j++)
      for (int i = 0; i < g.n[0]; i++) {
        long hidx = (i + ng) + ext[0] * ((j + ng) + ext[1] * (long)(k + ngz));
        buf[i + g.s[1] * j + g.s[2] * k] = host[hidx];
      }
  HIPCHK(c, hipMemcpyAsync(f->d + comp * g.N, buf.data(), sizeof(double) * g.N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LMM_OK;
}

int lmm_field_download_fab(const lmm_field *f, int comp, double *host, const int lo[3], const int hi[3], int ng) {
  if (!f || !host || !lo || !hi) return LMM_ERR_ARG;
  lmm_ctx *c = f->ctx;
  long ext[3];
  int nod[3];
  int rc = fab_check(f, comp, lo, hi, ng, ext, nod);
  if (rc) return rc;
  const Grid &g = c->g;
  std::vector<double> buf(g.N);
  HIPCHK(c, hipMemcpyAsync(buf.data(), f->d + comp * g.N, sizeof(double) * g.N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int ngz = g.dim == 3 ? ng : 0;
  auto wrap = [](int i, int n) { return ((i % n) + n) % n; };
  for (long k = -ngz; k < ext[2] - ngz; k++)
    for (long j = -ng; j < ext[1] - ng; j++)
      for (long i = -ng; i < ext[0] - ng; i++) {
        long src = wrap((int)i, g.n[0]) + g.s[1] * wrap((int)j, g.n[1]) + g.s[2] * wrap((int)k, g.n[2]);
        host[(i + ng) + ext[0] * ((j + ng) + ext[1] * (k + ngz))] = buf[src];
      }
  return LMM_OK;
}

int lmm_field_setval(lmm_field *f, double v) {
  if (!f) return LMM_ERR_ARG;
  launch_fill<double>(f->ctx->stream, f->d, kind_count(f->ctx) * f->ncomp, v);
  return LMM_OK;
}

int lmm_field_copy(lmm_field *dst, const lmm_field *src) {
  if (!dst || !src || dst->ctx != src->ctx || dst->ncomp != src->ncomp) return LMM_ERR_ARG;
  lmm_ctx *c = dst->ctx;
  HIPCHK(c, hipMemcpyAsync(dst->d, src->d, sizeof(double) * kind_count(c) * src->ncomp, hipMemcpyDeviceToDevice,
                           c->stream));
  return LMM_OK;
}

double *lmm_field_device_ptr(lmm_field *f) { return f ? f->d : nullptr; }

// ---------------- operators ----------------

static int finish(lmm_ctx *c) {
  if (c->defer_sync) {
    if (first_error()[0]) return set_err(c, LMM_ERR_DEVICE, "kernel %s", first_error());
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(c, LMM_ERR_DEVICE, "launch: %s", hipGetErrorString(e));
    return LMM_OK;
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  if (first_error()[0]) return set_err(c, LMM_ERR_DEVICE, "kernel %s", first_error());
  if (e != hipSuccess) return set_err(c, LMM_ERR_DEVICE, "stream: %s", hipGetErrorString(e));
  e = hipGetLastError();
  if (e != hipSuccess) return set_err(c, LMM_ERR_DEVICE, "launch: %s", hipGetErrorString(e));
  return LMM_OK;
}

static void copy_stokes_in(lmm_ctx *c, double *dst, lmm_field *const u[3], const lmm_field *p) {
  long N = c->g.N;
  for (int d = 0; d < c->g.dim; d++)
    hipMemcpyAsync(dst + d * N, u[d]->d, sizeof(double) * N, hipMemcpyDeviceToDevice, c->stream);
  hipMemcpyAsync(dst + c->g.dim * N, p->d, sizeof(double) * N, hipMemcpyDeviceToDevice, c->stream);
}

static void copy_stokes_out(lmm_ctx *c, const double *src, lmm_field *const u[3], lmm_field *p) {
  long N = c->g.N;
  for (int d = 0; d < c->g.dim; d++)
    hipMemcpyAsync(u[d]->d, src + d * N, sizeof(double) * N, hipMemcpyDeviceToDevice, c->stream);
  hipMemcpyAsync(p->d, src + c->g.dim * N, sizeof(double) * N, hipMemcpyDeviceToDevice, c->stream);
}

int lmm_gmres(lmm_ctx *c, lmm_field *const b_u[3], lmm_field *b_p, lmm_field *x_u[3], lmm_field *x_p,
              lmm_field *const alpha_fc[3], lmm_field *beta, lmm_field *const beta_ed[3], lmm_field *gamma,
              double theta_alpha, double *norm_pre_rhs, lmm_gmres_info *info) {
  if (!c) return LMM_ERR_ARG;
  TuningScope ts(&c->tune);
  int rc;
  if ((rc = check_stokes(c, b_u, b_p, "gmres b"))) return rc;
  if ((rc = check_stokes(c, x_u, x_p, "gmres x"))) return rc;
  if ((rc = check_coefs(c, alpha_fc, beta, beta_ed, gamma))) return rc;
  auto t0 = std::chrono::steady_clock::now();
  lmm_gmres_info loc;
  memset(&loc, 0, sizeof(loc));
  if (!info) info = &loc;
  memset(info, 0, sizeof(*info));
  const lmm_params &p = c->prm;
  if ((rc = ensure_krylov(c, p.gmres_max_inner))) return rc;
  bool mixed = p.precision != LMM_FP64;
  if (p.precision == LMM_LEARNED && !c->cnn)
    return set_err(c, LMM_ERR_ARG, "precision LMM_LEARNED needs a network (lmm_ctx_set_cnn)");
  if (mixed && (rc = ensure_lvl32(c))) return rc;
  double sf = p.scale_factor;
  double ta = theta_alpha * (sf != 1.0 ? sf : 1.0);
  if (sf != 1.0) theta_alpha = ta;
  c->vcycles = 0;
  prepare_coefs(c, coef_in(c, alpha_fc, beta, beta_ed, gamma), theta_alpha, sf, true, mixed);
  copy_stokes_in(c, c->b, b_u, b_p);
  copy_stokes_in(c, c->x, x_u, x_p);
  long N = c->g.N;
  if (sf != 1.0) {
    launch_scale<double>(c->stream, c->x + c->g.dim * N, N, sf);
    launch_scale<double>(c->stream, c->b, c->g.dim * N, sf);
  }
  rc = mixed ? gmres_flex(c, theta_alpha, info) : gmres_ref(c, theta_alpha, info);
  if (sf != 1.0) launch_scale<double>(c->stream, c->x + c->g.dim * N, N, 1.0 / sf);
  if (rc == LMM_OK) copy_stokes_out(c, c->x, x_u, x_p);
  int rc2 = finish(c);
  if (norm_pre_rhs) *norm_pre_rhs = info->norm_pre_rhs;
  info->vcycles = c->vcycles;
  info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rc ? rc : rc2;
}

int lmm_gmres_batched(int n, lmm_gmres_member *m, int nthreads) {
  if (n < 0 || (n > 0 && !m) || nthreads < 0) return LMM_ERR_ARG;
  for (int i = 0; i < n; i++) {
    if (!m[i].ctx) return LMM_ERR_ARG;
    for (int j = 0; j < i; j++)
      if (m[j].ctx == m[i].ctx) return LMM_ERR_ARG;  // a context is not re-entrant
  }
  int nt = nthreads > 0 ? nthreads : std::min(n, 16);
  nt = std::max(1, std::min(nt, n));
  std::atomic<int> next(0);
  int device = 0;
  hipGetDevice(&device);
  auto work = [&]() {
    hipSetDevice(device);  // the caller's device, where the contexts live
    for (int i; (i = next.fetch_add(1)) < n;) {
      lmm_gmres_member &a = m[i];
      a.status = lmm_gmres(a.ctx, a.b_u, a.b_p, a.x_u, a.x_p, a.alpha_fc, a.beta, a.beta_ed, a.gamma, a.theta_alpha,
                           &a.norm_pre_rhs, &a.info);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; t++) pool.emplace_back(work);
  work();
  for (auto &th : pool) th.join();
  for (int i = 0; i < n; i++)
    if (m[i].status != LMM_OK) return m[i].status;
  return LMM_OK;
}

int lmm_apply_matrix(lmm_ctx *c, lmm_field *b_u[3], lmm_field *b_p, lmm_field *const x_u[3], lmm_field *x_p,
                     lmm_field *const alpha_fc[3], lmm_field *beta, lmm_field *const beta_ed[3],
                     lmm_field *gamma, double theta_alpha) {
  if (!c) return LMM_ERR_ARG;
  TuningScope ts(&c->tune);
  int rc;
  if ((rc = check_stokes(c, b_u, b_p, "apply_matrix b"))) return rc;
  if ((rc = check_stokes(c, x_u, x_p, "apply_matrix x"))) return rc;
  if ((rc = check_coefs(c, alpha_fc, beta, beta_ed, gamma))) return rc;
  CoefIn ci = coef_in(c, alpha_fc, beta, beta_ed, gamma);
  for (int d = 0; d < c->g.dim; d++) launch_scale_copy(c->stream, c->s64[0].alpha[d], ci.alpha[d], c->g.N, theta_alpha);
  c->s64[0].beta = const_cast<double *>(ci.beta);
  c->s64[0].gamma = const_cast<double *>(ci.gamma);
  for (int e = 0; e < ne_of(c->g.dim); e++) c->s64[0].bed[e] = const_cast<double *>(ci.bed[e]);
  SV x, b;
  for (int d = 0; d < 3; d++) {
    x.u[d] = d < c->g.dim ? x_u[d]->d : nullptr;
    b.u[d] = d < c->g.dim ? b_u[d]->d : nullptr;
  }
  x.p = x_p->d;
  b.p = b_p->d;
  apply_matrix_dev(c, x, b);
  return finish(c);
}

int lmm_apply_precon(lmm_ctx *c, lmm_field *const b_u[3], lmm_field *b_p, lmm_field *x_u[3], lmm_field *x_p,
                     lmm_field *const alpha_fc[3], lmm_field *beta, lmm_field *const beta_ed[3],
                     lmm_field *gamma, double theta_alpha) {
  if (!c) return LMM_ERR_ARG;
  TuningScope ts(&c->tune);
  int rc;
  if ((rc = check_stokes(c, b_u, b_p, "apply_precon b"))) return rc;
  if ((rc = check_stokes(c, x_u, x_p, "apply_precon x"))) return rc;
  if ((rc = check_coefs(c, alpha_fc, beta, beta_ed, gamma))) return rc;
  c->vcycles = 0;
  prepare_coefs(c, coef_in(c, alpha_fc, beta, beta_ed, gamma), theta_alpha, 1.0, true, false);
  SV x, b;
  for (int d = 0; d < 3; d++) {
    x.u[d] = d < c->g.dim ? x_u[d]->d : nullptr;
    b.u[d] = d < c->g.dim ? b_u[d]->d : nullptr;
  }
  x.p = x_p->d;
  b.p = b_p->d;
  precon64(c, b, x, theta_alpha);
  return finish(c);
}

int lmm_stag_mg_solver(lmm_ctx *c, lmm_field *const alpha_fc[3], lmm_field *beta_cc, lmm_field *const beta_ed[3],
                       lmm_field *gamma_cc, double theta_alpha, lmm_field *phi_fc[3], lmm_field *const rhs_fc[3]) {
  if (!c) return LMM_ERR_ARG;
  TuningScope ts(&c->tune);
  int rc;
  if ((rc = check_stokes(c, phi_fc, nullptr, "stag_mg_solver phi"))) return rc;
  if ((rc = check_stokes(c, rhs_fc, nullptr, "stag_mg_solver rhs"))) return rc;
  if ((rc = check_coefs(c, alpha_fc, beta_cc, beta_ed, gamma_cc))) return rc;
  c->vcycles = 0;
  prepare_coefs(c, coef_in(c, alpha_fc, beta_cc, beta_ed, gamma_cc), theta_alpha, 1.0, false, false);
  double *ph[3] = {nullptr, nullptr, nullptr}, *rh[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < c->g.dim; d++) {
    ph[d] = phi_fc[d]->d;
    rh[d] = rhs_fc[d]->d;
  }
  stag_vcycles<double>(c, c->s64, ph, rh, false);
  return finish(c);
}

int lmm_macproject(lmm_ctx *c, lmm_field *phi, lmm_field *const umac[3], lmm_field *const alphainv_fc[3],
                   lmm_field *mac_rhs) {
  (void)umac;
  if (!c) return LMM_ERR_ARG;
  TuningScope ts(&c->tune);
  if (!field_ok(c, phi, LMM_CELL) || !field_ok(c, mac_rhs, LMM_CELL)) return set_err(c, LMM_ERR_ARG, "macproject: bad cell field");
  for (int d = 0; d < c->g.dim; d++)
    if (!field_ok(c, alphainv_fc[d], LMM_FACE_X + d)) return set_err(c, LMM_ERR_ARG, "macproject: bad alphainv_fc[%d]", d);
  auto &C = c->c64;
  for (int d = 0; d < c->g.dim; d++)
    hipMemcpyAsync(C[0].a[d], alphainv_fc[d]->d, sizeof(double) * c->g.N, hipMemcpyDeviceToDevice, c->stream);
  for (int n = 1; n < c->nl_c; n++) launch_stag_restrict<double>(c->stream, C[n].g, C[n - 1].a, C[n].a, 1);
  c->vcycles = 0;
  cc_vcycles<double>(c, C, phi->d, mac_rhs->d);
  return finish(c);
}

int lmm_stag_applyop(lmm_ctx *c, lmm_field *const phi_fc[3], lmm_field *Lphi_fc[3], lmm_field *const alpha_fc[3],
                     lmm_field *beta_cc, lmm_field *const beta_ed[3], lmm_field *gamma_cc, double theta_alpha,
                     int color) {
  if (!c) return LMM_ERR_ARG;
  TuningScope ts(&c->tune);
  int rc;
  if ((rc = check_stokes(c, phi_fc, nullptr, "stag_applyop phi"))) return rc;
  if ((rc = check_stokes(c, Lphi_fc, nullptr, "stag_applyop Lphi"))) return rc;
  if ((rc = check_coefs(c, alpha_fc, beta_cc, beta_ed, gamma_cc))) return rc;
  if (color < 0 || color > 2 * c->g.dim) return set_err(c, LMM_ERR_ARG, "color %d out of range", color);
  CoefIn ci = coef_in(c, alpha_fc, beta_cc, beta_ed, gamma_cc);
  auto &L = c->s64[0];
  for (int d = 0; d < c->g.dim; d++) launch_scale_copy(c->stream, L.alpha[d], ci.alpha[d], c->g.N, theta_alpha);
  L.beta = const_cast<double *>(ci.beta);
  L.gamma = const_cast<double *>(ci.gamma);
  for (int e = 0; e < ne_of(c->g.dim); e++) L.bed[e] = const_cast<double *>(ci.bed[e]);
  double *ph[3] = {nullptr, nullptr, nullptr}, *lp[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < c->g.dim; d++) {
    ph[d] = phi_fc[d]->d;
    lp[d] = Lphi_fc[d]->d;
  }
  launch_stag_applyop<double>(c->stream, c->g, c->prm.visc_type, color, ph, lp, L.K());
  return finish(c);
}

// diffusive_m_fluxdiv.f90:16-73: m_update = m_update - (0*I - L_eta)umac,
// stag_applyop_level with alpha_fc = 0 (kappa is the gamma of the operator)
int lmm_diffusive_m_fluxdiv(lmm_ctx *c, lmm_field *m_update[3], lmm_field *const umac[3], lmm_field *eta,
                            lmm_field *const eta_ed[3], lmm_field *kappa) {
  if (!c) return LMM_ERR_ARG;
  TuningScope ts(&c->tune);
  int rc;
  if ((rc = check_stokes(c, m_update, nullptr, "diffusive_m_fluxdiv m_update"))) return rc;
  if ((rc = check_stokes(c, umac, nullptr, "diffusive_m_fluxdiv umac"))) return rc;
  if (!field_ok(c, eta, LMM_CELL)) return set_err(c, LMM_ERR_ARG, "eta missing or wrong kind");
  if (!field_ok(c, kappa, LMM_CELL)) return set_err(c, LMM_ERR_ARG, "kappa missing or wrong kind");
  const int dim = c->g.dim;
  if (dim == 2) {
    if (!field_ok(c, eta_ed[0], LMM_NODE)) return set_err(c, LMM_ERR_ARG, "eta_ed[0] must be a node field in 2D");
  } else {
    for (int e = 0; e < 3; e++)
      if (!field_ok(c, eta_ed[e], LMM_EDGE_XY + e)) return set_err(c, LMM_ERR_ARG, "eta_ed[%d] missing or wrong kind", e);
  }
  for (int d = 0; d < dim; d++)
    if (m_update[d]->d == umac[d]->d) return set_err(c, LMM_ERR_ARG, "m_update and umac must be distinct fields");
  auto &L = c->s64[0];
  for (int d = 0; d < dim; d++) launch_fill<double>(c->stream, L.alpha[d], c->g.N, 0.0);
  L.beta = eta->d;
  L.gamma = kappa->d;
  for (int e = 0; e < ne_of(dim); e++) L.bed[e] = eta_ed[e]->d;
  double *ph[3] = {nullptr, nullptr, nullptr}, *mu[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < dim; d++) {
    ph[d] = umac[d]->d;
    mu[d] = m_update[d]->d;
  }
  launch_stag_applyop<double>(c->stream, c->g, c->prm.visc_type, 0, ph, mu, L.K(), true);
  return finish(c);
}

int lmm_stag_inner_prod(lmm_ctx *c, lmm_field *const m1[3], lmm_field *const m2[3], double prod_val[3]) {
  if (!c || !prod_val) return LMM_ERR_ARG;
  int rc;
  if ((rc = check_stokes(c, m1, nullptr, "stag_inner_prod m1")) || (rc = check_stokes(c, m2, nullptr, "stag_inner_prod m2")))
    return rc;
  long N = c->g.N;
  int dim = c->g.dim;
  // pack into the Stokes-vector layout with a zero pressure block
  for (int d = 0; d < dim; d++) {
    hipMemcpyAsync(c->r + d * N, m1[d]->d, sizeof(double) * N, hipMemcpyDeviceToDevice, c->stream);
    hipMemcpyAsync(c->w + d * N, m2[d]->d, sizeof(double) * N, hipMemcpyDeviceToDevice, c->stream);
  }
  launch_fill<double>(c->stream, c->r + dim * N, N, 0.0);
  launch_fill<double>(c->stream, c->w + dim * N, N, 0.0);
  launch_stokes_dot(c->stream, c->g, c->r, c->w, c->partial);
  launch_finalize(c->stream, c->g, c->partial, c->scal + 4100, 2, 1.0);
  double h[4];
  fetch(c, c->scal + 4100, 4, h);
  for (int d = 0; d < 3; d++) prod_val[d] = d < dim ? h[d] : 0.0;
  return finish(c);
}

int lmm_cc_inner_prod(lmm_ctx *c, lmm_field *m1, lmm_field *m2, double *prod_val) {
  if (!c || !prod_val || !field_ok(c, m1, LMM_CELL) || !field_ok(c, m2, LMM_CELL)) return LMM_ERR_ARG;
  long N = c->g.N;
  int dim = c->g.dim;
  launch_fill<double>(c->stream, c->r, dim * N, 0.0);
  launch_fill<double>(c->stream, c->w, dim * N, 0.0);
  hipMemcpyAsync(c->r + dim * N, m1->d, sizeof(double) * N, hipMemcpyDeviceToDevice, c->stream);
  hipMemcpyAsync(c->w + dim * N, m2->d, sizeof(double) * N, hipMemcpyDeviceToDevice, c->stream);
  launch_stokes_dot(c->stream, c->g, c->r, c->w, c->partial);
  launch_finalize(c->stream, c->g, c->partial, c->scal + 4100, 2, 1.0);
  double h[4];
  fetch(c, c->scal + 4100, 4, h);
  *prod_val = h[3];
  return finish(c);
}

int lmm_compute_div(lmm_ctx *c, lmm_field *const phi_fc[3], lmm_field *div, int increment) {
  if (!c) return LMM_ERR_ARG;
  int rc;
  if ((rc = check_stokes(c, phi_fc, div, "compute_div"))) return rc;
  double *u[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < c->g.dim; d++) u[d] = phi_fc[d]->d;
  launch_div_plus<double>(c->stream, c->g, u, nullptr, div->d, increment ? 3 : 0);
  return finish(c);
}

int lmm_compute_grad(lmm_ctx *c, lmm_field *phi, lmm_field *gradp[3]) {
  if (!c) return LMM_ERR_ARG;
  int rc;
  if ((rc = check_stokes(c, gradp, phi, "compute_grad"))) return rc;
  double *gp[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < c->g.dim; d++) gp[d] = gradp[d]->d;
  launch_grad(c->stream, c->g, phi->d, gp);
  return finish(c);
}

// Multi-component forms: component q of the input to component q of the
// output, q over [start, start + num_comp) (0-based).
static bool comps_ok(const lmm_field *f, int start, int num) {
  return start >= 0 && num >= 1 && start + num <= f->ncomp;
}

int lmm_compute_div_comps(lmm_ctx *c, lmm_field *const phi_fc[3], lmm_field *div, int start_incomp,
                          int start_outcomp, int num_comp, int increment) {
  if (!c) return LMM_ERR_ARG;
  int rc;
  if ((rc = check_stokes(c, phi_fc, div, "compute_div"))) return rc;
  for (int d = 0; d < c->g.dim; d++)
    if (!comps_ok(phi_fc[d], start_incomp, num_comp)) return set_err(c, LMM_ERR_ARG, "compute_div: input components");
  if (!comps_ok(div, start_outcomp, num_comp)) return set_err(c, LMM_ERR_ARG, "compute_div: output components");
  const long N = c->g.N;
  for (int q = 0; q < num_comp; q++) {
    double *u[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < c->g.dim; d++) u[d] = phi_fc[d]->d + (long)(start_incomp + q) * N;
    launch_div_plus<double>(c->stream, c->g, u, nullptr, div->d + (long)(start_outcomp + q) * N, increment ? 3 : 0);
  }
  return finish(c);
}

int lmm_compute_grad_comps(lmm_ctx *c, lmm_field *phi, lmm_field *gradp[3], int start_incomp, int start_outcomp,
                           int num_comp) {
  if (!c) return LMM_ERR_ARG;
  int rc;
  if ((rc = check_stokes(c, gradp, phi, "compute_grad"))) return rc;
  if (!comps_ok(phi, start_incomp, num_comp)) return set_err(c, LMM_ERR_ARG, "compute_grad: input components");
  for (int d = 0; d < c->g.dim; d++)
    if (!comps_ok(gradp[d], start_outcomp, num_comp)) return set_err(c, LMM_ERR_ARG, "compute_grad: output components");
  const long N = c->g.N;
  for (int q = 0; q < num_comp; q++) {
    double *gp[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < c->g.dim; d++) gp[d] = gradp[d]->d + (long)(start_outcomp + q) * N;
    launch_grad(c->stream, c->g, phi->d + (long)(start_incomp + q) * N, gp);
  }
  return finish(c);
}

int lmm_average_cc_to_face_comps(lmm_ctx *c, lmm_field *cc, lmm_field *face[3], int start_comp, int num_comp) {
  if (!c) return LMM_ERR_ARG;
  int rc;
  if ((rc = check_stokes(c, face, cc, "average_cc_to_face"))) return rc;
  if (!comps_ok(cc, start_comp, num_comp)) return set_err(c, LMM_ERR_ARG, "average_cc_to_face: components");
  for (int d = 0; d < c->g.dim; d++)
    if (!comps_ok(face[d], start_comp, num_comp)) return set_err(c, LMM_ERR_ARG, "average_cc_to_face: components");
  const long N = c->g.N;
  for (int q = 0; q < num_comp; q++) {
    double *f[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < c->g.dim; d++) f[d] = face[d]->d + (long)(start_comp + q) * N;
    launch_avg_cc_to_face(c->stream, c->g, cc->d + (long)(start_comp + q) * N, f);
  }
  return finish(c);
}

int lmm_average_cc_to_face(lmm_ctx *c, lmm_field *cc, lmm_field *face[3]) {
  if (!c) return LMM_ERR_ARG;
  int rc;
  if ((rc = check_stokes(c, face, cc, "average_cc_to_face"))) return rc;
  double *f[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < c->g.dim; d++) f[d] = face[d]->d;
  launch_avg_cc_to_face(c->stream, c->g, cc->d, f);
  return finish(c);
}

int lmm_average_cc_to_node(lmm_ctx *c, lmm_field *cc, lmm_field *node) {
  if (!c || c->g.dim != 2 || !field_ok(c, cc, LMM_CELL) || !field_ok(c, node, LMM_NODE)) return LMM_ERR_ARG;
  launch_avg_cc_to_node(c->stream, c->g, cc->d, node->d);
  return finish(c);
}

int lmm_average_cc_to_edge(lmm_ctx *c, lmm_field *cc, lmm_field *edge[3]) {
  if (!c || c->g.dim != 3 || !field_ok(c, cc, LMM_CELL)) return LMM_ERR_ARG;
  for (int e = 0; e < 3; e++)
    if (!field_ok(c, edge[e], LMM_EDGE_XY + e)) return LMM_ERR_ARG;
  double *ed[3] = {edge[0]->d, edge[1]->d, edge[2]->d};
  launch_avg_cc_to_edge(c->stream, c->g, cc->d, ed);
  return finish(c);
}

// ---------------- fluctuating fluxes ----------------

static void mt_seed(uint32_t *mt, uint32_t seed) {
  mt[0] = seed;
  for (int q = 1; q < 624; q++) mt[q] = 1812433253u * (mt[q - 1] ^ (mt[q - 1] >> 30)) + (uint32_t)q;
}

static uint32_t mt_next(uint32_t *mt, int &idx) {
  if (idx >= 624) {
    for (int q = 0; q < 624; q++) {
      uint32_t y = (mt[q] & 0x80000000u) | (mt[(q + 1) % 624] & 0x7fffffffu);
      uint32_t v = mt[(q + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      mt[q] = v;
    }
    idx = 0;
  }
  uint32_t y = mt[idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

// P. J. Acklam's inverse-normal rational approximation (published coefficients)
static double inv_normal(double p) {
  static const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                              1.383577518672690e+02,  -3.066479806614716e+01, 2.506628277459239e+00};
  static const double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                              6.680131188771972e+01, -1.328068155288572e+01};
  static const double cc[6] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                               -2.549732539343734e+00, 4.374664141464968e+00,  2.938163982698783e+00};
  static const double d[4] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                              3.754408661907416e+00};
  const double plow = 0.02425, phigh = 1.0 - plow;
  if (p < plow) {
    double q = std::sqrt(-2.0 * std::log(p));
    return (((((cc[0] * q + cc[1]) * q + cc[2]) * q + cc[3]) * q + cc[4]) * q + cc[5]) /
           ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
  }
  if (p > phigh) {
    double q = std::sqrt(-2.0 * std::log(1.0 - p));
    return -(((((cc[0] * q + cc[1]) * q + cc[2]) * q + cc[3]) * q + cc[4]) * q + cc[5]) /
           ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
  }
  double q = p - 0.5, r = q * q;
  return (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
         (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.0);
}

int lmm_fill_normal_host(uint32_t seed, int64_t count, double *out) {
  if (!out || count < 0) return LMM_ERR_ARG;
  std::vector<uint32_t> mt(624);
  mt_seed(mt.data(), seed);
  int idx = 624;
  for (int64_t q = 0; q < count; q++) {
    double u;
    do {
      u = (double)mt_next(mt.data(), idx) * (1.0 / 4294967296.0);
    } while (u == 0.0);
    out[q] = inv_normal(u);
  }
  return LMM_OK;
}

int lmm_init_m_stochastic(lmm_ctx *c, int n_rngs) {
  if (!c || n_rngs < 1 || n_rngs > 8) return LMM_ERR_ARG;
  dfree(c->Wcc);
  dfree(c->Wed);
  dfree(c->dweights);
  long N = c->g.N;
  int dim = c->g.dim, ne = ne_of(dim);
  c->Wcc = dalloc<double>(c->stream, (long)n_rngs * dim * N);
  c->Wed = dalloc<double>(c->stream, (long)n_rngs * ne * 2 * N);
  c->dweights = dalloc<double>(c->stream, 8);
  if (!c->Wcc || !c->Wed || !c->dweights) return set_err(c, LMM_ERR_NOMEM, "noise allocation failed");
  c->n_rngs = n_rngs;
  return LMM_OK;
}

// fill_m_stochastic (stochastic_m_fluxdiv.f90:556-609): per stage, W_cc,
// then each node/edge array.  Symmetric form (stoch_stress_form != 0): W_cc
// = sqrt(2)*N(0,1) (multifab_fill_random with variance 2, its :58-60 order
// sqrt(variance)*fp) and only component 1 of a node/edge array is drawn,
// component 2 being its copy (multifab_copy_c).
static bool symmetric_stress(const lmm_ctx *c) { return c->prm.stoch_stress_form != 0; }

int lmm_fill_m_stochastic(lmm_ctx *c, uint32_t seed) {
  if (!c || c->n_rngs < 1) return LMM_ERR_ARG;
  long N = c->g.N;
  int dim = c->g.dim, ne = ne_of(dim);
  const bool sym = symmetric_stress(c);
  const int edraw = sym ? 1 : 2;  // node/edge components drawn per array
  long per = (long)(dim + edraw * ne) * N;
  std::vector<double> h((size_t)per * c->n_rngs);
  lmm_fill_normal_host(seed, (int64_t)h.size(), h.data());
  const double s2 = std::sqrt(2.0);
  std::vector<double> ed((size_t)ne * 2 * N);
  for (int r = 0; r < c->n_rngs; r++) {
    double *src = h.data() + (size_t)r * per;
    if (sym)
      for (long q = 0; q < dim * N; q++) src[q] = s2 * src[q];
    for (int e = 0; e < ne; e++)
      for (int comp = 0; comp < 2; comp++) {
        const double *from = src + dim * N + ((long)e * edraw + (sym ? 0 : comp)) * N;
        memcpy(ed.data() + ((size_t)e * 2 + comp) * N, from, sizeof(double) * N);
      }
    HIPCHK(c, hipMemcpyAsync(c->Wcc + (long)r * dim * N, src, sizeof(double) * dim * N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->Wed + (long)r * ne * 2 * N, ed.data(), sizeof(double) * ne * 2 * N,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // ed is reused by the next stage
  }
  return LMM_OK;
}

int lmm_fill_m_stochastic_device(lmm_ctx *c, uint64_t seed, uint64_t step) {
  if (!c || c->n_rngs < 1) return LMM_ERR_ARG;
  const long N = c->g.N;
  const int dim = c->g.dim, ne = ne_of(dim);
  const bool sym = symmetric_stress(c);
  const int edraw = sym ? 1 : 2;
  const long per = (long)(dim + edraw * ne) * N;
  // the value order of lmm_fill_m_stochastic: stage, W_cc components, then
  // the node / edge arrays with their drawn components
  for (int r = 0; r < c->n_rngs; r++) {
    double *wcc = c->Wcc + (long)r * dim * N, *wed = c->Wed + (long)r * ne * 2 * N;
    launch_fill_normal_philox(c->stream, wcc, dim * N, r * per, seed, step);
    if (!sym) {
      launch_fill_normal_philox(c->stream, wed, ne * 2 * N, r * per + dim * N, seed, step);
      continue;
    }
    launch_scale_copy(c->stream, wcc, wcc, (long)dim * N, std::sqrt(2.0));
    for (int e = 0; e < ne; e++) {
      launch_fill_normal_philox(c->stream, wed + (long)e * 2 * N, N, r * per + (long)(dim + e) * N, seed, step);
      HIPCHK(c, hipMemcpyAsync(wed + ((long)e * 2 + 1) * N, wed + (long)e * 2 * N, sizeof(double) * N,
                               hipMemcpyDeviceToDevice, c->stream));
    }
  }
  return finish(c);
}

int lmm_field_fill_normal(lmm_field *f, uint64_t seed, uint64_t step, int64_t v0) {
  if (!f || v0 < 0) return LMM_ERR_ARG;
  lmm_ctx *c = f->ctx;
  launch_fill_normal_philox(c->stream, f->d, kind_count(c) * f->ncomp, (long)v0, seed, step);
  return finish(c);
}

static int stoch_common(lmm_ctx *c, lmm_field *force[3], const double *Wcc, double *const Wed[3], int nst,
                        const double *w_dev, long scc, long sed, lmm_field *eta, lmm_field *const eta_ed[3],
                        lmm_field *temp, lmm_field *const temp_ed[3], double variance) {
  int dim = c->g.dim, ne = ne_of(dim);
  int rc;
  if ((rc = check_stokes(c, force, eta, "stochastic_m_fluxdiv"))) return rc;
  if (!field_ok(c, temp, LMM_CELL)) return set_err(c, LMM_ERR_ARG, "temperature must be a cell field");
  int ek = dim == 2 ? LMM_NODE : LMM_EDGE_XY;
  for (int e = 0; e < ne; e++)
    if (!field_ok(c, eta_ed[e], ek + e) || !field_ok(c, temp_ed[e], ek + e))
      return set_err(c, LMM_ERR_ARG, "eta_ed/temperature_ed[%d] wrong kind", e);
  double *f[3] = {nullptr, nullptr, nullptr}, *ee[3] = {nullptr, nullptr, nullptr}, *te[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < dim; d++) f[d] = force[d]->d;
  for (int e = 0; e < ne; e++) {
    ee[e] = eta_ed[e]->d;
    te[e] = temp_ed[e]->d;
  }
  const int width = c->prm.filtering_width;
  if (width > 0 && !c->stoch_work) {
    c->stoch_work = dalloc<double>(c->stream, 2L * (dim + 2 * ne) * c->g.N);
    if (!c->stoch_work) return set_err(c, LMM_ERR_NOMEM, "cannot allocate the filtering work arrays");
  }
  launch_stoch_m_force(c->stream, c->g, f, Wcc, Wed, nst, w_dev, scc, sed, eta->d, ee, temp->d, te, variance, width,
                       c->stoch_work);
  return finish(c);
}

int lmm_stochastic_m_fluxdiv(lmm_ctx *c, lmm_field *stoch_m_force[3], lmm_field *eta, lmm_field *const eta_ed[3],
                             lmm_field *temperature, lmm_field *const temperature_ed[3], double dt,
                             const double *weights, double variance_coef_mom, double k_B) {
  if (!c || !weights || c->n_rngs < 1 || !(dt > 0)) return LMM_ERR_ARG;
  long N = c->g.N;
  int dim = c->g.dim, ne = ne_of(dim);
  double dv = 1.0;
  for (int d = 0; d < dim; d++) dv *= c->g.dx;
  // stochastic_m_fluxdiv.f90:135: product(dx(n,1:dm))
  double variance = std::sqrt(variance_coef_mom * 2.0 * k_B / (dv * dt));
  memcpy(c->hscal, weights, sizeof(double) * c->n_rngs);
  HIPCHK(c, hipMemcpyAsync(c->dweights, c->hscal, sizeof(double) * c->n_rngs, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double *Wed[3] = {nullptr, nullptr, nullptr};
  for (int e = 0; e < ne; e++) Wed[e] = c->Wed + (long)e * 2 * N;
  return stoch_common(c, stoch_m_force, c->Wcc, Wed, c->n_rngs, c->dweights, (long)dim * N, (long)ne * 2 * N, eta,
                      eta_ed, temperature, temperature_ed, variance);
}

int lmm_stoch_m_force(lmm_ctx *c, lmm_field *stoch_m_force[3], lmm_field *W_cc, lmm_field *const W_ed[3],
                      lmm_field *eta, lmm_field *const eta_ed[3], lmm_field *temperature,
                      lmm_field *const temperature_ed[3], double variance) {
  if (!c) return LMM_ERR_ARG;
  int dim = c->g.dim, ne = ne_of(dim);
  if (!field_ok(c, W_cc, LMM_CELL, dim)) return set_err(c, LMM_ERR_ARG, "W_cc must be a cell field with dim components");
  int ek = dim == 2 ? LMM_NODE : LMM_EDGE_XY;
  double *Wed[3] = {nullptr, nullptr, nullptr};
  for (int e = 0; e < ne; e++) {
    if (!field_ok(c, W_ed[e], ek + e, 2)) return set_err(c, LMM_ERR_ARG, "W_ed[%d] must have 2 components", e);
    Wed[e] = W_ed[e]->d;
  }
  c->hscal[0] = 1.0;
  HIPCHK(c, hipMemcpyAsync(c->scal + 4200, c->hscal, sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return stoch_common(c, stoch_m_force, W_cc->d, Wed, 1, c->scal + 4200, 0, 0, eta, eta_ed, temperature,
                      temperature_ed, variance);
}

// ---------------- species mass fluxes ----------------

void lmm_species_params_default(lmm_species_params *p, int nspecies) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->nspecies = nspecies;
  for (int q = 0; q < 8; q++) {
    p->molmass[q] = 1.0;
    p->rhobar[q] = 1.0;
  }
  for (int q = 0; q < 28; q++) p->Dbar[q] = 1.0;
  p->k_B = 1.0;
  p->variance_coef_mass = 1.0;
  p->fraction_tolerance = 1e-13;
  p->is_ideal_mixture = 1;
  p->is_nonisothermal = 1;
  p->use_lapack = 0;
  p->inverse_type = 1;
  p->chi_iterations = 10;
  p->correct_flux = 1;
  p->barodiffusion_type = 0;
}

int lmm_ctx_set_species(lmm_ctx *c, const lmm_species_params *p) {
  if (!c || !p) return LMM_ERR_ARG;
  if (p->nspecies < 2 || p->nspecies > SP_MAX)
    return set_err(c, LMM_ERR_UNSUPPORTED, "nspecies %d outside 2..%d", p->nspecies, SP_MAX);
  if (p->use_lapack && p->inverse_type != 1 && p->inverse_type != 2)
    return set_err(c, LMM_ERR_ARG, "inverse_type must be 1 or 2");
  if (p->chi_iterations < 0 || !(p->k_B > 0) || !(p->fraction_tolerance >= 0))
    return set_err(c, LMM_ERR_ARG, "chi_iterations, k_B or fraction_tolerance out of range");
  for (int q = 0; q < p->nspecies; q++)
    if (!(p->molmass[q] > 0) || !(p->rhobar[q] > 0)) return set_err(c, LMM_ERR_ARG, "molmass / rhobar must be positive");
  if (c->n_rngs_mass > 0 && c->species_set && p->nspecies != c->species.nspecies) {
    dfree(c->Wmass);
    c->Wmass = nullptr;
    c->n_rngs_mass = 0;  // the noise of another species count no longer fits
  }
  c->species = *p;
  c->species_set = true;
  return LMM_OK;
}

int lmm_ctx_get_species(const lmm_ctx *c, lmm_species_params *p) {
  if (!c || !p) return LMM_ERR_ARG;
  if (c->species_set)
    *p = c->species;
  else
    lmm_species_params_default(p, 2);
  return LMM_OK;
}

// Kernel-side knobs: the context's set, or the defaults for ns species.
static SpParams species_knobs(const lmm_ctx *c, int ns) {
  lmm_species_params p;
  if (c->species_set)
    p = c->species;
  else
    lmm_species_params_default(&p, ns);
  SpParams k;
  memset(&k, 0, sizeof(k));
  k.ns = p.nspecies;
  for (int q = 0; q < 8; q++) {
    k.molmass[q] = p.molmass[q];
    k.rhobar[q] = p.rhobar[q];
    k.Dtherm[q] = p.Dtherm[q];
    k.Hdiag[q] = p.H_diag[q];
  }
  for (int q = 0; q < 28; q++) {
    k.Dbar[q] = p.Dbar[q];
    k.Hoff[q] = p.H_offdiag[q];
  }
  k.k_B = p.k_B;
  k.vcm = p.variance_coef_mass;
  k.ftol = p.fraction_tolerance;
  k.ideal = p.is_ideal_mixture;
  k.noniso = p.is_nonisothermal;
  k.lapack = p.use_lapack;
  k.inverse_type = p.inverse_type;
  k.chi_iter = p.chi_iterations;
  k.correct = p.correct_flux;
  k.baro = p.barodiffusion_type;
  return k;
}

// The species count of the context's knobs (set, else 0 = not set).
static int species_ns(lmm_ctx *c) { return c->species_set ? c->species.nspecies : 0; }

// Species scratch of at least n doubles, kept by the context.
static double *sp_scratch(lmm_ctx *c, long n) {
  if (n > c->sp_work_n) {
    dfree(c->sp_work);
    c->sp_work = dalloc<double>(c->stream, n);
    c->sp_work_n = c->sp_work ? n : 0;
  }
  return c->sp_work;
}

static int sp_status(lmm_ctx *c, int rc, const char *what) {
  if (rc) return set_err(c, rc, "%s: nspecies outside the supported range", what);
  return finish(c);
}

#define SP_FIELD(f, kind, nc, what)                                                              \
  if (!field_ok(c, f, kind, nc)) return set_err(c, LMM_ERR_ARG, "%s: %s missing or with fewer than %d components", \
                                                what, #f, (int)(nc))

static int need_species(lmm_ctx *c, const char *what) {
  if (!c) return LMM_ERR_ARG;
  if (!c->species_set) return set_err(c, LMM_ERR_ARG, "%s: set the species knobs first (lmm_ctx_set_species)", what);
  return LMM_OK;
}

int lmm_correct_rho_with_drho(lmm_ctx *c, lmm_field *rho, lmm_field *drho) {
  if (int rc = need_species(c, "correct_rho_with_drho")) return rc;
  const int ns = species_ns(c);
  SP_FIELD(rho, LMM_CELL, ns, "correct_rho_with_drho");
  SP_FIELD(drho, LMM_CELL, ns, "correct_rho_with_drho");
  return sp_status(c, sp_correct_rho(c->stream, c->g, species_knobs(c, ns), rho->d, drho->d), "correct_rho_with_drho");
}

int lmm_convert_cons_to_prim(lmm_ctx *c, lmm_field *rho, lmm_field *rhotot, lmm_field *molarconc, lmm_field *molmtot) {
  if (int rc = need_species(c, "convert_cons_to_prim")) return rc;
  const int ns = species_ns(c);
  SP_FIELD(rho, LMM_CELL, ns, "convert_cons_to_prim");
  SP_FIELD(rhotot, LMM_CELL, 1, "convert_cons_to_prim");
  SP_FIELD(molarconc, LMM_CELL, ns, "convert_cons_to_prim");
  SP_FIELD(molmtot, LMM_CELL, 1, "convert_cons_to_prim");
  return sp_status(c, sp_cons_to_prim(c->stream, c->g, species_knobs(c, ns), rho->d, rhotot->d, molarconc->d, molmtot->d),
                   "convert_cons_to_prim");
}

int lmm_compute_mixture_properties(lmm_ctx *c, lmm_field *rho, lmm_field *rhotot, lmm_field *D_bar, lmm_field *D_therm,
                                   lmm_field *Hessian, lmm_field *Temp) {
  if (int rc = need_species(c, "compute_mixture_properties")) return rc;
  (void)rho;
  (void)rhotot;
  (void)Temp;
  const int ns = species_ns(c);
  SP_FIELD(D_bar, LMM_CELL, ns * ns, "compute_mixture_properties");
  SP_FIELD(D_therm, LMM_CELL, ns, "compute_mixture_properties");
  SP_FIELD(Hessian, LMM_CELL, ns * ns, "compute_mixture_properties");
  return sp_status(c, sp_mixture(c->stream, c->g, species_knobs(c, ns), D_bar->d, D_therm->d, Hessian->d),
                   "compute_mixture_properties");
}

int lmm_compute_Gama(lmm_ctx *c, lmm_field *rho, lmm_field *rhotot, lmm_field *molarconc, lmm_field *molmtot,
                     lmm_field *Hessian, lmm_field *Gama) {
  if (int rc = need_species(c, "compute_Gama")) return rc;
  (void)rho;
  (void)rhotot;
  (void)molmtot;
  const int ns = species_ns(c);
  SP_FIELD(molarconc, LMM_CELL, ns, "compute_Gama");
  SP_FIELD(Hessian, LMM_CELL, ns * ns, "compute_Gama");
  SP_FIELD(Gama, LMM_CELL, ns * ns, "compute_Gama");
  return sp_status(c, sp_gama(c->stream, c->g, species_knobs(c, ns), molarconc->d, Hessian->d, Gama->d), "compute_Gama");
}

int lmm_compute_chi(lmm_ctx *c, lmm_field *rho, lmm_field *rhotot, lmm_field *molarconc, lmm_field *chi,
                    lmm_field *D_bar, lmm_field *D_therm, lmm_field *Temp, lmm_field *zeta_by_Temp) {
  if (int rc = need_species(c, "compute_chi")) return rc;
  const int ns = species_ns(c);
  const SpParams sp = species_knobs(c, ns);
  SP_FIELD(rho, LMM_CELL, ns, "compute_chi");
  SP_FIELD(rhotot, LMM_CELL, 1, "compute_chi");
  SP_FIELD(molarconc, LMM_CELL, ns, "compute_chi");
  SP_FIELD(chi, LMM_CELL, ns * ns, "compute_chi");
  SP_FIELD(D_bar, LMM_CELL, ns * ns, "compute_chi");
  if (sp.noniso) {
    SP_FIELD(D_therm, LMM_CELL, ns, "compute_chi");
    SP_FIELD(Temp, LMM_CELL, 1, "compute_chi");
    SP_FIELD(zeta_by_Temp, LMM_CELL, ns, "compute_chi");
  }
  return sp_status(c, sp_chi(c->stream, c->g, sp, rho->d, rhotot->d, molarconc->d, D_bar->d,
                             sp.noniso ? D_therm->d : nullptr, sp.noniso ? Temp->d : nullptr, chi->d,
                             sp.noniso ? zeta_by_Temp->d : nullptr),
                   "compute_chi");
}

int lmm_compute_rhoWchi(lmm_ctx *c, lmm_field *rho, lmm_field *rhotot, lmm_field *chi, lmm_field *rhoWchi) {
  if (int rc = need_species(c, "compute_rhoWchi")) return rc;
  (void)rhotot;
  const int ns = species_ns(c);
  SP_FIELD(rho, LMM_CELL, ns, "compute_rhoWchi");
  SP_FIELD(chi, LMM_CELL, ns * ns, "compute_rhoWchi");
  SP_FIELD(rhoWchi, LMM_CELL, ns * ns, "compute_rhoWchi");
  return sp_status(c, sp_rhowchi(c->stream, c->g, species_knobs(c, ns), rho->d, chi->d, rhoWchi->d), "compute_rhoWchi");
}

int lmm_compute_Lonsager(lmm_ctx *c, lmm_field *rho, lmm_field *rhotot, lmm_field *molarconc, lmm_field *molmtot,
                         lmm_field *chi, lmm_field *Gama, lmm_field *Lonsager) {
  if (int rc = need_species(c, "compute_Lonsager")) return rc;
  (void)molarconc;
  (void)Gama;
  const int ns = species_ns(c);
  SP_FIELD(rho, LMM_CELL, ns, "compute_Lonsager");
  SP_FIELD(rhotot, LMM_CELL, 1, "compute_Lonsager");
  SP_FIELD(molmtot, LMM_CELL, 1, "compute_Lonsager");
  SP_FIELD(chi, LMM_CELL, ns * ns, "compute_Lonsager");
  SP_FIELD(Lonsager, LMM_CELL, ns * ns, "compute_Lonsager");
  return sp_status(c, sp_lonsager(c->stream, c->g, species_knobs(c, ns), rho->d, rhotot->d, molmtot->d, chi->d,
                                  Lonsager->d),
                   "compute_Lonsager");
}

// Arguments of diffusive_mass_fluxdiv common to the allocating and the
// allocation-free entry: checks the fields and resolves the optional pieces.
struct DiffArgs {
  SpParams sp;
  const double *temp = nullptr, *zeta = nullptr;
  const double *gpb[3] = {nullptr, nullptr, nullptr};
  bool baro = false;
  double *ft[3] = {nullptr, nullptr, nullptr};
};

static int diff_args(lmm_ctx *c, int ns, lmm_field *rho, lmm_field *molarconc, lmm_field *rhoWchi, lmm_field *Gama,
                     lmm_field *diff_fluxdiv, lmm_field *Temp, lmm_field *zeta_by_Temp, lmm_field *const gradp_baro[3],
                     lmm_field *flux_total[3], int correct_flux, DiffArgs *a) {
  if (ns < 1 || ns > SP_MAX) return set_err(c, LMM_ERR_UNSUPPORTED, "nspecies %d outside 1..%d", ns, SP_MAX);
  if (c->species_set && ns != c->species.nspecies)
    return set_err(c, LMM_ERR_ARG, "nspecies %d differs from the context's species knobs (%d)", ns,
                   c->species.nspecies);
  a->sp = species_knobs(c, ns);
  a->sp.ns = ns;
  a->sp.correct = correct_flux;
  if (!field_ok(c, molarconc, LMM_CELL, ns) || !field_ok(c, rhoWchi, LMM_CELL, ns * ns) ||
      !field_ok(c, Gama, LMM_CELL, ns * ns) || !field_ok(c, diff_fluxdiv, LMM_CELL, ns))
    return set_err(c, LMM_ERR_ARG, "diffusive_mass_fluxdiv: cell fields have wrong components");
  const int dim = c->g.dim;
  for (int d = 0; d < dim; d++) {
    if (!field_ok(c, flux_total[d], LMM_FACE_X + d, ns)) return set_err(c, LMM_ERR_ARG, "flux_total[%d] wrong", d);
    a->ft[d] = flux_total[d]->d;
  }
  if ((Temp == nullptr) != (zeta_by_Temp == nullptr))
    return set_err(c, LMM_ERR_ARG, "diffusive_mass_fluxdiv: Temp and zeta_by_Temp go together");
  if (a->sp.noniso && Temp) {
    if (!field_ok(c, Temp, LMM_CELL) || !field_ok(c, zeta_by_Temp, LMM_CELL, ns))
      return set_err(c, LMM_ERR_ARG, "diffusive_mass_fluxdiv: Temp / zeta_by_Temp wrong");
    a->temp = Temp->d;
    a->zeta = zeta_by_Temp->d;
  }
  if (a->sp.baro > 0) {
    if (dim != 2)
      return set_err(c, LMM_ERR_UNSUPPORTED, "barodiffusion is 2D only (compute_baro_coef_3d is not written)");
    if (!field_ok(c, rho, LMM_CELL, ns) || !field_ok(c, Temp, LMM_CELL) || !gradp_baro)
      return set_err(c, LMM_ERR_ARG, "barodiffusion needs rho, Temp and gradp_baro");
    for (int d = 0; d < dim; d++) {
      if (!field_ok(c, gradp_baro[d], LMM_FACE_X + d)) return set_err(c, LMM_ERR_ARG, "gradp_baro[%d] wrong", d);
      a->gpb[d] = gradp_baro[d]->d;
    }
    a->temp = Temp->d;
    a->baro = true;
  }
  return LMM_OK;
}

int lmm_diffusive_mass_fluxdiv(lmm_ctx *c, int ns, lmm_field *rho, lmm_field *rhotot, lmm_field *molarconc,
                               lmm_field *rhoWchi, lmm_field *Gama, lmm_field *diff_fluxdiv, lmm_field *Temp,
                               lmm_field *zeta_by_Temp, lmm_field *const gradp_baro[3], lmm_field *flux_total[3],
                               int correct_flux) {
  if (!c) return LMM_ERR_ARG;
  (void)rhotot;
  DiffArgs a;
  if (int rc = diff_args(c, ns, rho, molarconc, rhoWchi, Gama, diff_fluxdiv, Temp, zeta_by_Temp, gradp_baro,
                         flux_total, correct_flux, &a))
    return rc;
  const long N = c->g.N;
  double *work = sp_scratch(c, (long)(c->g.dim + (a.baro ? 1 : 0)) * ns * N);
  if (!work) return set_err(c, LMM_ERR_NOMEM, "species scratch");
  double *baro = a.baro ? work + (long)c->g.dim * ns * N : nullptr;
  return sp_status(c, sp_diffusive(c->stream, c->g, a.sp, a.baro ? rho->d : nullptr, molarconc->d, rhoWchi->d, Gama->d,
                                   a.temp, a.zeta, a.baro ? a.gpb : nullptr, baro, work, a.ft, diff_fluxdiv->d),
                   "diffusive_mass_fluxdiv");
}

int lmm_init_mass_stochastic(lmm_ctx *c, int n_rngs) {
  if (int rc = need_species(c, "init_mass_stochastic")) return rc;
  if (n_rngs < 1 || n_rngs > 8) return set_err(c, LMM_ERR_ARG, "n_rngs %d outside 1..8", n_rngs);
  dfree(c->Wmass);
  dfree(c->dweights_mass);
  const long cnt = (long)n_rngs * c->g.dim * c->species.nspecies * c->g.N;
  c->Wmass = dalloc<double>(c->stream, cnt);
  c->dweights_mass = dalloc<double>(c->stream, 8);
  c->n_rngs_mass = (c->Wmass && c->dweights_mass) ? n_rngs : 0;
  c->weights_mass_valid = false;
  if (!c->n_rngs_mass) return set_err(c, LMM_ERR_NOMEM, "mass noise allocation failed");
  return finish(c);
}

static long mass_noise_count(const lmm_ctx *c) {
  return (long)c->n_rngs_mass * c->g.dim * c->species.nspecies * c->g.N;
}

int lmm_fill_mass_stochastic(lmm_ctx *c, uint32_t seed) {
  if (!c || c->n_rngs_mass < 1) return c ? set_err(c, LMM_ERR_ARG, "fill_mass_stochastic before init") : LMM_ERR_ARG;
  std::vector<double> h((size_t)mass_noise_count(c));
  lmm_fill_normal_host(seed, (int64_t)h.size(), h.data());
  HIPCHK(c, hipMemcpyAsync(c->Wmass, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return finish(c);
}

int lmm_fill_mass_stochastic_device(lmm_ctx *c, uint64_t seed, uint64_t step) {
  if (!c || c->n_rngs_mass < 1) return c ? set_err(c, LMM_ERR_ARG, "fill_mass_stochastic before init") : LMM_ERR_ARG;
  launch_fill_normal_philox(c->stream, c->Wmass, mass_noise_count(c), 0, seed, step);
  return finish(c);
}

// stochastic_mass_fluxdiv.f90:69: sqrt(2 k_B variance_coef_mass / (product(dx) dt))
static double mass_variance(const lmm_ctx *c, double dt) {
  double dv = 1.0;
  for (int d = 0; d < c->g.dim; d++) dv = dv * c->g.dx;
  return std::sqrt(2.0 * c->species.k_B * c->species.variance_coef_mass / (dv * dt));
}

static int upload_mass_weights(lmm_ctx *c, const double *weights) {
  if (c->weights_mass_valid && !memcmp(c->hweights_mass, weights, sizeof(double) * c->n_rngs_mass)) return LMM_OK;
  memcpy(c->hweights_mass, weights, sizeof(double) * c->n_rngs_mass);
  c->weights_mass_valid = true;
  memcpy(c->hscal, weights, sizeof(double) * c->n_rngs_mass);
  HIPCHK(c, hipMemcpyAsync(c->dweights_mass, c->hscal, sizeof(double) * c->n_rngs_mass, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LMM_OK;
}

int lmm_stochastic_mass_fluxdiv(lmm_ctx *c, lmm_field *rho, lmm_field *rhotot, lmm_field *molarconc, lmm_field *molmtot,
                                lmm_field *chi, lmm_field *Gama, lmm_field *stoch_fluxdiv, lmm_field *flux_total[3],
                                double dt, const double *weights, int increment) {
  if (int rc = need_species(c, "stochastic_mass_fluxdiv")) return rc;
  (void)molarconc;
  (void)Gama;
  const int ns = species_ns(c), dim = c->g.dim;
  if (c->n_rngs_mass < 1 || !weights || !(dt > 0))
    return set_err(c, LMM_ERR_ARG, "stochastic_mass_fluxdiv: noise not initialised, no weights or dt <= 0");
  SP_FIELD(rho, LMM_CELL, ns, "stochastic_mass_fluxdiv");
  SP_FIELD(rhotot, LMM_CELL, 1, "stochastic_mass_fluxdiv");
  SP_FIELD(molmtot, LMM_CELL, 1, "stochastic_mass_fluxdiv");
  SP_FIELD(chi, LMM_CELL, ns * ns, "stochastic_mass_fluxdiv");
  SP_FIELD(stoch_fluxdiv, LMM_CELL, ns, "stochastic_mass_fluxdiv");
  double *ft[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < dim; d++) {
    if (!field_ok(c, flux_total[d], LMM_FACE_X + d, ns)) return set_err(c, LMM_ERR_ARG, "flux_total[%d] wrong", d);
    ft[d] = flux_total[d]->d;
  }
  const long N = c->g.N;
  double *work = sp_scratch(c, (long)(ns * ns + dim * ns) * N);
  if (!work) return set_err(c, LMM_ERR_NOMEM, "species scratch");
  if (int rc = upload_mass_weights(c, weights)) return rc;
  const SpParams sp = species_knobs(c, ns);
  double *L = work, *flux = work + (long)ns * ns * N;
  int rc = sp_lonsager(c->stream, c->g, sp, rho->d, rhotot->d, molmtot->d, chi->d, L);
  if (!rc)
    rc = sp_stochastic(c->stream, c->g, sp, L, c->Wmass, c->n_rngs_mass, c->dweights_mass, mass_variance(c, dt), flux,
                       ft, stoch_fluxdiv->d, increment);
  return sp_status(c, rc, "stochastic_mass_fluxdiv");
}

int lmm_compute_mass_fluxdiv(lmm_ctx *c, lmm_field *rho, lmm_field *const gradp_baro[3], lmm_field *diff_fluxdiv,
                             lmm_field *stoch_fluxdiv, lmm_field *Temp, lmm_field *flux_total[3], double dt,
                             double stage_time, const double *weights) {
  if (int rc = need_species(c, "compute_mass_fluxdiv")) return rc;
  (void)stage_time;
  const int ns = species_ns(c), dim = c->g.dim;
  const SpParams sp = species_knobs(c, ns);
  SP_FIELD(rho, LMM_CELL, ns, "compute_mass_fluxdiv");
  SP_FIELD(diff_fluxdiv, LMM_CELL, ns, "compute_mass_fluxdiv");
  SP_FIELD(stoch_fluxdiv, LMM_CELL, ns, "compute_mass_fluxdiv");
  double *ft[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < dim; d++) {
    if (!field_ok(c, flux_total[d], LMM_FACE_X + d, ns)) return set_err(c, LMM_ERR_ARG, "flux_total[%d] wrong", d);
    ft[d] = flux_total[d]->d;
  }
  if ((sp.noniso || sp.baro > 0) && !field_ok(c, Temp, LMM_CELL))
    return set_err(c, LMM_ERR_ARG, "compute_mass_fluxdiv: Temp is required (non-isothermal or barodiffusion)");
  const bool baro = sp.baro > 0;
  SpFaceIn in{};
  if (baro) {
    if (dim != 2)
      return set_err(c, LMM_ERR_UNSUPPORTED, "barodiffusion is 2D only (compute_baro_coef_3d is not written)");
    for (int d = 0; d < dim; d++) {
      if (!gradp_baro || !field_ok(c, gradp_baro[d], LMM_FACE_X + d))
        return set_err(c, LMM_ERR_ARG, "gradp_baro[%d] missing or wrong", d);
      in.gpb[d] = gradp_baro[d]->d;
    }
  }
  const bool stoch = sp.vcm != 0.0;
  if (stoch && (c->n_rngs_mass < 1 || !weights || !(dt > 0)))
    return set_err(c, LMM_ERR_ARG, "compute_mass_fluxdiv: mass noise not initialised, no weights or dt <= 0");
  // scratch: molarconc, rhoWchi, [Gama], [zeta], [L factor], [baro], face flux
  const long N = c->g.N;
  const long nlp = (long)ns * (ns + 1) / 2;
  const long need = (long)(ns + ns * ns + (sp.ideal ? 0 : ns * ns) + (sp.noniso ? ns : 0) + (stoch ? nlp : 0) +
                           (baro ? ns : 0) + (stoch ? 2 : 1) * dim * ns) * N;
  double *w = sp_scratch(c, need);
  if (!w) return set_err(c, LMM_ERR_NOMEM, "species scratch");
  if (stoch)
    if (int rc = upload_mass_weights(c, weights)) return rc;
  SpCellOut o{};
  o.rho = rho->d;
  o.xm = w;
  w += (long)ns * N;
  o.rw = w;
  w += (long)ns * ns * N;
  if (!sp.ideal) {
    o.gama = w;
    w += (long)ns * ns * N;
  }
  if (sp.noniso) {
    o.zeta = w;
    w += (long)ns * N;
  }
  if (stoch) {
    o.lpk = w;
    w += nlp * N;
  }
  if (baro) {
    o.baro = w;
    w += (long)ns * N;
  }
  double *flux = w;
  in.xm = o.xm;
  in.rw = o.rw;
  in.gama = o.gama;
  in.zeta = o.zeta;
  in.temp = Temp ? Temp->d : nullptr;
  in.baro = o.baro;
  in.lpk = o.lpk;
  in.W = stoch ? c->Wmass : nullptr;
  in.weights = c->dweights_mass;
  in.n_rngs = c->n_rngs_mass;
  in.variance = stoch ? mass_variance(c, dt) : 0.0;
  // algorithmic bytes (each input read once, each output written once):
  // the cell pass reads rho and Temp and writes rho back; the face passes
  // read the noise and write both divergences and flux_total
  const double bN = 8.0 * (double)N;
  int rc;
  {
    ProfScope ps(c, "species_cell", bN * (2 * ns + ((sp.noniso || baro) ? 1 : 0)));
    rc = sp_chain_cell(c->stream, c->g, sp, Temp ? Temp->d : nullptr, o);
  }
  if (!rc) {
    ProfScope ps(c, "species_face", bN * ((stoch ? c->n_rngs_mass * dim * ns : 0) + 2 * ns + dim * ns +
                                          (baro ? dim : 0)));
    rc = sp_chain_face(c->stream, c->g, sp, in, flux, diff_fluxdiv->d, stoch_fluxdiv->d, ft);
  }
  if (!rc && !stoch) launch_fill<double>(c->stream, stoch_fluxdiv->d, (long)ns * N, 0.0);
  return sp_status(c, rc, "compute_mass_fluxdiv");
}

// ---------------- inertial time step ----------------

void lmm_inertial_params_default(lmm_inertial_params *p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->variance_coef_mom = 1.0;  // probin_common.f90:213
  p->visc_coef = 1.0;          // :225
  p->noise_seed = 1;
  p->restart = -1;             // :205
}

int lmm_mk_advective_s_fluxdiv(lmm_ctx *c, lmm_field *const umac[3], lmm_field *const s_fc[3], lmm_field *s_update,
                               int start_comp, int num_comp) {
  if (!c) return LMM_ERR_ARG;
  if (int rc = check_stokes(c, umac, nullptr, "mk_advective_s_fluxdiv")) return rc;
  if (!s_update || s_update->ctx != c || s_update->kind != LMM_CELL || !comps_ok(s_update, start_comp, num_comp))
    return set_err(c, LMM_ERR_ARG, "mk_advective_s_fluxdiv: s_update components");
  AdvArgs a{};
  const long N = c->g.N;
  for (int d = 0; d < c->g.dim; d++) {
    if (!field_ok(c, s_fc[d], LMM_FACE_X + d) || !comps_ok(s_fc[d], start_comp, num_comp))
      return set_err(c, LMM_ERR_ARG, "mk_advective_s_fluxdiv: s_fc[%d] components", d);
    a.u[d] = umac[d]->d;
    a.s[d] = s_fc[d]->d + (long)start_comp * N;
  }
  a.dxinv = 1.0 / c->g.dx;
  RhoArgs r{};
  r.upd = s_update->d + (long)start_comp * N;
  launch_adv_s(c->stream, c->g, num_comp, a, r, 0);
  return finish(c);
}

int lmm_mk_advective_m_fluxdiv(lmm_ctx *c, lmm_field *const umac[3], lmm_field *const m[3], lmm_field *m_update[3]) {
  if (!c) return LMM_ERR_ARG;
  int rc;
  if ((rc = check_stokes(c, umac, nullptr, "mk_advective_m_fluxdiv umac")) ||
      (rc = check_stokes(c, m, nullptr, "mk_advective_m_fluxdiv m")) ||
      (rc = check_stokes(c, m_update, nullptr, "mk_advective_m_fluxdiv m_update")))
    return rc;
  double *u[3] = {nullptr, nullptr, nullptr}, *mm[3] = {nullptr, nullptr, nullptr}, *up[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < c->g.dim; d++) {
    u[d] = umac[d]->d;
    mm[d] = m[d]->d;
    up[d] = m_update[d]->d;
  }
  launch_adv_m(c->stream, c->g, u, mm, up);
  return finish(c);
}

int lmm_mk_grav_force(lmm_ctx *c, lmm_field *m_force[3], lmm_field *const s_fc_old[3], lmm_field *const s_fc_new[3],
                      const double grav[3]) {
  if (!c || !grav) return LMM_ERR_ARG;
  int rc;
  if ((rc = check_stokes(c, m_force, nullptr, "mk_grav_force m_force")) ||
      (rc = check_stokes(c, s_fc_old, nullptr, "mk_grav_force s_fc_old")) ||
      (rc = check_stokes(c, s_fc_new, nullptr, "mk_grav_force s_fc_new")))
    return rc;
  double *f[3] = {nullptr, nullptr, nullptr}, *so[3] = {nullptr, nullptr, nullptr}, *sn[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < c->g.dim; d++) {
    f[d] = m_force[d]->d;
    so[d] = s_fc_old[d]->d;
    sn[d] = s_fc_new[d]->d;
  }
  launch_grav(c->stream, c->g, grav, f, so, sn);
  return finish(c);
}

// Argument checks shared by the time steps (advance_timestep_inertial /
// _overdamped take the same fields); nrng: the noise stages the step uses.
static int check_step(lmm_ctx *c, const char *what, lmm_field *umac[3], lmm_field *rho_old, lmm_field *rho_new,
                      lmm_field *rhotot_old, lmm_field *rhotot_new, lmm_field *gradp_baro[3], lmm_field *pi,
                      lmm_field *eta, lmm_field *eta_ed[3], lmm_field *kappa, lmm_field *Temp,
                      lmm_field *Temp_ed[3], lmm_field *diff_mass_fluxdiv, lmm_field *stoch_mass_fluxdiv, double dt,
                      const lmm_inertial_params *ip, int nrng) {
  if (int rc = need_species(c, what)) return rc;
  if (!ip || !(dt > 0)) return set_err(c, LMM_ERR_ARG, "%s: no parameters or dt <= 0", what);
  if (ip->advection_type != 0)
    return set_err(c, LMM_ERR_UNSUPPORTED, "%s: only centred advection (advection_type 0)", what);
  if (ip->eta_model != 0 && ip->eta_model != 1)
    return set_err(c, LMM_ERR_ARG, "%s: eta_model is 0 (constant) or 1 (linear in rhotot)", what);
  const lmm_species_params &sp = c->species;
  const int ns = sp.nspecies, dim = c->g.dim, ne = ne_of(dim);
  const int baro = sp.barodiffusion_type;
  if (baro != 0 && baro != 2) return set_err(c, LMM_ERR_UNSUPPORTED, "%s: barodiffusion_type %d", what, baro);
  int rc;
  if ((rc = check_stokes(c, umac, pi, what))) return rc;
  SP_FIELD(rho_old, LMM_CELL, ns, what);
  SP_FIELD(rho_new, LMM_CELL, ns, what);
  SP_FIELD(diff_mass_fluxdiv, LMM_CELL, ns, what);
  SP_FIELD(stoch_mass_fluxdiv, LMM_CELL, ns, what);
  SP_FIELD(rhotot_old, LMM_CELL, 1, what);
  SP_FIELD(rhotot_new, LMM_CELL, 1, what);
  SP_FIELD(eta, LMM_CELL, 1, what);
  SP_FIELD(kappa, LMM_CELL, 1, what);
  SP_FIELD(Temp, LMM_CELL, 1, what);
  const int ek = dim == 2 ? LMM_NODE : LMM_EDGE_XY;
  for (int e = 0; e < ne; e++)
    if (!field_ok(c, eta_ed[e], ek + e) || !field_ok(c, Temp_ed[e], ek + e))
      return set_err(c, LMM_ERR_ARG, "%s: eta_ed / Temp_ed[%d] missing or wrong kind", what, e);
  if (baro == 2 && (rc = check_stokes(c, gradp_baro, nullptr, what))) return rc;
  if (ip->variance_coef_mom != 0.0 && c->n_rngs != nrng)
    return set_err(c, LMM_ERR_ARG, "%s: momentum noise needs lmm_init_m_stochastic(ctx, %d) and a first fill", what,
                   nrng);
  if (sp.variance_coef_mass != 0.0 && c->n_rngs_mass != nrng)
    return set_err(c, LMM_ERR_ARG, "%s: mass noise needs lmm_init_mass_stochastic(ctx, %d) and a first fill", what,
                   nrng);
  return LMM_OK;
}

// A time step's temporaries, views of the context's scratch c->iw: 10 face
// vectors, 2 face vectors of ns components, 5 cell arrays, ne
// node/edge arrays.
struct StepVec {
  lmm_field f[3];
  lmm_field *h[3];
  double *d[3];
};
struct StepWork {
  StepVec mold, mtemp, m_a, m_d, m_s, rv, dumac, rfo, gradpi, rf, rho_fc, flux_total;
  lmm_field bp, dpi, divu, eh, kh;
  lmm_field eeh[3];
  lmm_field *eehp[3];
};

static int step_work(lmm_ctx *c, const char *what, StepWork &W) {
  const int ns = c->species.nspecies, dim = c->g.dim, ne = ne_of(dim);
  const int ek = dim == 2 ? LMM_NODE : LMM_EDGE_XY;
  const long N = c->g.N;
  const long need = (long)(10 * dim + 2 * dim * ns + 5 + ne) * N;
  if (c->iw_n < need) {
    dfree(c->iw);
    c->iw = dalloc<double>(c->stream, need);
    c->iw_n = c->iw ? need : 0;
    if (!c->iw) return set_err(c, LMM_ERR_NOMEM, "%s: temporaries", what);
  }
  double *w = c->iw;
  auto take = [&](long n) {
    double *q = w;
    w += n;
    return q;
  };
  auto faces = [&](StepVec &v, int ncomp) {
    for (int d = 0; d < 3; d++) {
      v.d[d] = d < dim ? take((long)ncomp * N) : nullptr;
      v.f[d] = lmm_field{c, LMM_FACE_X + d, ncomp, v.d[d]};
      v.h[d] = d < dim ? &v.f[d] : nullptr;
    }
  };
  for (StepVec *v : {&W.mold, &W.mtemp, &W.m_a, &W.m_d, &W.m_s, &W.rv, &W.dumac, &W.rfo, &W.gradpi, &W.rf})
    faces(*v, 1);
  faces(W.rho_fc, ns);
  faces(W.flux_total, ns);
  W.bp = lmm_field{c, LMM_CELL, 1, take(N)};
  W.dpi = lmm_field{c, LMM_CELL, 1, take(N)};
  W.divu = lmm_field{c, LMM_CELL, 1, take(N)};
  W.eh = lmm_field{c, LMM_CELL, 1, take(N)};
  W.kh = lmm_field{c, LMM_CELL, 1, take(N)};
  for (int e = 0; e < 3; e++) {
    W.eeh[e] = lmm_field{c, ek + e, 1, e < ne ? take(N) : nullptr};
    W.eehp[e] = e < ne ? &W.eeh[e] : nullptr;
  }
  return LMM_OK;
}

// advance_timestep_inertial.f90:58-990 on a periodic box: the reference's
// multifab sequence, with runs of element-wise calls fused into single
// passes (lmm_inertial.hip) and the library's own routines for the
// operators, fluxes and solves.  Temporaries live in the context (c->iw),
// viewed as fields.
int lmm_advance_timestep_inertial(lmm_ctx *c, lmm_field *umac[3], lmm_field *rho_old, lmm_field *rho_new,
                                  lmm_field *rhotot_old, lmm_field *rhotot_new, lmm_field *gradp_baro[3],
                                  lmm_field *pi, lmm_field *eta, lmm_field *eta_ed[3], lmm_field *kappa,
                                  lmm_field *Temp, lmm_field *Temp_ed[3], lmm_field *diff_mass_fluxdiv,
                                  lmm_field *stoch_mass_fluxdiv, double dt, double time, int64_t istep,
                                  const lmm_inertial_params *ip, lmm_gmres_info info[2]) {
  const char *what = "advance_timestep_inertial";
  int rc;
  if ((rc = check_step(c, what, umac, rho_old, rho_new, rhotot_old, rhotot_new, gradp_baro, pi, eta, eta_ed, kappa,
                       Temp, Temp_ed, diff_mass_fluxdiv, stoch_mass_fluxdiv, dt, ip, 1)))
    return rc;
  const lmm_species_params &sp = c->species;
  const int ns = sp.nspecies, dim = c->g.dim, ne = ne_of(dim);
  const int baro = sp.barodiffusion_type;
  const bool noise_m = ip->variance_coef_mom != 0.0, noise_mass = sp.variance_coef_mass != 0.0;
  const long N = c->g.N;
  StepWork W;
  if ((rc = step_work(c, what, W))) return rc;
  StepVec &mold = W.mold, &mtemp = W.mtemp, &m_a = W.m_a, &m_d = W.m_d, &m_s = W.m_s, &rv = W.rv, &dumac = W.dumac,
          &rfo = W.rfo, &gradpi = W.gradpi, &rf = W.rf, &rho_fc = W.rho_fc, &flux_total = W.flux_total;
  lmm_field &bp_f = W.bp, &dpi_f = W.dpi, &divu_f = W.divu, &eh_f = W.eh, &kh_f = W.kh;
  lmm_field *eeh_f = W.eeh;
  lmm_field **eeh = W.eehp;
  double *u[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < dim; d++) u[d] = umac[d]->d;
  hipStream_t s = c->stream;
  const double inv_dt = 1.0 / dt;
  const double theta_alpha = 1.0 / dt;
  const double weights[8] = {1.0, 0, 0, 0, 0, 0, 0, 0};
  RhobarInv rbi{};
  for (int q = 0; q < ns; q++) rbi.v[q] = -1.0 / sp.rhobar[q];
  lmm_field *const *gb = baro ? gradp_baro : nullptr;
  AdvArgs adv{};
  for (int d = 0; d < dim; d++) {
    adv.u[d] = u[d];
    adv.s[d] = rho_fc.d[d];
  }
  adv.dxinv = 1.0 / c->g.dx;
#define ICHK(x)                  \
  do {                           \
    if ((rc = (x)) != LMM_OK) return rc; \
  } while (0)
  // the viscosity and bulk viscosity at half weight for the solves, and the
  // divergence of umac and gmres_rhs_p (:410-446, :883-902)
  auto solve = [&](lmm_gmres_info *inf, double *npr) -> int {
    launch_scale_copy(s, eh_f.d, eta->d, N, 0.5);
    launch_scale_copy(s, kh_f.d, kappa->d, N, 0.5);
    for (int e = 0; e < ne; e++) launch_scale_copy(s, eeh_f[e].d, eta_ed[e]->d, N, 0.5);
    for (int d = 0; d < dim; d++) launch_fill<double>(s, dumac.d[d], N, 0.0);
    launch_fill<double>(s, dpi_f.d, N, 0.0);
    int r = lmm_gmres(c, rv.h, &bp_f, dumac.h, &dpi_f, rf.h, &eh_f, eeh, &kh_f, theta_alpha, npr, inf);
    if (r) return r;
    // v = v + dumac; pi = pi + dpi
    for (int d = 0; d < dim; d++) launch_add(s, u[d], dumac.d[d], N);
    launch_add(s, pi->d, dpi_f.d, N);
    return LMM_OK;
  };
  // eta, eta_ed, kappa from the composition (compute_eta_kappa)
  auto eta_kappa = [&]() -> int {
    launch_eta(s, c->g, eta->d, kappa->d, rhotot_new->d, ip->visc_coef, ip->eta_model);
    return dim == 2 ? lmm_average_cc_to_node(c, eta, eta_ed[0]) : lmm_average_cc_to_edge(c, eta, eta_ed);
  };
  // the mass fluxes at the new state, negated, and gmres_rhs_p (:376-408)
  auto mass_and_rhs_p = [&]() -> int {
    int r = lmm_compute_mass_fluxdiv(c, rho_new, gb, diff_mass_fluxdiv, stoch_mass_fluxdiv, Temp, flux_total.h, dt,
                                     time, weights);
    if (r) return r;
    if ((r = lmm_compute_div(c, umac, &divu_f, 0))) return r;
    launch_rhs_p(s, c->g, ns, bp_f.d, diff_mass_fluxdiv->d, noise_mass ? stoch_mass_fluxdiv->d : nullptr, divu_f.d,
                 rbi);
    return LMM_OK;
  };
  lmm_gmres_info loc[2];
  if (!info) info = loc;
  memset(info, 0, 2 * sizeof(lmm_gmres_info));
  // one synchronisation at the end instead of one per inner call
  struct DeferSync {
    lmm_ctx *c;
    explicit DeferSync(lmm_ctx *x) : c(x) { c->defer_sync = true; }
    ~DeferSync() { c->defer_sync = false; }
  } defer(c);

  // ---- step 2: predictor Euler step for rho (:169-257)
  ICHK(lmm_average_cc_to_face_comps(c, rho_old, rho_fc.h, 0, ns));
  ICHK(lmm_average_cc_to_face(c, rhotot_old, rfo.h));
  {
    RhoArgs r{};
    r.diff = diff_mass_fluxdiv->d;
    r.stoch = noise_mass ? stoch_mass_fluxdiv->d : nullptr;
    r.rho_old = rho_old->d;
    r.rho_new = rho_new->d;
    r.dt = dt;
    launch_adv_s(s, c->g, ns, adv, r, 1);
  }
  launch_rhotot_conc(s, c->g, ns, rho_new->d, rhotot_new->d);
  ICHK(lmm_average_cc_to_face_comps(c, rho_new, rho_fc.h, 0, ns));
  ICHK(lmm_average_cc_to_face(c, rhotot_new, rf.h));

  // ---- step 4: predictor Crank-Nicolson step (:264-496)
  launch_mul_faces(s, c->g, mold.d, u, rfo.d);
  ICHK(lmm_compute_grad(c, pi, gradpi.h));
  if (baro == 2)
    for (int d = 0; d < dim; d++)
      hipMemcpyAsync(gradp_baro[d]->d, gradpi.d[d], sizeof(double) * N, hipMemcpyDeviceToDevice, s);
  for (int d = 0; d < dim; d++) {
    launch_fill<double>(s, m_a.d[d], N, 0.0);
    launch_fill<double>(s, m_d.d[d], N, 0.0);
    launch_fill<double>(s, m_s.d[d], N, 0.0);
  }
  launch_adv_m(s, c->g, u, mold.d, m_a.d);
  ICHK(lmm_diffusive_m_fluxdiv(c, m_d.h, umac, eta, eta_ed, kappa));
  if (noise_m)
    ICHK(lmm_stochastic_m_fluxdiv(c, m_s.h, eta, eta_ed, Temp, Temp_ed, dt, weights, ip->variance_coef_mom, sp.k_B));
  const bool grav = ip->grav[0] != 0.0 || ip->grav[1] != 0.0 || (dim == 3 && ip->grav[2] != 0.0);
  RhsVArgs rva{};
  for (int d = 0; d < dim; d++) {
    rva.v[d] = rv.d[d];
    rva.mold[d] = mold.d[d];
    rva.gradpi[d] = gradpi.d[d];
    rva.m_a[d] = m_a.d[d];
    rva.m_d[d] = m_d.d[d];
    rva.m_s[d] = m_s.d[d];
    rva.rfo[d] = rfo.d[d];
    rva.g[d] = ip->grav[d];
  }
  rva.inv_dt = inv_dt;
  rva.grav = grav;
  {
    RhsVArgs r = rva;
    for (int d = 0; d < dim; d++) r.rfn[d] = rfo.d[d];  // rho^n g at both ends
    r.half_md = 1;
    launch_rhs_v(s, c->g, r);
  }
  ICHK(eta_kappa());
  ICHK(mass_and_rhs_p());
  launch_v_sub_mt(s, c->g, rv.d, u, rf.d, inv_dt);
  for (int d = 0; d < dim; d++) launch_fill<double>(s, mtemp.d[d], N, 0.0);
  ICHK(lmm_diffusive_m_fluxdiv(c, mtemp.h, umac, eta, eta_ed, kappa));
  launch_v_add_half(s, c->g, rv.d, mtemp.d);
  double npr = 0.0;
  ICHK(solve(&info[0], &npr));
  launch_mul_faces(s, c->g, mtemp.d, u, rf.d);

  // ---- step 5: trapezoidal scalar corrector (:585-606)
  {
    RhoArgs r{};
    r.diff = diff_mass_fluxdiv->d;
    r.stoch = noise_mass ? stoch_mass_fluxdiv->d : nullptr;
    r.rho_old = rho_old->d;
    r.rho_new = rho_new->d;
    r.dt = dt / 2.0;
    launch_adv_s(s, c->g, ns, adv, r, 2);
  }
  launch_rhotot_conc(s, c->g, ns, rho_new->d, rhotot_new->d);
  ICHK(lmm_average_cc_to_face_comps(c, rho_new, rho_fc.h, 0, ns));
  ICHK(lmm_average_cc_to_face(c, rhotot_new, rf.h));
  ICHK(eta_kappa());

  // ---- step 7: corrector Crank-Nicolson step (:641-957)
  ICHK(lmm_compute_grad(c, pi, gradpi.h));
  if (baro == 2)
    for (int d = 0; d < dim; d++)
      hipMemcpyAsync(gradp_baro[d]->d, gradpi.d[d], sizeof(double) * N, hipMemcpyDeviceToDevice, s);
  launch_adv_m(s, c->g, u, mtemp.d, m_a.d);
  if (noise_m)
    ICHK(lmm_stochastic_m_fluxdiv(c, m_s.h, eta, eta_ed, Temp, Temp_ed, dt, weights, ip->variance_coef_mom, sp.k_B));
  {
    RhsVArgs r = rva;
    for (int d = 0; d < dim; d++) r.rfn[d] = rf.d[d];
    r.half_ma = 1;
    r.half_ms = noise_m ? 1 : 0;
    launch_rhs_v(s, c->g, r);
  }
  // fill_m_stochastic / fill_mass_stochastic: new noise for this stage and
  // the next step's predictor
  const uint64_t seed = ip->noise_seed, st = (uint64_t)istep;
  if (noise_m)
    ICHK(ip->noise ? lmm_fill_m_stochastic(c, (uint32_t)(seed + 2 * st))
                   : lmm_fill_m_stochastic_device(c, seed, 2 * st));
  if (noise_mass)
    ICHK(ip->noise ? lmm_fill_mass_stochastic(c, (uint32_t)(seed + 2 * st + 1))
                   : lmm_fill_mass_stochastic_device(c, seed, 2 * st + 1));
  ICHK(mass_and_rhs_p());
  launch_v_sub_mt(s, c->g, rv.d, u, rf.d, inv_dt);
  for (int d = 0; d < dim; d++) launch_fill<double>(s, m_d.d[d], N, 0.0);
  ICHK(lmm_diffusive_m_fluxdiv(c, m_d.h, umac, eta, eta_ed, kappa));
  launch_v_add_half(s, c->g, rv.d, m_d.d);
  // the corrector stops on the predictor's preconditioned right-hand side
  // (:480-486, :946)
  const double abs_in = c->prm.gmres_abs_tol;
  c->prm.gmres_abs_tol = std::max(abs_in, npr * c->prm.gmres_rel_tol);
  rc = solve(&info[1], nullptr);
  c->prm.gmres_abs_tol = abs_in;
  if (rc) return rc;
#undef ICHK
  c->defer_sync = false;
  return finish(c);
}

// advance_timestep_overdamped.f90:60-700 on a periodic box: the midpoint
// scheme with theta_alpha = 0, the full eta in the solves and the explicit
// viscous term, centred advection; the same temporaries as the inertial step.
int lmm_advance_timestep_overdamped(lmm_ctx *c, lmm_field *umac[3], lmm_field *rho_old, lmm_field *rho_new,
                                    lmm_field *rhotot_old, lmm_field *rhotot_new, lmm_field *gradp_baro[3],
                                    lmm_field *pi, lmm_field *eta, lmm_field *eta_ed[3], lmm_field *kappa,
                                    lmm_field *Temp, lmm_field *Temp_ed[3], lmm_field *diff_mass_fluxdiv,
                                    lmm_field *stoch_mass_fluxdiv, double dt, double time, int64_t istep,
                                    const lmm_inertial_params *ip, lmm_gmres_info info[2]) {
  const char *what = "advance_timestep_overdamped";
  if (!c) return LMM_ERR_ARG;
  if (!ip) return set_err(c, LMM_ERR_ARG, "%s: no parameters", what);
  const int alg = ip->algorithm_type;
  if (alg != 1 && alg != 2)
    return set_err(c, LMM_ERR_ARG, "%s: algorithm_type %d (the midpoint scheme takes 1 or 2)", what, alg);
  int rc;
  if ((rc = check_step(c, what, umac, rho_old, rho_new, rhotot_old, rhotot_new, gradp_baro, pi, eta, eta_ed, kappa,
                       Temp, Temp_ed, diff_mass_fluxdiv, stoch_mass_fluxdiv, dt, ip, alg)))
    return rc;
  const lmm_species_params &sp = c->species;
  const int ns = sp.nspecies, dim = c->g.dim;
  const int baro = sp.barodiffusion_type;
  const bool noise_m = ip->variance_coef_mom != 0.0, noise_mass = sp.variance_coef_mass != 0.0;
  const long N = c->g.N;
  StepWork W;
  if ((rc = step_work(c, what, W))) return rc;
  StepVec &rv = W.rv, &dumac = W.dumac, &gradpi = W.gradpi, &rf = W.rf, &rho_fc = W.rho_fc,
          &flux_total = W.flux_total;
  lmm_field &bp_f = W.bp, &dpi_f = W.dpi, &divu_f = W.divu;
  double *u[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < dim; d++) u[d] = umac[d]->d;
  hipStream_t s = c->stream;
  // weights(1:algorithm_type) = (1, 0), then 1/sqrt(2) each for the
  // corrector of algorithm_type 2 (:115-116, :459-461)
  double weights[8] = {1.0, 0, 0, 0, 0, 0, 0, 0};
  RhobarInv rbi{};
  for (int q = 0; q < ns; q++) rbi.v[q] = -1.0 / sp.rhobar[q];
  lmm_field *const *gb = baro ? gradp_baro : nullptr;
  AdvArgs adv{};
  for (int d = 0; d < dim; d++) {
    adv.u[d] = u[d];
    adv.s[d] = rho_fc.d[d];
  }
  adv.dxinv = 1.0 / c->g.dx;
  const bool grav = ip->grav[0] != 0.0 || ip->grav[1] != 0.0 || (dim == 3 && ip->grav[2] != 0.0);
#define ICHK(x)                          \
  do {                                   \
    if ((rc = (x)) != LMM_OK) return rc; \
  } while (0)
  lmm_gmres_info loc[2];
  if (!info) info = loc;
  memset(info, 0, 2 * sizeof(lmm_gmres_info));
  struct DeferSync {
    lmm_ctx *c;
    explicit DeferSync(lmm_ctx *x) : c(x) { c->defer_sync = true; }
    ~DeferSync() { c->defer_sync = false; }
  } defer(c);

  // one Stokes stage: gmres_rhs_v = -grad pi + div(Sigma) + rho g + A_0 v,
  // gmres_rhs_p = div v - S from the mass fluxes at rho_s, then the solve
  // and v += dv, pi += dpi (:166-336, :420-582)
  auto stage = [&](lmm_field *rho_s, double dt_s, lmm_gmres_info *inf, double *npr) -> int {
    ICHK(lmm_compute_grad(c, pi, gradpi.h));
    if (baro == 2)
      for (int d = 0; d < dim; d++)
        hipMemcpyAsync(gradp_baro[d]->d, gradpi.d[d], sizeof(double) * N, hipMemcpyDeviceToDevice, s);
    launch_zero_minus(s, c->g, rv.d, gradpi.d);
    if (noise_m)
      ICHK(lmm_stochastic_m_fluxdiv(c, rv.h, eta, eta_ed, Temp, Temp_ed, dt_s, weights, ip->variance_coef_mom,
                                    sp.k_B));
    if (grav) launch_grav(s, c->g, ip->grav, rv.d, rf.d, rf.d);
    ICHK(lmm_compute_mass_fluxdiv(c, rho_s, gb, diff_mass_fluxdiv, stoch_mass_fluxdiv, Temp, flux_total.h, dt_s,
                                  time, weights));
    ICHK(lmm_diffusive_m_fluxdiv(c, rv.h, umac, eta, eta_ed, kappa));
    ICHK(lmm_compute_div(c, umac, &divu_f, 0));
    launch_rhs_p(s, c->g, ns, bp_f.d, diff_mass_fluxdiv->d, noise_mass ? stoch_mass_fluxdiv->d : nullptr, divu_f.d,
                 rbi);
    for (int d = 0; d < dim; d++) launch_fill<double>(s, dumac.d[d], N, 0.0);
    launch_fill<double>(s, dpi_f.d, N, 0.0);
    ICHK(lmm_gmres(c, rv.h, &bp_f, dumac.h, &dpi_f, rf.h, eta, eta_ed, kappa, 0.0, npr, inf));
    for (int d = 0; d < dim; d++) launch_add(s, u[d], dumac.d[d], N);
    launch_add(s, pi->d, dpi_f.d, N);
    return LMM_OK;
  };
  // rho_new = rho_old + dt_s (F + A) with A from rho_fc, then rhotot, the
  // concentration round trip and the new faces (:340-418, :586-645)
  auto scalar = [&](double dt_s) {
    RhoArgs r{};
    r.diff = diff_mass_fluxdiv->d;
    r.stoch = noise_mass ? stoch_mass_fluxdiv->d : nullptr;
    r.rho_old = rho_old->d;
    r.rho_new = rho_new->d;
    r.dt = dt_s;
    launch_adv_s(s, c->g, ns, adv, r, 1);
    launch_rhotot_conc(s, c->g, ns, rho_new->d, rhotot_new->d);
  };
  auto eta_kappa = [&]() -> int {
    launch_eta(s, c->g, eta->d, kappa->d, rhotot_new->d, ip->visc_coef, ip->eta_model);
    return dim == 2 ? lmm_average_cc_to_node(c, eta, eta_ed[0]) : lmm_average_cc_to_edge(c, eta, eta_ed);
  };

  ICHK(lmm_average_cc_to_face_comps(c, rho_old, rho_fc.h, 0, ns));
  ICHK(lmm_average_cc_to_face(c, rhotot_old, rf.h));
  // new noise, except at the first step after initialisation or a restart
  // (:158-165)
  if (istep != 1 && istep != (int64_t)ip->restart + 1) {
    const uint64_t seed = ip->noise_seed, st = (uint64_t)istep;
    if (noise_mass)
      ICHK(ip->noise ? lmm_fill_mass_stochastic(c, (uint32_t)(seed + 2 * st + 1))
                     : lmm_fill_mass_stochastic_device(c, seed, 2 * st + 1));
    if (noise_m)
      ICHK(ip->noise ? lmm_fill_m_stochastic(c, (uint32_t)(seed + 2 * st))
                     : lmm_fill_m_stochastic_device(c, seed, 2 * st));
  }
  const double dt_pred = alg == 1 ? dt : 0.5 * dt;

  // ---- steps 1-2: predictor fluxes and Stokes solve (:166-336)
  double npr = 0.0;
  ICHK(stage(rho_old, dt_pred, &info[0], &npr));
  const double abs_in = c->prm.gmres_abs_tol;

  // ---- step 3: scalar predictor, half a step (:340-418)
  scalar(0.5 * dt);
  ICHK(lmm_average_cc_to_face_comps(c, rho_new, rho_fc.h, 0, ns));
  ICHK(lmm_average_cc_to_face(c, rhotot_new, rf.h));
  ICHK(eta_kappa());

  // ---- steps 4-5: corrector fluxes and Stokes solve (:420-582)
  if (alg == 2) weights[0] = weights[1] = 1.0 / sqrt(2.0);
  c->prm.gmres_abs_tol = std::max(abs_in, npr * c->prm.gmres_rel_tol);
  rc = stage(rho_new, dt, &info[1], nullptr);
  c->prm.gmres_abs_tol = abs_in;
  if (rc) return rc;

  // ---- step 6: midpoint scalar corrector (:586-648)
  scalar(dt);
  ICHK(eta_kappa());
#undef ICHK
  c->defer_sync = false;
  return finish(c);
}

// ---------------- allocation-free species diffusion ----------------

}  // extern "C"

// Device scratch of lmm_diffusive_mass_fluxdiv_noalloc: the face fluxes of
// every species (dim*nspecies*N doubles), at most NOALLOC_MAX doubles
// (32 MiB).  Each GPU holds its own copy until the process ends; host code
// uses only the address hipGetSymbolAddress returns on that device.
static constexpr long NOALLOC_MAX = 1L << 22;
static __device__ __align__(256) double scratch[NOALLOC_MAX];
static std::mutex g_scratch_mu;
static std::map<int, double *> g_scratch_ptr;  // device -> address of its scratch

// The scratch address on the current device, looked up once per device.
static int scratch_address(lmm_ctx *c, double **out) {
  int dev = 0;
  HIPCHK(c, hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) {
    *out = it->second;
    return LMM_OK;
  }
  double *p = nullptr;
  HIPCHK(c, hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[dev] = p;
  *out = p;
  return LMM_OK;
}

extern "C" {

int lmm_diffusive_mass_fluxdiv_noalloc(lmm_ctx *c, int ns, lmm_field *rho, lmm_field *rhotot, lmm_field *molarconc,
                                       lmm_field *rhoWchi, lmm_field *Gama, lmm_field *diff_fluxdiv, lmm_field *Temp,
                                       lmm_field *zeta_by_Temp, lmm_field *const gradp_baro[3],
                                       lmm_field *flux_total[3], int correct_flux) {
  if (!c) return LMM_ERR_ARG;
  if (Temp || zeta_by_Temp || (gradp_baro && gradp_baro[0]))
    return set_err(c, LMM_ERR_UNSUPPORTED, "diffusive_mass_fluxdiv_noalloc: isothermal, no barodiffusion");
  (void)rhotot;
  DiffArgs a;
  if (int rc = diff_args(c, ns, rho, molarconc, rhoWchi, Gama, diff_fluxdiv, nullptr, nullptr, nullptr, flux_total,
                         correct_flux, &a))
    return rc;
  if (a.baro) return set_err(c, LMM_ERR_UNSUPPORTED, "diffusive_mass_fluxdiv_noalloc: isothermal, no barodiffusion");
  const long need = (long)c->g.dim * ns * c->g.N;
  if (need > NOALLOC_MAX)
    return set_err(c, LMM_ERR_ARG,
                   "diffusive_mass_fluxdiv_noalloc: dim*nspecies*cells = %ld exceeds the %ld-value device scratch",
                   need, NOALLOC_MAX);
  double *flux = nullptr;
  int rc = scratch_address(c, &flux);
  if (rc) return rc;
  // every flux value is written before the divergence reads it: nothing to clear
  return sp_status(c, sp_diffusive(c->stream, c->g, a.sp, nullptr, molarconc->d, rhoWchi->d, Gama->d, nullptr, nullptr,
                                   nullptr, nullptr, flux, a.ft, diff_fluxdiv->d),
                   "diffusive_mass_fluxdiv_noalloc");
}

// ---------------- learned coarse-grid correction ----------------

void lmm_cnn_config_default(lmm_cnn_config *cfg) {
  if (!cfg) return;
  cfg->cin = 8;
  cfg->c1 = 16;
  cfg->c2 = 32;
  cfg->c3 = 64;
  cfg->cout = 4;
  cfg->level = 3;
  cfg->sample_steps = 3;
}

int lmm_cnn_create(lmm_ctx *c, const lmm_cnn_config *cfg, uint64_t seed, lmm_cnn **out) {
  if (!c || !out) return LMM_ERR_ARG;
  *out = nullptr;
  lmm_cnn_config d;
  lmm_cnn_config_default(&d);
  if (!cfg) cfg = &d;
  CnnConfig cc;
  cc.cin = cfg->cin;
  cc.c1 = cfg->c1;
  cc.c2 = cfg->c2;
  cc.c3 = cfg->c3;
  cc.cout = cfg->cout;
  std::string err;
  CnnNet *net = cnn_new(cc, seed, c->stream, &err);
  if (!net) return set_err(c, LMM_ERR_ARG, "%s", err.c_str());
  lmm_cnn *nn = new lmm_cnn();
  nn->ctx = c;
  nn->net = net;
  nn->cfg = *cfg;
  *out = nn;
  return LMM_OK;
}

int lmm_cnn_destroy(lmm_cnn *nn) {
  if (!nn) return LMM_ERR_ARG;
  if (nn->ctx && nn->ctx->cnn == nn) nn->ctx->cnn = nullptr;
  cnn_free(nn->net);
  dfree(nn->xin);
  dfree(nn->yout);
  dfree(nn->ss);
  delete nn;
  return LMM_OK;
}

int lmm_cnn_num_params(const lmm_cnn *nn, int64_t *n) {
  if (!nn || !n) return LMM_ERR_ARG;
  *n = cnn_nparams(nn->net);
  return LMM_OK;
}

int lmm_cnn_get_params(lmm_cnn *nn, float *host) {
  if (!nn || !host) return LMM_ERR_ARG;
  return cnn_get_params(nn->net, host) ? set_err(nn->ctx, LMM_ERR_DEVICE, "cnn_get_params") : LMM_OK;
}

int lmm_cnn_set_params(lmm_cnn *nn, const float *host) {
  if (!nn || !host) return LMM_ERR_ARG;
  return cnn_set_params(nn->net, host) ? set_err(nn->ctx, LMM_ERR_DEVICE, "cnn_set_params") : LMM_OK;
}

int lmm_cnn_forward_host(lmm_cnn *nn, int nb, const int n[3], const float *x, float *y) {
  if (!nn || !n || !x || !y) return LMM_ERR_ARG;
  int rc = cnn_forward_host(nn->net, nb, n, x, y);
  if (rc) return set_err(nn->ctx, rc, "cnn forward: grid extents must be multiples of 4 and nb*(n/4)^3 of 16");
  return finish(nn->ctx);
}

int lmm_cnn_loss_grad_host(lmm_cnn *nn, int nb, const int n[3], const float *x, const float *t, double *loss,
                           float *grad) {
  if (!nn || !n || !x || !t || !loss || !grad) return LMM_ERR_ARG;
  int rc = cnn_loss_grad_host(nn->net, nb, n, x, t, loss, grad);
  if (rc) return set_err(nn->ctx, rc, "cnn loss/gradient: grid extents must be multiples of 4 and nb*(n/4)^3 of 16");
  return finish(nn->ctx);
}

// the network's level must exist in both multigrids and its grid suit the
// network (3D; 8 input channels: residual (u, p), log beta, log theta*alpha;
// 4 output channels: error (u, p))
static int cnn_level_ok(lmm_ctx *c, lmm_cnn *nn) {
  const lmm_cnn_config &f = nn->cfg;
  if (c->g.dim != 3) return set_err(c, LMM_ERR_UNSUPPORTED, "the learned correction is 3D only");
  if (f.cin != 8 || f.cout != 4) return set_err(c, LMM_ERR_ARG, "the Stokes correction needs cin 8 and cout 4");
  if (f.sample_steps < 0 || f.sample_steps > 16) return set_err(c, LMM_ERR_ARG, "cnn sample_steps outside 0..16");
  const int nl = std::min(c->nl_s, c->nl_c);
  if (f.level < 1 || f.level >= nl) return set_err(c, LMM_ERR_ARG, "cnn level %d outside 1..%d", f.level, nl - 1);
  Grid gc = c->g;
  for (int n = 0; n < f.level; n++) gc = coarsen(gc);
  for (int d = 0; d < 3; d++)
    if (gc.n[d] % 4 != 0 || gc.n[d] < 4) return set_err(c, LMM_ERR_ARG, "cnn level grid must be a multiple of 4");
  if (((long)(gc.n[0] / 4) * (gc.n[1] / 4) * (gc.n[2] / 4)) % 16 != 0)
    return set_err(c, LMM_ERR_ARG, "cnn level grid: (n/4)^3 must be a multiple of 16");
  if (!nn->xin) {
    nn->xin = dalloc<uint16_t>(c->stream, gc.N * 8);
    nn->yout = dalloc<float>(c->stream, gc.N * CNN_OUT_STRIDE);
    nn->ss = dalloc<double>(c->stream, 4);  // sums of squares (u, p) and the prediction's pressure sum
    if (!nn->xin || !nn->yout || !nn->ss) return set_err(c, LMM_ERR_NOMEM, "cnn application buffers");
  }
  return LMM_OK;
}

int lmm_ctx_set_cnn(lmm_ctx *c, lmm_cnn *nn) {
  if (!c) return LMM_ERR_ARG;
  if (!nn) {
    c->cnn = nullptr;
    return LMM_OK;
  }
  if (nn->ctx != c) return set_err(c, LMM_ERR_ARG, "the network belongs to another context");
  int rc = cnn_level_ok(c, nn);
  if (rc) return rc;
  c->cnn = nn;
  return LMM_OK;
}

// One sample of the preconditioner's error on the coefficients the levels
// hold: phi* = M(white noise) (a Stokes vector of the solutions' kind),
// then `steps` times phi* <- phi* - M(A phi*) (the preconditioned error
// propagation, which leaves the modes M handles worst: the ones the flexible
// GMRES iterates on); b = A phi*, z = M(b) (FP32 projection preconditioner,
// no learned stage).  Leaves phi* in ps, b in c->b32 and z in z.
static void precon_sample(lmm_ctx *c, double theta_alpha, uint64_t seed, uint64_t q, int steps, const SVT<float> &ps,
                          const SVT<float> &z) {
  hipStream_t s = c->stream;
  const long N = c->g.N;
  auto &S = c->s32;
  auto &C = c->c32;
  const int visc = c->prm.visc_type;
  SVT<float> b = svt_of<float>(c, c->b32);
  launch_fill_normal_philox(s, c->r, 4 * N, 0, seed, q);
  launch_convert<float, double>(s, c->b32, c->r, 4 * N);
  apply_precon_t<float>(c, S, C, b.u, b.p, ps.u, ps.p, c->mac_rhs32, C[0].phi, theta_alpha);
  for (int k = 0; k < steps; k++) {
    launch_apply_matrix<float>(s, c->g, visc, ps.u, ps.p, b.u, b.p, S[0].K());
    apply_precon_t<float>(c, S, C, b.u, b.p, z.u, z.p, c->mac_rhs32, C[0].phi, theta_alpha);
    for (int d = 0; d < 3; d++) launch_sub_f32(s, ps.u[d], ps.u[d], z.u[d], N);
    launch_sub_f32(s, ps.p, ps.p, z.p, N);
  }
  launch_apply_matrix<float>(s, c->g, visc, ps.u, ps.p, b.u, b.p, S[0].K());
  apply_precon_t<float>(c, S, C, b.u, b.p, z.u, z.p, c->mac_rhs32, C[0].phi, theta_alpha);
}

// common set-up of pair generation and the ablation: coefficients staged
// (FP64 and FP32 hierarchies, both multigrids), work arrays, checks
static int learned_setup(lmm_ctx *c, lmm_cnn *nn, lmm_field *const alpha_fc[3], lmm_field *beta,
                         lmm_field *const beta_ed[3], lmm_field *gamma, double theta_alpha) {
  int rc;
  if ((rc = check_coefs(c, alpha_fc, beta, beta_ed, gamma))) return rc;
  if ((rc = cnn_level_ok(c, nn))) return rc;
  if ((rc = ensure_lvl32(c))) return rc;
  if (!c->gen32) {
    c->gen32 = dalloc<float>(c->stream, 8 * c->g.N);
    if (!c->gen32) return set_err(c, LMM_ERR_NOMEM, "sample work arrays");
  }
  prepare_coefs(c, coef_in(c, alpha_fc, beta, beta_ed, gamma), theta_alpha, 1.0, true, true);
  c->learned_on = false;
  return LMM_OK;
}

int lmm_cnn_generate_pairs(lmm_cnn *nn, lmm_field *const alpha_fc[3], lmm_field *beta, lmm_field *const beta_ed[3],
                           lmm_field *gamma, double theta_alpha, int npairs, uint64_t seed) {
  if (!nn || npairs < 1) return LMM_ERR_ARG;
  lmm_ctx *c = nn->ctx;
  TuningScope ts(&c->tune);
  int rc;
  if ((rc = learned_setup(c, nn, alpha_fc, beta, beta_ed, gamma, theta_alpha))) return rc;
  const int lc = nn->cfg.level;
  auto &S = c->s32;
  auto &C = c->c32;
  const int gn[3] = {S[lc].g.n[0], S[lc].g.n[1], S[lc].g.n[2]};
  const long have = cnn_dataset_count(nn->net);
  if ((rc = cnn_dataset_reserve(nn->net, have + npairs, gn)))
    return set_err(c, rc, "training pairs: out of memory or a grid other than the stored pairs'");
  hipStream_t s = c->stream;
  const long N = c->g.N;
  const SVT<float> ps = svt_of<float>(c, c->gen32), z = svt_of<float>(c, c->gen32 + 4 * N);
  const SVT<float> b = svt_of<float>(c, c->b32);
  float *ru[3] = {S[0].res[0], S[0].res[1], S[0].res[2]};
  for (int q = 0; q < npairs; q++) {
    precon_sample(c, theta_alpha, seed, (uint64_t)q, nn->cfg.sample_steps, ps, z);
    // residual r = b - A z, error e = phi* - z (into z)
    launch_apply_matrix<float>(s, c->g, c->prm.visc_type, z.u, z.p, ru, C[0].res, S[0].K());
    for (int d = 0; d < 3; d++) launch_sub_f32(s, ru[d], b.u[d], ru[d], N);
    launch_sub_f32(s, C[0].res, b.p, C[0].res, N);
    for (int d = 0; d < 3; d++) launch_sub_f32(s, z.u[d], ps.u[d], z.u[d], N);
    launch_sub_f32(s, z.p, ps.p, z.p, N);
    // the residual with the multigrid restrictions, the error with face / cell means
    launch_stag_restrict<float>(s, S[1].g, ru, S[1].rhs, 0);
    launch_cc_restrict<float>(s, C[1].g, C[0].res, C[1].rhs);
    launch_stag_restrict<float>(s, S[1].g, z.u, S[1].phi, 1);
    launch_cc_restrict<float>(s, C[1].g, z.p, C[1].phi);
    for (int n = 1; n < lc; n++) {
      launch_stag_restrict<float>(s, S[n + 1].g, S[n].rhs, S[n + 1].rhs, 0);
      launch_cc_restrict<float>(s, C[n + 1].g, C[n].rhs, C[n + 1].rhs);
      launch_stag_restrict<float>(s, S[n + 1].g, S[n].phi, S[n + 1].phi, 1);
      launch_cc_restrict<float>(s, C[n + 1].g, C[n].phi, C[n + 1].phi);
    }
    const long slot = have + q;
    launch_stokes_ss(s, S[lc].g, S[lc].rhs, C[lc].rhs, c->partial, nn->ss);
    launch_cnn_pack(s, S[lc].g, S[lc].rhs, C[lc].rhs, S[lc].beta, S[lc].alpha, nn->ss, cnn_dataset_x(nn->net, slot));
    launch_cnn_pack_target(s, S[lc].g, S[lc].phi, C[lc].phi, nn->ss, cnn_dataset_t(nn->net, slot));
    cnn_dataset_push(nn->net);
  }
  return finish(c);
}

int lmm_cnn_error_reduction(lmm_cnn *nn, lmm_field *const alpha_fc[3], lmm_field *beta, lmm_field *const beta_ed[3],
                            lmm_field *gamma, double theta_alpha, int npairs, uint64_t seed, double *err_plain,
                            double *err_learned) {
  if (!nn || npairs < 1 || !err_plain || !err_learned) return LMM_ERR_ARG;
  lmm_ctx *c = nn->ctx;
  TuningScope ts(&c->tune);
  int rc;
  if ((rc = learned_setup(c, nn, alpha_fc, beta, beta_ed, gamma, theta_alpha))) return rc;
  hipStream_t s = c->stream;
  const long N = c->g.N;
  const SVT<float> ps = svt_of<float>(c, c->gen32), z = svt_of<float>(c, c->gen32 + 4 * N);
  const SVT<float> b = svt_of<float>(c, c->b32);
  auto &S = c->s32;
  auto &C = c->c32;
  float *du[3] = {S[0].phi[0], S[0].phi[1], S[0].phi[2]};
  float *dp = C[0].rhs;
  lmm_cnn *saved = c->cnn;
  c->cnn = nn;
  double sp = 0.0, sl = 0.0;
  double *slot = c->scal + 4300;
  // |phi* - z| / |phi*| over velocity and pressure (both mean-free in pressure)
  auto err2 = [&](double *out) {
    for (int d = 0; d < 3; d++) launch_sub_f32(s, du[d], ps.u[d], z.u[d], N);
    launch_sub_f32(s, dp, ps.p, z.p, N);
    launch_stokes_ss(s, c->g, du, dp, c->partial, out);
  };
  for (int q = 0; q < npairs; q++) {
    precon_sample(c, theta_alpha, seed, (uint64_t)q, nn->cfg.sample_steps, ps, z);
    launch_stokes_ss(s, c->g, ps.u, ps.p, c->partial, slot);
    err2(slot + 2);
    learned_stage(c, b.u, b.p, z.u, z.p);
    err2(slot + 4);
    double h[6];
    fetch(c, slot, 6, h);
    const double n0 = h[0] + h[1];
    sp += std::sqrt((h[2] + h[3]) / n0);
    sl += std::sqrt((h[4] + h[5]) / n0);
  }
  c->cnn = saved;
  c->learned_applied = 0;
  *err_plain = sp / npairs;
  *err_learned = sl / npairs;
  return finish(c);
}

int lmm_cnn_num_pairs(const lmm_cnn *nn, int64_t *count) {
  if (!nn || !count) return LMM_ERR_ARG;
  *count = cnn_dataset_count(nn->net);
  return LMM_OK;
}

int lmm_cnn_train(lmm_cnn *nn, int steps, int batch, double lr, uint64_t seed, int64_t held_out, float *loss_history) {
  if (!nn || held_out < 0) return LMM_ERR_ARG;
  lmm_ctx *c = nn->ctx;
  const long cnt = cnn_dataset_count(nn->net);
  if (held_out >= cnt) return set_err(c, LMM_ERR_ARG, "no training pairs left after %lld held out", (long long)held_out);
  int rc = cnn_train(nn->net, steps, batch, lr, seed, cnt - held_out, loss_history);
  if (rc) return set_err(c, rc, "cnn training: bad steps/batch or out of memory");
  return finish(c);
}

int lmm_cnn_evaluate(lmm_cnn *nn, int64_t first, int64_t last, double *rel_err) {
  if (!nn || !rel_err) return LMM_ERR_ARG;
  int rc = cnn_eval(nn->net, (long)first, (long)last, rel_err);
  if (rc) return set_err(nn->ctx, rc, "cnn evaluation: pair range");
  return finish(nn->ctx);
}

}  // extern "C"
