// sparc_gpu.hip -- context, set-up and C ABI of the MI355X stepping engine.
// See include/sparc_gpu.h for the entry points and the reference seams they
// replace.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/sparc_gpu.h"
#include "kernels.h"

// Exact segmented z solve (k_zsolve_b: input and sweep results in LDS) for the
// epilogues that solve their own field (EPI_P3D, EPI_P2D, EPI_C2: bits of
// kZbEpis), when 2 x 64 x nz values fit one block's LDS; otherwise k_zsolve.
static constexpr int kZbEpis =
    (1 << sparc::EPI_P3D) | (1 << sparc::EPI_P2D) | (1 << sparc::EPI_C2);
// segments per column: k_zsolve_b 16 in FP32, 8 in FP64 (at 1024 threads its
// epilogues spill); k_zsolve with ud in LDS 8
static constexpr int kZbNseg = 16, kZbNseg64 = 8, kZsUldsNseg = 8;

using namespace sparc;

// single-precision literal of initialize.f90:38 (no -r8 in makefile:27)
static const double PI_R = (double)(float)3.14159265358979;

struct sparc_ctx {
  sparc_config cfg{};
  int fp32 = 0;
  size_t ts = 8;
  hipStream_t stream = nullptr;
  hipEvent_t xev = nullptr;  // orders this stream against another context's (cross-context calls)
  // captured step bodies (five stages + update), without / with the source plane
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  int device = 0;
  int lds_max = 65536;  // dynamic LDS bytes one block may take
  int cus = 256;        // compute units of the device
  long N = 0, nxy = 0, NP = 0;
  int nzpml = 0;
  void* a[SPARC_MAXFIELDS] = {};
  void* scr[SPARC_MAXFIELDS] = {};
  void* alo[SPARC_MAXFIELDS] = {};  // 16-bit low state words (compensated FP32 state)
  int nf = SPARC_NFIELDS;           // dimfive: 5 (option 2), 6 (displacement), 8 (option 8)
  bool comp = false;
  void* pmlvars[6] = {};
  void* scrpml[6] = {};
  int npv = 2;         // psi fields: 2, or 6 for option 10
  void* mw[12] = {};   // option 10/18 work arrays (fluxes, b, derivative pieces)
  void* dmul2 = nullptr;  // option 18: x-multipliers of the 2-D damping
  void* bg[SPARC_BG_COUNT] = {};
  int bg3d[SPARC_BG_COUNT] = {};
  bool bgset[SPARC_BG_COUNT] = {};
  void* prof[SPARC_PROF_COUNT] = {};
  bool profset[SPARC_PROF_COUNT] = {};
  std::vector<double> unstretch_h, z_h;
  void* planes_buf = nullptr;  // device storage behind src_planes
  unsigned char* scratch = nullptr;  // device address of this device's sample scratch
  void* W[8] = {};
  void* spec[3] = {};
  void* twx = nullptr;
  void* twy = nullptr;
  void* dmulx = nullptr;
  void* dmuly = nullptr;
  void* fmulx = nullptr;
  void* fmuly = nullptr;
  void* rates = nullptr;    // damping rates, transposed: [y + ny*m]
  void* rates_o = nullptr;  // damping rates as given: [m + (nx/2+1)*y]
  bool damping_set = false, filter_set = false;
  // option 4 on 256 x 256 FP32 planes: the damping by k_damp2d, one pass per
  // plane, instead of the forward x, y and inverse x passes (cfg.plane_damping)
  bool damp2d = false;
  void* tri[4] = {};  // per ibc 0..3: lowr | invb | gam | fcum, gcum (2 partitions), nz each
  double tc[14] = {};
  // per ibc and precision (0 fp64, 1 fp32): [kc0, kc1) rows with constant factors
  int kc[4][2] = {};
  double kcv[4][3] = {};
  FFTPlan px{}, py{};
  double betas[5] = {};
  // source
  int src_kind = SPARC_SRC_NONE, src_nt = 0, src_active_until = -1;
  void* src_mask = nullptr;
  int32_t* src_planes = nullptr;  // per-pixel 0-based source planes (MHD), null = uniform
  void* src_dw = nullptr;         // delta_width_2d per pixel
  std::vector<double> src_series;
  void* forcing = nullptr;
  // kernels
  int nslots = 0;
  void* snap = nullptr;
  double snap_scale = 1.0;
  double* kern[SPARC_KERN_COUNT] = {};
  double* grav = nullptr;
  int nckpt = 0;
  void* ckpt = nullptr;
  std::vector<void*> allocs;
  size_t bytes = 0;
  std::string err;
  // instrumentation
  bool profiling = false;
  struct Rec {
    int cls;
    double bytes;
    hipEvent_t e0, e1;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  double pms[SPARC_KC_COUNT] = {}, pbytes[SPARC_KC_COUNT] = {};
  int64_t plaunch[SPARC_KC_COUNT] = {};
};

// Event bracket around one launch (only while profiling).
struct ProfScope {
  sparc_ctx* c;
  int idx = -1;
  static hipEvent_t ev(sparc_ctx* c) {
    if (!c->pool.empty()) {
      hipEvent_t e = c->pool.back();
      c->pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  ProfScope(sparc_ctx* cc, int cls, double bytes) : c(cc) {
    if (!c->profiling) return;
    sparc_ctx::Rec r{cls, bytes, ev(c), ev(c)};
    if (!r.e0 || !r.e1) return;
    (void)hipEventRecord(r.e0, c->stream);
    c->recs.push_back(r);
    idx = (int)c->recs.size() - 1;
  }
  ~ProfScope() {
    if (idx >= 0) (void)hipEventRecord(c->recs[idx].e1, c->stream);
  }
};

static int fail(sparc_ctx* c, int code, const char* fmt, ...) {
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

#define HIPCHK(ctx, call)                                                             \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(ctx, SPARC_EHIP, "%s failed: %s", #call, hipGetErrorString(e_));    \
  } while (0)

#define LAUNCHCHK(ctx)                                                                \
  do {                                                                                \
    hipError_t e_ = hipGetLastError();                                                \
    if (e_ != hipSuccess)                                                             \
      return fail(ctx, SPARC_EHIP, "kernel launch failed at %s:%d: %s", __FILE__,     \
                  __LINE__, hipGetErrorString(e_));                                   \
  } while (0)

// Graph capture (the first replayed step of a context) must not overlap a
// device allocation or free made by another thread for another context: HIP
// refuses such calls while a stream captures.  Captures are short (one step's
// launches), so one process-wide lock covers capture, hipMalloc and hipFree.
static std::mutex g_capture_mu;

// Large allocations start on 2 MiB boundaries, so the same index of two
// field arrays shares its low address bits.  The n-th allocation is moved
// (n mod 16) * kStagger bytes past its start (a multiple of 256 keeps every
// alignment the kernels use); allocs keeps the pointer hipMalloc gave.
static constexpr size_t kStagger = 2304;
static int dalloc(sparc_ctx* c, void** p, size_t bytes) {
  static_assert(kStagger % 256 == 0, "stagger keeps 256-byte alignment");
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  const size_t off = (size_t)(c->allocs.size() % 16) * kStagger;
  void* raw = nullptr;
  hipError_t e;
  {
    std::lock_guard<std::mutex> lk(g_capture_mu);
    e = hipMalloc(&raw, bytes + off);
  }
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(c, SPARC_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  c->allocs.push_back(raw);
  c->bytes += bytes + off;
  *p = (char*)raw + off;
  return SPARC_OK;
}

static bool make_plan(int n, FFTPlan* p) {
  p->n = n;
  p->nrad = 0;
  int m = n;
  while (m % 4 == 0 && p->nrad < 20) {
    p->rad[p->nrad++] = 4;
    m /= 4;
  }
  while (m % 2 == 0 && p->nrad < 20) {
    p->rad[p->nrad++] = 2;
    m /= 2;
  }
  while (m % 3 == 0 && p->nrad < 20) {
    p->rad[p->nrad++] = 3;
    m /= 3;
  }
  return m == 1;
}

// host double -> device T
static int upload(sparc_ctx* c, void* dst, const double* h, size_t n) {
  if (c->fp32) {
    std::vector<float> t(n);
    for (size_t i = 0; i < n; ++i) t[i] = (float)h[i];
    HIPCHK(c, hipMemcpyAsync(dst, t.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // t is released on return
  } else {
    HIPCHK(c, hipMemcpyAsync(dst, h, n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return SPARC_OK;
}

static int download(sparc_ctx* c, double* h, const void* src, size_t n, double scale) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->fp32) {
    std::vector<float> t(n);
    HIPCHK(c, hipMemcpyAsync(t.data(), src, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; ++i) h[i] = (double)t[i] * scale;
  } else {
    HIPCHK(c, hipMemcpyAsync(h, src, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (scale != 1.0)
      for (size_t i = 0; i < n; ++i) h[i] *= scale;
  }
  return SPARC_OK;
}

static size_t spec_elems(const sparc_config* cfg) {
  // complex elements of the x-spectrum of nz/2 plane pairs
  return (size_t)((cfg->nz + 1) / 2) * cfg->nx * cfg->ny;
}

static bool is_mhd(int option) {
  return option == SPARC_OPT_MHD_PML_3D_DISPL || option == SPARC_OPT_MHD_PML_2D_DISPL ||
         option == SPARC_OPT_MHD_PML_3D;
}
// dimfive (initialize.f90:310-312)
static int nfields_of(int option) {
  return option == SPARC_OPT_QUIET_PML_3D ? 5 : (option == SPARC_OPT_MHD_PML_3D ? 8 : 6);
}
static bool is_2d(int option) {
  return option == SPARC_OPT_QUIET_PML_2D_DISPL || option == SPARC_OPT_MHD_PML_2D_DISPL;
}

static int check_config(const sparc_config* cfg, std::string* why) {
  char b[256];
  if (cfg->abi_version != SPARC_ABI_VERSION) {
    snprintf(b, sizeof b, "abi_version %d != %d", cfg->abi_version, SPARC_ABI_VERSION);
    *why = b;
    return SPARC_EINVAL;
  }
  if (cfg->option != SPARC_OPT_QUIET_PML_3D_DISPL && cfg->option != SPARC_OPT_QUIET_PML_2D_DISPL &&
      cfg->option != SPARC_OPT_QUIET_PML_3D && !is_mhd(cfg->option)) {
    snprintf(b, sizeof b, "option %d not supported (2, 4, 8, 10, 17 or 18)", cfg->option);
    *why = b;
    return SPARC_EINVAL;
  }
  if (cfg->precision != SPARC_FP64 && cfg->precision != SPARC_FP32) {
    *why = "precision must be SPARC_FP64 or SPARC_FP32";
    return SPARC_EINVAL;
  }
  if (cfg->nz < 12 || cfg->nz > 4096 || cfg->nx < 2 || cfg->ny < 1) {
    *why = "grid out of range (nz in [12,4096], nx >= 2)";
    return SPARC_EINVAL;
  }
  FFTPlan p;
  if (!make_plan(cfg->nx, &p)) {
    *why = "nx must be 2^a*3^b";
    return SPARC_EINVAL;
  }
  if (is_2d(cfg->option) && cfg->ny != 1) {
    *why = "the 2-D options need ny == 1";
    return SPARC_EINVAL;
  }
  if ((cfg->option == SPARC_OPT_QUIET_PML_3D || cfg->option == SPARC_OPT_MHD_PML_3D) &&
      cfg->mode != SPARC_FORWARD) {
    *why = "options 2 and 8 are forward-only (the kernels need the displacement formulation)";
    return SPARC_EINVAL;
  }
  if (!is_2d(cfg->option)) {
    if (!make_plan(cfg->ny, &p) || cfg->ny < 2) {
      *why = "ny must be 2^a*3^b >= 2 for the 3-D options";
      return SPARC_EINVAL;
    }
    if (cfg->nx % 32 != 0) {
      *why = "the 3-D options need nx % 32 == 0";
      return SPARC_EINVAL;
    }
  }
  if (cfg->npmlbot < 0 || cfg->npmltop < 0 || cfg->npmlbot + cfg->npmltop > cfg->nz / 2) {
    *why = "PML widths out of range";
    return SPARC_EINVAL;
  }
  if (cfg->npmlbot == 1 || cfg->npmltop == 1) {
    *why = "PML width 1 makes the profile length zero (z(npml)-z(1))";
    return SPARC_EINVAL;
  }
  if (cfg->deltat <= 0) {
    *why = "deltat must be > 0";
    return SPARC_EINVAL;
  }
  if (cfg->precision == SPARC_FP64 && cfg->field_scale != 1.0 && cfg->field_scale != 0.0) {
    *why = "field_scale is an FP32 option";
    return SPARC_EINVAL;
  }
  return SPARC_OK;
}

static size_t footprint_bytes(const sparc_config* cfg) {
  size_t ts = cfg->precision == SPARC_FP32 ? 4 : 8;
  size_t N = (size_t)cfg->nx * cfg->ny * cfg->nz;
  size_t nxy = (size_t)cfg->nx * cfg->ny;
  size_t NP = nxy * (size_t)std::max(cfg->npmlbot + cfg->npmltop, 1);
  size_t b = 0;
  const size_t nf = nfields_of(cfg->option);
  b += 2 * nf * N * ts;                     // a, scr
  if (cfg->precision == SPARC_FP32 && cfg->compensated) b += nf * N * 2;  // alo
  const size_t npv = is_mhd(cfg->option) ? 6 : 2;
  b += 2 * npv * NP * ts;                   // pmlvars, scrpml
  if (is_mhd(cfg->option)) b += 12 * N * ts + cfg->nx * ts;  // MHD work, 2-D damping
  b += SPARC_BG_COUNT * N * ts;             // backgrounds (upper bound: all 3-D)
  b += 8 * N * ts;                          // work
  if (!is_2d(cfg->option) && cfg->damp_waves)
    b += 3 * spec_elems(cfg) * 2 * ts;      // damping spectra
  b += (size_t)(cfg->nx + cfg->ny) * 6 * ts + 4 * 7 * cfg->nz * ts + SPARC_PROF_COUNT * cfg->nz * ts;
  b += 2 * (size_t)(cfg->nx / 2 + 1) * cfg->ny * ts + 2 * nxy * ts;  // rates (two layouts)
  b += 128 * 15 * kStagger;  // allocation stagger, up to 128 allocations
  return b;
}

extern "C" int sparc_footprint(const sparc_config* cfg, size_t* bytes) {
  if (!cfg || !bytes) return SPARC_EINVAL;
  std::string why;
  if (check_config(cfg, &why) != SPARC_OK) return SPARC_EINVAL;
  *bytes = footprint_bytes(cfg);
  return SPARC_OK;
}

// MTRIDAG ispecial=0 factors of DBYD2's matrix for one ibc (dbyd2.f90:151-196,
// mtridag.f90:94-99), FP64.
static void tri_factors(int n2, int ibc, std::vector<double>& lowr, std::vector<double>& invb,
                        std::vector<double>& gam) {
  std::vector<double> uppr(n2, 0.0), diag(n2, 1.0);
  lowr.assign(n2, 0.0);
  invb.assign(n2, 0.0);
  gam.assign(n2, 0.0);
  const double alpha = (double)(float)(1. / 3.);
  for (int j = 2; j < n2 - 2; ++j) uppr[j] = alpha;
  const double alpha2 = 3. / 4., gamma2 = 1. / 8., alpha1 = 6.0, beta1 = 3.;
  const double cst = 1. / (alpha2 - beta1 * gamma2);
  const double up1 = (alpha1 * alpha2 - beta1) * cst;
  uppr[0] = (ibc % 2 == 0) ? 0.0 : up1;
  uppr[1] = alpha2;
  uppr[n2 - 2] = gamma2;
  for (int i = 0; i < n2; ++i) lowr[i] = uppr[n2 - 1 - i];
  lowr[n2 - 1] = (ibc != 1 && ibc != 2) ? 0.0 : up1;
  std::vector<double> betad(n2);
  betad[0] = diag[0];
  for (int j = 1; j < n2; ++j) {
    gam[j] = uppr[j - 1] / betad[j - 1];
    betad[j] = diag[j] - lowr[j] * gam[j];
  }
  for (int j = 0; j < n2; ++j) invb[j] = 1.0 / betad[j];
}

// Staging of sparc_sample_noalloc: the point list and the gathered values of
// one call, in a file-scope device array so that the call allocates nothing.
// Each GPU holds its own copy until the process ends; host code reaches it
// only through the address hipGetSymbolAddress returns for that device.
#define SPARC_SAMPLE_SCRATCH_BYTES ((size_t)SPARC_SAMPLE_NOALLOC_MAX * (12 + 8))
static __device__ __align__(256) unsigned char scratch[SPARC_SAMPLE_SCRATCH_BYTES];
static_assert(SPARC_SAMPLE_NOALLOC_MAX % 32 == 0, "values start 256-byte aligned");
static std::mutex g_scratch_mu;  // guards the table and serialises the calls
static constexpr int kMaxDevices = 64;
static unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
static int scratch_address(sparc_ctx* c) {
  if (c->device < 0 || c->device >= kMaxDevices)
    return fail(c, SPARC_EINVAL, "device %d beyond the scratch table", c->device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[c->device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(c, SPARC_EHIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[c->device] = (unsigned char*)ptr;
  }
  c->scratch = g_scratch_ptr[c->device];
  return SPARC_OK;
}

extern "C" int sparc_create(const sparc_config* cfg, sparc_ctx** out) {
  if (!cfg || !out) return SPARC_EINVAL;
  *out = nullptr;
  sparc_ctx* c = new sparc_ctx();
  c->cfg = *cfg;
  c->damp2d = cfg->plane_damping != 0;
  if (c->cfg.field_scale == 0.0) c->cfg.field_scale = 1.0;
  std::string why;
  if (check_config(cfg, &why) != SPARC_OK) {
    c->err = why;
    *out = c;  // the caller can read the reason, then destroy
    return SPARC_EINVAL;
  }
  *out = c;
  c->fp32 = cfg->precision == SPARC_FP32;
  c->ts = c->fp32 ? 4 : 8;
  c->nxy = (long)cfg->nx * cfg->ny;
  c->N = c->nxy * cfg->nz;
  c->nzpml = cfg->npmlbot + cfg->npmltop;
  c->NP = c->nxy * std::max(c->nzpml, 1);
  HIPCHK(c, hipGetDevice(&c->device));
  HIPCHK(c, hipDeviceGetAttribute(&c->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                  c->device));
  HIPCHK(c, hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, c->device));
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(c, hipEventCreateWithFlags(&c->xev, hipEventDisableTiming));
  const size_t ts = c->ts, N = c->N;
  int rc;
  if ((rc = scratch_address(c))) return rc;
  c->comp = c->fp32 && cfg->compensated;
  c->nf = nfields_of(cfg->option);
  for (int f = 0; f < c->nf; ++f) {
    if ((rc = dalloc(c, &c->a[f], N * ts))) return rc;
    if ((rc = dalloc(c, &c->scr[f], N * ts))) return rc;
    if (c->comp && (rc = dalloc(c, &c->alo[f], N * 2))) return rc;
  }
  c->npv = is_mhd(cfg->option) ? 6 : 2;
  for (int f = 0; f < c->npv; ++f) {
    if ((rc = dalloc(c, &c->pmlvars[f], c->NP * ts))) return rc;
    if ((rc = dalloc(c, &c->scrpml[f], c->NP * ts))) return rc;
  }
  if (is_mhd(cfg->option)) {
    for (int w = 0; w < 12; ++w)
      if ((rc = dalloc(c, &c->mw[w], N * ts))) return rc;
    if ((rc = dalloc(c, &c->dmul2, cfg->nx * ts))) return rc;
  }
  for (int w = 0; w < 8; ++w)
    if ((rc = dalloc(c, &c->W[w], N * ts))) return rc;
  if (!is_2d(cfg->option) && cfg->damp_waves)
    for (int f = 0; f < 3; ++f)
      if ((rc = dalloc(c, &c->spec[f], spec_elems(cfg) * 2 * ts))) return rc;
  for (int p = 0; p < SPARC_PROF_COUNT; ++p)
    if ((rc = dalloc(c, &c->prof[p], cfg->nz * ts))) return rc;
  if ((rc = dalloc(c, &c->forcing, c->nxy * ts))) return rc;
  // FFT plans, twiddles, derivative multipliers (initialize.f90:204-210)
  make_plan(cfg->nx, &c->px);
  make_plan(cfg->ny, &c->py);
  auto twiddles = [&](int n, void** dst) -> int {
    std::vector<double> t(2 * n);
    for (int m = 0; m < n; ++m) {
      double ang = -2.0 * M_PI * (double)m / (double)n;
      t[2 * m] = cos(ang);
      t[2 * m + 1] = sin(ang);
    }
    int r = dalloc(c, dst, 2 * n * ts);
    if (r) return r;
    return upload(c, *dst, t.data(), 2 * n);
  };
  auto dmult = [&](int n, double stretch, void** dst) -> int {
    std::vector<double> t(n);
    for (int m = 0; m < n; ++m) {
      int ms = (m <= n / 2) ? m : m - n;
      if (2 * m == n) ms = 0;
      t[m] = (double)ms * 2.0 * PI_R * stretch / (double)n;
    }
    int r = dalloc(c, dst, n * ts);
    if (r) return r;
    return upload(c, *dst, t.data(), n);
  };
  if ((rc = twiddles(cfg->nx, &c->twx))) return rc;
  if ((rc = twiddles(cfg->ny, &c->twy))) return rc;
  if ((rc = dmult(cfg->nx, cfg->stretchx, &c->dmulx))) return rc;
  if ((rc = dmult(cfg->ny, cfg->stretchy, &c->dmuly))) return rc;
  if ((rc = dalloc(c, &c->fmulx, cfg->nx * ts))) return rc;
  if ((rc = dalloc(c, &c->fmuly, cfg->ny * ts))) return rc;
  if ((rc = dalloc(c, &c->rates, (size_t)(cfg->nx / 2 + 1) * cfg->ny * ts))) return rc;
  if ((rc = dalloc(c, &c->rates_o, (size_t)(cfg->nx / 2 + 1) * cfg->ny * ts))) return rc;
  // tridiagonal factors for ibc 0..3 and RHS coefficients (dbyd2.f90:88-146)
  for (int ibc = 0; ibc < 4; ++ibc) {
    std::vector<double> lowr, invb, gam;
    tri_factors(cfg->nz, ibc, lowr, invb, gam);
    std::vector<double> t(7 * cfg->nz);
    for (int k = 0; k < cfg->nz; ++k) {
      t[k] = lowr[k];
      t[cfg->nz + k] = invb[k];
      t[2 * cfg->nz + k] = gam[k];
    }
    // segment coupling products of the exact segmented solves for a partition
    // into nseg segments of L rows: fcum[k] = prod_{i=first..k} (-lowr_i
    // invb_i), gcum[k] = prod_{i=k+1..end} (-gam_i), 0 in the last segment;
    // rows 3-4 for k_zsolve_b, rows 5-6 for the exact mode of k_zsolve
    // (kZsUldsNseg segments)
    auto coupling = [&](int nseg, double* fc, double* gc) {
      const int n = cfg->nz, L = (n + nseg - 1) / nseg;
      for (int a = 0; a < n; a += L) {
        const int b = std::min(n, a + L);
        double p = 1.0;
        for (int k = a; k < b; ++k) fc[k] = (p *= -lowr[k] * invb[k]);
        p = 1.0;
        for (int k = b - 1; k >= a; --k) gc[k] = (b == n) ? 0.0 : (p *= -gam[k + 1]);
      }
    };
    coupling(c->fp32 ? kZbNseg : kZbNseg64, t.data() + 3 * cfg->nz,
             t.data() + 4 * cfg->nz);
    coupling(kZsUldsNseg, t.data() + 5 * cfg->nz, t.data() + 6 * cfg->nz);
    if ((rc = dalloc(c, &c->tri[ibc], 7 * cfg->nz * ts))) return rc;
    if ((rc = upload(c, c->tri[ibc], t.data(), 7 * cfg->nz))) return rc;
    // the widest range around the middle row where lowr, invb and gam (as
    // stored in the working precision) equal the middle row's values bitwise
    const int n = cfg->nz, mid = n / 2;
    auto same = [&](int k) {
      if (c->fp32)
        return (float)lowr[k] == (float)lowr[mid] && (float)invb[k] == (float)invb[mid] &&
               (float)gam[k] == (float)gam[mid];
      return lowr[k] == lowr[mid] && invb[k] == invb[mid] && gam[k] == gam[mid];
    };
    int k0 = mid, k1 = mid + 1;
    while (k0 - 1 >= 0 && same(k0 - 1)) --k0;
    while (k1 < n && same(k1)) ++k1;
    if (!same(mid)) k0 = k1 = 0;
    c->kc[ibc][0] = k0;
    c->kc[ibc][1] = k1;
    c->kcv[ibc][0] = lowr[mid];
    c->kcv[ibc][1] = invb[mid];
    c->kcv[ibc][2] = gam[mid];
  }
  const double d2i = (double)(cfg->nz - 1);
  const double alpha2 = 3. / 4., gamma2 = 1. / 8., beta1 = 3.;
  double tcv[14] = {d2i * 7.0 / 9.0,
                    d2i / 36.0,
                    (double)(float)(-43. / 96.) * d2i,
                    (double)(float)(-5. / 6.) * d2i,
                    9. / 8. * d2i,
                    (double)(float)(1. / 6.) * d2i,
                    (double)(float)(-1. / 96.) * d2i,
                    (double)(float)(-10 / 3.) * d2i,
                    -3.0 * d2i,
                    6. * d2i,
                    (double)(float)(1. / 3.) * d2i,
                    alpha2,
                    beta1,
                    1. / (alpha2 - beta1 * gamma2)};
  memcpy(c->tc, tcv, sizeof(tcv));
  // INITIALIZE_STEP (step.f90:24-41); optimals(3..5) are default-real literals
  double optimals[5] = {1.0, 0.5, (double)(float)0.1665579725791151184,
                        (double)(float)0.03950410250586729987,
                        (double)(float)0.007810706393337838236};
  c->betas[4] = optimals[1];
  c->betas[3] = optimals[2] / c->betas[4];
  c->betas[2] = optimals[3] / (c->betas[4] * c->betas[3]);
  c->betas[1] = optimals[4] / (c->betas[4] * c->betas[3] * c->betas[2]);
  c->betas[0] = 0.0;
  for (int s = 0; s < 5; ++s) c->betas[s] *= cfg->deltat;
  int r2 = sparc_reset(c);
  if (r2) return r2;
  c->err.clear();
  return SPARC_OK;
}

// Captured steps hold the kernel arguments of their capture; anything that
// can change them (backgrounds, profiles, damping, sources) drops them.
static void drop_graphs(sparc_ctx* c) {
  if ((c->gexec[0] || c->gexec[1]) && c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& g : c->gexec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

extern "C" void sparc_destroy(sparc_ctx* c) {
  if (c) drop_graphs(c);
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& r : c->recs) {
    (void)hipEventDestroy(r.e0);
    (void)hipEventDestroy(r.e1);
  }
  for (auto e : c->pool) (void)hipEventDestroy(e);
  {
    std::lock_guard<std::mutex> lk(g_capture_mu);
    for (void* p : c->allocs)
      if (p) (void)hipFree(p);
  }
  if (c->xev) (void)hipEventDestroy(c->xev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" const char* sparc_last_error(const sparc_ctx* c) {
  return c ? c->err.c_str() : "null context";
}

extern "C" int sparc_device_bytes(const sparc_ctx* c, size_t* bytes) {
  if (!c || !bytes) return SPARC_EINVAL;
  *bytes = c->bytes;
  return SPARC_OK;
}

extern "C" int sparc_reset(sparc_ctx* c) {
  if (!c || !c->stream) return SPARC_EINVAL;
  for (int f = 0; f < c->nf; ++f) {
    HIPCHK(c, hipMemsetAsync(c->a[f], 0, c->N * c->ts, c->stream));
    HIPCHK(c, hipMemsetAsync(c->scr[f], 0, c->N * c->ts, c->stream));
    if (c->comp) HIPCHK(c, hipMemsetAsync(c->alo[f], 0, c->N * 2, c->stream));
  }
  for (int f = 0; f < c->npv; ++f) {
    HIPCHK(c, hipMemsetAsync(c->pmlvars[f], 0, c->NP * c->ts, c->stream));
    HIPCHK(c, hipMemsetAsync(c->scrpml[f], 0, c->NP * c->ts, c->stream));
  }
  // W[1] is the y-derivative slot; it stays zero when the y part is absent
  for (int w = 0; w < 8; ++w) HIPCHK(c, hipMemsetAsync(c->W[w], 0, c->N * c->ts, c->stream));
  HIPCHK(c, hipMemsetAsync(c->forcing, 0, c->nxy * c->ts, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPARC_OK;
}

extern "C" int sparc_set_background(sparc_ctx* c, int id, const double* h, size_t n) {
  if (c) drop_graphs(c);
  if (!c || !h || id < 0 || id >= SPARC_BG_COUNT)
    return fail(c, SPARC_EINVAL, "bad background id %d", id);
  int is3d;
  if (n == (size_t)c->N)
    is3d = 1;
  else if (n == (size_t)c->cfg.nz)
    is3d = 0;
  else
    return fail(c, SPARC_EINVAL, "background %d: n=%zu is neither nx*ny*nz nor nz", id, n);
  if (c->bg[id] && c->bg3d[id] != is3d) {
    // shape changed: the old buffer stays in allocs, a new one is made
    c->bg[id] = nullptr;
  }
  if (!c->bg[id]) {
    int rc = dalloc(c, &c->bg[id], n * c->ts);
    if (rc) return rc;
  }
  c->bg3d[id] = is3d;
  c->bgset[id] = true;
  return upload(c, c->bg[id], h, n);
}

extern "C" int sparc_set_profile(sparc_ctx* c, int id, const double* h, int n) {
  if (c) drop_graphs(c);
  if (!c || !h || id < 0 || id >= SPARC_PROF_COUNT || n != c->cfg.nz)
    return fail(c, SPARC_EINVAL, "bad profile id %d or length %d", id, n);
  c->profset[id] = true;
  if (id == SPARC_PROF_UNSTRETCH) c->unstretch_h.assign(h, h + n);
  if (id == SPARC_PROF_Z) c->z_h.assign(h, h + n);
  return upload(c, c->prof[id], h, n);
}

extern "C" int sparc_set_damping(sparc_ctx* c, const double* r, size_t n) {
  if (c) drop_graphs(c);
  if (!c || !r || n != (size_t)(c->cfg.nx / 2 + 1) * c->cfg.ny)
    return fail(c, SPARC_EINVAL, "damping rates need (nx/2+1)*ny values");
  c->damping_set = true;
  if (c->dmul2 && c->cfg.ny == 1) {  // DAMP_VELOCITY's 2-D transform is along x only
    const int nx = c->cfg.nx;
    std::vector<double> t(nx);
    for (int m = 0; m < nx; ++m) t[m] = r[(m <= nx / 2) ? m : nx - m];
    int rc = upload(c, c->dmul2, t.data(), nx);
    if (rc) return rc;
  }
  // kept transposed, rates_t[y + ny*m], so the y kernels read it along y
  const int nh = c->cfg.nx / 2 + 1, ny = c->cfg.ny;
  std::vector<double> t((size_t)nh * ny);
  for (int y = 0; y < ny; ++y)
    for (int m = 0; m < nh; ++m) t[y + (size_t)ny * m] = r[m + (size_t)nh * y];
  int rc = upload(c, c->rates_o, r, n);
  if (rc) return rc;
  return upload(c, c->rates, t.data(), n);
}

extern "C" int sparc_set_filter(sparc_ctx* c, const double* dx, const double* dy) {
  if (c) drop_graphs(c);
  if (!c || !dx || !dy) return fail(c, SPARC_EINVAL, "null filter");
  // full-spectrum multipliers decay(min(m, n-m)); decay folds 1/n (initialize.f90:215-221)
  auto mk = [&](int n, const double* d, void* dst) -> int {
    std::vector<double> t(n);
    for (int m = 0; m < n; ++m) t[m] = d[(m <= n / 2) ? m : n - m];
    return upload(c, dst, t.data(), n);
  };
  int rc = mk(c->cfg.nx, dx, c->fmulx);
  if (rc) return rc;
  rc = mk(c->cfg.ny, dy, c->fmuly);
  if (rc) return rc;
  c->filter_set = true;
  return SPARC_OK;
}

extern "C" int sparc_set_source_planes(sparc_ctx* c, const int32_t* k_xy, size_t n) {
  if (c) drop_graphs(c);
  if (!c) return SPARC_EINVAL;
  if (n == 0) {
    c->src_planes = nullptr;  // the device buffers stay allocated for reuse
    return SPARC_OK;
  }
  if (!k_xy || n != (size_t)c->nxy) return fail(c, SPARC_EINVAL, "source planes need nx*ny values");
  if (!c->profset[SPARC_PROF_Z]) return fail(c, SPARC_ESTATE, "set the z profile first");
  const int nz = c->cfg.nz;
  std::vector<int32_t> k0(n);
  std::vector<double> dw(n);
  for (size_t i = 0; i < n; ++i) {
    const int k = k_xy[i];
    if (k < 2 || k > nz - 1) return fail(c, SPARC_EINVAL, "source plane %d outside 2..nz-1", k);
    k0[i] = k - 1;
    dw[i] = 2.0 / (c->z_h[k] - c->z_h[k - 2]);  // 2/(z(k+1)-z(k-1)) with 1-based k
  }
  int rc;
  void* p = c->planes_buf;
  if (!p) {
    if ((rc = dalloc(c, &p, n * sizeof(int32_t)))) return rc;
    c->planes_buf = p;
    if ((rc = dalloc(c, &c->src_dw, n * c->ts))) return rc;
  }
  HIPCHK(c, hipMemcpyAsync(p, k0.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if ((rc = upload(c, c->src_dw, dw.data(), n))) return rc;
  c->src_planes = (int32_t*)p;
  return SPARC_OK;
}

extern "C" int sparc_set_source(sparc_ctx* c, int kind, const double* mask, size_t nmask,
                                const double* series, int nt, int active_until) {
  if (c) drop_graphs(c);
  if (!c) return SPARC_EINVAL;
  c->src_kind = SPARC_SRC_NONE;
  if (kind == SPARC_SRC_NONE) return SPARC_OK;
  if (kind == SPARC_SRC_SERIES) {
    if (!mask || nmask != (size_t)c->nxy || !series || nt < 1)
      return fail(c, SPARC_EINVAL, "SERIES source needs mask(nx*ny) and a series");
    int rc = dalloc(c, &c->src_mask, c->nxy * c->ts);
    if (rc) return rc;
    if ((rc = upload(c, c->src_mask, mask, c->nxy))) return rc;
    c->src_series.assign(series, series + nt);
  } else if (kind == SPARC_SRC_LAGRANGE) {
    if (!mask || nt < 1 || nmask != (size_t)c->nxy * nt)
      return fail(c, SPARC_EINVAL, "LAGRANGE source needs vr(nx*ny*nt)");
    int rc = dalloc(c, &c->src_mask, nmask * c->ts);
    if (rc) return rc;
    if ((rc = upload(c, c->src_mask, mask, nmask))) return rc;
  } else {
    return fail(c, SPARC_EINVAL, "unknown source kind %d", kind);
  }
  c->src_kind = kind;
  c->src_nt = nt;
  c->src_active_until = active_until;
  return SPARC_OK;
}

// ---------------------------------------------------------------------------
// stepping
// ---------------------------------------------------------------------------
template <typename T>
static Fld<T> make_fld(sparc_ctx* c) {
  Fld<T> F{};
  for (int f = 0; f < SPARC_MAXFIELDS; ++f) {
    F.a[f] = (T*)c->a[f];
    F.scr[f] = (T*)c->scr[f];
  }
  for (int f = 0; f < c->npv; ++f) {
    F.pmlvars[f] = (T*)c->pmlvars[f];
    F.scrpml[f] = (T*)c->scrpml[f];
  }
  for (int w = 0; w < 12; ++w) F.M[w] = (T*)c->mw[w];
  for (int b = 0; b < SPARC_BG_COUNT; ++b) {
    F.bg[b] = (const T*)c->bg[b];
    F.bg3d[b] = c->bg3d[b];
  }
  F.stretch = (const T*)c->prof[SPARC_PROF_STRETCH];
  F.g = (const T*)c->prof[SPARC_PROF_G];
  for (int w = 0; w < 8; ++w) F.W[w] = (T*)c->W[w];
  F.forcing = (const T*)c->forcing;
  F.nx = c->cfg.nx;
  F.ny = c->cfg.ny;
  F.nz = c->cfg.nz;
  F.nxy = (int)c->nxy;
  F.npmlbot = c->cfg.npmlbot;
  F.npmltop = c->cfg.npmltop;
  F.src_plane = -1;
  F.delta_width = (T)c->cfg.delta_width;
  F.srcdep = (const T*)c->prof[SPARC_PROF_SOURCE_DEP];
  F.src_planes = c->src_planes;
  F.src_dw = (const T*)c->src_dw;
  return F;
}

template <typename T>
static TriCoef<T> make_tc(const sparc_ctx* c, int ibc) {
  const double* v = c->tc;
  TriCoef<T> t;
  t.kc0 = c->kc[ibc][0];
  t.kc1 = c->kc[ibc][1];
  t.lw = (T)c->kcv[ibc][0];
  t.ib = (T)c->kcv[ibc][1];
  t.gm = (T)c->kcv[ibc][2];
  t.ci1 = (T)v[0];
  t.ci2 = (T)v[1];
  t.c1 = (T)v[2];
  t.c2 = (T)v[3];
  t.c3 = (T)v[4];
  t.c4 = (T)v[5];
  t.c5 = (T)v[6];
  t.b1 = (T)v[7];
  t.b2 = (T)v[8];
  t.b3 = (T)v[9];
  t.b4 = (T)v[10];
  t.alpha2 = (T)v[11];
  t.beta1 = (T)v[12];
  t.cst = (T)v[13];
  return t;
}

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

static bool fast_size(int n) {
  return n == 32 || n == 64 || n == 96 || n == 128 || n == 192 || n == 256 || n == 384 ||
         n == 512;
}

// Length-256 lines: the x derivatives, the damping forward x pass and the
// damping inverse x pass run one wavefront per line with cross-lane shuffles
// (k_xw256); the y passes run the register four-step k_y256.  Row pairs per
// wavefront: 2 for the derivatives and the damping inverse, 4 for the
// damping forward pass.  Other sizes take the Stockham kernels of
// fft_reg.h.  Measured variants are listed in the design notes.
static constexpr int kXwLpw = 2;
template <typename T, int MODE, int N>
static void launch_xn(sparc_ctx* c, int nf, const XArgs<T>& X, T beta, const T* mult, int nrows,
                      int nlines) {
  if constexpr (N == 256 && (MODE == XM_DERIV || MODE == XM_DAMP_FWD || MODE == XM_DAMP_INV)) {
    constexpr int LPW = MODE == XM_DAMP_FWD ? 2 * kXwLpw : kXwLpw;
    hipLaunchKernelGGL((k_xw256<T, MODE, LPW>), dim3(cdiv(nlines, 4 * LPW), nf), dim3(256), 0,
                       c->stream, nrows, nlines, X, beta, mult, (const C2<T>*)c->twx,
                       c->cfg.ny, c->cfg.nz, c->nxy);
    return;
  }
  constexpr int LPB = XCfg<T, N>::LPB;
  if constexpr (MODE == XM_DAMP_INV) {
    if (X.pcp[0]) {  // with the next stage's damping forward pass
      hipLaunchKernelGGL((k_xlines_n<T, N, MODE, LPB, XCfg<T, N>::NT, true>),
                         dim3(cdiv(nlines, LPB), nf), dim3(XCfg<T, N>::NT), 0, c->stream, nrows,
                         nlines, X, beta, mult, (const C2<T>*)c->twx, c->cfg.ny, c->cfg.nz,
                         c->nxy);
      return;
    }
  }
  // 384-point derivative lines at 512 threads: 165 against 172 us (the
  // damping passes keep 256: 637 against 541 us with the next stage)
  constexpr int NT = (N == 384 && MODE == XM_DERIV) ? 512 : XCfg<T, N>::NT;
  hipLaunchKernelGGL((k_xlines_n<T, N, MODE, LPB, NT>), dim3(cdiv(nlines, LPB), nf), dim3(NT), 0,
                     c->stream, nrows, nlines, X, beta, mult, (const C2<T>*)c->twx, c->cfg.ny,
                     c->cfg.nz, c->nxy);
}

// x-lines over nf (1..3) fields that share the mode and beta.
template <typename T, int MODE>
static int launch_xm(sparc_ctx* c, int nf, const XArgs<T>& X, T beta, const T* mult) {
  const int n = c->cfg.nx;
  const long nrows = (long)c->cfg.ny * c->cfg.nz;
  long nlines;
  if (MODE == XM_DERIV || MODE == XM_FILTER)
    nlines = (nrows + 1) / 2;
  else
    nlines = (long)((c->cfg.nz + 1) / 2) * c->cfg.ny;
  // algorithmic arrays: inputs (a, s), outputs; the damping spectrum counts as one field
  const double fb = (double)c->N * sizeof(T);
  double nb;
  if (MODE == XM_DERIV || MODE == XM_FILTER || MODE == XM_DAMP_FWD)
    nb = (X.s[0] ? 2 : 1) + 1 + (X.cp[0] ? 1 : 0);
  else
    nb = X.pcp[0] ? 6 : 3;  // with the next stage: its a read, RHS_xi and spectrum written
  const int cls = MODE == XM_DAMP_FWD   ? SPARC_KC_DAMP
                  : MODE == XM_DAMP_INV ? SPARC_KC_DAMP_XI
                  : MODE == XM_FILTER   ? SPARC_KC_FILTER
                                        : SPARC_KC_XLINES;
  ProfScope ps(c, cls, nf * nb * fb);
  if (fast_size(n)) {
    switch (n) {
      case 32: launch_xn<T, MODE, 32>(c, nf, X, beta, mult, (int)nrows, (int)nlines); break;
      case 64: launch_xn<T, MODE, 64>(c, nf, X, beta, mult, (int)nrows, (int)nlines); break;
      case 96: launch_xn<T, MODE, 96>(c, nf, X, beta, mult, (int)nrows, (int)nlines); break;
      case 128: launch_xn<T, MODE, 128>(c, nf, X, beta, mult, (int)nrows, (int)nlines); break;
      case 192: launch_xn<T, MODE, 192>(c, nf, X, beta, mult, (int)nrows, (int)nlines); break;
      case 256: launch_xn<T, MODE, 256>(c, nf, X, beta, mult, (int)nrows, (int)nlines); break;
      case 384: launch_xn<T, MODE, 384>(c, nf, X, beta, mult, (int)nrows, (int)nlines); break;
      default: launch_xn<T, MODE, 512>(c, nf, X, beta, mult, (int)nrows, (int)nlines); break;
    }
    LAUNCHCHK(c);
    return SPARC_OK;
  }
  int LPB = std::max(1, (int)((sizeof(T) == 4 ? 2048 : 1024) / n));
  const size_t lds = 2 * (size_t)LPB * n * sizeof(C2<T>);
  for (int f = 0; f < nf; ++f) {
    hipLaunchKernelGGL((k_xlines<T, MODE>), dim3(cdiv(nlines, LPB)), dim3(256), lds, c->stream,
                       n, (int)nrows, (int)nlines, LPB, X.a[f], X.s[f], beta, X.out[f], mult,
                       c->px, (const C2<T>*)c->twx, X.spec[f], c->cfg.ny, c->cfg.nz, c->nxy,
                       X.edges);
    LAUNCHCHK(c);
    if (MODE == XM_DERIV && X.rhs) {
      hipLaunchKernelGGL((k_rhs_grad<T>), dim3(cdiv(c->N, 256)), dim3(256), 0, c->stream,
                         X.out[f], (const T*)X.out[f], X.rz, X.r3, c->nxy, c->cfg.nz, c->N);
      LAUNCHCHK(c);
    }
  }
  return SPARC_OK;
}

// the damping of three fields, one 256 x 256 plane per block (X.a the
// velocities, X.out the RHS they are subtracted from)
template <typename T>
static int launch_damp2d(sparc_ctx* c, const XArgs<T>& X) {
  if constexpr (sizeof(T) != 4) {  // an FP64 plane does not fit the LDS tile
    return fail(c, SPARC_EINVAL, "k_damp2d is FP32 only");
  } else {
    if (c->cfg.nx != 256 || c->cfg.ny != 256 || c->nxy != 65536)
      return fail(c, SPARC_EINVAL, "k_damp2d needs 256 x 256 planes");
    // the velocities read, the RHS read and written
    ProfScope ps(c, SPARC_KC_DAMP, 9.0 * (double)c->N * sizeof(T));
    hipLaunchKernelGGL((k_damp2d<T>), dim3(c->cfg.nz, 3), dim3(1024), 0, c->stream, X,
                       (const C2<T>*)c->twx, (const T*)c->rates_o, c->cfg.nz, c->nxy);
    LAUNCHCHK(c);
    return SPARC_OK;
  }
}

// the damping y pass of three fields (D) and the y derivative of p as the
// momentum RHS (G) in one launch (256-long y lines, nx a multiple of 32)
template <typename T>
static bool y_damp_deriv_fits(const sparc_ctx* c) {
  return c->cfg.ny == 256 && c->cfg.nx % 32 == 0;
}

template <typename T>
static int launch_y_damp_deriv(sparc_ctx* c, const XArgs<T>& D, const XArgs<T>& G, const T* mult) {
  const int nbd = (c->cfg.nx / 16) * ((c->cfg.nz + 1) / 2) * 3;
  const int nbg = (c->cfg.nx / 32) * c->cfg.nz;
  // three spectra read and written; p read, RHS_vy written
  ProfScope ps(c, SPARC_KC_DAMP_Y, 8.0 * (double)c->N * sizeof(T));
  hipLaunchKernelGGL((k_y256_damp_deriv<T>), dim3(nbd + nbg), dim3(256), 0, c->stream,
                     c->cfg.nx, c->cfg.nz, D, G, mult, (const C2<T>*)c->twy,
                     (const T*)c->rates_o, nbd);
  LAUNCHCHK(c);
  return SPARC_OK;
}

// the x derivative of p as the momentum RHS with the v_x damping inverse (H)
// and the damping inverse of two more fields (D) in one launch
template <typename T>
static int launch_x_rhs_dampinv(sparc_ctx* c, const XArgs<T>& H, const XArgs<T>& D,
                                const T* mult, bool prep) {
  const long nlines = (long)((c->cfg.nz + 1) / 2) * c->cfg.ny;
  const long nrows = (long)c->cfg.ny * c->cfg.nz;
  // p, the damping spectrum, the RHS written; two spectra, two RHS read and
  // written; prep: the three velocities read, their RHS_xi and spectra written
  ProfScope ps(c, SPARC_KC_DAMP_XI, (prep ? 18.0 : 9.0) * (double)c->N * sizeof(T));
  const dim3 grid(cdiv(nlines, 4 * kXwLpw), 3);
  if (prep)
    hipLaunchKernelGGL((k_xw256_rhs_dampinv<T, kXwLpw, true>), grid, dim3(256), 0, c->stream,
                       (int)nrows, (int)nlines, H, D, mult, (const C2<T>*)c->twx, c->cfg.ny,
                       c->cfg.nz, c->nxy);
  else
    hipLaunchKernelGGL((k_xw256_rhs_dampinv<T, kXwLpw, false>), grid, dim3(256), 0, c->stream,
                       (int)nrows, (int)nlines, H, D, mult, (const C2<T>*)c->twx, c->cfg.ny,
                       c->cfg.nz, c->nxy);
  LAUNCHCHK(c);
  return SPARC_OK;
}

// x derivative of X.a[0] and y derivative of Y.a[0] (each of a + beta s,
// stored or combined into out as X.acc / Y.acc say) in one launch on
// 256 x 256 planes: the blocks of either part fill the other's tail.
template <typename T>
static int launch_xy256(sparc_ctx* c, const XArgs<T>& X, const XArgs<T>& Y, T beta) {
  const long nrows = (long)c->cfg.ny * c->cfg.nz, nlines = (nrows + 1) / 2;
  const int nbx = cdiv(nlines, 4 * kXwLpw), nby = (c->cfg.nx / 32) * c->cfg.nz;
  auto arrays = [](const XArgs<T>& A) {
    return 2.0 + (A.s[0] ? 1.0 : 0.0) + (A.acc ? 1.0 : 0.0);
  };
  ProfScope ps(c, SPARC_KC_XLINES, (arrays(X) + arrays(Y)) * (double)c->N * sizeof(T));
  hipLaunchKernelGGL((k_xy256_div<T, kXwLpw>), dim3(nbx + nby), dim3(256), 0, c->stream,
                     (int)nrows, (int)nlines, X, Y, beta, (const T*)c->dmulx,
                     (const T*)c->dmuly, (const C2<T>*)c->twx, (const C2<T>*)c->twy, c->cfg.nx,
                     c->cfg.ny, c->cfg.nz, c->nxy, nbx);
  LAUNCHCHK(c);
  return SPARC_OK;
}

template <typename T, int MODE>
static int launch_x(sparc_ctx* c, const T* a, const T* s, T beta, T* out, const T* mult,
                    C2<T>* spec) {
  XArgs<T> X{};
  X.a[0] = a;
  X.s[0] = s;
  X.out[0] = out;
  X.spec[0] = spec;
  return launch_xm<T, MODE>(c, 1, X, beta, mult);
}

template <typename T, int MODE, int NY>
static void launch_yn(sparc_ctx* c, int nf, const XArgs<T>& X, T beta, const T* mult) {
  constexpr int TX = YCfg<T, NY, MODE>::TX;
  const int planes = (MODE == YM_DAMP) ? (c->cfg.nz + 1) / 2 : c->cfg.nz;
  if constexpr (NY == 256) {
    const int bx = MODE == YM_DAMP ? 16 : 32;
    if (c->cfg.nx % bx == 0) {
      hipLaunchKernelGGL((k_y256<T, MODE>), dim3(c->cfg.nx / bx, planes, nf), dim3(256), 0,
                         c->stream, c->cfg.nx, c->cfg.nz, X, beta, mult, (const C2<T>*)c->twy,
                         (const T*)c->rates_o);
      return;
    }
  }
  hipLaunchKernelGGL((k_ylines_n<T, NY, MODE>), dim3(c->cfg.nx / TX, planes, nf),
                     dim3(YCfg<T, NY, MODE>::NT), 0,
                     c->stream, c->cfg.nx, c->cfg.nz, X, beta, mult, (const C2<T>*)c->twy,
                     (const T*)c->rates);
}

template <typename T, int MODE, int NY>
static bool y_fits(int nx) {
  return nx % YCfg<T, NY, MODE>::TX == 0;
}

template <typename T, int MODE>
static int launch_ym(sparc_ctx* c, int nf, const XArgs<T>& X, T beta, const T* mult) {
  const int nx = c->cfg.nx, ny = c->cfg.ny;
  const double fb = (double)c->N * sizeof(T);
  ProfScope ps(c, MODE == YM_DAMP ? SPARC_KC_DAMP_Y : (MODE == YM_FILTER ? SPARC_KC_FILTER : SPARC_KC_YLINES),
               nf * (MODE == YM_DAMP ? 2.0 : (X.s[0] ? 3.0 : 2.0)) * fb);
  bool done = true;
  switch (ny) {
#define YCASE(NYV)                                              case NYV:                                                       if (y_fits<T, MODE, NYV>(nx))                                   launch_yn<T, MODE, NYV>(c, nf, X, beta, mult);              else                                                            done = false;                                               break;
    YCASE(32) YCASE(64) YCASE(96) YCASE(128) YCASE(192) YCASE(256) YCASE(384) YCASE(512)
#undef YCASE
    default:
      done = false;
  }
  if (done) {
    LAUNCHCHK(c);
    return SPARC_OK;
  }
  // generic path: tile width 128 bytes of x per row, LDS capped at 64 KiB
  int TX = (int)(128 / sizeof(T));
  if (MODE == YM_DAMP) TX /= 2;
  while (TX > 2 && (size_t)TX * ny * sizeof(C2<T>) * ((MODE == YM_DAMP) ? 2 : 1) > 65536) TX /= 2;
  const size_t lds = (MODE == YM_DAMP) ? 2 * (size_t)TX * ny * sizeof(C2<T>)
                                       : (size_t)TX * ny * sizeof(C2<T>);
  const int planes = (MODE == YM_DAMP) ? (c->cfg.nz + 1) / 2 : c->cfg.nz;
  for (int f = 0; f < nf; ++f) {
    hipLaunchKernelGGL((k_ylines<T, MODE>), dim3(nx / TX, planes), dim3(256), lds, c->stream, nx,
                       ny, c->cfg.nz, TX, X.a[f], X.s[f], beta, X.out[f], mult, c->py,
                       (const C2<T>*)c->twy, X.spec[f], (const T*)c->rates);
    LAUNCHCHK(c);
    if (MODE == YM_DERIV && X.rhs) {
      hipLaunchKernelGGL((k_rhs_grad<T>), dim3(cdiv(c->N, 256)), dim3(256), 0, c->stream,
                         X.out[f], (const T*)X.out[f], X.rz, X.r3, c->nxy, c->cfg.nz, c->N);
      LAUNCHCHK(c);
    }
  }
  return SPARC_OK;
}

template <typename T, int MODE>
static int launch_y(sparc_ctx* c, const T* a, const T* s, T beta, T* out, const T* mult,
                    C2<T>* spec) {
  XArgs<T> X{};
  X.a[0] = a;
  X.s[0] = s;
  X.out[0] = out;
  X.spec[0] = spec;
  return launch_ym<T, MODE>(c, 1, X, beta, mult);
}

// z-solve variants: k_zsolve_b for the epilogues in kZbEpis, k_zsolve with
// ud in LDS when it fits, else k_zsolve with a global ud and warm-ups.

template <typename T, int EPI>
static int launch_z(sparc_ctx* c, const Fld<T>& F, int ibc, const T* inA, const T* inS, T beta,
                    const T* prov, T provscale, T* out) {
  const T* tri = (const T*)c->tri[ibc];
  const int nz = c->cfg.nz;
  // algorithmic arrays of the solve + epilogue (the ud scratch is excluded)
  auto nbg = [&](std::initializer_list<int> ids) {
    double n = 0;
    for (int id : ids) n += c->bg3d[id] ? 1.0 : 0.0;
    return n;
  };
  const double in = inS ? 2.0 : 1.0;
  double nb;
  if (EPI == EPI_STORE)
    nb = in + 1;
  else if (EPI == EPI_P3D)
    nb = in + 2 + nbg({SPARC_BG_C2RHO0, SPARC_BG_GRADP0_Z, SPARC_BG_RHO0, SPARC_BG_GRADRHO0_Z}) + 2;
  else if (EPI == EPI_P2D)
    nb = in + 1 + nbg({SPARC_BG_C2RHO0, SPARC_BG_GRADP0_Z, SPARC_BG_RHO0, SPARC_BG_GRADRHO0_Z}) + 2;
  else if (EPI == EPI_RHS3D)  // rho temp, a (+ scr) of v; RHS of xi, v_z, rho
    nb = in + 1 + (F.beta != (T)0 ? 6 : 3) + nbg({SPARC_BG_RHOINV}) + 4;
  else if (EPI == EPI_RHS3N)  // rho temp; RHS of v_z
    nb = in + 1 + nbg({SPARC_BG_RHOINV}) + 1;
  else if (EPI == EPI_C2)  // v_z again (own plane), dvxdx, dvydy -> W4, W5
    nb = in + 2 + nbg({SPARC_BG_C2RHO0, SPARC_BG_GRADP0_Z, SPARC_BG_RHO0, SPARC_BG_GRADRHO0_Z}) + 2;
  else if (EPI == EPI_V2)  // gradp_x, gradp_y, W4, W5, rho temp -> scr of 5 fields
    nb = in + 4 + (F.beta != (T)0 ? 2 : 1) + nbg({SPARC_BG_RHOINV}) + 5;
  else
    nb = in + 2 + (F.beta != (T)0 ? 4 : 2) + nbg({SPARC_BG_RHOINV}) + 4;
  ProfScope ps(c, (EPI == EPI_P3D || EPI == EPI_P2D || EPI == EPI_C2) ? SPARC_KC_ZSOLVE_P
                                                                      : SPARC_KC_ZSOLVE,
               nb * (double)c->N * sizeof(T));
  // segments per column: L = nz/NSEG planes each, at least ~48 (>= 2x warm-up)
  const int nseg = nz >= 192 ? 4 : (nz >= 96 ? 2 : 1);
  const TriCoef<T> tcv = make_tc<T>(c, ibc);
  const dim3 grid(cdiv(c->nxy, 64));
  const size_t lds_b = 2 * 64 * (size_t)nz * sizeof(T);
  if (((kZbEpis >> EPI) & 1) && nz >= 5 && lds_b <= (size_t)c->lds_max) {
    constexpr int NS = sizeof(T) == 4 ? kZbNseg : kZbNseg64;
    const bool bg3 = c->bg3d[SPARC_BG_C2RHO0] || c->bg3d[SPARC_BG_GRADP0_Z] ||
                     c->bg3d[SPARC_BG_RHO0] || c->bg3d[SPARC_BG_GRADRHO0_Z];
#define ZB(G)                                                                                    \
  hipLaunchKernelGGL((k_zsolve_b<T, EPI, NS, G>), grid, dim3(64 * NS), lds_b, c->stream, F, tcv, \
                     tri, tri + nz, tri + 2 * nz, tri + 3 * nz, tri + 4 * nz, ibc, inA, inS, beta, \
                     prov, provscale, out)
    if (bg3)
      ZB(false);
    else
      ZB(true);
#undef ZB
    LAUNCHCHK(c);
    return SPARC_OK;
  }
  // ud in LDS when two blocks of 64 columns fit in one CU's LDS; with the
  // exact coupling any segment of at least 3 planes works, the warm-up
  // needs long columns (nz >= 192)
  const size_t lds_ud = 64 * (size_t)nz * sizeof(T);
  const bool seg_ok = (nz + kZsUldsNseg - 1) / kZsUldsNseg >= 3;
  const bool ulds = seg_ok && lds_ud <= 81920 && lds_ud <= (size_t)c->lds_max;
  const size_t lds = ulds ? lds_ud : 5 * (size_t)nz * sizeof(T);
#define ZL(NS, U)                                                                                 \
  hipLaunchKernelGGL((k_zsolve<T, EPI, NS, U>), grid, dim3(64 * NS), lds, c->stream, F, tcv, tri, \
                     tri + nz, tri + 2 * nz, ibc, inA, inS, beta, (T*)c->W[7], prov, provscale,     \
                     out, tri + 5 * nz, tri + 6 * nz)
  if (ulds)
    ZL(kZsUldsNseg, true);
  else if (nseg == 4)
    ZL(4, false);
  else if (nseg == 2)
    ZL(2, false);
  else
    ZL(1, false);
#undef ZL
  LAUNCHCHK(c);
  return SPARC_OK;
}

// The pressure and its z derivative in one launch (k_zsolve_pv): FP32
// columns whose two tiles of 64 columns fit the LDS, at most kPvLmax planes
// per segment.  Writes W3 = p, rvz = RHS_vz before the damping (zero
// planes: RHS_vz itself) and both PML psi.
static constexpr int kPvLmax = 20;
// The streaming form (k_zstream_pv): whole blocks of 64-column waves, its
// plane table and rings within the LDS, every background of the two
// epilogues a z profile.
static bool zstream_fits(const sparc_ctx* c) {
  for (int id : {SPARC_BG_C2RHO0, SPARC_BG_GRADP0_Z, SPARC_BG_RHO0, SPARC_BG_GRADRHO0_Z,
                 SPARC_BG_RHOINV, SPARC_BG_AZ, SPARC_BG_BZPML})
    if (c->bg3d[id] || !c->bg[id]) return false;
  // one block per 64 * kZsWaves columns: below one block per CU the
  // segmented column kernel runs faster (128 x 128 x 150: 35 against 65 us)
  return c->fp32 && c->nxy % (64 * kZsWaves) == 0 && c->nxy / (64 * kZsWaves) >= c->cus &&
         c->cfg.nz >= 8 &&
         (size_t)c->N * 4 < ((size_t)1 << 31) &&
         zstream_lds_bytes(c->cfg.nz) + kZsRingBytes <= (size_t)c->lds_max;
}
template <typename T>
static bool zpv_fits(const sparc_ctx* c) {
  const int nz = c->cfg.nz, L = (nz + kZbNseg - 1) / kZbNseg;
  return sizeof(T) == 4 && nz >= 5 &&
         (zstream_fits(c) ||
          (L <= kPvLmax && 2 * 64 * (size_t)nz * sizeof(T) <= (size_t)c->lds_max));
}

template <typename T>
static int launch_zpv(sparc_ctx* c, const Fld<T>& F, const T* inA, const T* inS, T beta, T* rvz) {
  if constexpr (sizeof(T) != 4) {
    return fail(c, SPARC_EINVAL, "k_zsolve_pv is FP32 only");
  } else {
    const int nz = c->cfg.nz;
    auto nbg = [&](std::initializer_list<int> ids) {
      double n = 0;
      for (int id : ids) n += c->bg3d[id] ? 1.0 : 0.0;
      return n;
    };
    // xi_z temp, the x and y divergence parts, backgrounds; p and RHS_vz
    const double nb = (inS ? 2.0 : 1.0) + 2 +
                      nbg({SPARC_BG_C2RHO0, SPARC_BG_GRADP0_Z, SPARC_BG_RHO0,
                           SPARC_BG_GRADRHO0_Z, SPARC_BG_RHOINV}) + 2;
    ProfScope ps(c, SPARC_KC_ZSOLVE_P, nb * (double)c->N * sizeof(T));
    const bool bg3 = c->bg3d[SPARC_BG_C2RHO0] || c->bg3d[SPARC_BG_GRADP0_Z] ||
                     c->bg3d[SPARC_BG_RHO0] || c->bg3d[SPARC_BG_GRADRHO0_Z];
    const T* t1 = (const T*)c->tri[1];
    const T* t0 = (const T*)c->tri[0];
    if (zstream_fits(c)) {
      ZsArgs Z{};
      Z.a = inA;
      Z.s = inS;
      Z.w0 = F.W[0];
      Z.w1 = F.W[1];
      Z.p = F.W[3];
      Z.rvz = rvz;
      Z.scrvz = F.scr[SPARC_V_Z];
      for (int w = 0; w < 2; ++w) {
        Z.pv[w] = F.pmlvars[w];
        Z.ps[w] = F.scrpml[w];
      }
      Z.tab[ZT_ST] = F.stretch;
      Z.tab[ZT_C2R] = F.bg[SPARC_BG_C2RHO0];
      Z.tab[ZT_GP] = F.bg[SPARC_BG_GRADP0_Z];
      Z.tab[ZT_R0] = F.bg[SPARC_BG_RHO0];
      Z.tab[ZT_GR] = F.bg[SPARC_BG_GRADRHO0_Z];
      Z.tab[ZT_G] = F.g;
      Z.tab[ZT_AZ] = F.bg[SPARC_BG_AZ];
      Z.tab[ZT_BZ] = F.bg[SPARC_BG_BZPML];
      Z.tab[ZT_LOWR1] = t1;
      Z.tab[ZT_INVB1] = t1 + nz;
      Z.tab[ZT_LOWR0] = t0;
      Z.tab[ZT_INVB0] = t0 + nz;
      Z.tab[ZT_GAM1] = t1 + 2 * nz;
      Z.tab[ZT_GAM0] = t0 + 2 * nz;
      Z.tab[ZT_RI] = F.bg[SPARC_BG_RHOINV];
      Z.tab[ZT_ST2] = F.stretch;
      Z.forcing = F.forcing;
      Z.nbytes = (unsigned)(c->N * 4);
      Z.pbytes = (unsigned)(c->nxy * (long)c->nzpml * 4);
      Z.nxy = (int)c->nxy;
      Z.n = nz;
      Z.npmlbot = F.npmlbot;
      Z.npmltop = F.npmltop;
      Z.src_plane = F.src_plane;
      Z.beta = F.beta;
      Z.beta_in = beta;
      Z.delta_width = F.delta_width;
      for (int ibc = 0; ibc < 2; ++ibc) {
        const TriCoef<T> tq = make_tc<T>(c, ibc);
        Z.kc0[ibc] = tq.kc0;
        Z.kc1[ibc] = tq.kc1;
        Z.lw[ibc] = tq.lw;
        Z.ib[ibc] = tq.ib;
        Z.gm[ibc] = tq.gm;
      }
      // columns long enough for two waves each (one from each end)
      Z.split = nz >= 96;
      hipLaunchKernelGGL(k_zstream_pv, dim3(cdiv(c->nxy, 64 * kZsWaves)), dim3(128 * kZsWaves),
                         zstream_lds_bytes(nz), c->stream, Z, make_tc<T>(c, 1));
      LAUNCHCHK(c);
      return SPARC_OK;
    }
    const size_t lds = 2 * 64 * (size_t)nz * sizeof(T);
#define ZPV(G)                                                                                \
  hipLaunchKernelGGL((k_zsolve_pv<kZbNseg, G, kPvLmax>), dim3(cdiv(c->nxy, 64)),              \
                     dim3(64 * kZbNseg), lds, c->stream, F, make_tc<T>(c, 1), t1, t0, inA, inS, \
                     beta, rvz)
    if (bg3)
      ZPV(false);
    else
      ZPV(true);
#undef ZPV
    LAUNCHCHK(c);
    return SPARC_OK;
  }
}

// FILTER_VAR_Z alone (kernels.f90:845-881) on one field, in place
template <typename T>
static void zfilter_only(sparc_ctx* c, T* f) {
  UpdArgs<T> A{};
  A.a[0] = f;
  hipLaunchKernelGGL((k_update_filter<T>), dim3(cdiv(c->nxy, 64), 1), dim3(64), 0, c->stream, A,
                     (T)0, (int)c->nxy, c->cfg.nz, 2);
}

// Lagrange weights of LAGRANGE_INTERP (physics.f90:680-734) for one time.
static LagArgs lagrange_args(const sparc_ctx* c, int time) {
  LagArgs la{};
  const double timestep = c->cfg.timestep, cad = c->cfg.cadforcing;
  const int num_steps = c->cfg.num_steps;
  double xt = (double)time * timestep / cad;
  const int time1 = (int)floor(xt) + 1;
  xt = time * timestep;
  double xs[7];
  for (int q = 0; q < 7; ++q) xs[q] = (time1 - 4 + q) * cad;
  bool use[7];
  for (int q = 0; q < 4; ++q) use[q] = xs[q] >= 0;
  use[4] = true;
  use[5] = time1 < num_steps;
  use[6] = time1 < num_steps - 1;
  for (int q = 0; q < 7; ++q) {
    if (!use[q]) continue;
    const int samp = time1 - 3 + q;  // 1-based sample of vr
    if (samp < 1 || samp > c->src_nt) continue;
    double den = 1.0, num = 1.0;
    for (int r = 0; r < 7; ++r)
      if (r != q) {
        den *= (xs[q] - xs[r]);
        num *= (xt - xs[r]);
      }
    la.idx[la.nq] = samp - 1;
    la.w[la.nq] = num / den;
    la.nq++;
  }
  return la;
}

template <typename T>
static int run_stage(sparc_ctx* c, Fld<T>& F, int s) {
  const T beta = (T)c->betas[s];
  F.beta = beta;
  const bool first = (s == 0);
  // Option 4 with the damping on x lines of a fast length and the fused z
  // solve: the last pass of stages 0-3 also does the next stage's damping
  // forward pass (chain). It writes RHS_xi of stage s + 1
  // while the divergence passes of stage s + 1 have still to read RHS_xi of
  // stage s, so RHS_xi alternates between scr (even stages) and W2, W5, W6
  // (odd stages, idle in this path); stage 4's, which the update reads,
  // lands in scr.
  bool fuse = false, pv = false, ydd = false, chain = false;
  if (c->cfg.option == SPARC_OPT_QUIET_PML_3D_DISPL) {
    const bool damp = c->cfg.damp_waves && c->damping_set;
    const bool d2d = damp && c->damp2d && sizeof(T) == 4 && c->cfg.nx == 256 &&
                     c->cfg.ny == 256;
    fuse = damp && !d2d && fast_size(c->cfg.nx);
    pv = fuse && zpv_fits<T>(c);
    ydd = pv && y_damp_deriv_fits<T>(c);
    chain = pv;
  }
  T* const xi_alt[3] = {F.W[2], F.W[5], F.W[6]};
  auto xi_buf = [&](int stage, int f) -> T* { return (stage & 1) ? xi_alt[f] : F.scr[f]; };
  auto S = [&](int f) -> const T* {
    if (first) return nullptr;
    return (chain && f < 3) ? xi_buf(s - 1, f) : (const T*)c->scr[f];
  };
  const T* dmx = (const T*)c->dmulx;
  const T* dmy = (const T*)c->dmuly;
  int rc;
  if (c->cfg.option == SPARC_OPT_QUIET_PML_3D_DISPL) {
    // ddxyz(xi_x, xi_y, xi_z, bc=1) + PML + div, p, rho (displacement.f90:491-535)
    if (c->cfg.nx == 256 && c->cfg.ny == 256) {
      // the x and y derivatives in one launch
      XArgs<T> X{}, Y{};
      X.a[0] = F.a[0];
      X.s[0] = S(0);
      X.out[0] = F.W[0];
      Y.a[0] = F.a[1];
      Y.s[0] = S(1);
      Y.out[0] = F.W[1];
      if ((rc = launch_xy256<T>(c, X, Y, beta))) return rc;
    } else {
      if ((rc = launch_x<T, XM_DERIV>(c, F.a[0], S(0), beta, F.W[0], dmx, nullptr))) return rc;
      if ((rc = launch_y<T, YM_DERIV>(c, F.a[1], S(1), beta, F.W[1], dmy, nullptr))) return rc;
    }
    const bool damp = c->cfg.damp_waves && c->damping_set;
    // FP32 on 256 x 256 planes with cfg.plane_damping: the whole damping in
    // one pass per plane
    const bool d2d = damp && c->damp2d && sizeof(T) == 4 && c->cfg.nx == 256 &&
                     c->cfg.ny == 256;
    // fuse (damping on x lines of a fast length): the damping forward pass
    // reads the stage temp v = a + beta scr of the velocities itself and
    // also stores it as RHS_xi (once the divergence passes above have read
    // the previous RHS_xi), so the z solve of p below writes only RHS_vz; on
    // 256-wide lines RHS_vx is written once, by the pass that also
    // subtracts its damping.
    // pv (fuse, FP32 and columns that fit): p, rho' and the z part of grad p
    // in one z-solve launch (k_zstream_pv), RHS_vz staged in W4 for the
    // damping inverse; the damping forward pass then follows it
    XArgs<T> DX{};
    for (int f = 0; f < 3; ++f) {
      DX.out[f] = F.scr[3 + f];
      DX.spec[f] = (C2<T>*)c->spec[f];
    }
    if (pv) {
      if ((rc = launch_zpv<T>(c, F, F.a[2], S(2), beta, F.W[4]))) return rc;
    } else {
      if ((rc = launch_z<T, EPI_P3D>(c, F, 1, F.a[2], S(2), beta, nullptr, 0, nullptr))) return rc;
    }
    if (fuse) {
      // DAMP_VELOCITY (damping.f90:125-185) forward x and y passes; with
      // chain, stages 1-4 have the x pass done by the previous stage's last
      // pass
      for (int f = 0; f < 3; ++f) {
        DX.a[f] = F.a[3 + f];
        DX.s[f] = S(3 + f);
        DX.cp[f] = F.scr[f];
      }
      if ((!chain || first) && (rc = launch_xm<T, XM_DAMP_FWD>(c, 3, DX, beta, nullptr)))
        return rc;
      // with the fused z solve the damping y pass runs with the y derivative
      // of p below
      if (!ydd && (rc = launch_ym<T, YM_DAMP>(c, 3, DX, 0, nullptr))) return rc;
      // ddxyz(p, bc=0) (displacement.f90:547): the z part with RHS_vz
      if (!pv &&
          (rc = launch_z<T, EPI_RHS3N>(c, F, 0, F.W[3], nullptr, 0, nullptr, 0, nullptr)))
        return rc;
    } else {
      // ddxyz(p, bc=0) (displacement.f90:547): the z part with the RHS
      // epilogue (RHS_xi = temp v, RHS_vz; reads the previous RHS_v)
      if ((rc = launch_z<T, EPI_RHS3D>(c, F, 0, F.W[3], nullptr, 0, nullptr, 0, nullptr))) return rc;
    }
    // the x and y parts of grad p written straight into RHS_vx, RHS_vy as
    // -rhoinv*dp
    XArgs<T> G{};
    G.a[0] = F.W[3];
    G.rhs = 1;
    G.rz = F.bg3d[SPARC_BG_RHOINV] ? nullptr : F.bg[SPARC_BG_RHOINV];
    G.r3 = F.bg3d[SPARC_BG_RHOINV] ? F.bg[SPARC_BG_RHOINV] : nullptr;
    const bool x256 = fuse && c->cfg.nx == 256;
    if (!x256) {
      G.out[0] = F.scr[SPARC_V_X];
      if ((rc = launch_xm<T, XM_DERIV>(c, 1, G, 0, dmx))) return rc;
    }
    G.out[0] = F.scr[SPARC_V_Y];
    if (ydd) {
      if ((rc = launch_y_damp_deriv<T>(c, DX, G, dmy))) return rc;
    } else if ((rc = launch_ym<T, YM_DERIV>(c, 1, G, 0, dmy))) {
      return rc;
    }
    if (fuse && !x256) {
      // DAMP_VELOCITY's inverse x pass on the RHS the passes above wrote
      // (RHS_vz from W4 when the fused z solve staged it there)
      XArgs<T> DI{};
      for (int f = 0; f < 3; ++f) {
        DI.out[f] = DX.out[f];
        DI.spec[f] = DX.spec[f];
      }
      if (pv) DI.rin[2] = F.W[4];
      DI.edges = DX.edges;
      if (chain && s < 4) {
        // the next stage's temp v = a + beta' RHS_v, its x spectrum and RHS_xi
        for (int f = 0; f < 3; ++f) {
          DI.pa[f] = F.a[3 + f];
          DI.pcp[f] = xi_buf(s + 1, f);
        }
        DI.pbeta = (T)c->betas[s + 1];
      }
      if ((rc = launch_xm<T, XM_DAMP_INV>(c, 3, DI, 0, nullptr))) return rc;
    } else if (fuse) {
      XArgs<T> H = G;
      H.out[0] = F.scr[SPARC_V_X];
      H.spec[0] = DX.spec[0];
      H.edges = DX.edges;
      XArgs<T> D2{};
      for (int f = 0; f < 2; ++f) {
        D2.out[f] = DX.out[f + 1];
        D2.spec[f] = DX.spec[f + 1];
      }
      if (pv) D2.rin[1] = F.W[4];  // RHS_vz before the damping
      D2.edges = DX.edges;
      const bool prep = chain && s < 4;
      if (prep) {
        // the next stage's temp v = a + beta' RHS_v, its x spectrum and RHS_xi
        H.pa[0] = F.a[3];
        H.pcp[0] = xi_buf(s + 1, 0);
        H.pbeta = (T)c->betas[s + 1];
        for (int f = 0; f < 2; ++f) {
          D2.pa[f] = F.a[4 + f];
          D2.pcp[f] = xi_buf(s + 1, 1 + f);
        }
        D2.pbeta = H.pbeta;
      }
      if ((rc = launch_x_rhs_dampinv<T>(c, H, D2, dmx, prep))) return rc;
    } else if (damp) {
      // DAMP_VELOCITY (damping.f90:125-185) on the stage's temp v, which the
      // epilogue above has just stored as RHS_xi = v (zero on planes 1 and nz,
      // whose damping output is discarded)
      for (int f = 0; f < 3; ++f) DX.a[f] = F.scr[f];
      if (d2d) return launch_damp2d<T>(c, DX);
      if ((rc = launch_xm<T, XM_DAMP_FWD>(c, 3, DX, 0, nullptr))) return rc;
      if ((rc = launch_ym<T, YM_DAMP>(c, 3, DX, 0, nullptr))) return rc;
      if ((rc = launch_xm<T, XM_DAMP_INV>(c, 3, DX, 0, nullptr))) return rc;
    }
  } else if (c->cfg.option == SPARC_OPT_MHD_PML_3D_DISPL) {
    // MP_MHD_PML_3D_DISPL (displacement.f90:663-849): 18 derivative passes
    // (ddx, ddy, compact ddz) joined by pointwise kernels
    const long N = c->N;
    const dim3 gN(cdiv(N, 256)), b256(256);
    const long NB = c->nxy * (long)c->cfg.npmlbot;
    auto pmlb = [&](T* d, int w) -> int {
      if (NB <= 0) return SPARC_OK;
      ProfScope ps(c, SPARC_KC_OTHER, 5.0 * (double)NB * sizeof(T));
      hipLaunchKernelGGL((k_pml_bottom<T>), dim3(cdiv(NB, 256)), b256, 0, c->stream, F, d, w, NB);
      LAUNCHCHK(c);
      return SPARC_OK;
    };
    auto dz = [&](const T* in, const T* s, T b, int ibc, T* out) -> int {
      return launch_z<T, EPI_STORE>(c, F, ibc, in, s, b, nullptr, 0, out);
    };
    T** M = F.M;
    // on 256 x 256 planes independent x and y derivative passes share a
    // launch (launch_xy256)
    const bool xy = c->cfg.nx == 256 && c->cfg.ny == 256;
    auto xypair = [&](const T* ax, const T* sx, T* ox, T aldx, T goutx, const T* ay,
                      const T* sy, T* oy, T aldy, T gouty, T b) -> int {
      XArgs<T> X{}, Y{};
      X.a[0] = ax;
      X.s[0] = sx;
      X.out[0] = ox;
      X.acc = aldx != (T)0;
      X.ald = aldx;
      X.gout = goutx;
      Y.a[0] = ay;
      Y.s[0] = sy;
      Y.out[0] = oy;
      Y.acc = aldy != (T)0;
      Y.ald = aldy;
      Y.gout = gouty;
      return launch_xy256<T>(c, X, Y, b);
    };
    // div xi (bc=1) with psivz; p, rho and flux = xi x B0 in one pass
    if (xy) {
      if ((rc = xypair(F.a[0], S(0), F.W[0], 0, 0, F.a[1], S(1), F.W[1], 0, 0, beta))) return rc;
    } else {
      if ((rc = launch_x<T, XM_DERIV>(c, F.a[0], S(0), beta, F.W[0], dmx, nullptr))) return rc;
      if ((rc = launch_y<T, YM_DERIV>(c, F.a[1], S(1), beta, F.W[1], dmy, nullptr))) return rc;
    }
    if ((rc = dz(F.a[2], S(2), beta, 1, F.W[2]))) return rc;
    if ((rc = pmlb(F.W[2], 0))) return rc;
    {
      ProfScope ps(c, SPARC_KC_OTHER, 14.0 * (double)N * sizeof(T));
      hipLaunchKernelGGL((k_mhd_pr<T>), gN, b256, 0, c->stream, F, N);
      LAUNCHCHK(c);
    }
    // grad p (bc=0) with psip
    if (xy) {
      if ((rc = xypair(F.W[3], nullptr, M[3], 0, 0, F.W[3], nullptr, M[4], 0, 0, 0))) return rc;
    } else {
      if ((rc = launch_x<T, XM_DERIV>(c, F.W[3], nullptr, 0, M[3], dmx, nullptr))) return rc;
      if ((rc = launch_y<T, YM_DERIV>(c, F.W[3], nullptr, 0, M[4], dmy, nullptr))) return rc;
    }
    if ((rc = dz(F.W[3], nullptr, 0, 0, M[5]))) return rc;
    if ((rc = pmlb(M[5], 1))) return rc;
    // b = curl(flux): dx f3, dy f3, dz f2 (psiinductionbx), dx f2, dy f1, dz f1
    // (psiinductionby).  With the register line kernels the pieces are
    // combined as they are produced (out = ald d + gout out) into M9..M11,
    // after the z pieces and their PML: bx = dy f3 - dz f2, by = dz f1 -
    // dx f3, bz = dx f2 - dy f1.  The curl of b then goes to the freed slots
    // and k_mhd_rhs reads them through a remapped table.  Otherwise the six
    // pieces go to M6..M11 and k_mhd_b combines them into M0..M2.
    Fld<T> FR = F;  // the arrays k_mhd_rhs reads as M0..M11
    const bool bacc = fast_size(c->cfg.nx) && fast_size(c->cfg.ny);
    auto xacc = [&](const T* in, T* out, T ald, T gout) -> int {
      XArgs<T> X{};
      X.a[0] = in;
      X.out[0] = out;
      X.acc = 1;
      X.ald = ald;
      X.gout = gout;
      return launch_xm<T, XM_DERIV>(c, 1, X, 0, dmx);
    };
    auto yacc = [&](const T* in, T* out, T ald, T gout) -> int {
      XArgs<T> X{};
      X.a[0] = in;
      X.out[0] = out;
      X.acc = 1;
      X.ald = ald;
      X.gout = gout;
      return launch_ym<T, YM_DERIV>(c, 1, X, 0, dmy);
    };
    if (bacc && xy) {
      // the same passes, the independent x/y pairs in one launch each
      if ((rc = dz(M[1], nullptr, 0, 1, M[9]))) return rc;
      if ((rc = pmlb(M[9], 4))) return rc;  // psiinductionbx
      if ((rc = dz(M[0], nullptr, 0, 1, M[10]))) return rc;
      if ((rc = pmlb(M[10], 5))) return rc;  // psiinductionby
      // by = dz f1 - dx f3 (x part), bx = dy f3 - dz f2 (y part)
      if ((rc = xypair(M[2], nullptr, M[10], -1, 1, M[2], nullptr, M[9], 1, -1, 0))) return rc;
      if ((rc = launch_x<T, XM_DERIV>(c, M[1], nullptr, 0, M[11], dmx, nullptr))) return rc;
      if ((rc = yacc(M[0], M[11], -1, 1))) return rc;  // bz = dx f2 - dy f1
      // J = curl(b) pieces dx by, dy bz, dz bx, dx bz, dy bx, dz by into M0,
      // M1, M2, M6, M7, M8 (flux no longer needed)
      if ((rc = xypair(M[10], nullptr, M[0], 0, 0, M[11], nullptr, M[1], 0, 0, 0))) return rc;
      if ((rc = dz(M[9], nullptr, 0, 1, M[2]))) return rc;
      if ((rc = pmlb(M[2], 2))) return rc;  // psidzbx
      if ((rc = xypair(M[11], nullptr, M[6], 0, 0, M[9], nullptr, M[7], 0, 0, 0))) return rc;
      if ((rc = dz(M[10], nullptr, 0, 1, M[8]))) return rc;
      if ((rc = pmlb(M[8], 3))) return rc;  // psidzby
      T* const perm[12] = {M[9], M[10], M[11], M[3], M[4], M[5], M[0], M[1], M[2], M[6], M[7], M[8]};
      for (int w = 0; w < 12; ++w) FR.M[w] = perm[w];
    } else if (bacc) {
      if ((rc = dz(M[1], nullptr, 0, 1, M[9]))) return rc;
      if ((rc = pmlb(M[9], 4))) return rc;  // psiinductionbx
      if ((rc = yacc(M[2], M[9], 1, -1))) return rc;   // bx = dy f3 - dz f2
      if ((rc = dz(M[0], nullptr, 0, 1, M[10]))) return rc;
      if ((rc = pmlb(M[10], 5))) return rc;  // psiinductionby
      if ((rc = xacc(M[2], M[10], -1, 1))) return rc;  // by = dz f1 - dx f3
      if ((rc = launch_x<T, XM_DERIV>(c, M[1], nullptr, 0, M[11], dmx, nullptr))) return rc;
      if ((rc = yacc(M[0], M[11], -1, 1))) return rc;  // bz = dx f2 - dy f1
      // J = curl(b) pieces dx by, dy bz, dz bx, dx bz, dy bx, dz by into M0,
      // M1, M2, M6, M7, M8 (flux no longer needed)
      if ((rc = launch_x<T, XM_DERIV>(c, M[10], nullptr, 0, M[0], dmx, nullptr))) return rc;
      if ((rc = launch_y<T, YM_DERIV>(c, M[11], nullptr, 0, M[1], dmy, nullptr))) return rc;
      if ((rc = dz(M[9], nullptr, 0, 1, M[2]))) return rc;
      if ((rc = pmlb(M[2], 2))) return rc;  // psidzbx
      if ((rc = launch_x<T, XM_DERIV>(c, M[11], nullptr, 0, M[6], dmx, nullptr))) return rc;
      if ((rc = launch_y<T, YM_DERIV>(c, M[9], nullptr, 0, M[7], dmy, nullptr))) return rc;
      if ((rc = dz(M[10], nullptr, 0, 1, M[8]))) return rc;
      if ((rc = pmlb(M[8], 3))) return rc;  // psidzby
      T* const perm[12] = {M[9], M[10], M[11], M[3], M[4], M[5], M[0], M[1], M[2], M[6], M[7], M[8]};
      for (int w = 0; w < 12; ++w) FR.M[w] = perm[w];
    } else {
      if ((rc = launch_x<T, XM_DERIV>(c, M[2], nullptr, 0, M[6], dmx, nullptr))) return rc;
      if ((rc = launch_y<T, YM_DERIV>(c, M[2], nullptr, 0, M[7], dmy, nullptr))) return rc;
      if ((rc = dz(M[1], nullptr, 0, 1, M[8]))) return rc;
      if ((rc = pmlb(M[8], 4))) return rc;  // psiinductionbx
      if ((rc = launch_x<T, XM_DERIV>(c, M[1], nullptr, 0, M[9], dmx, nullptr))) return rc;
      if ((rc = launch_y<T, YM_DERIV>(c, M[0], nullptr, 0, M[10], dmy, nullptr))) return rc;
      if ((rc = dz(M[0], nullptr, 0, 1, M[11]))) return rc;
      if ((rc = pmlb(M[11], 5))) return rc;  // psiinductionby
      {
        ProfScope ps(c, SPARC_KC_OTHER, 9.0 * (double)N * sizeof(T));
        hipLaunchKernelGGL((k_mhd_b<T>), gN, b256, 0, c->stream, F, N);
        LAUNCHCHK(c);
      }
      // J = curl(b): dx by, dy bz, dz bx (psidzbx), dx bz, dy bx, dz by (psidzby)
      if ((rc = launch_x<T, XM_DERIV>(c, M[1], nullptr, 0, M[6], dmx, nullptr))) return rc;
      if ((rc = launch_y<T, YM_DERIV>(c, M[2], nullptr, 0, M[7], dmy, nullptr))) return rc;
      if ((rc = dz(M[0], nullptr, 0, 1, M[8]))) return rc;
      if ((rc = pmlb(M[8], 2))) return rc;  // psidzbx
      if ((rc = launch_x<T, XM_DERIV>(c, M[2], nullptr, 0, M[9], dmx, nullptr))) return rc;
      if ((rc = launch_y<T, YM_DERIV>(c, M[0], nullptr, 0, M[10], dmy, nullptr))) return rc;
      if ((rc = dz(M[1], nullptr, 0, 1, M[11]))) return rc;
      if ((rc = pmlb(M[11], 3))) return rc;  // psidzby
    }
    const bool damp = c->cfg.damp_waves && c->damping_set;
    XArgs<T> DX{};
    for (int f = 0; f < 3; ++f) {
      DX.a[f] = F.a[3 + f];
      DX.s[f] = S(3 + f);
      DX.out[f] = F.scr[3 + f];
      DX.spec[f] = (C2<T>*)c->spec[f];
    }
    // on 256-wide x lines (k_xw256) the damping forward pass also stores the
    // stage temp v as RHS_xi (every read of the previous RHS_xi is above),
    // which k_mhd_rhs then reads
    const bool vin = damp && c->cfg.nx == 256;
    if (damp) {
      if (vin)
        for (int f = 0; f < 3; ++f) DX.cp[f] = F.scr[f];
      if ((rc = launch_xm<T, XM_DAMP_FWD>(c, 3, DX, beta, nullptr))) return rc;
      if ((rc = launch_ym<T, YM_DAMP>(c, 3, DX, 0, nullptr))) return rc;
    }
    {
      ProfScope ps(c, SPARC_KC_OTHER, (vin ? 19.0 : 25.0) * (double)N * sizeof(T));
      hipLaunchKernelGGL((k_mhd_rhs<T>), gN, b256, 0, c->stream, FR,
                         (const T*)c->prof[SPARC_PROF_SPONGE], N, vin ? 1 : 0);
      LAUNCHCHK(c);
    }
    if (damp && (rc = launch_xm<T, XM_DAMP_INV>(c, 3, DX, 0, nullptr))) return rc;
  } else if (c->cfg.option == SPARC_OPT_MHD_PML_3D) {
    // MP_MHD_PML_3D (pml.f90:80-208): six ddxyz calls, PML top and bottom on
    // six psi fields, then one pointwise pass for all eight right-hand sides
    const long N = c->N;
    const dim3 gN(cdiv(N, 256)), b256(256);
    T** M = F.M;
    auto pmla = [&](T* d, int w) -> int {
      if (c->nzpml <= 0) return SPARC_OK;
      ProfScope ps(c, SPARC_KC_OTHER, 5.0 * (double)c->NP * sizeof(T));
      hipLaunchKernelGGL((k_pml_planes<T>), gN, b256, 0, c->stream, F, d, w, N);
      LAUNCHCHK(c);
      return SPARC_OK;
    };
    auto X = [&](const T* in, const T* s, T b, T* out) {
      return launch_x<T, XM_DERIV>(c, in, s, b, out, dmx, nullptr);
    };
    auto Y = [&](const T* in, const T* s, T b, T* out) {
      return launch_y<T, YM_DERIV>(c, in, s, b, out, dmy, nullptr);
    };
    auto Z = [&](const T* in, const T* s, T b, T* out) {
      return launch_z<T, EPI_STORE>(c, F, 1, in, s, b, nullptr, 0, out);
    };
    {
      ProfScope ps(c, SPARC_KC_OTHER, 12.0 * (double)N * sizeof(T));
      hipLaunchKernelGGL((k_mhd8_flux<T>), gN, b256, 0, c->stream, F, N);
      LAUNCHCHK(c);
    }
    // ddxyz(v_x, v_y, v_z); (flux3, p, flux1); (p, flux3, flux2); (flux2, bz, bx);
    // (by, bx, p); (bz, flux1, by) -- all with bc = 1
    if ((rc = X(F.a[1], S(1), beta, F.W[0])) || (rc = Y(F.a[2], S(2), beta, F.W[1])) ||
        (rc = Z(F.a[3], S(3), beta, F.W[2])))
      return rc;
    if ((rc = X(M[2], nullptr, 0, M[3])) || (rc = Y(F.a[4], S(4), beta, F.W[3])) ||
        (rc = Z(M[0], nullptr, 0, M[4])))
      return rc;
    if ((rc = X(F.a[4], S(4), beta, F.W[4])) || (rc = Y(M[2], nullptr, 0, M[5])) ||
        (rc = Z(M[1], nullptr, 0, M[6])))
      return rc;
    if ((rc = X(M[1], nullptr, 0, M[7])) || (rc = Y(F.a[7], S(7), beta, M[8])) ||
        (rc = Z(F.a[5], S(5), beta, M[9])))
      return rc;
    if ((rc = X(F.a[6], S(6), beta, M[10])) || (rc = Y(F.a[5], S(5), beta, M[11])) ||
        (rc = Z(F.a[4], S(4), beta, F.W[5])))
      return rc;
    if ((rc = X(F.a[7], S(7), beta, F.W[6])) || (rc = Y(M[0], nullptr, 0, M[1])) ||
        (rc = Z(F.a[6], S(6), beta, M[2])))
      return rc;
    // PML_BOUNDARY_ACCOUNTING_3D: psip, psivz, psiinductionbx, psiinductionby, psidzbx, psidzby
    if ((rc = pmla(F.W[5], 1)) || (rc = pmla(F.W[2], 0)) || (rc = pmla(M[6], 4)) ||
        (rc = pmla(M[4], 5)) || (rc = pmla(M[9], 2)) || (rc = pmla(M[2], 3)))
      return rc;
    const bool damp = c->cfg.damp_waves && c->damping_set;
    XArgs<T> DX{};
    DX.edges = 1;  // zeroing of planes 1, nz precedes DAMP_VELOCITY (pml.f90:193-206)
    for (int f = 0; f < 3; ++f) {
      DX.a[f] = F.a[1 + f];
      DX.s[f] = S(1 + f);
      DX.out[f] = F.scr[1 + f];
      DX.spec[f] = (C2<T>*)c->spec[f];
    }
    if (damp) {
      if ((rc = launch_xm<T, XM_DAMP_FWD>(c, 3, DX, beta, nullptr))) return rc;
      if ((rc = launch_ym<T, YM_DAMP>(c, 3, DX, 0, nullptr))) return rc;
    }
    {
      ProfScope ps(c, SPARC_KC_OTHER, 45.0 * (double)N * sizeof(T));
      hipLaunchKernelGGL((k_mhd8_rhs<T>), gN, b256, 0, c->stream, F, N);
      LAUNCHCHK(c);
    }
    if (damp && (rc = launch_xm<T, XM_DAMP_INV>(c, 3, DX, 0, nullptr))) return rc;
  } else if (c->cfg.option == SPARC_OPT_MHD_PML_2D_DISPL) {
    // MP_MHD_PML_2D_DISPL (physics2d.f90:724-943), ny = 1
    const long N = c->N;
    const dim3 gN(cdiv(N, 256)), b256(256);
    const int nz = c->cfg.nz;
    T** M = F.M;
    auto pmla = [&](T* d, int w) -> int {
      if (c->nzpml <= 0) return SPARC_OK;
      ProfScope ps(c, SPARC_KC_OTHER, 5.0 * (double)c->NP * sizeof(T));
      hipLaunchKernelGGL((k_pml_planes<T>), gN, b256, 0, c->stream, F, d, w, N);
      LAUNCHCHK(c);
      return SPARC_OK;
    };
    auto dz = [&](const T* in, const T* s, T b, int ibc, const T* prov, T ps, T* out) -> int {
      return launch_z<T, EPI_STORE>(c, F, ibc, in, s, b, prov, ps, out);
    };
    // ddx(xi_x); ddz(xi_z, 3) with dxizdz(nz) = -dxixdx(nz)*unstretch(nz); psivz
    if ((rc = launch_x<T, XM_DERIV>(c, F.a[0], S(0), beta, F.W[0], dmx, nullptr))) return rc;
    const T us = (T)(-c->unstretch_h[nz - 1]);
    if ((rc = dz(F.a[2], S(2), beta, 3, F.W[0] + c->nxy * (nz - 1), us, F.W[2]))) return rc;
    if ((rc = pmla(F.W[2], 0))) return rc;
    {
      ProfScope ps(c, SPARC_KC_OTHER, 10.0 * (double)N * sizeof(T));
      hipLaunchKernelGGL((k_mhd2_pr<T>), gN, b256, 0, c->stream, F, N);
      LAUNCHCHK(c);
    }
    // ddx(p); ddz(p, 2) with gradp_z(:,:,1) = 0; psip
    if ((rc = launch_x<T, XM_DERIV>(c, F.W[3], nullptr, 0, M[6], dmx, nullptr))) return rc;
    if ((rc = dz(F.W[3], nullptr, 0, 2, nullptr, 0, M[7]))) return rc;
    if ((rc = pmla(M[7], 1))) return rc;
    // bx = -(ddz(flux2, 0) + psiinductionbx), bz = ddx(flux2)
    if ((rc = dz(M[0], nullptr, 0, 0, nullptr, 0, M[1]))) return rc;
    if ((rc = launch_x<T, XM_DERIV>(c, M[0], nullptr, 0, M[2], dmx, nullptr))) return rc;
    if ((rc = pmla(M[1], 4))) return rc;  // psiinductionbx
    hipLaunchKernelGGL((k_negate<T>), gN, b256, 0, c->stream, M[1], N);
    LAUNCHCHK(c);
    // dzbx = ddz(bx, 0) + psidzbx; ddx(bz)
    if ((rc = dz(M[1], nullptr, 0, 0, nullptr, 0, M[3]))) return rc;
    if ((rc = launch_x<T, XM_DERIV>(c, M[2], nullptr, 0, M[4], dmx, nullptr))) return rc;
    if ((rc = pmla(M[3], 2))) return rc;  // psidzbx
    // DAMP_VELOCITY 2-D (v_x, v_z along x) on the stage's temp v, before RHSv is written
    const bool damp = c->cfg.damp_waves && c->damping_set;
    if (damp) {
      if ((rc = launch_x<T, XM_FILTER>(c, F.a[3], S(3), beta, F.W[5], (const T*)c->dmul2,
                                       nullptr)))
        return rc;
      if ((rc = launch_x<T, XM_FILTER>(c, F.a[5], S(5), beta, F.W[6], (const T*)c->dmul2,
                                       nullptr)))
        return rc;
    }
    {
      ProfScope ps(c, SPARC_KC_OTHER, 16.0 * (double)N * sizeof(T));
      hipLaunchKernelGGL((k_mhd2_rhs<T>), gN, b256, 0, c->stream, F, N);
      LAUNCHCHK(c);
    }
    if (damp) {
      hipLaunchKernelGGL((k_sub_above_plane1<T>), gN, b256, 0, c->stream, F.scr[3],
                         (const T*)F.W[5], c->nxy, N);
      hipLaunchKernelGGL((k_sub_above_plane1<T>), gN, b256, 0, c->stream, F.scr[5],
                         (const T*)F.W[6], c->nxy, N);
      LAUNCHCHK(c);
    }
  } else if (c->cfg.option == SPARC_OPT_QUIET_PML_3D) {
    // ddxyz(v_x, v_y, v_z, bc=1) (pml.f90:27); the z solve's epilogue forms
    // div v, RHScont and RHSp into W4, W5 (pml.f90:36-53)
    if ((rc = launch_x<T, XM_DERIV>(c, F.a[1], S(1), beta, F.W[0], dmx, nullptr))) return rc;
    if ((rc = launch_y<T, YM_DERIV>(c, F.a[2], S(2), beta, F.W[1], dmy, nullptr))) return rc;
    if ((rc = launch_z<T, EPI_C2>(c, F, 1, F.a[3], S(3), beta, nullptr, 0, nullptr))) return rc;
    // ddxyz(p, bc=0) (pml.f90:29)
    if ((rc = launch_x<T, XM_DERIV>(c, F.a[4], S(4), beta, F.W[0], dmx, nullptr))) return rc;
    if ((rc = launch_y<T, YM_DERIV>(c, F.a[4], S(4), beta, F.W[1], dmy, nullptr))) return rc;
    const bool damp = c->cfg.damp_waves && c->damping_set;
    XArgs<T> DX{};
    DX.edges = 1;  // zeroing of planes 1, nz precedes DAMP_VELOCITY here (pml.f90:56-74)
    for (int f = 0; f < 3; ++f) {
      DX.a[f] = F.a[1 + f];
      DX.s[f] = S(1 + f);
      DX.out[f] = F.scr[1 + f];
      DX.spec[f] = (C2<T>*)c->spec[f];
    }
    if (damp) {
      if ((rc = launch_xm<T, XM_DAMP_FWD>(c, 3, DX, beta, nullptr))) return rc;
      if ((rc = launch_ym<T, YM_DAMP>(c, 3, DX, 0, nullptr))) return rc;
    }
    // gradp_z + RHSv; RHScont, RHSp moved from W4, W5 into scr
    if ((rc = launch_z<T, EPI_V2>(c, F, 0, F.a[4], S(4), beta, nullptr, 0, nullptr))) return rc;
    if (damp && (rc = launch_xm<T, XM_DAMP_INV>(c, 3, DX, 0, nullptr))) return rc;
  } else {
    // option 17 (physics2d.f90:959-1069)
    if ((rc = launch_x<T, XM_DERIV>(c, F.a[0], S(0), beta, F.W[0], dmx, nullptr))) return rc;
    const int nz = c->cfg.nz;
    const T us = (T)(-c->unstretch_h[nz - 1]);
    if ((rc = launch_z<T, EPI_P2D>(c, F, 3, F.a[2], S(2), beta, F.W[0] + c->nxy * (nz - 1), us,
                                   nullptr)))
      return rc;
    if ((rc = launch_x<T, XM_DERIV>(c, F.W[3], nullptr, 0, F.W[0], dmx, nullptr))) return rc;
    // ibc=2: gradp_z(:,:,1) is the caller's value, which stays zero
    if ((rc = launch_z<T, EPI_RHS2D>(c, F, 2, F.W[3], nullptr, 0, nullptr, 0, nullptr))) return rc;
  }
  return SPARC_OK;
}

// The five stages of one step and the update (step.f90:49-118 without the
// spectral filter): the part of a step whose launches do not change from step
// to step, replayed as a captured graph when cfg.use_graphs is set.
template <typename T>
static int step_body(sparc_ctx* c, Fld<T>& F) {
  const int nxy = (int)c->nxy;
  const bool is2d = is_2d(c->cfg.option);
  int rc;
  for (int s = 0; s < 5; ++s)
    if ((rc = run_stage<T>(c, F, s))) return rc;
  // a += deltat*scr with the z filter (step.f90:114-118)
  const T dt = (T)c->cfg.deltat;
  const int nfld = c->nf;  // dimfive
  {
    // all fields in one launch (blockIdx.y = field)
    // a, scr read and a written; 2-byte low words read and written
    ProfScope ps(c, SPARC_KC_UPDATE,
                 nfld * (3.0 * sizeof(T) + (c->comp ? 4.0 : 0.0)) * (double)c->N);
    UpdArgs<T> A{};
    for (int f = 0; f < nfld; ++f) {
      A.a[f] = F.a[f];
      A.scr[f] = (const T*)F.scr[f];
    }
    if (c->comp) {
      UpdArgs<float> Af{};
      for (int f = 0; f < nfld; ++f) {
        Af.a[f] = (float*)F.a[f];
        Af.lo[f] = (short*)c->alo[f];
        Af.scr[f] = (const float*)F.scr[f];
      }
      hipLaunchKernelGGL(k_update_filter_c, dim3(cdiv(nxy, 64), nfld), dim3(64), 0, c->stream,
                         Af, c->cfg.deltat, nxy, c->cfg.nz, is2d ? 6 : 2);
    } else {
      hipLaunchKernelGGL((k_update_filter<T>), dim3(cdiv(nxy, 64), nfld), dim3(64), 0,
                         c->stream, A, dt, nxy, c->cfg.nz, is2d ? 6 : 2);
    }
    LAUNCHCHK(c);
  }
  if (c->nzpml > 0)
    for (int f = 0; f < c->npv; ++f) {
      ProfScope ps(c, SPARC_KC_OTHER, 3.0 * (double)c->NP * sizeof(T));
      hipLaunchKernelGGL((k_axpy<T>), dim3(cdiv(c->NP, 256)), dim3(256), 0, c->stream,
                         F.pmlvars[f], (const T*)F.scrpml[f], dt, c->NP);
      LAUNCHCHK(c);
    }
  return SPARC_OK;
}

template <typename T>
static int step_impl(sparc_ctx* c, int t0, int nsteps) {
  Fld<T> F = make_fld<T>(c);
  const int nxy = (int)c->nxy;
  const bool is2d = is_2d(c->cfg.option);
  const T scale = (T)c->cfg.field_scale;
  int rc;
  for (int i = t0; i < t0 + nsteps; ++i) {
    // source for this step (constant over the 5 stages)
    F.src_plane = -1;
    if (c->src_kind != SPARC_SRC_NONE && i <= c->src_active_until) {
      F.src_plane = (c->cfg.mode == SPARC_ADJOINT) ? c->cfg.o_rad - 1 : c->cfg.e_rad - 1;
      if (c->src_kind == SPARC_SRC_SERIES) {
        const double sv = (i < c->src_nt) ? c->src_series[i] : 0.0;
        hipLaunchKernelGGL((k_forcing_series<T>), dim3(cdiv(nxy, 256)), dim3(256), 0, c->stream,
                           (T*)c->forcing, (const T*)c->src_mask, (T)(sv * c->cfg.field_scale),
                           nxy);
      } else {
        LagArgs la = lagrange_args(c, i);
        hipLaunchKernelGGL((k_forcing_lagrange<T>), dim3(cdiv(nxy, 256)), dim3(256), 0,
                           c->stream, (T*)c->forcing, (const T*)c->src_mask, nxy, la, scale);
      }
      LAUNCHCHK(c);
    }
    if (c->cfg.use_graphs && !c->profiling) {
      const int gi = F.src_plane >= 0 ? 1 : 0;
      if (!c->gexec[gi]) {
        // relaxed: other threads driving other contexts (the serving loop)
        // keep making stream-ordered calls meanwhile; allocations and frees
        // wait for the lock
        std::lock_guard<std::mutex> lk(g_capture_mu);
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        rc = step_body<T>(c, F);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc || e != hipSuccess) {
          if (g) (void)hipGraphDestroy(g);
          return rc ? rc : fail(c, SPARC_EHIP, "graph capture: %s", hipGetErrorString(e));
        }
        const hipError_t ei = hipGraphInstantiate(&c->gexec[gi], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
          c->gexec[gi] = nullptr;
          return fail(c, SPARC_EHIP, "graph instantiate: %s", hipGetErrorString(ei));
        }
      }
      HIPCHK(c, hipGraphLaunch(c->gexec[gi], c->stream));
    } else if ((rc = step_body<T>(c, F))) {
      return rc;
    }
    // spectral low-pass (physics.f90:792-830; physics2d.f90:563-585)
    const int ff = c->cfg.freq_filtering;
    if (ff > 0 && c->filter_set) {
      if (!is2d && i % ff == 0) {
        for (int f = 0; f < 5; ++f)
          if ((rc = launch_x<T, XM_FILTER>(c, F.a[f], nullptr, 0, F.a[f], (const T*)c->fmulx,
                                           nullptr)))
            return rc;
        for (int f = 0; f < 5; ++f)
          if ((rc = launch_y<T, YM_FILTER>(c, F.a[f], nullptr, 0, F.a[f], (const T*)c->fmuly,
                                           nullptr)))
            return rc;
      } else if (is2d && (i + 1) % ff == 0) {
        for (int f = 0; f < SPARC_NFIELDS; ++f)
          if ((rc = launch_x<T, XM_FILTER>(c, F.a[f], nullptr, 0, F.a[f], (const T*)c->fmulx,
                                           nullptr)))
            return rc;
      }
    }
  }
  return SPARC_OK;
}

extern "C" int sparc_step(sparc_ctx* c, int t0, int nsteps) {
  if (!c || !c->stream) return SPARC_EINVAL;
  if (nsteps < 0) return fail(c, SPARC_EINVAL, "nsteps < 0");
  static const int need[] = {SPARC_BG_RHO0,       SPARC_BG_RHOINV, SPARC_BG_C2RHO0,
                             SPARC_BG_GRADP0_Z,   SPARC_BG_GRADRHO0_Z, SPARC_BG_AZ,
                             SPARC_BG_BZPML};
  for (int id : need)
    if (!c->bgset[id]) return fail(c, SPARC_ESTATE, "background %d not set", id);
  if (!c->profset[SPARC_PROF_STRETCH] || !c->profset[SPARC_PROF_G] ||
      (is_2d(c->cfg.option) && !c->profset[SPARC_PROF_UNSTRETCH]))
    return fail(c, SPARC_ESTATE, "stretch/g/unstretch profiles not set");
  if (c->cfg.option != SPARC_OPT_QUIET_PML_2D_DISPL && c->cfg.damp_waves && !c->damping_set)
    return fail(c, SPARC_ESTATE, "damp_waves set but damping rates not given");
  if (c->cfg.freq_filtering > 0 && !c->filter_set)
    return fail(c, SPARC_ESTATE, "freq_filtering > 0 but filter not given");
  if ((c->cfg.option == SPARC_OPT_QUIET_PML_3D || c->cfg.option == SPARC_OPT_MHD_PML_3D) &&
      c->src_kind != SPARC_SRC_NONE && !c->profset[SPARC_PROF_SOURCE_DEP])
    return fail(c, SPARC_ESTATE, "options 2 and 8 with a source need the source_dep profile");
  if (c->cfg.option == SPARC_OPT_MHD_PML_3D)
    for (int id = SPARC_BG_BX; id <= SPARC_BG_GRADRHO0_Y; ++id)
      if (!c->bgset[id]) return fail(c, SPARC_ESTATE, "option 8 needs background %d", id);
  if (c->cfg.option == SPARC_OPT_MHD_PML_3D_DISPL) {
    for (int id = SPARC_BG_BX; id <= SPARC_BG_GRADRHO0_Y; ++id)
      if (!c->bgset[id]) return fail(c, SPARC_ESTATE, "option 10 needs background %d", id);
    if (!c->profset[SPARC_PROF_SPONGE])
      return fail(c, SPARC_ESTATE, "option 10 needs the sponge profile");
  }
  if (c->cfg.option == SPARC_OPT_MHD_PML_2D_DISPL) {
    static const int need2[] = {SPARC_BG_BX, SPARC_BG_BZ, SPARC_BG_CURLBY, SPARC_BG_GRADP0_X,
                                SPARC_BG_GRADRHO0_X};
    for (int id : need2)
      if (!c->bgset[id]) return fail(c, SPARC_ESTATE, "option 18 needs background %d", id);
    if (!c->profset[SPARC_PROF_UNSTRETCH])
      return fail(c, SPARC_ESTATE, "option 18 needs the unstretch profile");
  }
  return c->fp32 ? step_impl<float>(c, t0, nsteps) : step_impl<double>(c, t0, nsteps);
}

extern "C" int sparc_sync(sparc_ctx* c) {
  if (!c) return SPARC_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPARC_OK;
}

// ---------------------------------------------------------------------------
// state access
// ---------------------------------------------------------------------------
static void* arr_ptr(sparc_ctx* c, int arr, int field, size_t* n) {
  if ((arr == SPARC_ARR_A || arr == SPARC_ARR_SCR) && field >= 0 && field < c->nf) {
    *n = c->N;
    return arr == SPARC_ARR_A ? c->a[field] : c->scr[field];
  }
  if ((arr == SPARC_ARR_PMLVARS || arr == SPARC_ARR_SCRPML) && field >= 0 && field < c->npv) {
    *n = c->NP;
    return arr == SPARC_ARR_PMLVARS ? c->pmlvars[field] : c->scrpml[field];
  }
  if (arr == SPARC_ARR_ALO && c->comp && field >= 0 && field < c->nf) {
    *n = c->N;
    return c->alo[field];
  }
  return nullptr;
}

// raw hi words and 16-bit low words of field f of the compensated state
static int download_hilo(sparc_ctx* c, int f, std::vector<float>& hi, std::vector<short>& lo) {
  hi.resize(c->N);
  lo.resize(c->N);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpyAsync(hi.data(), c->a[f], c->N * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(lo.data(), c->alo[f], c->N * 2, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPARC_OK;
}

static int upload_raw(sparc_ctx* c, void* dst, const void* h, size_t bytes) {
  HIPCHK(c, hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));  // h is released by the caller
  return SPARC_OK;
}

extern "C" int sparc_get_state(sparc_ctx* c, int arr, int field, double* h) {
  if (!c || !h) return SPARC_EINVAL;
  size_t n;
  void* p = arr_ptr(c, arr, field, &n);
  if (!p) return fail(c, SPARC_EINVAL, "bad array %d / field %d", arr, field);
  const double s = 1.0 / c->cfg.field_scale;
  if (!(c->comp && (arr == SPARC_ARR_A || arr == SPARC_ARR_ALO))) return download(c, h, p, n, s);
  // hi + lo of the compensated state (or lo alone), decoded in FP64
  std::vector<float> hi;
  std::vector<short> lo;
  int rc = download_hilo(c, field, hi, lo);
  if (rc) return rc;
  for (size_t i = 0; i < n; ++i) {
    const double l = lo_decode(hi[i], lo[i]);
    h[i] = (arr == SPARC_ARR_A ? (double)hi[i] + l : l) * s;
  }
  return SPARC_OK;
}

extern "C" int sparc_put_state(sparc_ctx* c, int arr, int field, const double* h) {
  if (!c || !h) return SPARC_EINVAL;
  size_t n;
  void* p = arr_ptr(c, arr, field, &n);
  if (!p) return fail(c, SPARC_EINVAL, "bad array %d / field %d", arr, field);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<double> t(h, h + n);
  if (c->cfg.field_scale != 1.0)
    for (auto& v : t) v *= c->cfg.field_scale;
  if (!(c->comp && (arr == SPARC_ARR_A || arr == SPARC_ARR_ALO))) return upload(c, p, t.data(), n);
  std::vector<float> hi(n);
  std::vector<short> lo(n);
  int rc;
  if (arr == SPARC_ARR_A) {
    // split each FP64 value into hi = fl32(v) and the low word of v - hi
    for (size_t i = 0; i < n; ++i) {
      hi[i] = (float)t[i];
      lo[i] = lo_encode(hi[i], t[i]);
    }
    if ((rc = upload_raw(c, c->a[field], hi.data(), n * 4))) return rc;
  } else {
    // new low parts under the high words already held
    if ((rc = download_hilo(c, field, hi, lo))) return rc;
    for (size_t i = 0; i < n; ++i) lo[i] = lo_encode(hi[i], (double)hi[i] + t[i]);
  }
  return upload_raw(c, c->alo[field], lo.data(), n * 2);
}

extern "C" int sparc_get_plane(sparc_ctx* c, int field, int k, double* h) {
  if (!c || !h || field < 0 || field >= c->nf || k < 1 || k > c->cfg.nz)
    return fail(c, SPARC_EINVAL, "bad field/plane");
  return download(c, h, (char*)c->a[field] + (size_t)(k - 1) * c->nxy * c->ts, c->nxy,
                  1.0 / c->cfg.field_scale);
}

extern "C" int sparc_norm(sparc_ctx* c, int field, double* out) {
  if (!c || !out || field < 0 || field >= c->nf) return SPARC_EINVAL;
  std::vector<double> h(c->N);
  int rc = download(c, h.data(), c->a[field], c->N, 1.0 / c->cfg.field_scale);
  if (rc) return rc;
  double s = 0.0;
  for (double v : h) s += v * v;
  *out = sqrt(s / ((double)c->cfg.nx * (double)c->cfg.ny * (double)c->cfg.nz));
  return SPARC_OK;
}

template <typename T>
__global__ void k_gather(double* out, const T* __restrict__ f, const int32_t* __restrict__ ijk,
                         int npts, int nx, long nxy, double scale) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const long q = ijk[3 * p] + (long)nx * ijk[3 * p + 1] + nxy * (ijk[3 * p + 2] - 1);
  out[p] = (double)f[q] * scale;
}

static int sample_check(sparc_ctx* c, int field, int npts, const int32_t* ijk, const double* h) {
  if (!c || !h || !ijk || npts < 0 || field < 0 || field >= c->nf) return SPARC_EINVAL;
  for (int p = 0; p < npts; ++p)
    if (ijk[3 * p] < 0 || ijk[3 * p] >= c->cfg.nx || ijk[3 * p + 1] < 0 ||
        ijk[3 * p + 1] >= c->cfg.ny || ijk[3 * p + 2] < 1 || ijk[3 * p + 2] > c->cfg.nz)
      return fail(c, SPARC_EINVAL, "sample point %d out of the grid", p);
  return SPARC_OK;
}

// Stages ijk into dijk, gathers into dout and copies the values to h, on the
// context's stream; returns after the copy is done.
static int sample_on(sparc_ctx* c, int field, int npts, const int32_t* ijk, double* h,
                     int32_t* dijk, double* dout) {
  HIPCHK(c, hipMemcpyAsync(dijk, ijk, 12 * (size_t)npts, hipMemcpyHostToDevice, c->stream));
  const double sc = 1.0 / c->cfg.field_scale;
  if (c->fp32)
    hipLaunchKernelGGL(k_gather<float>, dim3(cdiv(npts, 64)), dim3(64), 0, c->stream, dout,
                       (const float*)c->a[field], dijk, npts, c->cfg.nx, c->nxy, sc);
  else
    hipLaunchKernelGGL(k_gather<double>, dim3(cdiv(npts, 64)), dim3(64), 0, c->stream, dout,
                       (const double*)c->a[field], dijk, npts, c->cfg.nx, c->nxy, sc);
  LAUNCHCHK(c);
  HIPCHK(c, hipMemcpyAsync(h, dout, 8 * (size_t)npts, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPARC_OK;
}

extern "C" int sparc_sample(sparc_ctx* c, int field, int npts, const int32_t* ijk, double* h) {
  int rc = sample_check(c, field, npts, ijk, h);
  if (rc || npts == 0) return rc;
  int32_t* dijk;
  double* dout;
  HIPCHK(c, hipMallocAsync((void**)&dijk, 12 * (size_t)npts, c->stream));
  HIPCHK(c, hipMallocAsync((void**)&dout, 8 * (size_t)npts, c->stream));
  rc = sample_on(c, field, npts, ijk, h, dijk, dout);
  HIPCHK(c, hipFreeAsync(dijk, c->stream));
  HIPCHK(c, hipFreeAsync(dout, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return rc;
}

extern "C" int sparc_sample_noalloc(sparc_ctx* c, int field, int npts, const int32_t* ijk,
                                    double* h) {
  int rc = sample_check(c, field, npts, ijk, h);
  if (rc || npts == 0) return rc;
  if (npts > SPARC_SAMPLE_NOALLOC_MAX)
    return fail(c, SPARC_EINVAL, "sparc_sample_noalloc: %d points exceed the maximum %d", npts,
                SPARC_SAMPLE_NOALLOC_MAX);
  if (!c->scratch) return fail(c, SPARC_ESTATE, "sample scratch not set up");
  // one call at a time per process: every context on a device shares its array
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  int32_t* dijk = (int32_t*)c->scratch;
  double* dout = (double*)(c->scratch + 12 * (size_t)SPARC_SAMPLE_NOALLOC_MAX);
  return sample_on(c, field, npts, ijk, h, dijk, dout);
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// Cross-context ordering without blocking the host: `consumer` waits for the
// work already queued on `producer`, and afterwards `producer` waits for what
// `consumer` has queued that reads the producer's memory, so a later store
// into the producer's snapshot bank cannot overtake that read.
static int order_before(sparc_ctx* consumer, const sparc_ctx* producer) {
  HIPCHK(consumer, hipEventRecord(producer->xev, producer->stream));
  HIPCHK(consumer, hipStreamWaitEvent(consumer->stream, producer->xev, 0));
  return SPARC_OK;
}
static int order_after(sparc_ctx* consumer, const sparc_ctx* producer) {
  HIPCHK(consumer, hipEventRecord(consumer->xev, consumer->stream));
  HIPCHK(consumer, hipStreamWaitEvent(producer->stream, consumer->xev, 0));
  return SPARC_OK;
}

// fields per snapshot slot: xi, acc (+ v when the flow kernels are on)
static int snap_fields(const sparc_ctx* c) { return c->cfg.flow_kernels ? 9 : 6; }

extern "C" int sparc_snapshots_alloc(sparc_ctx* c, int nslots) {
  if (!c || nslots < 0) return SPARC_EINVAL;
  if (nslots <= c->nslots) return SPARC_OK;
  void* p;
  int rc = dalloc(c, &p, (size_t)nslots * snap_fields(c) * c->N * c->ts);
  if (rc) return rc;
  c->snap = p;
  c->nslots = nslots;
  c->snap_scale = c->cfg.field_scale;
  return SPARC_OK;
}

extern "C" int sparc_snapshot_store(sparc_ctx* c, int slot) {
  if (!c || slot < 0 || slot >= c->nslots) return fail(c, SPARC_EINVAL, "bad slot %d", slot);
  if (c->cfg.option == SPARC_OPT_QUIET_PML_3D || c->cfg.option == SPARC_OPT_MHD_PML_3D)
    return fail(c, SPARC_ESTATE, "snapshots hold xi and acc: displacement options only");
  char* s = (char*)c->snap + (size_t)slot * snap_fields(c) * c->N * c->ts;
  const size_t fb = c->N * c->ts;
  for (int f = 0; f < 3; ++f) {
    HIPCHK(c, hipMemcpyAsync(s + f * fb, c->a[f], fb, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s + (3 + f) * fb, c->scr[3 + f], fb, hipMemcpyDeviceToDevice,
                             c->stream));
    if (c->cfg.flow_kernels)
      HIPCHK(c, hipMemcpyAsync(s + (6 + f) * fb, c->a[3 + f], fb, hipMemcpyDeviceToDevice,
                               c->stream));
  }
  c->snap_scale = c->cfg.field_scale;
  return SPARC_OK;
}

extern "C" int sparc_snapshots_copy(sparc_ctx* dst, const sparc_ctx* src) {
  if (!dst || !src) return SPARC_EINVAL;
  if (dst->N != src->N || dst->fp32 != src->fp32 || snap_fields(dst) != snap_fields(src))
    return fail(dst, SPARC_EINVAL, "snapshot banks differ in shape, precision or fields");
  int rc = sparc_snapshots_alloc(dst, src->nslots);
  if (rc) return rc;
  if ((rc = order_before(dst, src))) return rc;
  HIPCHK(dst, hipMemcpyAsync(dst->snap, src->snap,
                             (size_t)src->nslots * snap_fields(src) * src->N * src->ts,
                             hipMemcpyDeviceToDevice, dst->stream));
  dst->snap_scale = src->snap_scale;
  return order_after(dst, src);
}

extern "C" int sparc_kernel_begin(sparc_ctx* c, const double* grav, int nz) {
  if (!c || !grav || nz != c->cfg.nz) return fail(c, SPARC_EINVAL, "grav needs nz values");
  if (c->cfg.option == SPARC_OPT_QUIET_PML_3D || c->cfg.option == SPARC_OPT_MHD_PML_3D)
    return fail(c, SPARC_ESTATE, "kernels need the displacement formulation");
  int rc;
  const bool mag = c->cfg.option == SPARC_OPT_MHD_PML_3D_DISPL ||
                   c->cfg.option == SPARC_OPT_MHD_PML_2D_DISPL;
  auto needed = [&](int q) {
    return q < SPARC_KERN_VX || (c->cfg.flow_kernels && q < SPARC_KERN_BX) ||
           (mag && q >= SPARC_KERN_BX);
  };
  if (!c->kern[0]) {
    for (int q = 0; q < SPARC_KERN_COUNT; ++q)
      if (needed(q) && (rc = dalloc(c, (void**)&c->kern[q], c->N * 8))) return rc;
    if ((rc = dalloc(c, (void**)&c->grav, nz * 8))) return rc;
  }
  for (int q = 0; q < SPARC_KERN_COUNT; ++q)
    if (c->kern[q]) HIPCHK(c, hipMemsetAsync(c->kern[q], 0, c->N * 8, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->grav, grav, nz * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPARC_OK;
}

// MAGNETIC_KERNELS (kernels.f90:336-430) for one snapshot pair: six bilinear
// terms in xi_f and v_a built from cross products with B0, curl B0 and seven
// CURL_KERNs (their z derivatives with ibc 0 + FILTER_VAR_Z: bcz is an
// unassigned module integer, initialize.f90:145).  Work: W0..W6, M0..M11.
template <typename T>
static int magnetic_kernels(sparc_ctx* c, Fld<T>& F, const T* fx, const T* fy, const T* fz,
                            const T* ax, const T* ay, const T* az, double scale) {
  const long N = c->N;
  const dim3 gN(cdiv(N, 256)), b256(256);
  const bool has_y = c->cfg.ny > 1;
  const T* dmx = (const T*)c->dmulx;
  const T* dmy = (const T*)c->dmuly;
  double *kx = c->kern[SPARC_KERN_BX], *ky = c->kern[SPARC_KERN_BY], *kz = c->kern[SPARC_KERN_BZ];
  T** M = F.M;
  T** W = F.W;
  int rc;
  auto arr = [](const T* x, const T* y, const T* z, const T* s = nullptr) {
    Vec3<T> v{};
    v.p[0] = x;
    v.p[1] = y;
    v.p[2] = z;
    v.bg = -1;
    v.s = s;
    return v;
  };
  auto bgt = [](int id) {
    Vec3<T> v{};
    v.bg = id;
    return v;
  };
  auto cross = [&](const Vec3<T>& a, const Vec3<T>& b, T* ox, T* oy, T* oz) {
    hipLaunchKernelGGL((k_cross3<T, false>), gN, b256, 0, c->stream, F, a, b, ox, oy, oz,
                       (double*)nullptr, (double*)nullptr, (double*)nullptr, 0.0, N);
  };
  auto acc_cross = [&](const Vec3<T>& a, const Vec3<T>& b) {
    hipLaunchKernelGGL((k_cross3<T, true>), gN, b256, 0, c->stream, F, a, b, (T*)nullptr,
                       (T*)nullptr, (T*)nullptr, kx, ky, kz, scale, N);
  };
  auto acc = [&](const Vec3<T>& v) {
    hipLaunchKernelGGL((k_acc_vec<T>), gN, b256, 0, c->stream, F, v, kx, ky, kz, scale, N);
  };
  auto dzk = [&](const T* in, T* out) -> int {  // ddzkern with ibc 0
    int r = launch_z<T, EPI_STORE>(c, F, 0, in, nullptr, 0, nullptr, 0, out);
    if (r) return r;
    zfilter_only<T>(c, out);
    return SPARC_OK;
  };
  auto curl = [&](const T* x, const T* y, const T* z, T* ox, T* oy, T* oz) -> int {
    int r;
    if ((r = launch_x<T, XM_DERIV>(c, y, nullptr, 0, W[0], dmx, nullptr))) return r;
    if (has_y && (r = launch_y<T, YM_DERIV>(c, x, nullptr, 0, W[1], dmy, nullptr))) return r;
    if (has_y && (r = launch_y<T, YM_DERIV>(c, z, nullptr, 0, W[2], dmy, nullptr))) return r;
    if ((r = dzk(y, W[3]))) return r;
    if ((r = dzk(x, W[4]))) return r;
    if ((r = launch_x<T, XM_DERIV>(c, z, nullptr, 0, W[5], dmx, nullptr))) return r;
    hipLaunchKernelGGL((k_curl_combine<T>), gN, b256, 0, c->stream, ox, oy, oz,
                       (const T*)W[0], has_y ? (const T*)W[1] : nullptr,
                       has_y ? (const T*)W[2] : nullptr, (const T*)W[3], (const T*)W[4],
                       (const T*)W[5], N);
    return SPARC_OK;
  };
  const Vec3<T> xi = arr(fx, fy, fz), va = arr(ax, ay, az);
  const Vec3<T> B0 = bgt(SPARC_BG_BX), J0 = bgt(SPARC_BG_CURLBX);
  // v_a x curl(curl(xi x B0)); fb = curl(xi x B0) kept in M3..M5
  cross(xi, B0, M[0], M[1], M[2]);
  if ((rc = curl(M[0], M[1], M[2], M[3], M[4], M[5]))) return rc;
  if ((rc = curl(M[3], M[4], M[5], M[6], M[7], M[8]))) return rc;
  acc_cross(va, arr(M[6], M[7], M[8]));
  // curl(fb x v_a)
  cross(arr(M[3], M[4], M[5]), va, M[0], M[1], M[2]);
  if ((rc = curl(M[0], M[1], M[2], M[9], M[10], M[11]))) return rc;
  acc(arr(M[9], M[10], M[11]));
  // xi x curl(curl(v_a x B0))
  cross(va, B0, M[0], M[1], M[2]);
  if ((rc = curl(M[0], M[1], M[2], M[3], M[4], M[5]))) return rc;
  if ((rc = curl(M[3], M[4], M[5], M[6], M[7], M[8]))) return rc;
  acc_cross(xi, arr(M[6], M[7], M[8]));
  // curl(v_a x J0) x xi
  cross(va, J0, M[0], M[1], M[2]);
  if ((rc = curl(M[0], M[1], M[2], M[3], M[4], M[5]))) return rc;
  acc_cross(arr(M[3], M[4], M[5]), xi);
  // (xi x J0) div(v_a), div as COMPUTE_DEL (no y part when ny <= 4)
  if ((rc = launch_x<T, XM_DERIV>(c, ax, nullptr, 0, W[0], dmx, nullptr))) return rc;
  const bool div_y = c->cfg.ny > 4;
  if (div_y && (rc = launch_y<T, YM_DERIV>(c, ay, nullptr, 0, W[1], dmy, nullptr))) return rc;
  if ((rc = launch_z<T, EPI_STORE>(c, F, 1, az, nullptr, 0, nullptr, 0, W[2]))) return rc;
  zfilter_only<T>(c, W[2]);
  hipLaunchKernelGGL((k_sum3<T>), gN, b256, 0, c->stream, W[6], (const T*)W[0],
                     div_y ? (const T*)W[1] : nullptr, (const T*)W[2], N);
  cross(xi, J0, M[0], M[1], M[2]);
  acc(arr(M[0], M[1], M[2], W[6]));
  // curl(B0 x (xi div(v_a)))
  cross(B0, arr(fx, fy, fz, W[6]), M[0], M[1], M[2]);
  if ((rc = curl(M[0], M[1], M[2], M[3], M[4], M[5]))) return rc;
  acc(arr(M[3], M[4], M[5]));
  LAUNCHCHK(c);
  return SPARC_OK;
}

template <typename T>
static int kernel_acc_impl(sparc_ctx* c, const sparc_ctx* src, int slot) {
  const long N = c->N, nxy = c->nxy;
  Fld<T> F = make_fld<T>(c);
  T* s = (T*)src->snap + (size_t)slot * snap_fields(src) * N;
  const T *fxx = s, *fxy = s + N, *fxz = s + 2 * N, *fax = s + 3 * N, *fay = s + 4 * N,
          *faz = s + 5 * N;
  const T *avx = F.a[3], *avy = F.a[4], *avz = F.a[5];
  const T* dmx = (const T*)c->dmulx;
  const T* dmy = (const T*)c->dmuly;
  const bool has_y = c->cfg.ny > 4;
  int rc;
  // COMPUTE_DIV (kernels.f90:798-816), z part through ddzkern + FILTER_VAR_Z
  auto div = [&](const T* fx, const T* fy, const T* fz, T* outdiv) -> int {
    int r;
    if ((r = launch_x<T, XM_DERIV>(c, fx, nullptr, 0, F.W[0], dmx, nullptr))) return r;
    if (has_y && (r = launch_y<T, YM_DERIV>(c, fy, nullptr, 0, F.W[1], dmy, nullptr))) return r;
    if ((r = launch_z<T, EPI_STORE>(c, F, 1, fz, nullptr, 0, nullptr, 0, F.W[2]))) return r;
    zfilter_only<T>(c, F.W[2]);
    LAUNCHCHK(c);
    hipLaunchKernelGGL((k_sum3<T>), dim3(cdiv(N, 256)), dim3(256), 0, c->stream, outdiv,
                       (const T*)F.W[0], has_y ? (const T*)F.W[1] : nullptr, (const T*)F.W[2], N);
    LAUNCHCHK(c);
    return SPARC_OK;
  };
  if ((rc = div(fxx, fxy, fxz, F.W[5]))) return rc;
  if ((rc = div(avx, avy, avz, F.W[6]))) return rc;
  // grad(a_vel_z * grav) (kernels.f90:294-301)
  hipLaunchKernelGGL((k_scale_z<T>), dim3(cdiv(N, 256)), dim3(256), 0, c->stream, F.W[3], avz,
                     (const double*)c->grav, nxy, N);
  LAUNCHCHK(c);
  if ((rc = launch_x<T, XM_DERIV>(c, F.W[3], nullptr, 0, F.W[0], dmx, nullptr))) return rc;
  const bool has_y2 = c->cfg.ny > 1;
  if (has_y2 && (rc = launch_y<T, YM_DERIV>(c, F.W[3], nullptr, 0, F.W[1], dmy, nullptr)))
    return rc;
  if ((rc = launch_z<T, EPI_STORE>(c, F, 1, F.W[3], nullptr, 0, nullptr, 0, F.W[2]))) return rc;
  zfilter_only<T>(c, F.W[2]);
  LAUNCHCHK(c);
  const double sff = 1.0 / src->snap_scale, sfa = 1.0 / c->cfg.field_scale;
  hipLaunchKernelGGL((k_kern_accum<T>), dim3(cdiv(N, 256)), dim3(256), 0, c->stream,
                     c->kern[SPARC_KERN_HESSIAN], c->kern[SPARC_KERN_C2], c->kern[SPARC_KERN_RHO],
                     fxx, fxy, fxz, fax, fay, faz, avx, avy, avz, (const T*)F.W[5],
                     (const T*)F.W[6], (const T*)F.W[0], has_y2 ? (const T*)F.W[1] : nullptr,
                     (const T*)F.W[2], (const T*)c->bg[SPARC_BG_C2], c->bg3d[SPARC_BG_C2],
                     (const double*)c->grav, sff, sfa, nxy, N);
  LAUNCHCHK(c);
  if (c->cfg.option == SPARC_OPT_MHD_PML_3D_DISPL || c->cfg.option == SPARC_OPT_MHD_PML_2D_DISPL)
    if ((rc = magnetic_kernels<T>(c, F, fxx, fxy, fxz, avx, avy, avz, sff * sfa))) return rc;
  if (c->cfg.flow_kernels) {
    // COMPUTE_DEL + kernelv (kernels.f90:317-332, 821-841): for each forward
    // velocity component j, K_v(i) += d_i v_f,j * v_a,j (no y part when ny <= 4)
    const T* av[3] = {avx, avy, avz};
    for (int j = 0; j < 3; ++j) {
      const T* fv = s + (size_t)(6 + j) * N;
      if ((rc = launch_x<T, XM_DERIV>(c, fv, nullptr, 0, F.W[0], dmx, nullptr))) return rc;
      if (has_y && (rc = launch_y<T, YM_DERIV>(c, fv, nullptr, 0, F.W[1], dmy, nullptr)))
        return rc;
      if ((rc = launch_z<T, EPI_STORE>(c, F, 1, fv, nullptr, 0, nullptr, 0, F.W[2]))) return rc;
      zfilter_only<T>(c, F.W[2]);
      LAUNCHCHK(c);
      hipLaunchKernelGGL((k_kern_flow<T>), dim3(cdiv(N, 256)), dim3(256), 0, c->stream,
                         c->kern[SPARC_KERN_VX], c->kern[SPARC_KERN_VY], c->kern[SPARC_KERN_VZ],
                         (const T*)F.W[0], has_y ? (const T*)F.W[1] : nullptr,
                         (const T*)F.W[2], av[j], sff, sfa, N);
      LAUNCHCHK(c);
    }
  }
  return SPARC_OK;
}

extern "C" int sparc_kernel_accumulate_from(sparc_ctx* c, const sparc_ctx* src, int slot) {
  if (!c || !src) return SPARC_EINVAL;
  if (slot < 0 || slot >= src->nslots) return fail(c, SPARC_EINVAL, "bad slot %d", slot);
  if (src->N != c->N || src->fp32 != c->fp32)
    return fail(c, SPARC_EINVAL, "snapshot bank differs in shape or precision");
  if (c->cfg.flow_kernels && !src->cfg.flow_kernels)
    return fail(c, SPARC_EINVAL, "flow kernels need snapshots that keep v (flow_kernels)");
  if (!c->kern[0]) return fail(c, SPARC_ESTATE, "sparc_kernel_begin not called");
  if (!c->bgset[SPARC_BG_C2]) return fail(c, SPARC_ESTATE, "background C2 not set");
  int rc;
  if (src != c && (rc = order_before(c, src))) return rc;
  rc = c->fp32 ? kernel_acc_impl<float>(c, src, slot) : kernel_acc_impl<double>(c, src, slot);
  if (rc || src == c) return rc;
  return order_after(c, src);
}

extern "C" int sparc_kernel_accumulate(sparc_ctx* c, int slot) {
  return sparc_kernel_accumulate_from(c, c, slot);
}

static size_t ckpt_slot_bytes(const sparc_ctx* c) {
  return c->nf * (size_t)c->N * c->ts + (c->comp ? c->nf * (size_t)c->N * 2 : 0) +
         c->npv * (size_t)c->NP * c->ts;
}

extern "C" int sparc_checkpoints_alloc(sparc_ctx* c, int nslots) {
  if (!c || nslots < 0) return SPARC_EINVAL;
  if (nslots <= c->nckpt) return SPARC_OK;
  void* p;
  int rc = dalloc(c, &p, (size_t)nslots * ckpt_slot_bytes(c));
  if (rc) return rc;
  c->ckpt = p;
  c->nckpt = nslots;
  return SPARC_OK;
}

static int ckpt_copy(sparc_ctx* c, int slot, bool save) {
  if (!c || slot < 0 || slot >= c->nckpt) return fail(c, SPARC_EINVAL, "bad checkpoint %d", slot);
  // slot layout: a(:,:,:,f) for every field, their 16-bit low words when the
  // state is compensated, then pmlvars
  char* pp = (char*)c->ckpt + (size_t)slot * ckpt_slot_bytes(c);
  auto move = [&](void* dev, size_t bytes) -> int {
    void* d = save ? (void*)pp : dev;
    const void* s = save ? (const void*)dev : (const void*)pp;
    HIPCHK(c, hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, c->stream));
    pp += bytes;
    return SPARC_OK;
  };
  int rc;
  for (int f = 0; f < c->nf; ++f)
    if ((rc = move(c->a[f], c->N * c->ts))) return rc;
  for (int f = 0; c->comp && f < c->nf; ++f)
    if ((rc = move(c->alo[f], c->N * 2))) return rc;
  for (int f = 0; f < c->npv; ++f)
    if ((rc = move(c->pmlvars[f], c->NP * c->ts))) return rc;
  return SPARC_OK;
}

extern "C" int sparc_checkpoint_save(sparc_ctx* c, int slot) { return ckpt_copy(c, slot, true); }
extern "C" int sparc_checkpoint_load(sparc_ctx* c, int slot) { return ckpt_copy(c, slot, false); }

// kernel_vectorpsi (kernels.f90:508-509): CURL_KERN(K_Bx, 0, K_Bz), y component,
// d_z with ibc 0 + FILTER_VAR_Z, at the state precision over the FP64 sums.
template <typename T>
static int psi_kernel(sparc_ctx* c) {
  Fld<T> F = make_fld<T>(c);
  const long N = c->N;
  const dim3 gN(cdiv(N, 256)), b256(256);
  T** W = F.W;
  int rc;
  hipLaunchKernelGGL((k_from_double<T>), gN, b256, 0, c->stream, W[0],
                     (const double*)c->kern[SPARC_KERN_BX], N);
  hipLaunchKernelGGL((k_from_double<T>), gN, b256, 0, c->stream, W[1],
                     (const double*)c->kern[SPARC_KERN_BZ], N);
  if ((rc = launch_z<T, EPI_STORE>(c, F, 0, W[0], nullptr, 0, nullptr, 0, W[2]))) return rc;
  zfilter_only<T>(c, W[2]);
  if ((rc = launch_x<T, XM_DERIV>(c, W[1], nullptr, 0, W[3], (const T*)c->dmulx, nullptr)))
    return rc;
  hipLaunchKernelGGL((k_curl_y<T>), gN, b256, 0, c->stream, c->kern[SPARC_KERN_PSI],
                     (const T*)W[2], (const T*)W[3], N);
  HIPCHK(c, hipGetLastError());
  return SPARC_OK;
}

extern "C" int sparc_kernel_get(sparc_ctx* c, int which, double* h) {
  if (!c || !h || which < 0 || which >= SPARC_KERN_COUNT || !c->kern[0] || !c->kern[which])
    return fail(c, SPARC_EINVAL, "bad kernel id, kernels not begun, or flow kernels off");
  if (which == SPARC_KERN_PSI) {
    const int rc = c->fp32 ? psi_kernel<float>(c) : psi_kernel<double>(c);
    if (rc) return rc;
  }
  HIPCHK(c, hipMemcpyAsync(h, c->kern[which], c->N * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPARC_OK;
}

// ---------------------------------------------------------------------------
// single operators
// ---------------------------------------------------------------------------
template <typename T>
static int apply_op_impl(sparc_ctx* c, int op, int ibc, const double* in, double* out) {
  const long N = c->N, nxy = c->nxy;
  const int nz = c->cfg.nz;
  Fld<T> F = make_fld<T>(c);
  T* src = F.W[5];
  T* dst = F.W[6];
  int rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if ((rc = upload(c, src, in, N))) return rc;
  switch (op) {
    case SPARC_OP_DDX:
      rc = launch_x<T, XM_DERIV>(c, src, nullptr, 0, dst, (const T*)c->dmulx, nullptr);
      break;
    case SPARC_OP_DDY:
      if (c->cfg.ny < 2) return fail(c, SPARC_EINVAL, "DDY needs ny >= 2");
      rc = launch_y<T, YM_DERIV>(c, src, nullptr, 0, dst, (const T*)c->dmuly, nullptr);
      break;
    case SPARC_OP_DDZ: {
      if (ibc < 0 || ibc > 3) return fail(c, SPARC_EINVAL, "ibc %d not supported (0..3)", ibc);
      if (!c->profset[SPARC_PROF_STRETCH]) return fail(c, SPARC_ESTATE, "stretch not set");
      const T* prov = nullptr;
      if (ibc == 2 || ibc == 3) {
        if ((rc = upload(c, F.W[4], out, N))) return rc;
        prov = F.W[4] + (ibc == 2 ? 0 : nxy * (nz - 1));
      }
      rc = launch_z<T, EPI_STORE>(c, F, ibc, src, nullptr, 0, prov, (T)1, dst);
      break;
    }
    case SPARC_OP_ZFILTER:
      HIPCHK(c, hipMemcpyAsync(dst, src, N * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
      zfilter_only<T>(c, dst);
      LAUNCHCHK(c);
      rc = SPARC_OK;
      break;
    case SPARC_OP_DAMP:
      if (!c->spec[0] || !c->damping_set)
        return fail(c, SPARC_ESTATE, "DAMP needs option 4, damp_waves and damping rates");
      HIPCHK(c, hipMemsetAsync(dst, 0, N * sizeof(T), c->stream));
      if ((rc = launch_x<T, XM_DAMP_FWD>(c, src, nullptr, 0, nullptr, nullptr, (C2<T>*)c->spec[0])))
        return rc;
      if ((rc = launch_y<T, YM_DAMP>(c, nullptr, nullptr, 0, nullptr, nullptr, (C2<T>*)c->spec[0])))
        return rc;
      rc = launch_x<T, XM_DAMP_INV>(c, nullptr, nullptr, 0, dst, nullptr, (C2<T>*)c->spec[0]);
      break;
    case SPARC_OP_LOWPASS_X:
      if (!c->filter_set) return fail(c, SPARC_ESTATE, "filter not set");
      rc = launch_x<T, XM_FILTER>(c, src, nullptr, 0, dst, (const T*)c->fmulx, nullptr);
      break;
    case SPARC_OP_LOWPASS_Y:
      if (!c->filter_set || c->cfg.ny < 2) return fail(c, SPARC_ESTATE, "filter not set / ny < 2");
      rc = launch_y<T, YM_FILTER>(c, src, nullptr, 0, dst, (const T*)c->fmuly, nullptr);
      break;
    default:
      return fail(c, SPARC_EINVAL, "unknown op %d", op);
  }
  if (rc) return rc;
  // DAMP leaves -term in dst
  return download(c, out, dst, N, op == SPARC_OP_DAMP ? -1.0 : 1.0);
}

extern "C" int sparc_apply_op(sparc_ctx* c, int op, int ibc, const double* in, double* out) {
  if (!c || !in || !out) return SPARC_EINVAL;
  return c->fp32 ? apply_op_impl<float>(c, op, ibc, in, out)
                 : apply_op_impl<double>(c, op, ibc, in, out);
}

// ---------------------------------------------------------------------------
// instrumentation
// ---------------------------------------------------------------------------
static int prof_collect(sparc_ctx* c) {
  if (c->recs.empty()) return SPARC_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto& r : c->recs) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, r.e0, r.e1));
    c->pms[r.cls] += ms;
    c->pbytes[r.cls] += r.bytes;
    c->plaunch[r.cls] += 1;
    c->pool.push_back(r.e0);
    c->pool.push_back(r.e1);
  }
  c->recs.clear();
  return SPARC_OK;
}

extern "C" int sparc_profile_enable(sparc_ctx* c, int on) {
  if (!c || !c->stream) return SPARC_EINVAL;
  int rc = prof_collect(c);
  if (rc) return rc;
  if (on) {
    for (int k = 0; k < SPARC_KC_COUNT; ++k) {
      c->pms[k] = 0;
      c->pbytes[k] = 0;
      c->plaunch[k] = 0;
    }
  }
  c->profiling = on != 0;
  return SPARC_OK;
}

extern "C" int sparc_profile_read(sparc_ctx* c, int cls, double* ms, double* bytes,
                                  int64_t* launches) {
  if (!c || cls < 0 || cls >= SPARC_KC_COUNT) return SPARC_EINVAL;
  int rc = prof_collect(c);
  if (rc) return rc;
  if (ms) *ms = c->pms[cls];
  if (bytes) *bytes = c->pbytes[cls];
  if (launches) *launches = c->plaunch[cls];
  return SPARC_OK;
}
