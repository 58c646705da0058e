// ocfd_api.hip -- implementation of the C ABI declared in include/ocfd_hip.h.
//
// One handle owns one block's device arrays (the reference's flow_data module,
// OCFD_parameters.f90:99-108, plus the work arrays of the inviscid and viscous
// modules) on one GPU and one HIP stream.  Host arrays are only copied in/out.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <functional>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ocfd_hip.h"
#include "ocfd_kernels.h"

using namespace ocfd;

// Scratch of the entry points that allocate nothing (see noalloc_scratch)
constexpr int64_t NOALLOC_SCRATCH_DOUBLES = 2 * 1024 * 1024;  // 16 MiB
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
thread_local std::string g_err = "";

int fail(const std::string &msg) {
  g_err = msg;
  return -1;
}

struct ProfRec {
  std::string name;
  hipEvent_t a, b;
};
}  // namespace

struct ocfd_handle {
  ocfd_config cfg;
  int device = 0;
  hipStream_t st = nullptr;
  Block B;
  double *met[M_N] = {};
  double *fbuf = nullptr;
  double *work = nullptr;      // reduction partials
  double *wout = nullptr;      // reduction results (32 doubles)
  int *neg = nullptr;          // [0] stage count, [1..64] cells, [65] total, [66] max per stage
  unsigned long long *counts = nullptr;
  int ibc = 0;
  double tw = 0.0;
  int ghost[6] = {};
  int istep = 0;
  double tt = 0.0;
  bool prof = false;
  bool use_graph = true;           // replay whole steps from a captured hipGraph
  int tile[3] = {0, 0, 0};         // sweep tile shape per axis (SweepTile)
  BLArgs bl = {};                  // boundary-layer / swept-corner data (ibc 2, 3)
  bool bl_set = false;
  std::vector<void *> bl_mem;      // its device arrays
  // one time step (three stages), captured on first use; [tpar + 2 uvpar]
  // starts with the temperature buffers in parity tpar, which every stage
  // flips, and the u, v, w sets in parity uvpar, which every fused stage flips
  hipGraphExec_t gexec[4] = {nullptr, nullptr, nullptr, nullptr};
  int tpar = 0;
  int uvpar = 0;
  bool fused = false;              // stages run k_visc_fused (Block::u2, v2, w2 allocated)
  // per-routine calls merged into the whole-step stage kernels
  // (ocfd_set_stage_fusion); pend holds what the stage so far left to do
  bool stage_fusion = false;
  int pend = 0;
  std::vector<ProfRec> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, long long>> acc;
  size_t halo_bytes() const { return (size_t)B.g.nh * sizeof(double); }
  size_t f_bytes() const { return (size_t)5 * B.g.ni * sizeof(double); }
};

namespace {

#define HIPCHK(expr)                                                                                   \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess) return fail(std::string(#expr) + ": " + hipGetErrorString(e_));            \
  } while (0)

// ---------------------------------------------------------------------------
// Scratch of the entry points that allocate nothing (ocfd_show_flow_msg_noalloc):
// a file-scope device array that every GPU the library runs on holds, at its
// own address, until the process ends.  Host code must not pass `scratch`
// itself to a kernel (in host code it names a host-side stand-in); it uses
// the device address hipGetSymbolAddress returns with that device current,
// looked up once per device and kept in g_scratch_ptr.  Calls on one GPU
// share the array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// 3 partial sums per workgroup of the show_flow_msg reduction, then the 3 totals
inline int64_t noalloc_flow_msg_doubles(const Geom &g) { return 3 * flow_msg_partials(g) + 3; }

// the device address of `scratch` on `dev`, looked up with dev current on
// the first call for that device
int noalloc_scratch(int dev, double **out) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail("no-allocation scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return fail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

hipEvent_t ev_get(ocfd_handle *h) {
  if (!h->pool.empty()) {
    hipEvent_t e = h->pool.back();
    h->pool.pop_back();
    return e;
  }
  // timing only: no system-scope fence (cache writeback) around the kernel,
  // which would add to the time measured; flush_prof synchronizes the stream
  // before reading the events
  hipEvent_t e;
  (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
  return e;
}

// run one kernel launch, timing it on the handle's stream when profiling
template <typename F>
int launch(ocfd_handle *h, const char *name, F &&fn) {
  hipEvent_t a = nullptr, b = nullptr;
  if (h->prof) {
    a = ev_get(h);
    b = ev_get(h);
    (void)hipEventRecord(a, h->st);
  }
  hipError_t e = fn();
  if (h->prof) {
    (void)hipEventRecord(b, h->st);
    h->pending.push_back({name, a, b});
  }
  if (e != hipSuccess) return fail(std::string(name) + ": " + hipGetErrorString(e));
  return 0;
}

int flush_prof(ocfd_handle *h) {
  if (h->pending.empty()) return 0;
  HIPCHK(hipStreamSynchronize(h->st));
  for (auto &r : h->pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    auto &slot = h->acc[r.name];
    slot.first += ms;
    slot.second += 1;
    h->pool.push_back(r.a);
    h->pool.push_back(r.b);
  }
  h->pending.clear();
  return 0;
}

bool valid_scheme(int s) { return s == 1 || s == 2 || s == 3 || s == 4 || s == 10; }

int set_params(ocfd_handle *h) {
  const ocfd_config &c = h->cfg;
  if (c.lap != LAP) return fail("LAP must be 4");
  if (c.nx < 8 || c.ny < 8 || c.nz < 8) return fail("each block extent must be >= 8");
  if (!valid_scheme(c.scheme_invis)) return fail("Scheme_Invis not supported");
  if (c.scheme_vis != SCH_CD6 && c.scheme_vis != SCH_CD8) return fail("Scheme_Vis must be CD6 or CD8");
  if (c.flux_splitting != 1 && c.flux_splitting != 2) return fail("Flux_Splitting must be 1 or 2");
  if (c.ibc != 0 && c.ibc != 1 && c.ibc != 2 && c.ibc != 3 && c.ibc != 99)
    return fail("BoundaryCondition not supported");
  if (c.nfiltering < 0 || c.nfiltering > 10) return fail("NFiltering must be <= 10");
  if (c.precision != OCFD_PREC_FP64 && c.precision != OCFD_PREC_MIXED)
    return fail("precision must be OCFD_PREC_FP64 or OCFD_PREC_MIXED");
  for (int q = 0; q < 6; q++)
    if (c.scheme_boundary[q] < -1 || c.scheme_boundary[q] > 1) return fail("Scheme_boundary must be -1, 0 or 1");
  Block &B = h->B;
  B.prec = c.precision == OCFD_PREC_MIXED ? OCFD_PREC_MIXED_CODE : OCFD_PREC_FP64_CODE;
  B.kseg = 0;  // set by a batch
  Geom &g = B.g;
  g.nx = c.nx; g.ny = c.ny; g.nz = c.nz;
  g.sx = c.nx + 2 * LAP; g.sy = c.ny + 2 * LAP; g.sz = c.nz + 2 * LAP;
  g.nh = (int64_t)g.sx * g.sy * g.sz;
  g.ni = (int64_t)g.nx * g.ny * g.nz;
  // hybrid face lists hold 2 * face index + class in an int
  if (c.scheme_invis == SCH_HYBRID &&
      2 * (g.ni + std::max({(int64_t)g.ny * g.nz, (int64_t)g.nx * g.nz, (int64_t)g.nx * g.ny})) >= INT32_MAX)
    return fail("HYBRID needs 2 * (faces per axis) < 2^31 on one block");
  // set_parameters (OCFD_readpara.f90:261-341)
  SchemeParams &S = B.S;
  S.scheme_invis = c.scheme_invis;
  S.scheme_vis = c.scheme_vis;
  S.flux_split = c.flux_splitting;
  S.is_char = c.if_scheme_character;
  switch (c.scheme_invis) {
    case SCH_WENO5: S.ka1 = -2; S.kb1 = 2; S.ka2 = -1; S.kb2 = 3; break;
    case SCH_WENO7: S.ka1 = -3; S.kb1 = 3; S.ka2 = -2; S.kb2 = 4; break;
    default: S.ka1 = -3; S.kb1 = 4; S.ka2 = -3; S.kb2 = 4; break;
  }
  for (int a = 0; a < 3; a++) {
    S.iperiodic[a] = c.iperiodic[a];
    S.bidx[a][0] = S.bidx[a][1] = c.iperiodic[a] == 0 ? 1 : 0;
  }
  for (int q = 0; q < 6; q++) S.sb[q] = c.scheme_boundary[q];
  const double ud7[8] = {-3.0, 25.0, -101.0, 319.0, 214.0, -38.0, 4.0, 0.0};
  const double cd8[8] = {-3.0, 29.0, -139.0, 533.0, 533.0, -139.0, 29.0, -3.0};
  for (int q = 0; q < 8; q++) S.ud7l[q] = c.ud7l_diss * ud7[q] / 420.0 + (1.0 - c.ud7l_diss) * cd8[q] / 840.0;
  S.hyb2 = c.hybrid_para[1];
  S.hyb3 = c.hybrid_para[2];
  Phys &P = B.P;
  P.gamma = c.gamma; P.ma = c.ma; P.re = c.re; P.pr = c.pr; P.dt = c.dt; P.ref_amu_t0 = c.ref_amu_t0;
  P.h[0] = 1.0 / ((double)c.nx - 1.0);
  P.h[1] = 1.0 / ((double)c.ny - 1.0);
  P.h[2] = 1.0 / ((double)c.nz - 1.0);
  for (int a = 0; a < 3; a++) P.hinv[a] = 1.0 / P.h[a];
  for (int a = 0; a < 3; a++) P.dc[a] = make_dcoef(P.h[a]);
  P.cv = 1.0 / (c.gamma * (c.gamma - 1.0) * c.ma * c.ma);
  P.cp = P.cv * c.gamma;
  P.if_mass_force = c.if_mass_force;
  for (int a = 0; a < 3; a++) P.mass_force[a] = c.mass_force[a];
  h->ibc = c.ibc;
  h->tw = c.bc_para[0];
  for (int q = 0; q < 6; q++) h->ghost[q] = c.ghost_cell[q];
  return 0;
}

template <typename E>
__global__ void k_fill(E *a, int64_t n, E v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}
hipError_t fill(double *a, int64_t n, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n, v);
  return hipGetLastError();
}
hipError_t fill(const PrimField &a, int64_t n, double v, hipStream_t st) {
  if (!a.f32) return fill(a.as<double>(), n, v, st);
  hipLaunchKernelGGL(k_fill<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a.as<float>(), n,
                     float(v));
  return hipGetLastError();
}

__global__ void k_fill_int(int *a, int64_t n, int v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

__global__ void k_zero_int(int *a, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = 0;
}

// forced-WENO patches of the sensor (OCFD_SetHybrid.f90:87-110), with the j
// range read from Hybrid_para(7:8) as the reference's comments intend
__global__ void k_patch(const Geom g, double *r, int ib, int ie, int jb, int je, int kb, int ke) {
  const int64_t ni = (int64_t)(ie - ib + 1) * (je - jb + 1) * (ke - kb + 1);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ni) return;
  const int w = ie - ib + 1, hgt = je - jb + 1;
  const int i = ib + (int)(t % w), j = jb + (int)((t / w) % hgt), k = kb + (int)(t / ((int64_t)w * hgt));
  const int64_t q = g.hi(i, j, k);
  r[q] = r[q] + 100.0;
}

int alloc(double **p, size_t bytes) {
  if (hipMalloc((void **)p, bytes) != hipSuccess) {
    *p = nullptr;
    return fail("hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  return 0;
}
// a halo'd primitive field of nh points: FP32 in the mixed-precision form
int alloc_prim(PrimField *p, int64_t nh, bool f32) {
  p->f32 = f32 ? 1 : 0;
  const size_t bytes = (size_t)nh * (f32 ? sizeof(float) : sizeof(double));
  if (hipMalloc(&p->p, bytes) != hipSuccess) {
    p->p = nullptr;
    return fail("hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  return 0;
}

PrimField prims_of(ocfd_handle *h, int q) {
  const PrimField p[5] = {h->B.d, h->B.u, h->B.v, h->B.w, h->B.T};
  return p[q];
}

// primitive q, halo'd, to / from nh doubles on the host (converted from and
// to FP32 in the mixed-precision form)
int prim_download(ocfd_handle *h, int q, double *dst) {
  const PrimField p = prims_of(h, q);
  const size_t n = (size_t)h->B.g.nh;
  if (!p.f32) {
    HIPCHK(hipMemcpy(dst, p.p, n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  }
  std::vector<float> tmp(n);
  HIPCHK(hipMemcpy(tmp.data(), p.p, n * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; i++) dst[i] = (double)tmp[i];
  return 0;
}
int prim_upload(ocfd_handle *h, int q, const double *src) {
  const PrimField p = prims_of(h, q);
  const size_t n = (size_t)h->B.g.nh;
  if (!p.f32) {
    HIPCHK(hipMemcpy(p.p, src, n * sizeof(double), hipMemcpyHostToDevice));
    return 0;
  }
  std::vector<float> tmp(n);
  for (size_t i = 0; i < n; i++) tmp[i] = (float)src[i];
  HIPCHK(hipMemcpy(p.p, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

// the primitives from nh doubles per field on the host; with make_f, f from
// those FP64 values as at start-up (init3d, OCFD_init.f90:10-20).  In the
// mixed-precision form the FP64 values pass through device scratch fields, so
// that the conserved variables are formed from them and not from their FP32
// rounding.
int upload_prims(ocfd_handle *h, const double *const src[5], bool make_f) {
  HIPCHK(hipStreamSynchronize(h->st));
  if (!h->B.d.f32) {
    for (int q = 0; q < 5; q++)
      if (prim_upload(h, q, src[q])) return -1;
    if (make_f && launch(h, "init_f", [&] { return launch_init_f(h->B, h->st); })) return -1;
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
  }
  int rc = 0;
  if (make_f) {
    Block B64 = h->B;
    PrimField *f64[5] = {&B64.d, &B64.u, &B64.v, &B64.w, &B64.T};
    for (int q = 0; q < 5; q++) f64[q]->p = nullptr;
    for (int q = 0; q < 5 && !rc; q++) {
      rc = alloc_prim(f64[q], B64.g.nh, false);
      if (!rc && hipMemcpy(f64[q]->p, src[q], h->halo_bytes(), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail("copy of the primitives failed");
    }
    if (!rc && launch(h, "init_f", [&] { return launch_init_f(B64, h->st); })) rc = -1;
    if (!rc && hipStreamSynchronize(h->st) != hipSuccess) rc = fail("init_f failed");
    for (int q = 0; q < 5; q++)
      if (f64[q]->p) (void)hipFree(f64[q]->p);
  }
  for (int q = 0; q < 5 && !rc; q++) rc = prim_upload(h, q, src[q]);
  return rc;
}

int inviscid(ocfd_handle *h, bool ch) {
  Block B = h->B;
  B.S.is_char = ch ? 1 : 0;
  if (B.hcount) {
    // hybrid: the WENO faces of the three sweeps first
    if ((h->cfg.if_scheme_character != 0) != ch)
      return fail("hybrid scheme: the inviscid routine must match IF_Scheme_Character");
    for (int a = 0; a < 3; a++)
      if (launch(h, "weno_faces", [&] { return launch_weno_faces(B, a, h->st); })) return -1;
  }
  static const char *nm[2][3] = {{"inviscid_x", "inviscid_y", "inviscid_z"},
                                 {"inviscid_char_x", "inviscid_char_y", "inviscid_char_z"}};
  const int modes[3] = {INV_ASSIGN, INV_ACCUM, INV_FINAL};
  for (int a = 0; a < 3; a++)
    if (launch(h, nm[ch][a], [&] { return launch_inviscid(B, a, modes[a], SweepTile{h->tile[a]}, h->st); }))
      return -1;
  return 0;
}

int exchange_prims(ocfd_handle *h) {
  const PrimField f[5] = {h->B.d, h->B.u, h->B.v, h->B.w, h->B.T};
  return launch(h, "halo_fill", [&] { return launch_halo_prims(h->B, f, 5, h->st); });
}

int sensor(ocfd_handle *h) {
  const ocfd_config &c = h->cfg;
  const int kind = (int)std::lround(c.hybrid_para[0]);
  if (kind == 1) {
    if (launch(h, "sensor", [&] { return launch_sensor(h->B, h->st); })) return -1;
  } else {
    if (launch(h, "sensor", [&] { return launch_sensor2(h->B, h->work, h->st); })) return -1;
  }
  const int mpatch = (int)std::lround(c.hybrid_para[3]);
  for (int m = 1; m <= mpatch; m++) {
    const double *hp = c.hybrid_para;
    int ib = (int)std::lround(hp[4 + (m - 1) * 6]), ie = (int)std::lround(hp[5 + (m - 1) * 6]);
    int jb = (int)std::lround(hp[6 + (m - 1) * 6]), je = (int)std::lround(hp[7 + (m - 1) * 6]);
    int kb = (int)std::lround(hp[8 + (m - 1) * 6]), ke = (int)std::lround(hp[9 + (m - 1) * 6]);
    ib = ib < 1 ? 1 : ib; jb = jb < 1 ? 1 : jb; kb = kb < 1 ? 1 : kb;
    ie = ie > c.nx ? c.nx : ie; je = je > c.ny ? c.ny : je; ke = ke > c.nz ? c.nz : ke;
    if (ie < ib || je < jb || ke < kb) continue;
    const int64_t n = (int64_t)(ie - ib + 1) * (je - jb + 1) * (ke - kb + 1);
    hipLaunchKernelGGL(k_patch, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->st, h->B.g, h->B.rhyb, ib, ie,
                       jb, je, kb, ke);
  }
  double *r[1] = {h->B.rhyb};
  if (launch(h, "halo_fill", [&] { return launch_halo_fill(h->B, r, 1, h->st); })) return -1;
  // the faces that use WENO7 / WENO5 until the next sensor evaluation
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(128), 0, h->st, h->B.hcount, 4);
  return launch(h, "hyb_list", [&] { return launch_hybrid_lists(h->B, h->st); });
}

int filtering(ocfd_handle *h) {
  const ocfd_config &c = h->cfg;
  for (int m = 0; m < c.nfiltering; m++) {
    const double *fp = c.filter_para[m];
    const int per = (int)std::lround(fp[0]);
    if (per == 0 || h->istep % per != 0) continue;
    int rng[6];
    for (int q = 0; q < 6; q++) rng[q] = (int)std::lround(fp[4 + q]);
    if (rng[0] < 1) rng[0] = 1;
    if (rng[1] > c.nx) rng[1] = c.nx;
    if (rng[2] < 1) rng[2] = 1;
    if (rng[3] > c.ny) rng[3] = c.ny;
    if (rng[4] < 1) rng[4] = 1;
    if (rng[5] > c.nz) rng[5] = c.nz;
    const int scheme = (int)std::lround(fp[10]);
    if (scheme != 1 && scheme != 2) continue;
    for (int a = 0; a < 3; a++) {
      if ((int)std::lround(fp[1 + a]) != 1) continue;
      if (launch(h, "filter", [&] {
            return launch_filter(h->B, a, scheme, fp[11], fp[12], rng, h->fbuf, nullptr, h->st);
          }))
        return -1;
      std::swap(h->B.f, h->fbuf);
    }
  }
  return 0;
}

// after every stage's T <= 0 repair: this stage's temperature (T2) becomes
// the current one (Block::T2)
void swap_T(ocfd_handle *h) {
  std::swap(h->B.T, h->B.T2);
  h->tpar ^= 1;
}

// after every fused stage: its new u, v, w (in u2, v2, w2) become current
void swap_uvw(ocfd_handle *h) {
  std::swap(h->B.u, h->B.u2);
  std::swap(h->B.v, h->B.v2);
  std::swap(h->B.w, h->B.w2);
  h->uvpar ^= 1;
}

// ocfd_handle::pend bits (stage fusion)
constexpr int PEND_VISC = 1, PEND_UPDATE = 2, PEND_REFRESHED = 4, PEND_UVW = 8;

int read_neg(ocfd_handle *h, int out[3]) {
  HIPCHK(hipMemcpyAsync(out, h->neg, sizeof(int), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipMemcpyAsync(out + 1, h->neg + 65, 2 * sizeof(int), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}

}  // namespace

// ===========================================================================
extern "C" {

const char *ocfd_last_error(void) { return g_err.c_str(); }

int ocfd_create(const ocfd_config *cfg, int device, ocfd_handle **out) {
  if (!cfg || !out) return fail("null argument");
  ocfd_handle *h = new ocfd_handle();
  h->cfg = *cfg;
  if (set_params(h)) { delete h; return -1; }
  if (cfg->metric_mode != OCFD_METRIC_FIELD && cfg->metric_mode != OCFD_METRIC_UNIFORM) {
    delete h;
    return fail("metric_mode must be FIELD or UNIFORM");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { delete h; return fail("no HIP device"); }
  if (device < 0 || device >= ndev) { delete h; return fail("device index out of range"); }
  h->device = device;
  h->use_graph = std::getenv("OCFD_NO_GRAPH") == nullptr;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&h->st) != hipSuccess) {
    delete h;
    return fail("cannot create HIP stream");
  }
  Block &B = h->B;
  const size_t hb = h->halo_bytes(), fb = h->f_bytes();
  int rc = 0;
  // the primitives are FP32 in the mixed-precision form
  const bool p32 = B.prec == OCFD_PREC_MIXED_CODE;
  for (PrimField *p : {&B.d, &B.u, &B.v, &B.w, &B.T, &B.T2, &B.u2, &B.v2, &B.w2}) p->p = nullptr;
  for (PrimField *p : {&B.d, &B.u, &B.v, &B.w, &B.T, &B.T2}) rc |= alloc_prim(p, B.g.nh, p32);
  // the viscous terms fused into the stage update where the block allows it
  // (OCFD_NO_FUSED keeps the two kernels, for comparisons)
  B.M.uniform = cfg->metric_mode == OCFD_METRIC_UNIFORM;
  h->fused = cfg->if_viscous == 1 && cfg->ibc == 0 && visc_fused_supported(B) &&
             std::getenv("OCFD_NO_FUSED") == nullptr;
  for (int q = 0; q < 6; q++) h->fused = h->fused && cfg->ghost_cell[q] == 0;
  if (h->fused)
    for (PrimField *p : {&B.u2, &B.v2, &B.w2}) rc |= alloc_prim(p, B.g.nh, p32);
  B.halt = nullptr;
  rc |= alloc(&B.f, fb) | alloc(&B.fn, fb) | alloc(&B.du, fb);
  if (cfg->scheme_invis == SCH_HYBRID) rc |= alloc(&B.rhyb, hb);
  if (cfg->if_viscous == 1) rc |= alloc(&B.ev, 12 * (size_t)B.g.ni * sizeof(double));
  if (cfg->nfiltering > 0) rc |= alloc(&h->fbuf, fb);
  rc |= alloc(&B.trash, 16 * 256 * sizeof(double));
  for (int a = 0; a < 3; a++) {
    B.hlist[a] = B.hgbase[a] = nullptr;
    B.hbits[a] = nullptr;
    B.wflux[a] = nullptr;
  }
  B.hcount = nullptr;
  if (cfg->scheme_invis == SCH_HYBRID) {
    // worst case: every face of every axis uses WENO
    const int nh = cfg->if_scheme_character ? 5 : 10;
    const int64_t nfa[3] = {TileIndex<0>::nfaces(B.g), TileIndex<1>::nfaces(B.g), TileIndex<2>::nfaces(B.g)};
    for (int a = 0; a < 3; a++) {
      rc |= alloc((double **)&B.hlist[a], (size_t)nfa[a] * sizeof(int));
      const int64_t ng = (nfa[a] + 63) / 64;
      rc |= alloc((double **)&B.hbits[a], (size_t)ng * sizeof(unsigned long long));
      rc |= alloc((double **)&B.hgbase[a], (size_t)ng * sizeof(int));
      rc |= alloc(&B.wflux[a], (size_t)nfa[a] * nh * sizeof(double));
    }
    rc |= alloc((double **)&B.hcount, 4 * sizeof(int));
  }
  B.M.uniform = cfg->metric_mode == OCFD_METRIC_UNIFORM;
  if (!B.M.uniform)
    for (int m = 0; m < M_N; m++) rc |= alloc(&h->met[m], hb);
  const int64_t nblk = (B.g.ni + 255) / 256;
  rc |= alloc(&h->work, (size_t)(nblk * 14 + 64) * sizeof(double));
  rc |= alloc(&h->wout, 32 * sizeof(double));
  if (hipMalloc((void **)&h->neg, 72 * sizeof(int)) != hipSuccess) rc = -1;
  if (hipMalloc((void **)&h->counts, 4 * sizeof(unsigned long long)) != hipSuccess) rc = -1;
  if (rc) {
    std::string msg = g_err;
    ocfd_destroy(h);
    return fail(msg.empty() ? "allocation failed" : msg);
  }
  for (int m = 0; m < M_N; m++) {
    B.M.p[m] = h->met[m];
    B.M.u[m] = 1.0;
  }
  // allocate_flow_data initial values (OCFD_NS_Solver.f90:116-118, 140-147)
  for (const PrimField *p : {&B.d, &B.u, &B.v, &B.w, &B.T, &B.T2}) (void)fill(*p, B.g.nh, 1.0, h->st);
  if (h->fused)
    for (const PrimField *p : {&B.u2, &B.v2, &B.w2}) (void)fill(*p, B.g.nh, 1.0, h->st);
  if (!B.M.uniform)
    for (int m = 0; m < M_N; m++) (void)fill(h->met[m], B.g.nh, 1.0, h->st);
  if (B.rhyb) (void)fill(B.rhyb, B.g.nh, 0.0, h->st);
  (void)fill(B.du, 5 * B.g.ni, 0.0, h->st);
  (void)fill(B.f, 5 * B.g.ni, 0.0, h->st);
  (void)fill(B.fn, 5 * B.g.ni, 0.0, h->st);
  if (B.ev) (void)fill(B.ev, 12 * B.g.ni, 0.0, h->st);
  if (B.hcount) {
    // no face listed until the first comput_Rhybrid
    for (int a = 0; a < 3; a++) {
      const int64_t ng = a == 0 ? TileIndex<0>::ngroups(B.g) : (a == 1 ? TileIndex<1>::ngroups(B.g) : TileIndex<2>::ngroups(B.g));
      hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((2 * ng + 255) / 256)), dim3(256), 0, h->st, (int *)B.hbits[a],
                         2 * ng, 0);
      hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, h->st, B.hgbase[a], ng, 0);
    }
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(128), 0, h->st, B.hcount, 4);
  }
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(128), 0, h->st, h->neg, 72);
  if (hipStreamSynchronize(h->st) != hipSuccess) {
    ocfd_destroy(h);
    return fail("initialisation failed");
  }
  *out = h;
  return 0;
}

void invalidate_graph(ocfd_handle *h);

int ocfd_destroy(ocfd_handle *h) {
  if (!h) return 0;
  (void)hipSetDevice(h->device);
  if (h->st) (void)hipStreamSynchronize(h->st);
  invalidate_graph(h);
  Block &B = h->B;
  for (const PrimField *p : {&B.d, &B.u, &B.v, &B.w, &B.T, &B.T2, &B.u2, &B.v2, &B.w2})
    if (p->p) (void)hipFree(p->p);
  double *ps[] = {B.f, B.fn, B.du, B.rhyb, B.ev, B.trash, h->fbuf, h->work, h->wout};
  for (double *p : ps)
    if (p) (void)hipFree(p);
  for (int a = 0; a < 3; a++) {
    if (B.hlist[a]) (void)hipFree(B.hlist[a]);
    if (B.hbits[a]) (void)hipFree(B.hbits[a]);
    if (B.hgbase[a]) (void)hipFree(B.hgbase[a]);
    if (B.wflux[a]) (void)hipFree(B.wflux[a]);
  }
  if (B.hcount) (void)hipFree(B.hcount);
  for (int m = 0; m < M_N; m++)
    if (h->met[m]) (void)hipFree(h->met[m]);
  if (h->neg) (void)hipFree(h->neg);
  if (h->counts) (void)hipFree(h->counts);
  for (void *p : h->bl_mem) (void)hipFree(p);
  for (auto &r : h->pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : h->pool) (void)hipEventDestroy(e);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
  return 0;
}

int ocfd_synchronize(ocfd_handle *h) {
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}

// ---- grid / metrics --------------------------------------------------------
int ocfd_set_grid(ocfd_handle *h, const double *x, const double *y, const double *z, const double spans[9]) {
  if (h->B.M.uniform) return fail("ocfd_set_grid needs OCFD_METRIC_FIELD");
  const Geom &g = h->B.g;
  std::vector<double> stage((size_t)g.nh);
  double *dc[3] = {};
  const double *src[3] = {x, y, z};
  int rc = 0;
  for (int c = 0; c < 3 && !rc; c++) {
    rc = alloc(&dc[c], h->halo_bytes());
    if (rc) break;
    std::fill(stage.begin(), stage.end(), 1.0);
    for (int k = 1; k <= g.nz; k++)
      for (int j = 1; j <= g.ny; j++)
        std::memcpy(&stage[(size_t)g.hi(1, j, k)], &src[c][(size_t)g.ii(1, j, k)], (size_t)g.nx * sizeof(double));
    if (hipMemcpy(dc[c], stage.data(), h->halo_bytes(), hipMemcpyHostToDevice) != hipSuccess) rc = fail("copy");
  }
  if (!rc) {
    for (int m = 0; m < M_N; m++) (void)fill(h->met[m], g.nh, 1.0, h->st);
    rc = launch(h, "metrics", [&] { return launch_metrics(h->B, h->ibc, dc, h->met, spans, h->ghost, nullptr, h->st); });
    if (!rc && hipStreamSynchronize(h->st) != hipSuccess) rc = fail("metrics kernels failed");
  }
  for (int c = 0; c < 3; c++)
    if (dc[c]) (void)hipFree(dc[c]);
  return rc;
}

int ocfd_upload_metrics(ocfd_handle *h, const double *const metrics[19]) {
  if (h->B.M.uniform) return fail("ocfd_upload_metrics needs OCFD_METRIC_FIELD");
  for (int m = 0; m < M_N; m++) HIPCHK(hipMemcpy(h->met[m], metrics[m], h->halo_bytes(), hipMemcpyHostToDevice));
  return 0;
}

int ocfd_set_uniform_metrics(ocfd_handle *h, const double m[10]) {
  if (!h->B.M.uniform) return fail("ocfd_set_uniform_metrics needs OCFD_METRIC_UNIFORM");
  for (int q = 0; q < 10; q++) h->B.M.u[q] = m[q];
  for (int q = 0; q < 9; q++) h->B.M.u[M_AKX1 + q] = m[q] / m[M_AJAC];
  invalidate_graph(h);  // the constants travel in the kernel arguments
  return 0;
}

int ocfd_download_metrics(ocfd_handle *h, double *const metrics[19]) {
  const Geom &g = h->B.g;
  for (int m = 0; m < M_N; m++) {
    if (!metrics[m]) continue;
    if (h->B.M.uniform) {
      for (int64_t q = 0; q < g.nh; q++) metrics[m][q] = h->B.M.u[m];
    } else {
      HIPCHK(hipMemcpy(metrics[m], h->met[m], h->halo_bytes(), hipMemcpyDeviceToHost));
    }
  }
  return 0;
}

// ---- state -----------------------------------------------------------------
int ocfd_upload_primitives(ocfd_handle *h, const double *d, const double *u, const double *v, const double *w,
                           const double *T, int32_t istep, double tt) {
  const double *const src[5] = {d, u, v, w, T};
  if (upload_prims(h, src, true)) return -1;
  h->istep = istep;
  h->tt = tt;
  return 0;
}

// ---- restart / snapshot files ------------------------------------------------
// Fortran unformatted sequential records: int32 length, payload, int32 length.
namespace {
bool write_rec(FILE *fp, const void *p, int64_t n) {
  const int32_t len = (int32_t)n;
  return fwrite(&len, 4, 1, fp) == 1 && fwrite(p, 1, (size_t)n, fp) == (size_t)n && fwrite(&len, 4, 1, fp) == 1;
}
// 1: record read; 0: clean end of file before it; -1: malformed
int read_rec(FILE *fp, void *p, int64_t n) {
  int32_t a = 0, b = 0;
  const size_t got = fread(&a, 1, 4, fp);
  if (got == 0 && feof(fp)) return 0;
  if (got != 4 || a != (int32_t)n) return -1;
  if (fread(p, 1, (size_t)n, fp) != (size_t)n || fread(&b, 4, 1, fp) != 1 || b != a) return -1;
  return 1;
}
struct FileCloser {
  FILE *fp;
  ~FileCloser() {
    if (fp) fclose(fp);
  }
};
}  // namespace

int ocfd_save_flow_data(ocfd_handle *h, const char *path, int32_t with_f) {
  char name[32];
  if (!path) {
    std::snprintf(name, sizeof name, "OCFD%08d.dat", h->istep);
    path = name;
  }
  HIPCHK(hipStreamSynchronize(h->st));
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return fail(std::string("cannot open ") + path + " for writing");
  FileCloser fc{fp};
  const Geom &g = h->B.g;
  unsigned char head[12];
  const int32_t istep = h->istep;
  std::memcpy(head, &istep, 4);
  std::memcpy(head + 4, &h->tt, 8);
  if (!write_rec(fp, head, 12)) return fail("write error");
  // save_flow_data (OCFD_IO.f90:44-64) / write_3d1: per field, one record per
  // k plane of the nx x ny interior
  std::vector<double> buf((size_t)g.nh), plane((size_t)g.nx * g.ny);
  for (int q = 0; q < 5; q++) {
    if (prim_download(h, q, buf.data())) return -1;
    for (int k = 1; k <= g.nz; k++) {
      for (int j = 1; j <= g.ny; j++)
        std::memcpy(&plane[(size_t)(j - 1) * g.nx], &buf[(size_t)g.hi(1, j, k)], (size_t)g.nx * sizeof(double));
      if (!write_rec(fp, plane.data(), (int64_t)plane.size() * 8)) return fail("write error");
    }
  }
  if (with_f) {
    // the conserved variables, appended after the reference's records (its
    // reader stops before them): f(:,:,k,m) per record
    std::vector<double> f((size_t)5 * g.ni);
    HIPCHK(hipMemcpy(f.data(), h->B.f, h->f_bytes(), hipMemcpyDeviceToHost));
    for (int64_t r = 0; r < 5 * (int64_t)g.nz; r++)
      if (!write_rec(fp, &f[(size_t)(r * g.nx * g.ny)], (int64_t)g.nx * g.ny * 8)) return fail("write error");
  }
  if (std::fflush(fp) != 0) return fail("write error");
  return 0;
}

int ocfd_read_flow_data(ocfd_handle *h, const char *path, int32_t *has_f) {
  FILE *fp = std::fopen(path ? path : "opencfd.dat", "rb");
  if (!fp) return fail(std::string("cannot open ") + (path ? path : "opencfd.dat"));
  FileCloser fc{fp};
  const Geom &g = h->B.g;
  unsigned char head[12];
  if (read_rec(fp, head, 12) != 1) return fail("malformed flow-data file (first record)");
  int32_t istep;
  double tt;
  std::memcpy(&istep, head, 4);
  std::memcpy(&tt, head + 4, 8);
  // read_flow_data (OCFD_IO.f90:5-39) / read_3d1: only the interior of each
  // primitive is replaced; the halos keep their values
  HIPCHK(hipStreamSynchronize(h->st));
  std::vector<std::vector<double>> bufs(5, std::vector<double>((size_t)g.nh));
  std::vector<double> plane((size_t)g.nx * g.ny);
  for (int q = 0; q < 5; q++) {
    std::vector<double> &buf = bufs[q];
    if (prim_download(h, q, buf.data())) return -1;
    for (int k = 1; k <= g.nz; k++) {
      if (read_rec(fp, plane.data(), (int64_t)plane.size() * 8) != 1)
        return fail("malformed flow-data file (field " + std::to_string(q + 1) + ", plane " + std::to_string(k) +
                    "): expected records of nx*ny doubles");
      for (int j = 1; j <= g.ny; j++)
        std::memcpy(&buf[(size_t)g.hi(1, j, k)], &plane[(size_t)(j - 1) * g.nx], (size_t)g.nx * sizeof(double));
    }
  }
  h->istep = istep;
  h->tt = tt;
  // f from the primitives, as at start-up (OCFD_init.f90), unless the file
  // carries f
  std::vector<double> f((size_t)5 * g.ni);
  int rc = read_rec(fp, f.data(), (int64_t)g.nx * g.ny * 8);
  if (rc < 0) return fail("malformed conserved-variable records");
  if (rc == 1)
    for (int64_t r = 1; r < 5 * (int64_t)g.nz; r++)
      if (read_rec(fp, &f[(size_t)(r * g.nx * g.ny)], (int64_t)g.nx * g.ny * 8) != 1)
        return fail("malformed conserved-variable records");
  const double *const src[5] = {bufs[0].data(), bufs[1].data(), bufs[2].data(), bufs[3].data(), bufs[4].data()};
  if (upload_prims(h, src, rc != 1)) return -1;
  if (rc == 1) HIPCHK(hipMemcpy(h->B.f, f.data(), h->f_bytes(), hipMemcpyHostToDevice));
  HIPCHK(hipStreamSynchronize(h->st));
  if (has_f) *has_f = rc == 1 ? 1 : 0;
  return 0;
}

int ocfd_upload_f(ocfd_handle *h, const double *f) {
  HIPCHK(hipMemcpy(h->B.f, f, h->f_bytes(), hipMemcpyHostToDevice));
  return 0;
}
int ocfd_download_f(ocfd_handle *h, double *f) {
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(hipMemcpy(f, h->B.f, h->f_bytes(), hipMemcpyDeviceToHost));
  return 0;
}
int ocfd_download_fn(ocfd_handle *h, double *f) {
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(hipMemcpy(f, h->B.fn, h->f_bytes(), hipMemcpyDeviceToHost));
  return 0;
}
// du is FP32 on the device in the mixed-precision form: converted on the host
int ocfd_download_du(ocfd_handle *h, double *f) {
  HIPCHK(hipStreamSynchronize(h->st));
  if (h->B.prec != OCFD_PREC_MIXED_CODE) {
    HIPCHK(hipMemcpy(f, h->B.du, h->f_bytes(), hipMemcpyDeviceToHost));
    return 0;
  }
  const size_t n = (size_t)5 * h->B.g.ni;
  std::vector<float> tmp(n);
  HIPCHK(hipMemcpy(tmp.data(), h->B.du, n * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t q = 0; q < n; q++) f[q] = (double)tmp[q];
  return 0;
}
int ocfd_upload_du(ocfd_handle *h, const double *f) {
  if (h->B.prec != OCFD_PREC_MIXED_CODE) {
    HIPCHK(hipMemcpy(h->B.du, f, h->f_bytes(), hipMemcpyHostToDevice));
    return 0;
  }
  const size_t n = (size_t)5 * h->B.g.ni;
  std::vector<float> tmp(n);
  for (size_t q = 0; q < n; q++) tmp[q] = (float)f[q];
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(hipMemcpy(h->B.du, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}
int ocfd_download_primitives(ocfd_handle *h, double *d, double *u, double *v, double *w, double *T) {
  HIPCHK(hipStreamSynchronize(h->st));
  double *dst[5] = {d, u, v, w, T};
  for (int q = 0; q < 5; q++)
    if (dst[q] && prim_download(h, q, dst[q])) return -1;
  return 0;
}
int ocfd_download_rhybrid(ocfd_handle *h, double *r) {
  if (!h->B.rhyb) return fail("Rhybrid exists only for the HYBRID scheme");
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(hipMemcpy(r, h->B.rhyb, h->halo_bytes(), hipMemcpyDeviceToHost));
  return 0;
}
int ocfd_get_time(ocfd_handle *h, int32_t *istep, double *tt) {
  *istep = h->istep;
  *tt = h->tt;
  return 0;
}

// ---- one reference routine each ---------------------------------------------
int ocfd_exchange_primitives(ocfd_handle *h) { return exchange_prims(h); }

int ocfd_comput_rhybrid(ocfd_handle *h, int64_t counts[4]) {
  if (!h->B.rhyb) return fail("comput_Rhybrid needs the HYBRID scheme");
  if (sensor(h)) return -1;
  if (counts) {
    HIPCHK(hipMemsetAsync(h->counts, 0, 4 * sizeof(unsigned long long), h->st));
    if (launch(h, "hybrid_counts", [&] { return launch_hybrid_counts(h->B, h->counts, h->st); })) return -1;
    unsigned long long c[3];
    HIPCHK(hipMemcpyAsync(c, h->counts, sizeof c, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    for (int q = 0; q < 3; q++) counts[q] = (int64_t)c[q];
    counts[3] = (int64_t)(h->cfg.nx + 1) * h->cfg.ny * h->cfg.nz;
  }
  return 0;
}

int ocfd_du_inviscous(ocfd_handle *h) { return inviscid(h, false); }
int ocfd_du_inviscous_character(ocfd_handle *h) { return inviscid(h, true); }

int ocfd_du_viscous(ocfd_handle *h) {
  if (!h->B.ev) return fail("du_viscous needs IF_Viscous = 1");
  if (h->stage_fusion) {
    // the fluxes now (inside the update for a fused block); their
    // divergence joins the update in ocfd_time_adv_rk3
    if (!h->fused && launch(h, "visc_flux", [&] { return launch_visc_flux(h->B, h->st); })) return -1;
    h->pend |= PEND_VISC;
    return 0;
  }
  if (launch(h, "visc_flux", [&] { return launch_visc_flux(h->B, h->st); })) return -1;
  return launch(h, "visc_div", [&] { return launch_visc_div(h->B, h->st); });
}

int ocfd_time_adv_rk3(ocfd_handle *h, int32_t krk) {
  if (krk < 1 || krk > 3) return fail("KRK must be 1, 2 or 3");
  if (h->stage_fusion) {
    // ocfd_step's stage update: viscous divergence (when du_viscous ran),
    // RK3 (fn = f at KRK = 1), primitives, boundary-plane f on periodic
    // blocks, and the list of T <= 0 cells for ocfd_comput_duvwt
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(128), 0, h->st, h->neg, NEG_STRIDE);
    const int refresh = h->ibc == 0 ? 1 : 0;
    const bool visc = (h->pend & PEND_VISC) != 0;
    int pend = PEND_UPDATE | (refresh ? PEND_REFRESHED : 0);
    if (visc && h->fused) {
      if (launch(h, "stage_update_fused", [&] { return launch_visc_fused(h->B, krk, refresh, h->neg, h->st); }))
        return -1;
      swap_uvw(h);
      pend |= PEND_UVW;
    } else {
      Block B = h->B;
      if (!visc) B.ev = nullptr;
      if (launch(h, "stage_update", [&] { return launch_visc_div_update(B, krk, refresh, h->neg, h->st); }))
        return -1;
    }
    h->pend = pend;
    return 0;
  }
  return launch(h, "rk3", [&] { return launch_rk3(h->B, krk, h->st); });
}

int ocfd_save_fn(ocfd_handle *h) {
  if (h->stage_fusion) return 0;  // ocfd_time_adv_rk3(1) stores fn = f
  return launch(h, "copy_fn", [&] { return launch_copy_f(h->B, h->st); });
}

int ocfd_set_stage_fusion(ocfd_handle *h, int32_t on) {
  h->stage_fusion = on != 0;
  h->pend = 0;
  return 0;
}

int ocfd_comput_duvwt(ocfd_handle *h) {
  if (h->stage_fusion && (h->pend & PEND_UPDATE)) {
    // primitives already recovered by the stage update: repair and count
    const int uvw = (h->pend & PEND_UVW) ? 1 : 0;
    if (launch(h, "negt_repair", [&] { return launch_negt_repair(h->B, h->neg, h->st, uvw); })) return -1;
    swap_T(h);
    h->pend &= ~(PEND_UPDATE | PEND_UVW);
    int c[3];
    if (read_neg(h, c)) return -1;
    return c[1];
  }
  h->pend = 0;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(128), 0, h->st, h->neg, NEG_STRIDE);
  if (launch(h, "duvwt", [&] { return launch_duvwt(h->B, h->neg, h->st); })) return -1;
  if (launch(h, "negt_repair", [&] { return launch_negt_repair(h->B, h->neg, h->st); })) return -1;
  swap_T(h);
  int c[3];
  if (read_neg(h, c)) return -1;
  return c[1];  // cells with T <= 0 found (and repaired) in this call
}

int ocfd_bc(ocfd_handle *h) {
  if ((h->ibc == 2 || h->ibc == 3) && !h->bl_set)
    return fail("BC_BoundaryLayer / BC_SweptCorner: call ocfd_set_bc_boundarylayer first");
  const bool refreshed = h->stage_fusion && (h->pend & PEND_REFRESHED);
  h->pend = 0;
  if (refreshed) {
    // the stage update refreshed f on the boundary planes: ghost cells only
    bool any = false;
    for (int q = 0; q < 6; q++) any |= (h->ghost[q] == 1 || h->ghost[q] == 2);
    if (!any) return 0;
    return launch(h, "bc", [&] { return launch_bc(h->B, -1, h->tw, nullptr, h->ghost, 0, h->st); });
  }
  return launch(h, "bc", [&] { return launch_bc(h->B, h->ibc, h->tw, &h->bl, h->ghost, 1, h->st); });
}

int ocfd_filtering(ocfd_handle *h) { return filtering(h); }

int ocfd_set_bc_boundarylayer(ocfd_handle *h, double aoa, const double *inlet, const double *upper,
                              const int32_t *nonref, const double *pertb, const int32_t *wall) {
  const ocfd_config &c = h->cfg;
  if (c.ibc != 2 && c.ibc != 3) return fail("ocfd_set_bc_boundarylayer needs BC_BoundaryLayer or BC_SweptCorner");
  const Geom &g = h->B.g;
  const int64_t nyz = (int64_t)g.ny * g.nz, nxz = (int64_t)g.nx * g.nz;
  const double *bp = c.bc_para;
  BLArgs a = {};
  int inlet_mode, upper_mode;
  if (c.ibc == 2) {  // ocfd_bc_BoundaryLayer, BC_para(1:14) (OCFD_BC_Boundarylayers.f90:26-43)
    inlet_mode = (int)std::lround(bp[0]);
    upper_mode = (int)std::lround(bp[1]);
    a.outlet = (int)std::lround(bp[3]);
    a.Tw = bp[4];
    a.A = bp[7];
  } else {  // ocfd_bc_swept_corner, BC_para(1:12) (:452-472): 1-d inlet data, free-stream upper
    inlet_mode = 1;
    upper_mode = 0;
    a.outlet = (int)std::lround(bp[0]);
    a.Tw = bp[1];
    a.A = bp[4];
    a.z1 = (int)std::lround(bp[10]);
    a.z2 = (int)std::lround(bp[11]);
  }
  if (inlet_mode != 0 && !inlet) return fail("this boundary condition reads inlet data: pass inlet[5][nz][ny]");
  if (upper_mode != 0 && !upper) return fail("this boundary condition reads upper-boundary data: pass upper[5][nz][nx]");
  a.u0 = std::cos(aoa);
  a.v0 = std::sin(aoa);
  a.inletT = 0.0;  // the free-stream inlet sets T = 0 (OCFD_BC_Boundarylayers.f90:59)
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->st));
  for (void *p : h->bl_mem) (void)hipFree(p);
  h->bl_mem.clear();
  auto put = [&](const void *src, size_t bytes, const void **dst) -> int {
    void *d = nullptr;
    HIPCHK(hipMalloc(&d, bytes));
    h->bl_mem.push_back(d);
    if (src) HIPCHK(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice));
    else HIPCHK(hipMemset(d, 0, bytes));
    *dst = d;
    return 0;
  };
  const void *p;
  if (inlet_mode != 0) {
    if (put(inlet, 5 * nyz * sizeof(double), &p)) return -1;
    a.inlet = (const double *)p;
  }
  if (upper_mode != 0) {
    if (put(upper, 5 * nxz * sizeof(double), &p)) return -1;
    a.upper = (const double *)p;
  }
  if (put(nonref, nxz * sizeof(int32_t), &p)) return -1;
  a.nonref = (const int *)p;
  if (put(pertb, nxz * sizeof(double), &p)) return -1;
  a.pertb = (const double *)p;
  if (wall) {
    if (put(wall, nxz * sizeof(int32_t), &p)) return -1;
    a.wall = (const int *)p;
  }
  const double one = 1.0;
  if (put(&one, sizeof(double), &p)) return -1;
  a.ht = (const double *)p;
  h->bl = a;
  h->bl_set = true;
  invalidate_graph(h);
  return 0;
}

int ocfd_set_bc_time_factor(ocfd_handle *h, double ht) {
  if (!h->bl_set) return fail("call ocfd_set_bc_boundarylayer first");
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(hipMemcpy((void *)h->bl.ht, &ht, sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int ocfd_cfl_dt(ocfd_handle *h, double cfl, double *dt) {
  if (!dt || !(cfl > 0.0)) return fail("ocfd_cfl_dt: cfl must be > 0 and dt non-NULL");
  unsigned long long *w = reinterpret_cast<unsigned long long *>(h->wout);
  hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(64), 0, h->st, h->wout, (int64_t)1, 0.0);
  if (launch(h, "cfl_rate", [&] { return launch_cfl_rate(h->B, w, h->st); })) return -1;
  unsigned long long bits = 0;
  HIPCHK(hipMemcpyAsync(&bits, w, sizeof bits, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  double rate;
  std::memcpy(&rate, &bits, sizeof rate);
  if (!(rate > 0.0) || !std::isfinite(rate)) return fail("ocfd_cfl_dt: no finite positive rate in the block");
  *dt = cfl / rate;
  return 0;
}

int ocfd_advance_clock(ocfd_handle *h) {
  h->istep += 1;
  h->tt = h->tt + h->cfg.dt;
  return 0;
}

// ---- whole steps -----------------------------------------------------------
namespace {
// the three stages of one time step (OCFD_NS_Solver.f90:33-60)
int step_body(ocfd_handle *h) {
  const ocfd_config &c = h->cfg;
  const bool hyb = c.scheme_invis == SCH_HYBRID;
  const bool ch = c.if_scheme_character != 0;
  for (int krk = 1; krk <= 3; krk++) {
    if (exchange_prims(h)) return -1;
    if (krk == 1 && hyb && sensor(h)) return -1;
    if (inviscid(h, ch)) return -1;
    const int refresh = h->ibc == 0 ? 1 : 0;
    if (h->fused) {
      // viscous terms, divergence and update in one march (k_visc_fused)
      if (launch(h, "stage_update_fused", [&] { return launch_visc_fused(h->B, krk, refresh, h->neg, h->st); }))
        return -1;
      swap_uvw(h);
    } else {
      if (c.if_viscous == 1 && launch(h, "visc_flux", [&] { return launch_visc_flux(h->B, h->st); })) return -1;
      Block B = h->B;
      if (c.if_viscous != 1) B.ev = nullptr;
      if (launch(h, "stage_update", [&] { return launch_visc_div_update(B, krk, refresh, h->neg, h->st); }))
        return -1;
    }
    const int uvw = h->fused ? 1 : 0;
    if (launch(h, "negt_repair", [&] { return launch_negt_repair(h->B, h->neg, h->st, uvw); })) return -1;
    swap_T(h);
    if (h->ibc != 0) {
      if ((h->ibc == 2 || h->ibc == 3) && !h->bl_set)
        return fail("BC_BoundaryLayer / BC_SweptCorner: call ocfd_set_bc_boundarylayer first");
      if (launch(h, "bc", [&] { return launch_bc(h->B, h->ibc, h->tw, &h->bl, h->ghost, 1, h->st); })) return -1;
    } else {
      bool any = false;
      for (int q = 0; q < 6; q++) any |= (h->ghost[q] == 1 || h->ghost[q] == 2);
      if (any && launch(h, "bc", [&] { return launch_bc(h->B, -1, h->tw, nullptr, h->ghost, 0, h->st); })) return -1;
    }
  }
  return 0;
}

// Records step_body into a graph (kernel launches only: no memset, copy or
// host synchronisation happens inside it).
int capture_step(ocfd_handle *h) {
  hipGraph_t graph = nullptr;
  hipGraphExec_t *slot = &h->gexec[h->tpar + 2 * h->uvpar];  // step_body flips the parities
  HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
  const int rc = step_body(h);
  const hipError_t e = hipStreamEndCapture(h->st, &graph);
  if (rc || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc ? -1 : fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  const hipError_t ei = hipGraphInstantiate(slot, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    *slot = nullptr;
    return fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  return 0;
}
}  // namespace

void invalidate_graph(ocfd_handle *h) {
  for (hipGraphExec_t &g : h->gexec) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

int ocfd_step(ocfd_handle *h, int32_t nsteps) {
  const ocfd_config &c = h->cfg;
  h->pend = 0;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(128), 0, h->st, h->neg, NEG_STRIDE);
  // filtering decides per step on the host; per-kernel timing needs eager launches
  const bool graph = h->use_graph && !h->prof && c.nfiltering == 0;
  const int istep0 = h->istep;
  const double tt0 = h->tt;
  // the first stage with more than NEG_MAX cells T <= 0 halts the run on the
  // device (Block::halt); the host looks every HALT_POLL steps and stops
  constexpr int HALT_POLL = 8;
  h->B.halt = h->neg + NEG_HALT;
  int rc = 0;
  for (int n = 0; n < nsteps && !rc; n++) {
    if (graph) {
      const int p = h->tpar + 2 * h->uvpar;
      if (!h->gexec[p]) {
        rc = capture_step(h);  // records the three stages, flipping the parities
      } else {
        // what the three replayed stages do to the buffers
        swap_T(h);
        if (h->fused) swap_uvw(h);
      }
      if (!rc && hipGraphLaunch(h->gexec[p], h->st) != hipSuccess) rc = fail("hipGraphLaunch failed");
    } else {
      rc = step_body(h);
    }
    if (rc) break;
    h->istep += 1;
    h->tt = h->tt + c.dt;
    rc = filtering(h);
    if (!rc && (n + 1) % HALT_POLL == 0 && n + 1 < nsteps) {
      int flag = 0;
      if (hipMemcpyAsync(&flag, h->neg + NEG_HALT, sizeof(int), hipMemcpyDeviceToHost, h->st) != hipSuccess ||
          hipStreamSynchronize(h->st) != hipSuccess)
        rc = fail("reading the T <= 0 counters failed");
      if (flag) break;
    }
  }
  h->B.halt = nullptr;
  if (rc) return -1;
  int cnt[NEG_STRIDE];
  HIPCHK(hipMemcpyAsync(cnt, h->neg, sizeof cnt, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  if (cnt[NEG_HALT]) {
    // the failing stage's update and repair stand; Istep / tt count the
    // steps completed before it
    const int stage = cnt[NEG_HALT_STAGE], whole = (stage - 1) / 3;
    h->istep = istep0 + whole;
    h->tt = tt0;
    for (int q = 0; q < whole; q++) h->tt = h->tt + c.dt;
    return fail("more than 10 cells with T <= 0 in stage " + std::to_string((stage - 1) % 3 + 1) + " of step " +
                std::to_string(h->istep + 1) + " (" + std::to_string(cnt[66]) + " cells); the run stops there");
  }
  return cnt[65];
}

// ---- sweep tiles ---------------------------------------------------------------
int ocfd_set_sweep_tiles(ocfd_handle *h, const int32_t shape[3]) {
  for (int a = 0; a < 3; a++)
    if (shape[a] < 0 || shape[a] >= NSWEEP_SHAPES) return fail("sweep tile shape must be 0 .. 4");
  for (int a = 0; a < 3; a++) h->tile[a] = shape[a];
  invalidate_graph(h);
  return 0;
}

int ocfd_get_sweep_tiles(ocfd_handle *h, int32_t shape[3]) {
  for (int a = 0; a < 3; a++) shape[a] = h->tile[a];
  return 0;
}

int ocfd_autotune(ocfd_handle *h, int32_t reps, int32_t shape_out[3]) {
  if (reps < 1) reps = 5;
  const ocfd_config &c = h->cfg;
  const bool ch = c.if_scheme_character != 0;
  const bool prof = h->prof;
  h->prof = false;
  // the state the sweeps see in a stage: halos, and for the hybrid scheme the
  // sensor, the face lists and the WENO faces
  int rc = exchange_prims(h);
  if (!rc && c.scheme_invis == SCH_HYBRID) rc = sensor(h);
  if (!rc) rc = inviscid(h, ch);
  h->prof = prof;
  if (rc) return -1;
  Block B = h->B;
  B.S.is_char = ch ? 1 : 0;
  const int modes[3] = {INV_ASSIGN, INV_ACCUM, INV_FINAL};
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  rc = 0;
  for (int a = 0; a < 3 && !rc; a++) {
    // the alternative shapes exist for periodic axes with uniform metrics
    // (the marches for the y and z sweeps of some schemes)
    const bool fast = c.iperiodic[a] == 1 && B.M.uniform;
    float best = 0.f;
    int bests = 0;
    for (int sh = 0; sh < NSWEEP_SHAPES && !rc; sh++) {
      if (!sweep_shape_supported(a, sh, fast, c.scheme_invis, ch)) continue;
      hipError_t e = launch_inviscid(B, a, modes[a], SweepTile{sh}, h->st);  // warm-up
      e = e != hipSuccess ? e : hipEventRecord(e0, h->st);
      for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_inviscid(B, a, modes[a], SweepTile{sh}, h->st);
      e = e != hipSuccess ? e : hipEventRecord(e1, h->st);
      e = e != hipSuccess ? e : hipEventSynchronize(e1);
      float ms = 0.f;
      e = e != hipSuccess ? e : hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        rc = fail(std::string("autotune: ") + hipGetErrorString(e));
        break;
      }
      if (sh == 0 || ms < best) {
        best = ms;
        bests = sh;
      }
    }
    h->tile[a] = bests;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return -1;
  invalidate_graph(h);
  if (shape_out)
    for (int a = 0; a < 3; a++) shape_out[a] = h->tile[a];
  return 0;
}

// ---- diagnostics -------------------------------------------------------------
int ocfd_show_flow_msg(ocfd_handle *h, double out[3]) {
  if (launch(h, "flow_msg", [&] { return launch_reduce_flow_msg(h->B, h->work, h->wout, h->st); })) return -1;
  double s[3];
  HIPCHK(hipMemcpyAsync(s, h->wout, sizeof s, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  const double n = (double)(h->cfg.nx * h->cfg.ny * h->cfg.nz);
  for (int q = 0; q < 3; q++) out[q] = s[q] / n;
  return 0;
}

// show_flow_msg with the partial sums in the library's file-scope device array
// (no allocation).  k_flow_msg writes every partial and k_sum_partials every
// total, so nothing in the array has to start at zero.
int ocfd_show_flow_msg_noalloc(ocfd_handle *h, double out[3]) {
  if (!h || !out) return fail("ocfd_show_flow_msg_noalloc: NULL argument");
  const int64_t need = noalloc_flow_msg_doubles(h->B.g);
  if (need > NOALLOC_SCRATCH_DOUBLES)
    return fail("ocfd_show_flow_msg_noalloc: a block of " + std::to_string(h->B.g.ni) + " cells needs " +
                std::to_string(need) + " doubles of scratch, more than the " +
                std::to_string(NOALLOC_SCRATCH_DOUBLES) + " it has (at most about 178.9 million cells)");
  double *s = nullptr;
  if (noalloc_scratch(h->device, &s)) return -1;
  double *tot = s + 3 * flow_msg_partials(h->B.g);
  HIPCHK(launch_reduce_flow_msg(h->B, s, tot, h->st));
  double r[3];
  HIPCHK(hipMemcpyAsync(r, tot, sizeof r, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  const double n = (double)(h->cfg.nx * h->cfg.ny * h->cfg.nz);
  for (int q = 0; q < 3; q++) out[q] = r[q] / n;
  return 0;
}

int ocfd_ana_box(ocfd_handle *h, double out[20]) {
  const PrimField uvw[3] = {h->B.u, h->B.v, h->B.w};
  if (launch(h, "halo_fill", [&] { return launch_halo_prims(h->B, uvw, 3, h->st); })) return -1;
  if (launch(h, "ana_box", [&] { return launch_ana_box(h->B, nullptr, nullptr, h->work, h->wout, h->st); }))
    return -1;
  double s[15];
  HIPCHK(hipMemcpyAsync(s, h->wout, sizeof s, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  // statistics of ana_box (OCFD_ANA.f90:255-293)
  const double n = (double)(h->cfg.nx * h->cfg.ny * h->cfg.nz);
  for (int q = 0; q < 14; q++) s[q] = s[q] / n;
  double u_rms = s[0], v_rms = s[1], w_rms = s[2], Ux_rms = s[3];
  double Ux_sk = s[4], Ux_fl = s[5], d_aver = s[6], c_aver = s[7], Amu_aver = s[8], E_k = s[9];
  double T_aver = s[10], u_sk = s[11], u_fl = s[12], ens = 0.5 * s[13];
  double VV_rms = std::sqrt((u_rms + v_rms + w_rms) / 3.0);
  double Amt = std::sqrt(u_rms + v_rms + w_rms) / c_aver;
  u_rms = std::sqrt(u_rms); v_rms = std::sqrt(v_rms); w_rms = std::sqrt(w_rms); Ux_rms = std::sqrt(Ux_rms);
  double Alamda = VV_rms / Ux_rms, Alamdax = u_rms / Ux_rms;
  double Re_l = VV_rms * Alamda * d_aver / Amu_aver, Re_lx = VV_rms * Alamdax * d_aver / Amu_aver;
  Ux_sk = Ux_sk / (Ux_rms * Ux_rms * Ux_rms);
  Ux_fl = Ux_fl / ((Ux_rms * Ux_rms) * (Ux_rms * Ux_rms));
  u_sk = u_sk / (u_rms * u_rms * u_rms);
  u_fl = u_fl / ((u_rms * u_rms) * (u_rms * u_rms));
  const double d_rms = std::sqrt(s[14] / n);
  const double o[20] = {h->tt, Re_l, Re_lx, Amt, VV_rms, u_rms, v_rms, w_rms, u_sk, u_fl, Ux_sk, Ux_fl,
                        E_k, ens, d_rms, T_aver, Amu_aver, d_aver, Ux_rms, Alamdax};
  std::memcpy(out, o, sizeof o);
  return 0;
}

// ---- instrumentation -------------------------------------------------------
int ocfd_profile_enable(ocfd_handle *h, int32_t on) {
  if (flush_prof(h)) return -1;
  h->prof = on != 0;
  return 0;
}

int ocfd_profile_read(ocfd_handle *h, int32_t max_entries, char names[][32], double *ms, int64_t *launches) {
  if (flush_prof(h)) return -1;
  int n = 0;
  for (auto &kv : h->acc) {
    if (n >= max_entries) break;
    std::snprintf(names[n], 32, "%s", kv.first.c_str());
    ms[n] = kv.second.first;
    launches[n] = kv.second.second;
    n++;
  }
  return n;
}

int ocfd_profile_reset(ocfd_handle *h) {
  if (flush_prof(h)) return -1;
  h->acc.clear();
  return 0;
}

}  // extern "C"

// ===========================================================================
// Batches of independent periodic boxes (ocfd_batch_*).  Every box is an
// ocfd_handle of its own (set-up, transfers and diagnostics go through the
// single-block entry points); a batched time step launches each kernel phase
// once over the tiles of all boxes, from a device table of the boxes' Block
// descriptors, and is replayed from a captured hipGraph.
// ===========================================================================
struct ocfd_batch {
  int device = 0;
  hipStream_t st = nullptr;
  int prec = OCFD_PREC_FP64_CODE;
  std::vector<ocfd_handle *> box;
  // device copies of the boxes' Block descriptors: [0] as the boxes hold
  // them when a call starts, [1] with T and T2 swapped (and u, v, w with
  // u2, v2, w2 for the boxes whose stages run k_visc_fused); stage s of step
  // n reads table (n + s) % 2 (every stage swaps those buffers).  drep[q]
  // is table q with the fused boxes' u, v, w sets swapped: the T <= 0
  // repair after a stage sees their new u, v, w beside the old T
  Block *dblk[2] = {nullptr, nullptr};
  Block *drep[2] = {nullptr, nullptr};
  int *dtab = nullptr;        // box lists and tile offsets of every launch
  int *neg = nullptr;         // NEG_STRIDE ints per box
  std::vector<int> tab;       // host image of dtab
  struct Launch {             // one batched launch: boxes ids, offsets off
    int ids = 0, off = 0, m = 0;
    int64_t ntiles = 0;
  };
  struct SweepGroup {         // boxes sharing scheme and characteristic flag
    int sch = 0;
    bool ch = false;
    std::vector<int> ids;
    Launch l[3][NSWEEP_SHAPES];
    int shape[3] = {0, 0, 0};
  };
  struct ViscGroup {          // boxes sharing IF_Viscous, the CD order and the fused update
    bool visc = false, cd8 = false;
    bool fused = false;         // k_visc_fused (the boxes' handles run it alone too)
    std::vector<int> ids;
    Launch l[NBATCH_KSEG];     // per segment length BATCH_KSEGS[q]
  };
  int kq = NBATCH_KSEG - 1;    // chosen segment length: BATCH_KSEGS[kq]
  std::vector<SweepGroup> sweep;
  std::vector<ViscGroup> vis;
  Launch halo;
  Launch hcopy;               // the T -> T2 halo copies after each stage update
  bool use_graph = true;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};  // one step starting from table [q]
  int64_t cells = 0;
  // one persistent launch per step (launch mode 1): the queue of every
  // stage's phases (k_batch_queue), rebuilt when the segment length changes
  int mode = 0;
  bool queue_ok = false;      // every group runs in the queue kernel
  std::string queue_why;      // why not
  Launch qrep;                // repair items: 1 + the halo-copy tiles of each box
  int qkq = -1;               // segment length the plan was built for
  std::vector<QSeg> qseg;
  int qitems = 0, qgrid = 0;
  QSeg *dqseg = nullptr;
  int *dqneed = nullptr;      // [box][Q_NKIND]
  int *dqstate = nullptr;     // [0] ticket counter, [1] error flag, then the done counters
  hipGraphExec_t qexec[2] = {nullptr, nullptr};
};

namespace {

BoxSet boxset(const ocfd_batch *b, const ocfd_batch::Launch &l, int table = 0, bool repair = false) {
  return BoxSet{repair ? b->drep[table] : b->dblk[table], b->dtab + l.ids, b->dtab + l.off, l.m};
}

// box list + tile offsets of one launch, appended to the host table
ocfd_batch::Launch add_launch(ocfd_batch *b, const std::vector<int> &ids,
                              const std::function<int64_t(const Geom &)> &tiles) {
  ocfd_batch::Launch l;
  l.m = (int)ids.size();
  l.ids = (int)b->tab.size();
  for (int i : ids) b->tab.push_back(i);
  l.off = (int)b->tab.size();
  int64_t t = 0;
  for (int i : ids) {
    b->tab.push_back((int)t);
    t += tiles(b->box[i]->B.g);
  }
  b->tab.push_back((int)t);
  l.ntiles = t;
  return l;
}

void batch_invalidate(ocfd_batch *b) {
  for (hipGraphExec_t &g : b->gexec) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
  for (hipGraphExec_t &g : b->qexec) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

void batch_free(ocfd_batch *b) {
  batch_invalidate(b);
  for (ocfd_handle *h : b->box) ocfd_destroy(h);
  for (Block *t : b->dblk)
    if (t) (void)hipFree(t);
  for (Block *t : b->drep)
    if (t) (void)hipFree(t);
  if (b->dtab) (void)hipFree(b->dtab);
  if (b->neg) (void)hipFree(b->neg);
  if (b->dqseg) (void)hipFree(b->dqseg);
  if (b->dqneed) (void)hipFree(b->dqneed);
  if (b->dqstate) (void)hipFree(b->dqstate);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

int batch_sweep(ocfd_batch *b, const ocfd_batch::SweepGroup &g, int a, int shape, int table = 0) {
  static const int modes[3] = {INV_ASSIGN, INV_ACCUM, INV_FINAL};
  const ocfd_batch::Launch &l = g.l[a][shape];
  const int eff = sweep_shape_effective(a, shape, true, g.sch, g.ch);
  const hipError_t e = launch_inviscid_batch(boxset(b, l, table), b->prec, a, g.sch, g.ch, modes[a], eff,
                                             (int)l.ntiles, b->st);
  return e == hipSuccess ? 0 : fail(std::string("batched sweep: ") + hipGetErrorString(e));
}

int batch_halo(ocfd_batch *b, int table = 0) {
  const hipError_t e = launch_halo_batch(boxset(b, b->halo, table), (int)b->halo.ntiles, b->st);
  return e == hipSuccess ? 0 : fail(std::string("batched halo fill: ") + hipGetErrorString(e));
}

// the three stages of one time step of every box (OCFD_NS_Solver.f90:33-60),
// the first reading descriptor table q0
int batch_step_body(ocfd_batch *b, int q0) {
  for (int krk = 1; krk <= 3; krk++) {
    const int q = q0 ^ ((krk - 1) & 1);
    if (batch_halo(b, q)) return -1;
    for (const auto &g : b->sweep)
      for (int a = 0; a < 3; a++)
        if (batch_sweep(b, g, a, g.shape[a], q)) return -1;
    for (const auto &v : b->vis) {
      if (!v.visc || v.fused) continue;
      const auto &l = v.l[b->kq];
      const hipError_t e = launch_visc_flux_batch(boxset(b, l, q), b->prec, v.cd8, (int)l.ntiles, b->st);
      if (e != hipSuccess) return fail(std::string("batched viscous flux: ") + hipGetErrorString(e));
    }
    for (const auto &v : b->vis) {
      const auto &l = v.l[b->kq];
      const hipError_t e =
          v.fused ? launch_visc_fused_batch(boxset(b, l, q), krk, b->neg, (int)l.ntiles, b->st)
                  : launch_visc_div_update_batch(boxset(b, l, q), b->prec, v.visc, v.cd8, krk, b->neg,
                                                 (int)l.ntiles, b->st);
      if (e != hipSuccess) return fail(std::string("batched stage update: ") + hipGetErrorString(e));
    }
    const hipError_t e = launch_negt_repair_batch(boxset(b, b->hcopy, q, true), (int)b->box.size(), b->neg,
                                                  (int)b->hcopy.ntiles, b->st);
    if (e != hipSuccess) return fail(std::string("batched T <= 0 repair: ") + hipGetErrorString(e));
  }
  return 0;
}

// The queue of one step (launch mode 1): per stage, the halo fill, the
// sweeps of every scheme group along x, y, z, the viscous fluxes, the stage
// updates and the repairs, each a segment over the tiles of its boxes (the
// tables of the per-phase launches), with the items each box has of every
// kind.  Built for the current segment length of the k-marching kernels.
int batch_queue_plan(ocfd_batch *b) {
  if (b->qkq == b->kq) return 0;
  const int nbox = (int)b->box.size();
  const int kseg = BATCH_KSEGS[b->kq];
  std::vector<QSeg> segs;
  int64_t first = 0;
  auto add = [&](int kind, int stage, const ocfd_batch::Launch &l, int sch, int visc) {
    QSeg q = {};
    q.kind = kind;
    q.stage = stage;
    q.sch = sch;
    q.visc = visc;
    q.ids = l.ids;
    q.off = l.off;
    q.m = l.m;
    q.first = (int)first;
    q.ntiles = (int)l.ntiles;
    first += l.ntiles;
    segs.push_back(q);
  };
  for (int s = 1; s <= 3; s++) {
    add(Q_HALO, s, b->halo, 0, 0);
    for (int a = 0; a < 3; a++)
      for (const auto &g : b->sweep) add(Q_X + a, s, g.l[a][queue_shape(a)], g.sch, 0);
    for (const auto &v : b->vis)
      if (v.visc) add(Q_VFLUX, s, v.l[b->kq], 0, 1);
    for (const auto &v : b->vis) add(Q_UPDATE, s, v.l[b->kq], 0, v.visc ? 1 : 0);
    add(Q_REPAIR, s, b->qrep, 0, 0);
  }
  if (first >= INT32_MAX / 2) return fail("ocfd_batch: too many queue items for one step");
  std::vector<int> need((size_t)nbox * Q_NKIND, 0);
  for (int i = 0; i < nbox; i++) {
    const Geom &g = b->box[i]->B.g;
    int *n = &need[(size_t)i * Q_NKIND];
    n[Q_HALO] = (int)halo_tiles(g);
    for (int a = 0; a < 3; a++) n[Q_X + a] = (int)sweep_tiles(g, a, queue_shape(a));
    n[Q_VFLUX] = b->box[i]->cfg.if_viscous == 1 ? (int)march_tiles(g, kseg) : 0;
    n[Q_UPDATE] = (int)march_tiles(g, kseg);
    n[Q_REPAIR] = 1 + (int)halo_copy_tiles(g);
  }
  // every box's items in each segment are the count its counters wait for
  std::vector<int> seen((size_t)nbox * 3 * Q_NKIND, 0);
  for (const QSeg &q : segs)
    for (int j = 0; j < q.m; j++) {
      const int box = b->tab[q.ids + j], cnt = b->tab[q.off + j + 1] - b->tab[q.off + j];
      if (box < 0 || box >= nbox || cnt != need[(size_t)box * Q_NKIND + q.kind])
        return fail("ocfd_batch: queue plan inconsistent (segment kind " + std::to_string(q.kind) + ")");
      seen[((size_t)box * 3 + q.stage - 1) * Q_NKIND + q.kind] += 1;
    }
  for (int i = 0; i < nbox; i++)
    for (int s = 0; s < 3; s++)
      for (int k = 0; k < Q_NKIND; k++)
        if (seen[((size_t)i * 3 + s) * Q_NKIND + k] != (need[(size_t)i * Q_NKIND + k] > 0 ? 1 : 0))
          return fail("ocfd_batch: queue plan misses box " + std::to_string(i));
  HIPCHK(hipSetDevice(b->device));
  if (!b->dqseg) {
    HIPCHK(hipMalloc(&b->dqseg, 64 * sizeof(QSeg)));
    HIPCHK(hipMalloc(&b->dqneed, need.size() * sizeof(int)));
    HIPCHK(hipMalloc(&b->dqstate, (2 + (size_t)nbox * 3 * Q_NKIND) * sizeof(int)));
  }
  if (segs.size() > 64) return fail("ocfd_batch: too many queue segments");
  HIPCHK(hipMemcpy(b->dqseg, segs.data(), segs.size() * sizeof(QSeg), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->dqneed, need.data(), need.size() * sizeof(int), hipMemcpyHostToDevice));
  // captured queue steps hold the old item counts
  for (hipGraphExec_t &gx : b->qexec) {
    if (gx) (void)hipGraphExecDestroy(gx);
    gx = nullptr;
  }
  b->qseg = segs;
  b->qitems = (int)first;
  b->qgrid = batch_queue_grid(b->prec, b->device);
  if (b->qgrid <= 0) return fail("ocfd_batch: cannot size the queue kernel's grid");
  b->qkq = b->kq;
  return 0;
}

// one step of every box as one persistent launch (after clearing the
// ticket and completion counters with a fill kernel)
int batch_queue_body(ocfd_batch *b, int q0) {
  const int nbox = (int)b->box.size();
  const int ncnt = nbox * 3 * Q_NKIND;
  hipLaunchKernelGGL(k_zero_int, dim3((unsigned)((ncnt + 127) / 128)), dim3(128), 0, b->st, b->dqstate + 2, ncnt);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(128), 0, b->st, b->dqstate, 1);
  QPlan P = {};
  P.blk0 = b->dblk[0];
  P.blk1 = b->dblk[1];
  P.q0 = q0;
  P.tab = b->dtab;
  P.seg = b->dqseg;
  P.nseg = (int)b->qseg.size();
  P.nitems = b->qitems;
  P.head = b->dqstate;
  P.err = b->dqstate + 1;
  P.done = b->dqstate + 2;
  P.need = b->dqneed;
  P.neg = b->neg;
  const hipError_t e = launch_batch_queue(P, b->prec, b->qgrid, b->st);
  return e == hipSuccess ? 0 : fail(std::string("batched queue step: ") + hipGetErrorString(e));
}

int batch_capture(ocfd_batch *b, int q0) {
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  const int rc = b->mode == 1 ? batch_queue_body(b, q0) : batch_step_body(b, q0);
  const hipError_t e = hipStreamEndCapture(b->st, &graph);
  if (rc || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc ? -1 : fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  hipGraphExec_t *slot = b->mode == 1 ? &b->qexec[q0] : &b->gexec[q0];
  const hipError_t ei = hipGraphInstantiate(slot, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    *slot = nullptr;
    return fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  return 0;
}

// the boxes' descriptors and the launch tables to the device; the boxes'
// own streams are drained first (set-up work may still run there)
int batch_sync_tables(ocfd_batch *b) {
  HIPCHK(hipSetDevice(b->device));
  std::vector<Block> blk[2], rep[2];
  for (size_t i = 0; i < b->box.size(); i++) {
    ocfd_handle *h = b->box[i];
    HIPCHK(hipStreamSynchronize(h->st));
    Block B = h->B;
    B.kseg = BATCH_KSEGS[b->kq];
    B.halt = b->neg + NEG_STRIDE * i + NEG_HALT;
    auto swap_uvw_of = [&](Block &x) {
      if (!h->fused) return;
      std::swap(x.u, x.u2);
      std::swap(x.v, x.v2);
      std::swap(x.w, x.w2);
    };
    blk[0].push_back(B);
    swap_uvw_of(B);
    rep[0].push_back(B);
    std::swap(B.T, B.T2);
    blk[1].push_back(B);
    swap_uvw_of(B);
    rep[1].push_back(B);
  }
  for (int q = 0; q < 2; q++) {
    HIPCHK(hipMemcpy(b->dblk[q], blk[q].data(), blk[q].size() * sizeof(Block), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->drep[q], rep[q].data(), rep[q].size() * sizeof(Block), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(b->dtab, b->tab.data(), b->tab.size() * sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

extern "C" {

int ocfd_batch_create(const ocfd_config *cfgs, int32_t nbox, int32_t device, ocfd_batch **out) {
  if (!cfgs || !out || nbox < 1) return fail("ocfd_batch_create: need at least one box");
  for (int i = 0; i < nbox; i++) {
    const ocfd_config &c = cfgs[i];
    const std::string at = "box " + std::to_string(i) + ": ";
    if (c.iperiodic[0] != 1 || c.iperiodic[1] != 1 || c.iperiodic[2] != 1)
      return fail(at + "batched boxes must be periodic on every axis");
    if (c.metric_mode != OCFD_METRIC_UNIFORM) return fail(at + "batched boxes need OCFD_METRIC_UNIFORM");
    if (c.scheme_invis == SCH_HYBRID) return fail(at + "the HYBRID scheme is not batched");
    if (c.ibc != 0 || c.nfiltering != 0) return fail(at + "batched boxes take no boundary hook and no filter");
    for (int q = 0; q < 6; q++)
      if (c.ghost_cell[q] != 0) return fail(at + "batched boxes take no ghost cells");
    if (c.precision != cfgs[0].precision) return fail(at + "all boxes of a batch share one precision");
  }
  ocfd_batch *b = new ocfd_batch();
  b->device = device;
  b->prec = cfgs[0].precision == OCFD_PREC_MIXED ? OCFD_PREC_MIXED_CODE : OCFD_PREC_FP64_CODE;
  for (int i = 0; i < nbox; i++) {
    ocfd_handle *h = nullptr;
    if (ocfd_create(&cfgs[i], device, &h)) {
      const std::string msg = "box " + std::to_string(i) + ": " + g_err;
      batch_free(b);
      return fail(msg);
    }
    b->box.push_back(h);
    b->cells += (int64_t)cfgs[i].nx * cfgs[i].ny * cfgs[i].nz;
  }
  // launch groups
  for (int i = 0; i < nbox; i++) {
    const ocfd_config &c = cfgs[i];
    const bool ch = c.if_scheme_character != 0;
    auto it = std::find_if(b->sweep.begin(), b->sweep.end(),
                           [&](const ocfd_batch::SweepGroup &g) { return g.sch == c.scheme_invis && g.ch == ch; });
    if (it == b->sweep.end()) {
      b->sweep.emplace_back();
      b->sweep.back().sch = c.scheme_invis;
      b->sweep.back().ch = ch;
      it = b->sweep.end() - 1;
    }
    it->ids.push_back(i);
    const bool visc = c.if_viscous == 1, cd8 = c.scheme_vis == SCH_CD8;
    // a box that steps alone with k_visc_fused (mixed precision, whole 32 x 8
    // tiles, CD6; its second u, v, w buffers exist) runs it in the batch too
    const bool fused = b->box[i]->fused;
    auto iv = std::find_if(b->vis.begin(), b->vis.end(), [&](const ocfd_batch::ViscGroup &g) {
      return g.visc == visc && g.cd8 == cd8 && g.fused == fused;
    });
    if (iv == b->vis.end()) {
      b->vis.emplace_back();
      b->vis.back().visc = visc;
      b->vis.back().cd8 = cd8;
      b->vis.back().fused = fused;
      iv = b->vis.end() - 1;
    }
    iv->ids.push_back(i);
  }
  std::vector<int> all(nbox);
  for (int i = 0; i < nbox; i++) all[i] = i;
  b->halo = add_launch(b, all, [](const Geom &g) { return halo_tiles(g); });
  b->hcopy = add_launch(b, all, [](const Geom &g) { return halo_copy_tiles(g); });
  b->qrep = add_launch(b, all, [](const Geom &g) { return 1 + halo_copy_tiles(g); });
  // the queue kernel runs the non-characteristic sweeps and CD6
  b->queue_ok = true;
  for (int i = 0; i < nbox && b->queue_ok; i++) {
    if (cfgs[i].if_scheme_character != 0) b->queue_why = "box " + std::to_string(i) + " is characteristic";
    else if (cfgs[i].if_viscous == 1 && cfgs[i].scheme_vis == SCH_CD8) b->queue_why = "box " + std::to_string(i) + " uses CD8";
    else if (b->box[i]->fused) b->queue_why = "box " + std::to_string(i) + " runs the fused viscous stage update";
    b->queue_ok = b->queue_why.empty();
  }
  for (auto &g : b->sweep)
    for (int a = 0; a < 3; a++)
      for (int sh = 0; sh < NSWEEP_SHAPES; sh++)
        g.l[a][sh] = add_launch(b, g.ids, [a, sh, &g](const Geom &gg) {
          return sweep_tiles(gg, a, sweep_shape_effective(a, sh, true, g.sch, g.ch));
        });
  for (auto &v : b->vis)
    for (int q = 0; q < NBATCH_KSEG; q++)
      v.l[q] = add_launch(b, v.ids, [q, &v](const Geom &g) {
        return v.fused ? fused_batch_tiles(g, BATCH_KSEGS[q]) : march_tiles(g, BATCH_KSEGS[q]);
      });
  int64_t maxt = 0;
  for (int x : b->tab) maxt = std::max<int64_t>(maxt, x);
  if (maxt >= INT32_MAX / 2) {
    batch_free(b);
    return fail("ocfd_batch_create: too many workgroups for one launch");
  }
  int rc = 0;
  rc |= hipSetDevice(device) != hipSuccess;
  rc |= hipStreamCreate(&b->st) != hipSuccess;
  for (int q = 0; q < 2; q++) rc |= hipMalloc(&b->dblk[q], nbox * sizeof(Block)) != hipSuccess;
  for (int q = 0; q < 2; q++) rc |= hipMalloc(&b->drep[q], nbox * sizeof(Block)) != hipSuccess;
  rc |= hipMalloc(&b->dtab, b->tab.size() * sizeof(int)) != hipSuccess;
  rc |= hipMalloc(&b->neg, (size_t)NEG_STRIDE * nbox * sizeof(int)) != hipSuccess;
  if (rc) {
    batch_free(b);
    return fail("ocfd_batch_create: device allocation failed");
  }
  b->use_graph = std::getenv("OCFD_NO_GRAPH") == nullptr;
  *out = b;
  return 0;
}

int ocfd_batch_destroy(ocfd_batch *b) {
  if (b) batch_free(b);
  return 0;
}

int ocfd_batch_box(ocfd_batch *b, int32_t i, ocfd_handle **h) {
  if (i < 0 || i >= (int)b->box.size()) return fail("ocfd_batch_box: box index out of range");
  *h = b->box[i];
  return 0;
}

int ocfd_batch_step(ocfd_batch *b, int32_t nsteps) {
  const int nbox = (int)b->box.size();
  if (batch_sync_tables(b)) return -1;
  hipLaunchKernelGGL(k_zero_int, dim3((unsigned)((NEG_STRIDE * nbox + 127) / 128)), dim3(128), 0, b->st, b->neg,
                     NEG_STRIDE * nbox);
  std::vector<int> istep0(nbox);
  std::vector<double> tt0(nbox);
  for (int i = 0; i < nbox; i++) {
    istep0[i] = b->box[i]->istep;
    tt0[i] = b->box[i]->tt;
  }
  std::vector<int> neg((size_t)NEG_STRIDE * nbox);
  auto read_counters = [&]() -> int {
    HIPCHK(hipMemcpyAsync(neg.data(), b->neg, neg.size() * sizeof(int), hipMemcpyDeviceToHost, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    return 0;
  };
  const bool queue = b->mode == 1;
  if (queue) {
    if (batch_queue_plan(b)) return -1;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(128), 0, b->st, b->dqstate + 1, 1);  // error flag
  }
  int n = 0, rc = 0;
  for (; n < nsteps && !rc; n++) {
    const int q0 = n & 1;
    hipGraphExec_t *slot = queue ? &b->qexec[q0] : &b->gexec[q0];
    if (b->use_graph) {
      if (!*slot) rc = batch_capture(b, q0);
      if (!rc && hipGraphLaunch(*slot, b->st) != hipSuccess) rc = fail("hipGraphLaunch failed");
    } else {
      rc = queue ? batch_queue_body(b, q0) : batch_step_body(b, q0);
    }
    // stop early once every box has stopped at too many T <= 0 cells
    if (!rc && (n + 1) % 8 == 0 && n + 1 < nsteps) {
      rc = read_counters();
      bool all = true;
      for (int i = 0; i < nbox && all; i++) all = neg[(size_t)NEG_STRIDE * i + NEG_HALT] != 0;
      if (all) {
        n++;
        break;
      }
    }
  }
  if (rc) return -1;
  if (read_counters()) return -1;
  if (queue) {
    int qerr = 0;
    HIPCHK(hipMemcpy(&qerr, b->dqstate + 1, sizeof(int), hipMemcpyDeviceToHost));
    if (qerr) return fail("batched queue step: a wait for a box's previous phase ran out; results are invalid");
  }
  // the boxes' temperature buffers (and the fused boxes' u, v, w sets)
  // swapped once per stage
  for (ocfd_handle *h : b->box)
    if ((3 * n) & 1) {
      swap_T(h);
      if (h->fused) swap_uvw(h);
    }
  int total = 0;
  std::string err;
  for (int i = 0; i < nbox; i++) {
    ocfd_handle *h = b->box[i];
    const int *c = &neg[(size_t)NEG_STRIDE * i];
    int whole = n;
    if (c[NEG_HALT]) {
      whole = (c[NEG_HALT_STAGE] - 1) / 3;
      if (err.empty())
        err = "box " + std::to_string(i) + ": more than 10 cells with T <= 0 in stage " +
              std::to_string((c[NEG_HALT_STAGE] - 1) % 3 + 1) + " of step " + std::to_string(istep0[i] + whole + 1) +
              "; that box stops there";
    }
    h->istep = istep0[i] + whole;
    h->tt = tt0[i];
    for (int q = 0; q < whole; q++) h->tt = h->tt + h->cfg.dt;
    total += c[65];
  }
  if (!err.empty()) return fail(err);
  return total;
}

int ocfd_batch_set_sweep_tiles(ocfd_batch *b, int32_t group, const int32_t shape[3]) {
  if (group < 0 || group >= (int)b->sweep.size()) return fail("sweep group out of range");
  for (int a = 0; a < 3; a++)
    if (shape[a] < 0 || shape[a] >= NSWEEP_SHAPES) return fail("sweep tile shape must be 0 .. 4");
  for (int a = 0; a < 3; a++) b->sweep[group].shape[a] = shape[a];
  batch_invalidate(b);
  return 0;
}

int ocfd_batch_set_launch_mode(ocfd_batch *b, int32_t mode) {
  if (mode != 0 && mode != 1) return fail("launch mode must be 0 (per phase) or 1 (one persistent launch per step)");
  if (mode == 1 && !b->queue_ok) return fail("one persistent launch per step: " + b->queue_why);
  // The queue kernel (k_batch_queue) did not finish on the GPU in its first
  // two runs: its loop, split by the compiler, repeated each workgroup's
  // first item.  The rewritten loop has not run on the GPU yet, so the mode
  // stays off and every batch runs the per-phase launches.
  if (mode == 1) return fail("one persistent launch per step is switched off: its kernel did not finish on the GPU");
  b->mode = mode;
  return 0;
}

int ocfd_batch_info(ocfd_batch *b, int32_t *ngroups, int32_t *shapes, int64_t *cells, int32_t *kseg) {
  if (kseg) *kseg = BATCH_KSEGS[b->kq];
  if (ngroups) *ngroups = (int)b->sweep.size();
  if (shapes)
    for (size_t g = 0; g < b->sweep.size(); g++)
      for (int a = 0; a < 3; a++) shapes[3 * g + a] = b->sweep[g].shape[a];
  if (cells) *cells = b->cells;
  return 0;
}

int ocfd_batch_autotune(ocfd_batch *b, int32_t reps) {
  if (reps < 1) reps = 5;
  if (batch_sync_tables(b)) return -1;
  if (batch_halo(b)) return -1;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int rc = 0;
  for (auto &g : b->sweep) {
    for (int a = 0; a < 3 && !rc; a++) {
      float best = 0.f;
      int bests = 0;
      for (int sh = 0; sh < NSWEEP_SHAPES && !rc; sh++) {
        if (!sweep_shape_supported(a, sh, true, g.sch, g.ch)) continue;
        rc = batch_sweep(b, g, a, sh);  // warm-up
        if (!rc && hipEventRecord(e0, b->st) != hipSuccess) rc = fail("hipEventRecord failed");
        for (int r = 0; r < reps && !rc; r++) rc = batch_sweep(b, g, a, sh);
        float ms = 0.f;
        if (!rc && (hipEventRecord(e1, b->st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
          rc = fail("autotune: event timing failed");
        if (!rc && (sh == 0 || ms < best)) {
          best = ms;
          bests = sh;
        }
      }
      g.shape[a] = bests;
    }
  }
  // segment length of the k-marching kernels, timed on the viscous flux
  // kernel (it writes only Ev); the descriptors are re-sent per candidate
  bool anyvisc = false;
  // (the fused groups' update changes the state, so only the viscous
  // flux of the others is timed; the fused march takes at least 16 planes)
  for (const auto &v : b->vis) anyvisc |= v.visc && !v.fused;
  if (!rc && anyvisc) {
    float best = 0.f;
    int bestq = b->kq;
    for (int q = 0; q < NBATCH_KSEG && !rc; q++) {
      b->kq = q;
      if (batch_sync_tables(b)) {
        rc = -1;
        break;
      }
      float tot = 0.f;
      for (const auto &v : b->vis) {
        if (!v.visc || v.fused) continue;
        const auto &l = v.l[q];
        hipError_t e = launch_visc_flux_batch(boxset(b, l), b->prec, v.cd8, (int)l.ntiles, b->st);
        e = e != hipSuccess ? e : hipEventRecord(e0, b->st);
        for (int r = 0; r < reps && e == hipSuccess; r++)
          e = launch_visc_flux_batch(boxset(b, l), b->prec, v.cd8, (int)l.ntiles, b->st);
        float ms = 0.f;
        e = e != hipSuccess ? e : hipEventRecord(e1, b->st);
        e = e != hipSuccess ? e : hipEventSynchronize(e1);
        e = e != hipSuccess ? e : hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
          rc = fail(std::string("autotune: ") + hipGetErrorString(e));
          break;
        }
        tot += ms;
      }
      if (!rc && (q == 0 || tot < best)) {
        best = tot;
        bestq = q;
      }
    }
    b->kq = bestq;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  batch_invalidate(b);
  return rc ? -1 : 0;
}

}  // extern "C"
