// batch.hip -- batched ensemble: many independent grids (realisations)
// advanced by one stage-kernel launch per RK4 stage, the step sequence
// captured in a HIP graph and replayed.
//
// Each realisation is an ordinary context (its own registers, FRW
// integrator, sources and diagnostics), so every single-grid entry point
// works on it through cg_batch_member.  What the batch adds is the launch:
// the tiles of all grids are numbered grid after grid, and one launch of
// stage_kernel_batch covers them, each block looking up its grid's
// arguments and FRW values in device tables.  The FRW values of every stage
// of a call are integrated on the host in advance (the background does not
// depend on the fields) and uploaded before the graph is replayed, so the
// graph itself never changes between calls of the same length.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cosmo_gpu.h"
#include "bssn_stage.h"
#include "runtime.h"

using namespace cg;

constexpr int kTables = 4;  // sets of four stage tables (see build_tables)

struct cg_batch {
  std::vector<cg_ctx*> m;
  int prec = 0, order = 4, H = 2;
  hipStream_t stream = nullptr;
  const StageVariant* var = nullptr;
  int req_ti = 0, req_tj = 0, req_tk = 0;
  bool dma_allowed = true, dma_pick = true;  // LDS-DMA halo staging (as cg_ctx)
  bool incr_pick = true;                     // RK4 form picked by the autotuner (as cg_ctx)
  // device tables: stage arguments [4][n], first tile of each grid [n],
  // FRW values [steps][4][n][FRW_N]
  StageArgs* tab = nullptr;
  int* tstart = nullptr;
  double* frw = nullptr;
  int frw_cap = 0;  // steps the FRW table holds
  // pinned staging of the tables
  StageArgs* tab_h = nullptr;
  int* tstart_h = nullptr;
  double* frw_h = nullptr;
  int total = 0;
  // captured step sequence
  bool use_graph = true;
  hipGraphExec_t graph = nullptr;
  int graph_steps = 0;
  const StageVariant* graph_var = nullptr;
  bool graph_dma = false, graph_incr = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double last_ms = 0.0;
  unsigned* nan_scratch = nullptr;  // NaN count of the autotuner's stage-4 launches
  std::string err;
};

static int bfail(cg_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}
#define BCHK(b, expr)                                                                                    \
  do {                                                                                                   \
    hipError_t e_ = (expr);                                                                              \
    if (e_ != hipSuccess) return bfail(b, CG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

static bool batch_live(const cg_batch* b) {
  for (auto* c : b->m)
    if (alpha_live(c)) return true;
  return false;
}

static bool batch_fits(const cg_batch* b, const StageVariant& v) {
  if (!v.launch_batch || v.prec != b->prec || v.order != b->order || v.halo != b->H || v.live_alpha != batch_live(b))
    return false;
  for (auto* c : b->m)
    if (!variant_fits(c, v)) return false;
  return true;
}

// the LDS-DMA twin of v applies to every grid of the batch
static bool batch_dma_fits(const cg_batch* b, const StageVariant& v) {
  if (!b->dma_allowed || !b->dma_pick || !v.launch_batch_dma) return false;
  for (auto* c : b->m)
    if (!dma_fits(c, v)) return false;
  return true;
}

// stage 1 may use the DMA twin only when no grid normalises its metric there
static bool batch_norm_stage1(const cg_batch* b) {
  for (auto* c : b->m)
    if (c->p.normalize_metric) return true;
  return false;
}

static const StageVariant* batch_pick(const cg_batch* b) {
  const StageVariant* best = nullptr;
  int best_score = -1;
  for (const auto& v : stage_variants()) {
    if (!batch_fits(b, v)) continue;
    int score = 0;
    if (b->req_ti) {
      if (v.ti != b->req_ti || v.tj != b->req_tj || v.tk != b->req_tk) continue;
      score += 100;
    } else if (v.ti == 4 && v.tj == 8 && v.tk == 8) {
      score += 100;
    }
    score += (v.wpe == (b->prec == CG_FP64 ? 2 : 3)) ? 10 : (v.wpe == 1 ? 5 : 0);
    if (score > best_score) { best = &v; best_score = score; }
  }
  return best;
}

static void drop_graph(cg_batch* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  b->graph_steps = 0;
  b->graph_var = nullptr;
}

const char* cg_batch_last_error(const cg_batch* b) { return b ? b->err.c_str() : "null batch"; }

int cg_batch_destroy(cg_batch* b) {
  if (!b) return CG_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (auto* c : b->m) cg_destroy(c);
  if (b->tab) (void)hipFree(b->tab);
  if (b->tstart) (void)hipFree(b->tstart);
  if (b->frw) (void)hipFree(b->frw);
  if (b->tab_h) (void)hipHostFree(b->tab_h);
  if (b->nan_scratch) (void)hipFree(b->nan_scratch);
  if (b->tstart_h) (void)hipHostFree(b->tstart_h);
  if (b->frw_h) (void)hipHostFree(b->frw_h);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return CG_OK;
}

int cg_batch_create(const cg_params* params, int n, cg_batch** out) {
  if (!params || !out || n < 1) return CG_ERR_ARG;
  *out = nullptr;
  const cg_params& p0 = params[0];
  for (int r = 0; r < n; ++r) {
    const cg_params& p = params[r];
    if (p.precision != p0.precision || p.stencil_order != p0.stencil_order || p.device != p0.device ||
        (p.ko_coeff != 0.0) != (p0.ko_coeff != 0.0))
      return CG_ERR_ARG;  // one kernel variant serves every grid
    if (lapse_needs_kavg(p.lapse)) return CG_ERR_ARG;  // per-grid K_avg every stage: single grids only
    if (p.matter == CG_MATTER_DUST) return CG_ERR_ARG;           // the dust fluid runs on single grids
    if (p.precision == CG_FP32_ALL || p.precision == CG_FP32_FIX) return CG_ERR_ARG;  // single grids only
    if (p.shift != CG_SHIFT_NONE || p.lapse == CG_LAPSE_TEST_K_DRIVER) return CG_ERR_ARG;  // gauge pass: single grids
    if (p.ko_coeff != 0.0 && p.normalize_metric && (p.stencil_order == 2 || p.stencil_order == 8))
      return CG_ERR_ARG;  // needs the materialised stepInit of the single-grid path
  }
  cg_batch* b = new cg_batch();
  *out = b;
  b->prec = p0.precision;
  b->order = p0.stencil_order;
  b->H = halo_for(p0.stencil_order, p0.ko_coeff);
  if (p0.tile_i > 0) { b->req_ti = p0.tile_i; b->req_tj = p0.tile_j; b->req_tk = p0.tile_k; }
  for (int r = 0; r < n; ++r) {
    cg_ctx* c = nullptr;
    int rc = cg_create(&params[r], &c);
    if (rc) {
      b->err = std::string("realisation ") + std::to_string(r) + ": " + (c ? cg_last_error(c) : "invalid parameters");
      if (c) cg_destroy(c);
      return rc;
    }
    b->m.push_back(c);
  }
  BCHK(b, hipSetDevice(p0.device));
  BCHK(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  BCHK(b, hipMalloc((void**)&b->tab, sizeof(StageArgs) * 4 * kTables * n));
  BCHK(b, hipMalloc((void**)&b->tstart, sizeof(int) * n));
  BCHK(b, hipHostMalloc((void**)&b->tab_h, sizeof(StageArgs) * 4 * kTables * n));
  BCHK(b, hipHostMalloc((void**)&b->tstart_h, sizeof(int) * n));
  BCHK(b, hipEventCreate(&b->e0));
  BCHK(b, hipEventCreate(&b->e1));
  return CG_OK;
}

int cg_batch_size(const cg_batch* b) { return b ? (int)b->m.size() : 0; }

int cg_batch_member(cg_batch* b, int r, cg_ctx** out) {
  if (!b || !out || r < 0 || r >= (int)b->m.size()) return CG_ERR_ARG;
  *out = b->m[r];
  return CG_OK;
}

int cg_batch_set_dma(cg_batch* b, int on) {
  if (!b) return CG_ERR_ARG;
  b->dma_allowed = on != 0;
  drop_graph(b);
  return CG_OK;
}

int cg_batch_set_graph(cg_batch* b, int on) {
  if (!b) return CG_ERR_ARG;
  b->use_graph = on != 0;
  return CG_OK;
}

// every member may step in the accumulator-free form (StageArgs::incr)
static bool batch_incr_allowed(const cg_batch* b) {
  for (auto* c : b->m)
    if (!fused_incr(c)) return false;
  return true;
}
static bool batch_incr(const cg_batch* b) { return b->incr_pick && batch_incr_allowed(b); }

// stage slots of the reference's form: 1: P -> X (normalised on the fly),
// 2: X -> Y, 3: Y -> X, 4: X -> P; _p = slot 0, _f = slot 1
static const int kIn[4] = {0, 2, 3, 2}, kOut[4] = {2, 3, 2, 0};

// argument tables of the stages (kTables sets of four, [set*4 + stage-1][n])
// and the tile offsets for variant v.  With LDS-DMA staging, stage 4 also
// writes the normalised copy of the new state (stepInit's _a) into a dead
// slot, and stage 1 of every step but the first reads DIFFgamma_ij and A_ij
// from there instead of normalising halos on the fly.
//   Accumulator-free form: set 0 is the first step (a1, a2, a3 in slots 1,
//   2, 3); set m = 1..3 a step whose stage 1 reads the copy in slot m
//   (slots from incr_slots).
//   Reference form: set 0 (P -> X -> Y -> X -> P, _f in slot 1; the copy
//   goes to slot 3), and set 1's stage 1 reads the copy.
static int build_tables(cg_batch* b, const StageVariant* v) {
  const int n = (int)b->m.size();
  const bool dma = batch_dma_fits(b, *v), incr = batch_incr(b);
  int total = 0;
  for (int r = 0; r < n; ++r) {
    cg_ctx* c = b->m[r];
    const int ntx = (c->p.nx + v->ti - 1) / v->ti, nty = (c->p.ny + v->tj - 1) / v->tj,
              ntz = (c->p.nz + v->tk - 1) / v->tk;
    b->tstart_h[r] = total;
    const int norm = c->p.normalize_metric;
    for (int set = 0; set < kTables; ++set)
      for (int st = 1; st <= 4; ++st) {
        StageArgs a;
        if (incr) {
          int s1, s2, s3;
          incr_slots(set, s1, s2, s3);
          const int in[4] = {0, s1, s2, s3}, out[4] = {s1, s2, s3, 0};
          a = make_args(c, st, in[st - 1], 0, st == 4 ? s1 : 1, out[st - 1], st == 1 && set == 0 ? norm : 0);
          a.incr = 1;
          if (st == 1 && set > 0 && norm) a.s_in_norm = set;
          if (st == 4) {
            a.s_a2 = s2;
            if (dma && norm) a.s_norm_out = s1;
          }
        } else {
          a = make_args(c, st, kIn[st - 1], 0, 1, kOut[st - 1], st == 1 && set == 0 ? norm : 0);
          if (st == 1 && set > 0 && norm) a.s_in_norm = 3;
          if (st == 4 && dma && norm) a.s_norm_out = 3;
        }
        a.ntx = ntx; a.nty = nty; a.ntz = ntz;
        a.per_xcd = 0;          // unused: the batch kernel maps blocks to tiles itself
        a.ext_part = nullptr;   // averages of a batch step are formed on demand per grid
        b->tab_h[(set * 4 + st - 1) * n + r] = a;
      }
    const long t = (long)ntx * nty * ntz;
    if (total + t > (1L << 30)) return bfail(b, CG_ERR_ARG, "too many tiles in one batch");
    total += (int)t;
  }
  b->total = total;
  BCHK(b, hipMemcpyAsync(b->tab, b->tab_h, sizeof(StageArgs) * 4 * kTables * n, hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipMemcpyAsync(b->tstart, b->tstart_h, sizeof(int) * n, hipMemcpyHostToDevice, b->stream));
  return CG_OK;
}

static int ensure_frw(cg_batch* b, int nsteps) {
  if (nsteps <= b->frw_cap) return CG_OK;
  const size_t words = (size_t)nsteps * 4 * b->m.size() * FRW_N;
  if (b->frw) { (void)hipFree(b->frw); b->frw = nullptr; }
  if (b->frw_h) { (void)hipHostFree(b->frw_h); b->frw_h = nullptr; }
  drop_graph(b);  // the captured launches point into the old table
  BCHK(b, hipMalloc((void**)&b->frw, words * sizeof(double)));
  BCHK(b, hipHostMalloc((void**)&b->frw_h, words * sizeof(double)));
  b->frw_cap = nsteps;
  return CG_OK;
}

// the table set of step s (see build_tables) and the stage-4 input slot
// of the last step
static int step_set(bool dma, bool incr, int s) {
  if (!dma || s == 0) return 0;
  return incr ? (s - 1) % 3 + 1 : 1;  // the copy of step s-1 sits over its a1
}
static int last_stage4_input(bool dma, bool incr, int nsteps) {
  if (!incr) return kIn[3];
  int s1, s2, s3;
  incr_slots(step_set(dma, incr, nsteps - 1), s1, s2, s3);
  return s3;
}

static int enqueue_steps(cg_batch* b, const StageVariant* v, int nsteps) {
  const int n = (int)b->m.size();
  const bool dma = batch_dma_fits(b, *v), norm1 = batch_norm_stage1(b), incr = batch_incr(b);
  for (int s = 0; s < nsteps; ++s)
    for (int st = 1; st <= 4; ++st) {
      // stage 1 after the first step reads the normalised copy
      const bool pre = dma && s > 0 && st == 1;
      BatchArgs ba{};
      ba.tab = b->tab + (step_set(dma, incr, s) * 4 + st - 1) * n;
      ba.tstart = b->tstart;
      ba.frw = b->frw + ((size_t)s * 4 + (st - 1)) * n * FRW_N;
      ba.nreal = n;
      ba.total = b->total;
      ba.per_xcd = (b->total + 7) / 8;
      int rc = (dma && (pre || !(st == 1 && norm1)) ? v->launch_batch_dma : v->launch_batch)(ba, b->stream);
      if (rc) return bfail(b, CG_ERR_HIP, std::string("batch launch: ") + hipGetErrorString((hipError_t)rc));
    }
  return CG_OK;
}

int cg_batch_step(cg_batch* b, int nsteps) {
  if (!b || nsteps < 0) return CG_ERR_ARG;
  if (nsteps == 0) return CG_OK;
  const int n = (int)b->m.size();
  for (auto* c : b->m) {
    if (!c->a_is_p) return bfail(b, CG_ERR_STATE, "a realisation is between method-level stages");
    BCHK(b, hipStreamSynchronize(c->stream));  // work queued on the member streams first
  }
  if (!b->var || !batch_fits(b, *b->var)) b->var = batch_pick(b);
  const StageVariant* v = b->var;
  if (!v) return bfail(b, CG_ERR_ARG, "no batched stage kernel variant fits every grid");
  if (int rc = build_tables(b, v)) return rc;
  if (int rc = ensure_frw(b, nsteps)) return rc;

  // FRW values of every stage, integrated ahead on the host (utils/FRW.h):
  // stage st sees the getters after stage st-1, and its averages are
  // weighted with phi after its own FRW step
  std::vector<SrcArgs> src4(n);
  for (int r = 0; r < n; ++r) {
    cg_ctx* c = b->m[r];
    for (int s = 0; s < nsteps; ++s)
      for (int st = 1; st <= 4; ++st) {
        double* f = b->frw_h + (((size_t)s * 4 + (st - 1)) * n + r) * FRW_N;
        f[0] = c->frw.g_phi; f[1] = c->frw.g_K; f[2] = c->frw.g_rho; f[3] = c->frw.g_S;
        f[4] = 2.0 / 3.0 * c->frw.g_K * c->frw.g_K - 16.0 * M_PI * c->frw.g_rho;
        if (st == 4) src4[r] = SrcArgs{c->DIFFD, c->p.matter, c->frw.g_phi, c->frw.g_rho, c->lambda};
        c->frw.stage(st, c->p.dt);
        f[5] = c->frw.g_phi;
      }
  }
  BCHK(b, hipMemcpyAsync(b->frw, b->frw_h, (size_t)nsteps * 4 * n * FRW_N * sizeof(double), hipMemcpyHostToDevice,
                         b->stream));

  BCHK(b, hipEventRecord(b->e0, b->stream));
  if (b->use_graph) {
    const bool dma = batch_dma_fits(b, *v);
    const bool incr = batch_incr(b);
    if (!b->graph || b->graph_steps != nsteps || b->graph_var != v || b->graph_dma != dma || b->graph_incr != incr) {
      drop_graph(b);
      hipGraph_t g = nullptr;
      BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = enqueue_steps(b, v, nsteps);
      hipError_t ec = hipStreamEndCapture(b->stream, &g);
      if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
      BCHK(b, ec);
      hipError_t ei = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      BCHK(b, ei);
      b->graph_steps = nsteps;
      b->graph_var = v;
      b->graph_dma = dma;
      b->graph_incr = incr;
    }
    BCHK(b, hipGraphLaunch(b->graph, b->stream));
  } else {
    if (int rc = enqueue_steps(b, v, nsteps)) return rc;
  }
  BCHK(b, hipEventRecord(b->e1, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  float ms = 0.f;
  BCHK(b, hipEventElapsedTime(&ms, b->e0, b->e1));
  b->last_ms = ms;

  // member state as after a fused step: _a, _f alias _p; the averages of
  // K4Finalize are formed on demand from _p and the stage-4 input
  const int in4 = last_stage4_input(batch_dma_fits(b, *v), batch_incr(b), nsteps);
  std::string nan_list;
  for (int r = 0; r < n; ++r) {
    cg_ctx* c = b->m[r];
    c->sP = 0; c->sF = 1; c->sA = 0; c->sC = in4; c->a_is_p = true;
    c->ext_fresh = false;
    c->ext_lazy = true;
    c->ext_lazy_in = in4;
    c->ext_lazy_src = src4[r];
    c->ext_lazy_phiF = c->frw.g_phi;
    unsigned nn = 0;
    BCHK(b, hipMemcpy(&nn, c->nan_dev, sizeof(unsigned), hipMemcpyDeviceToHost));
    if (nn != c->nan_seen) {
      c->nan_seen = nn;
      nan_list += (nan_list.empty() ? "" : ",") + std::to_string(r);
    }
  }
  if (!nan_list.empty()) return bfail(b, CG_ERR_NAN, "NAN detected in DIFFphi of realisation(s) " + nan_list);
  return CG_OK;
}

int cg_batch_step_time_ms(const cg_batch* b, double* ms) {
  if (!b || !ms) return CG_ERR_ARG;
  *ms = b->last_ms;
  return CG_OK;
}

int cg_batch_constraints(cg_batch* b, double* H, double* M, double* G, double* A, double* S) {
  if (!b) return CG_ERR_ARG;
  for (size_t r = 0; r < b->m.size(); ++r) {
    int rc = cg_constraints(b->m[r], H ? H + 8 * r : nullptr, M ? M + 8 * r : nullptr, G ? G + 7 * r : nullptr,
                            A ? A + 7 * r : nullptr, S ? S + 7 * r : nullptr);
    if (rc) return bfail(b, rc, std::string("realisation ") + std::to_string(r) + ": " + cg_last_error(b->m[r]));
  }
  return CG_OK;
}

int cg_batch_autotune(cg_batch* b, int reps, char* report, size_t report_len) {
  if (!b || reps < 1) return CG_ERR_ARG;
  const int n = (int)b->m.size();
  for (auto* c : b->m) {
    if (!c->a_is_p) return bfail(b, CG_ERR_STATE, "a realisation is between method-level stages");
    if (int rc = ext_resolve(c)) return bfail(b, rc, cg_last_error(c));  // before slot 3 is written
    BCHK(b, hipStreamSynchronize(c->stream));
  }
  if (int rc = ensure_frw(b, 1)) return rc;
  if (!b->nan_scratch) BCHK(b, hipMalloc((void**)&b->nan_scratch, 256));
  for (int r = 0; r < n; ++r) {
    cg_ctx* c = b->m[r];
    double* f = b->frw_h + (size_t)r * FRW_N;
    f[0] = c->frw.g_phi; f[1] = c->frw.g_K; f[2] = c->frw.g_rho; f[3] = c->frw.g_S;
    f[4] = 2.0 / 3.0 * c->frw.g_K * c->frw.g_K - 16.0 * M_PI * c->frw.g_rho;
    f[5] = c->frw.g_phi;
  }
  BCHK(b, hipMemcpyAsync(b->frw, b->frw_h, (size_t)n * FRW_N * sizeof(double), hipMemcpyHostToDevice, b->stream));
  std::string rep = "[";
  const StageVariant* best = nullptr;
  double best_ms = 1e30;
  int best_order = 0;
  bool best_dma = false, best_incr = false;
  for (const auto& v : stage_variants()) {
    if (!batch_fits(b, v)) continue;
    for (int d = 0; d < 2; ++d) {
      b->dma_pick = d == 1;
      if (d == 1 && !batch_dma_fits(b, v)) continue;
      for (int order = 0; order < 6; ++order) {
        if (order == 3) {
          bool ok = true;
          for (auto* c : b->m) ok = ok && morton_fits((c->p.nx + v.ti - 1) / v.ti, (c->p.ny + v.tj - 1) / v.tj);
          if (!ok) continue;
        }
        for (auto* c : b->m) c->tile_order = order;
        auto go = d ? v.launch_batch_dma : v.launch_batch;
        for (int form = batch_incr_allowed(b) ? 1 : 0; form >= 0; --form) {
          b->incr_pick = form == 1;
          if (int rc = build_tables(b, &v)) return rc;
          // stage-1, -2 and -4 tables of the first set retargeted to write
          // the spare slot 3 and read dead slots (stage 1: stencils of _p
          // and slot 2 as the normalised copy with DMA staging, as in a
          // later step; stages 2 and 4: stencils of slot 2, stage 4's a1 in
          // slot 1 and a2 in slot 3; NaN counts to scratch), so the evolved
          // state is untouched
          for (int q = 0; q <= 3; ++q)
            for (int r = 0; r < n; ++r) {
              StageArgs& a = b->tab_h[q * n + r];
              a.s_in = q == 0 ? 0 : 2; a.s_out = 3;
              a.s_f = 1; a.s_a2 = 3;
              a.s_norm_out = -1;
              a.nan_count = b->nan_scratch;
              if (q == 0 && d == 1 && b->m[r]->p.normalize_metric) { a.normalize = 0; a.s_in_norm = 2; }
            }
          BCHK(b, hipMemcpyAsync(b->tab, b->tab_h, sizeof(StageArgs) * 4 * n, hipMemcpyHostToDevice, b->stream));
          double tq[3] = {0.0, 0.0, 0.0};
          for (int q = 0; q < 3; ++q) {
            BatchArgs ba{};
            ba.tab = b->tab + (q == 0 ? 0 : q == 1 ? 1 : 3) * n;
            ba.tstart = b->tstart;
            ba.frw = b->frw;
            ba.nreal = n;
            ba.total = b->total;
            ba.per_xcd = (b->total + 7) / 8;
            // with DMA, stage 1 reads the stand-in copy instead of normalising
            int rc = go(ba, b->stream);  // warm-up
            if (rc) return bfail(b, CG_ERR_HIP, std::string("batch launch: ") + hipGetErrorString((hipError_t)rc));
            BCHK(b, hipEventRecord(b->e0, b->stream));
            for (int k = 0; k < reps; ++k)
              if ((rc = go(ba, b->stream)))
                return bfail(b, CG_ERR_HIP, std::string("batch launch: ") + hipGetErrorString((hipError_t)rc));
            BCHK(b, hipEventRecord(b->e1, b->stream));
            BCHK(b, hipEventSynchronize(b->e1));
            float ms = 0.f;
            BCHK(b, hipEventElapsedTime(&ms, b->e0, b->e1));
            tq[q] = ms / reps;
          }
          const double avg = 0.25 * (tq[0] + 2.0 * tq[1] + tq[2]);  // mean stage of a step
          int vg = 0, sc = 0, ld = 0;
          (d ? v.info_dma : v.info)(&vg, &sc, &ld);
          char buf[320];
          snprintf(buf, sizeof buf,
                   "%s{\"tile\":[%d,%d,%d],\"wpe\":%d,\"dma\":%d,\"order\":%d,\"incr\":%d,\"ms\":%.4f,"
                   "\"ms1\":%.4f,\"ms2\":%.4f,\"ms4\":%.4f,\"vgprs\":%d,\"scratch\":%d}",
                   rep.size() > 1 ? "," : "", v.ti, v.tj, v.tk, v.wpe, d, order, form, avg, tq[0], tq[1], tq[2], vg,
                   sc);
          rep += buf;
          if (avg < best_ms) {
            best_ms = avg; best = &v; best_order = order; best_dma = d == 1; best_incr = form == 1;
          }
        }
      }
    }
  }
  for (auto* c : b->m) c->tile_order = best_order;
  b->dma_pick = best_dma || !best;
  b->incr_pick = best_incr || !best;
  rep += "]";
  if (report && report_len) {
    strncpy(report, rep.c_str(), report_len - 1);
    report[report_len - 1] = 0;
  }
  if (!best) return bfail(b, CG_ERR_ARG, "no batched stage kernel variant fits every grid");
  b->var = best;
  b->req_ti = best->ti; b->req_tj = best->tj; b->req_tk = best->tk;
  drop_graph(b);
  return CG_OK;
}

int cg_batch_kernel_info(const cg_batch* bc, int* vgprs, int* scratch, int* ti, int* tj, int* tk) {
  cg_batch* b = const_cast<cg_batch*>(bc);
  if (!b || !vgprs || !scratch || !ti || !tj || !tk) return CG_ERR_ARG;
  if (!b->var || !batch_fits(b, *b->var)) b->var = batch_pick(b);
  if (!b->var) return bfail(b, CG_ERR_ARG, "no batched stage kernel variant fits every grid");
  int lds = 0;
  if ((batch_dma_fits(b, *b->var) ? b->var->info_dma : b->var->info)(vgprs, scratch, &lds)) return CG_ERR_HIP;
  *ti = b->var->ti; *tj = b->var->tj; *tk = b->var->tk;
  return CG_OK;
}
